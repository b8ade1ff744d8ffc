// pvrun -- runs the layers, connections and probes of an OpenPV .params file
// on the GPU through libpvhip (the drop-in for OpenPV's run of the same file
// for the delivery + LCA path).
//
//   pvrun -p file.params [-o outputPath] [--steps N] [--autotune]
//
// Supported objects (the LCA column of tests/GPULCATest and LCATest):
//   HyPerCol; PvpLayer, ANNErrorLayer, HyPerLCALayer, MomentumLCALayer, ANNLayer, HyPerLayer;
//   IdentConn, HyPerConn (FileWeight, normalizeL2), TransposeConn, CloneConn;
//   AdaptiveTimeScaleProbe, ColumnEnergyProbe, L2NormProbe, L1NormProbe,
//   L1NormLCAProbe, FirmThresholdCostFnLCAProbe, RequireAllZeroActivityProbe.
// Delivery selection follows HyPerDeliveryCreator (HyPerDeliveryCreator.cpp:41-126):
// updateGSynFromPostPerspective picks the post-perspective gather or the
// pre-perspective scatter; receiveGpu = false gives the EXACT form (the
// reference's CPU summation order, bit for bit), receiveGpu = true the form
// named by the HyPerCol parameter gpuArithmetic ("fp32" default, "bf16",
// "fp8": the gather in FP8 with a source scale from fp8SourceBound or the
// input range; scatters and channel GEMMs run FP8 requests in BF16).
// A GPU scatter runs as a gather with transposed weights where the geometry
// allows (as the Python runtime does), else as the ordered scatter in FP32.
// updateGpu selects the transfer form (false: applyVThresh/applyAMax,
// true: the vertex form of the reference's GPU kernel).
// The step follows HyPerCol::advanceTime (HyPerCol.cpp:659-791): per phase,
// every layer of the phase receives, then every layer of the phase updates
// and publishes (with delay 0 a layer reads the others' last published state).
// Activity files are written dense (type 4), sparse layers included.
#include "pvhip.h"
#include "pvhost.h"
#include "pvparams.h"

#include <algorithm>
#include <cctype>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

[[noreturn]] void fatal(const std::string &msg) { throw std::runtime_error(msg); }

void check(int st, const char *what) {
   if (st != PVHIP_OK) {
      fatal(std::string(what) + " failed: " + pvhip_last_error());
   }
}

// ------------------------------------------------------------------ PVP I/O
// Headers as in src/utils/BufferUtilsPvp.hpp:13-99.
struct PvpFrames {
   int nx = 0, ny = 0, nf = 0;
   std::vector<double> times;
   std::vector<std::vector<float>> frames;
};

PvpFrames readActivity(const std::string &path) {
   std::ifstream f(path, std::ios::binary);
   if (!f) {
      fatal("cannot open " + path);
   }
   int32_t h[18];
   double t0;
   f.read(reinterpret_cast<char *>(h), sizeof h);
   f.read(reinterpret_cast<char *>(&t0), sizeof t0);
   if (h[2] != 4 || h[9] != 3) {
      fatal(path + ": only dense float activity files (type 4) are read");
   }
   PvpFrames r;
   r.nx = h[3];
   r.ny = h[4];
   r.nf = h[5];
   f.seekg(h[0]);
   size_t n = (size_t)r.nx * r.ny * r.nf;
   while (true) {
      double t;
      std::vector<float> v(n);
      if (!f.read(reinterpret_cast<char *>(&t), sizeof t)) {
         break;
      }
      if (!f.read(reinterpret_cast<char *>(v.data()), n * 4)) {
         break;
      }
      r.times.push_back(t);
      r.frames.push_back(std::move(v));
   }
   return r;
}

// Shared weights (type 5), one arbor: returns [numPatches][nyp][nxp][nfp].
std::vector<float> readWeights(const std::string &path, int nxp, int nyp, int nfp, int numPatches) {
   std::ifstream f(path, std::ios::binary);
   if (!f) {
      fatal("cannot open " + path);
   }
   int32_t h[26];
   f.read(reinterpret_cast<char *>(h), sizeof h);
   if (h[2] != 5) {
      fatal(path + ": not a shared-weight file (type 5)");
   }
   if (h[20] != nxp || h[21] != nyp || h[22] != nfp || h[25] != numPatches) {
      fatal(path + ": patch size or patch count differs from the connection");
   }
   f.seekg(h[0]);
   size_t ps = (size_t)nxp * nyp * nfp;
   std::vector<float> w(ps * numPatches);
   for (int p = 0; p < numPatches; p++) {
      char geo[8];
      f.read(geo, 8);
      f.read(reinterpret_cast<char *>(w.data() + p * ps), ps * 4);
   }
   if (!f) {
      fatal(path + ": short file");
   }
   return w;
}

void writeActivity(const std::string &path, int nx, int ny, int nf,
      const std::vector<std::vector<float>> &frames, const std::vector<double> &times) {
   std::ofstream f(path, std::ios::binary);
   int32_t h[18] = {80, 20, 4, nx, ny, nf, 1, nx * ny * nf, 4, 3, 1, 1, nx, ny, 0, 0, 1,
         (int32_t)frames.size()};
   double t0 = times.empty() ? 0.0 : times[0];
   f.write(reinterpret_cast<const char *>(h), sizeof h);
   f.write(reinterpret_cast<const char *>(&t0), sizeof t0);
   for (size_t i = 0; i < frames.size(); i++) {
      f.write(reinterpret_cast<const char *>(&times[i]), sizeof(double));
      f.write(reinterpret_cast<const char *>(frames[i].data()), frames[i].size() * 4);
   }
}

// ------------------------------------------------------------------ device memory
struct Device {
   pvhip_device *dev = nullptr;
   pvhip_stream stream = nullptr;
   std::vector<pvhip_buffer *> buffers;
   void *alloc(size_t bytes) {
      pvhip_buffer *b = nullptr;
      check(pvhip_buffer_alloc(dev, std::max<size_t>(bytes, 4), &b), "pvhip_buffer_alloc");
      buffers.push_back(b);
      void *p = pvhip_buffer_ptr(b);
      check(pvhip_fill_zero(p, std::max<size_t>(bytes, 4), stream), "pvhip_fill_zero");
      return p;
   }
   // Setup and frame changes only: synchronous, so the (pageable, often
   // temporary) host data may be released on return.
   void h2d(void *d, const void *h, size_t bytes) {
      // copies go through a buffer handle: find the one holding d
      for (auto *b : buffers) {
         char *base = static_cast<char *>(pvhip_buffer_ptr(b));
         if ((char *)d >= base && (char *)d + bytes <= base + pvhip_buffer_size(b)) {
            check(pvhip_buffer_h2d(b, h, bytes, (char *)d - base, stream), "pvhip_buffer_h2d");
            check(pvhip_stream_sync(stream), "pvhip_stream_sync");
            return;
         }
      }
      fatal("h2d into unknown device memory");
   }
   void d2h(void *h, const void *d, size_t bytes) {
      for (auto *b : buffers) {
         char *base = static_cast<char *>(pvhip_buffer_ptr(b));
         if ((const char *)d >= base && (const char *)d + bytes <= base + pvhip_buffer_size(b)) {
            check(pvhip_buffer_d2h(b, h, bytes, (const char *)d - base, stream), "pvhip_buffer_d2h");
            check(pvhip_stream_sync(stream), "pvhip_stream_sync");
            return;
         }
      }
      fatal("d2h from unknown device memory");
   }
   int32_t *table(const std::vector<int32_t> &v) {
      auto *d = static_cast<int32_t *>(alloc(v.size() * 4));
      if (!v.empty()) {
         h2d(d, v.data(), v.size() * 4);
      }
      return d;
   }
   ~Device() {
      for (auto *b : buffers) {
         pvhip_buffer_free(b);
      }
      if (dev) {
         pvhip_device_destroy(dev);
      }
   }
};

// ------------------------------------------------------------------ network
struct Params {
   const pvparams *p;
   int g;
   double num(const char *k, double dflt) const {
      double v;
      return pvparams_number(p, g, k, &v) ? v : dflt;
   }
   bool has(const char *k) const {
      double v;
      const char *s;
      return pvparams_number(p, g, k, &v) || pvparams_string(p, g, k, &s);
   }
   std::string str(const char *k, const char *dflt = "") const {
      const char *s = nullptr;
      if (pvparams_string(p, g, k, &s)) {
         return s ? s : "";
      }
      return dflt;
   }
};

enum LayerKind { L_INPUT, L_ERROR, L_LCA, L_ANN };

struct Layer {
   std::string name;
   LayerKind kind;
   Params prm;
   int phase = 0;
   pvhip_loc loc{};
   int halo   = 0;
   int numChannels = 1;
   float *V = nullptr, *A = nullptr, *gsyn = nullptr;
   pvhip_transfer transfer{};
   bool hasTransfer = false;
   float errScale   = 1.0f;
   // LCA
   float tau = 1.0f;
   bool selfInteract = true;
   bool momentum = false;          // MomentumLCALayer
   float momentumRate = 0.0f;      // LCAMomentumRate
   float *prevDrive = nullptr;     // restricted, zero at start
   std::string adaptiveProbe;
   float *expTau = nullptr;
   std::vector<double> dt;
   // input
   PvpFrames input;
   int displayPeriod = 0, frameIndex = 0;
   double lastUpdate = 1.0;
   // sparse publish
   bool sparse = false, needActive = false;
   pvhip_sparse_entry *entries = nullptr;
   int64_t *numActive = nullptr;
   uint32_t *posStart = nullptr, *compactScratch = nullptr;
   // output
   int writeStep = -1;
   double initialWriteTime = 0;
   std::vector<std::vector<float>> outFrames;
   std::vector<double> outTimes;
};

enum ConnMode { C_NONE, C_IDENT, C_GATHER, C_SCATTER };

struct Conn {
   std::string name, keyword;
   Params prm;
   int pre = -1, post = -1, channel = 0;
   int nxp = 0, nyp = 0, nfp = 0; // pre-perspective patch of the weights this conn uses
   std::string original;
   ConnMode mode = C_NONE;
   int arith     = PVHIP_ARITH_EXACT;
   bool routed   = false;    // a GPU scatter run as a gather with transposed weights
   std::vector<float> hostW; // pre-perspective weights (originals only)
   float *W = nullptr;       // device weights the delivery reads
   void *Wlp = nullptr;
   int nxq = 0, nyq = 0;     // post-perspective patch of a gather
   int nkx = 1, nky = 1;
   int32_t *tab[8] = {};
   int32_t *inv[2] = {};
   int win[4] = {};
   void *scratch = nullptr; // channel-GEMM split partials (pvhip_recv_post_scratch_bytes)
   size_t scratchBytes = 0;
   int tile = 0, splits = 0; // channel-GEMM launch shape (0: the library's default)
   float srcScale = 0.0f;    // FP8 gather: power-of-two source scale
};

struct Probe {
   std::string name, keyword;
   Params prm;
   std::vector<double> values, prev;
   // norm probes
   int target = -1;
   double coefficient = 1.0;
   std::string energyProbe;
   // adaptive
   std::string targetName, triggerLayer;
   double baseMax = 1, baseMin = 1, tauFactor = 1, growthFactor = 1;
   std::vector<double> ts, tsMax, tsTrue;
   // RequireAllZeroActivityProbe
   double nnzThreshold = 0.0;
   bool exitOnFailure = true, immediateExit = false, failed = false;
   double firstFailure = 0.0;
   std::FILE *out = nullptr;
};

struct Net {
   const pvparams *P;
   Device D;
   int nxGlobal = 0, nyGlobal = 0, nbatch = 1;
   double stopTime = 1, dtCol = 1;
   std::string outputPath, gpuArith = "fp32"; // outputPath: -o, else the HyPerCol's
   std::vector<Layer> layers;
   std::vector<Conn> conns;
   std::vector<Probe> probes;
   double *energyDev = nullptr;

   int layerIndex(const std::string &n) const {
      for (size_t i = 0; i < layers.size(); i++) {
         if (layers[i].name == n) {
            return (int)i;
         }
      }
      fatal("no layer named \"" + n + "\"");
   }
   int connIndex(const std::string &n) const {
      for (size_t i = 0; i < conns.size(); i++) {
         if (conns[i].name == n) {
            return (int)i;
         }
      }
      fatal("no connection named \"" + n + "\"");
   }
   int probeIndex(const std::string &n) const {
      for (size_t i = 0; i < probes.size(); i++) {
         if (probes[i].name == n) {
            return (int)i;
         }
      }
      return -1;
   }

   void build();
   bool transposedRoute(Conn &c);
   void buildConn(Conn &c);
   void step(double t);
   pvhip_recv_post_args gatherArgs(Conn &c);
   double sourceBound(const Layer &pre);
   void autotune();
   void deliver(Conn &c);
   void receive(Layer &L);
   void updateLayer(Layer &L, double t);
   void checkAllZero(Probe &p, double t);
   std::vector<double> probeEnergy(Probe &p);
   void output(double t);
   int finish();
};

// ANNActivityBuffer::ioParam_VThresh .. ioParam_VWidth (ANNActivityBuffer.cpp:136-164):
// AMin defaults to VThresh; the defaults of the others give the identity.
pvhip_transfer makeTransfer(const Params &q, bool vertices) {
   if (q.has("verticesV") || q.has("verticesA")) {
      fatal("verticesV/verticesA lists are not supported");
   }
   pvhip_transfer t{};
   float VT  = (float)q.num("VThresh", -FLT_MAX);
   float AMn = (float)q.num("AMin", VT);
   float AMx = (float)q.num("AMax", FLT_MAX);
   float ASh = (float)q.num("AShift", 0.0);
   float VW  = (float)q.num("VWidth", 0.0);
   if (pvhost_make_transfer(VT, AMn, AMx, ASh, VW, vertices ? PVHIP_TRANSFER_VERTICES : PVHIP_TRANSFER_VTHRESH,
             &t)
       != 0) {
      fatal("transfer parameters rejected");
   }
   return t;
}

void Net::build() {
   int col = -1;
   for (int g = 0; g < pvparams_num_groups(P); g++) {
      if (std::string(pvparams_group_keyword(P, g)) == "HyPerCol") {
         col = g;
      }
   }
   if (col < 0) {
      fatal("no HyPerCol group");
   }
   Params c{P, col};
   nxGlobal = (int)c.num("nx", 0);
   nyGlobal = (int)c.num("ny", 0);
   nbatch   = (int)c.num("nbatch", 1);
   dtCol    = c.num("dt", 1.0);
   stopTime = c.num("stopTime", 1.0);
   if (outputPath.empty()) {
      outputPath = c.has("outputPath") ? c.str("outputPath") : "output/";
   }
   std::error_code ec;
   std::filesystem::create_directories(outputPath, ec);
   if (ec) {
      fatal("cannot create the output directory " + outputPath);
   }
   if (c.has("gpuArithmetic")) {
      gpuArith = c.str("gpuArithmetic");
   }
   if (dtCol != 1.0) {
      fatal("only dt = 1 is supported");
   }
   for (int g = 0; g < pvparams_num_groups(P); g++) {
      std::string kw = pvparams_group_keyword(P, g);
      Params q{P, g};
      if (kw == "PvpLayer" || kw == "ANNErrorLayer" || kw == "HyPerLCALayer"
          || kw == "MomentumLCALayer" || kw == "ANNLayer" || kw == "HyPerLayer") {
         Layer L;
         L.name  = pvparams_group_name(P, g);
         L.prm   = q;
         L.kind  = kw == "PvpLayer" ? L_INPUT : kw == "ANNErrorLayer" ? L_ERROR
                 : kw == "HyPerLCALayer" || kw == "MomentumLCALayer" ? L_LCA : L_ANN;
         L.momentum = kw == "MomentumLCALayer";
         L.phase = (int)q.num("phase", 0);
         L.loc.nbatch = nbatch;
         L.loc.nx     = (int)std::lround(nxGlobal * q.num("nxScale", 1.0));
         L.loc.ny     = (int)std::lround(nyGlobal * q.num("nyScale", 1.0));
         L.loc.nf     = (int)q.num("nf", 1);
         L.sparse     = q.num("sparseLayer", 0) != 0;
         L.writeStep  = (int)q.num("writeStep", -1);
         L.initialWriteTime = q.num("initialWriteTime", 0);
         if (q.num("mirrorBCflag", 0) != 0 || q.num("valueBC", 0) != 0) {
            fatal(L.name + ": mirrorBCflag and a nonzero valueBC are not supported");
         }
         layers.push_back(std::move(L));
      }
      else if (kw == "IdentConn" || kw == "HyPerConn" || kw == "TransposeConn" || kw == "CloneConn") {
         Conn C;
         C.name    = pvparams_group_name(P, g);
         C.keyword = kw;
         C.prm     = q;
         conns.push_back(std::move(C));
      }
      else if (kw == "AdaptiveTimeScaleProbe" || kw == "ColumnEnergyProbe" || kw == "L2NormProbe"
               || kw == "L1NormProbe" || kw == "L1NormLCAProbe" || kw == "FirmThresholdCostFnLCAProbe"
               || kw == "RequireAllZeroActivityProbe") {
         Probe p;
         p.name    = pvparams_group_name(P, g);
         p.keyword = kw;
         p.prm     = q;
         probes.push_back(std::move(p));
      }
      else if (kw == "HyPerCol") {
      }
      else {
         std::fprintf(stderr, "pvrun: ignoring %s \"%s\" (not supported)\n", kw.c_str(),
               pvparams_group_name(P, g));
      }
   }
   // connections: endpoints, geometry, halo requirements
   for (auto &C : conns) {
      C.pre     = layerIndex(C.prm.str("preLayerName"));
      C.post    = layerIndex(C.prm.str("postLayerName"));
      C.channel = (int)C.prm.num("channelCode", 0);
      if (C.channel >= PVHIP_MAX_CHANNELS) {
         fatal(C.name + ": channelCode beyond the supported channels");
      }
      Layer &post      = layers[C.post];
      post.numChannels = std::max(post.numChannels, C.channel + 1);
      double d[4] = {0};
      if (pvparams_array(P, C.prm.g, "delay", d, 4) > 0 && d[0] != 0.0) {
         fatal(C.name + ": delays are not supported");
      }
   }
   for (auto &C : conns) {
      Layer &pre = layers[C.pre], &post = layers[C.post];
      if (C.keyword == "IdentConn") {
         C.mode = C_IDENT;
         continue;
      }
      bool postPersp = C.prm.num("updateGSynFromPostPerspective", 0) != 0;
      bool gpu       = C.prm.num("receiveGpu", 0) != 0;
      // HyPerDeliveryCreator::ioParam_pvpatchAccumulateType (HyPerDeliveryCreator.cpp:60-100)
      std::string acc = C.prm.str("pvpatchAccumulateType", "convolve");
      std::transform(acc.begin(), acc.end(), acc.begin(), [](unsigned char ch) { return (char)std::tolower(ch); });
      if (acc == "stochastic") {
         fatal(gpu ? C.name + " sets receiveGpu to true and pvpatchAccumulateType to stochastic, but stochastic "
                              "release has not been implemented on the GPU"
                   : C.name + ": stochastic release is not supported");
      }
      if (acc != "convolve") {
         fatal(C.name + ": pvpatchAccumulateType \"" + acc + "\" is unrecognized (\"convolve\" or \"stochastic\")");
      }
      if (C.prm.num("convertRateToSpikeCount", 0) != 0) {
         fatal(C.name + ": convertRateToSpikeCount is not supported");
      }
      if (C.prm.num("numAxonalArbors", 1) != 1) {
         fatal(C.name + ": only one arbor is supported");
      }
      if (C.keyword == "HyPerConn" && C.prm.num("sharedWeights", 1) == 0) {
         fatal(C.name + ": only shared weights are supported");
      }
      std::printf("pvrun: %s \"%s\": %ssynapticPerspective%sDelivery\n", C.keyword.c_str(), C.name.c_str(),
            postPersp ? "Post" : "Pre", gpu ? "GPU" : "Convolve");
      C.arith        = !gpu                 ? PVHIP_ARITH_EXACT
                       : gpuArith == "bf16" ? PVHIP_ARITH_BF16
                       : gpuArith == "fp8"  ? PVHIP_ARITH_FP8
                                            : PVHIP_ARITH_FP32;
      if (gpu && gpuArith != "fp32" && gpuArith != "bf16" && gpuArith != "fp8") {
         fatal("gpuArithmetic \"" + gpuArith + "\" is not fp32, bf16 or fp8");
      }
      if (C.arith == PVHIP_ARITH_FP8 && !postPersp) {
         C.arith = PVHIP_ARITH_BF16; // the scatter's MFMA route has BF16 and FP32 forms
      }
      if (C.keyword == "HyPerConn") {
         C.nxp = (int)C.prm.num("nxp", 1);
         C.nyp = (int)C.prm.num("nyp", 1);
         C.nfp = (int)C.prm.num("nfp", post.loc.nf);
      }
      else {
         C.original   = C.prm.str("originalConnName");
         Conn &O      = conns[connIndex(C.original)];
         int onxp     = (int)O.prm.num("nxp", 1), onyp = (int)O.prm.num("nyp", 1);
         if (C.keyword == "CloneConn") {
            C.nxp = onxp;
            C.nyp = onyp;
            C.nfp = layers[O.post].loc.nf;
         }
         else { // TransposeConn: pre/post of the original swapped
            C.nxp = pvhost_post_patch_size(onxp, layers[O.pre].loc.nx, layers[O.post].loc.nx);
            C.nyp = pvhost_post_patch_size(onyp, layers[O.pre].loc.ny, layers[O.post].loc.ny);
            C.nfp = layers[O.pre].loc.nf;
         }
      }
      if (C.channel < 0) {
         C.mode = C_NONE;
      }
      else {
         C.mode = postPersp ? C_GATHER : C_SCATTER;
      }
      // every connection sets the margin its pre layer needs (HyPerConn too)
      int mx = pvhost_required_margin(pre.loc.nx, post.loc.nx, C.nxp);
      int my = pvhost_required_margin(pre.loc.ny, post.loc.ny, C.nyp);
      pre.halo = std::max(pre.halo, std::max(mx, my));
   }
   // parameters checked: from here on the device is used
   check(pvhip_device_create(0, &D.dev), "pvhip_device_create");
   D.stream = pvhip_device_stream(D.dev);
   for (auto &L : layers) {
      L.loc.halo = {L.halo, L.halo, L.halo, L.halo};
      size_t nres = (size_t)nbatch * L.loc.nx * L.loc.ny * L.loc.nf;
      size_t next = (size_t)nbatch * (L.loc.nx + 2 * L.halo) * (L.loc.ny + 2 * L.halo) * L.loc.nf;
      L.V    = static_cast<float *>(D.alloc(nres * 4));
      L.A    = static_cast<float *>(D.alloc(next * 4));
      L.gsyn = static_cast<float *>(D.alloc((size_t)L.numChannels * nres * 4));
      bool updGpu = L.prm.num("updateGpu", 0) != 0;
      if (L.kind == L_INPUT) {
         L.input         = readActivity(L.prm.str("inputPath"));
         L.displayPeriod = (int)L.prm.num("displayPeriod", 0);
         if (L.input.nx != L.loc.nx || L.input.ny != L.loc.ny || L.input.nf != L.loc.nf) {
            fatal(L.name + ": input file size differs from the layer (autoResizeFlag is not supported)");
         }
      }
      else {
         L.transfer = makeTransfer(L.prm, updGpu);
         L.errScale = (float)L.prm.num("errScale", 1.0);
      }
      if (L.kind == L_LCA) {
         L.tau          = (float)(L.prm.num("timeConstantTau", 1.0) / dtCol);
         L.selfInteract = L.prm.num("selfInteract", 1) != 0;
         if (L.momentum) { // MomentumLCAInternalStateBuffer.cpp: LCAMomentumRate, the prevDrive buffer
            L.momentumRate = (float)L.prm.num("LCAMomentumRate", 0.0);
            L.prevDrive    = static_cast<float *>(D.alloc(nres * 4));
         }
         L.adaptiveProbe = L.prm.str("adaptiveTimeScaleProbe");
         L.expTau        = static_cast<float *>(D.alloc(nbatch * 4));
         L.dt.assign(nbatch, dtCol);
         std::string init = L.prm.str("InitVType", "ZeroV");
         if (init == "InitVFromFile") {
            PvpFrames v = readActivity(L.prm.str("Vfilename"));
            if (v.nx != L.loc.nx || v.ny != L.loc.ny || v.nf != L.loc.nf || v.frames.empty()) {
               fatal(L.name + ": Vfilename size differs from the layer");
            }
            for (int b = 0; b < nbatch; b++) {
               const auto &fr = v.frames[std::min<size_t>(b, v.frames.size() - 1)];
               D.h2d(L.V + (size_t)b * fr.size(), fr.data(), fr.size() * 4);
            }
         }
         else if (init != "ZeroV") {
            fatal(L.name + ": InitVType " + init + " is not supported (ZeroV, InitVFromFile)");
         }
      }
      if (L.kind != L_INPUT) {
         check(pvhip_ann_activity(&L.loc, &L.transfer, L.V, L.A, D.stream), "initial activity");
      }
   }
   // weights of the originals, then each delivery's tables
   for (auto &C : conns) {
      if (C.keyword != "HyPerConn") {
         continue;
      }
      Layer &pre = layers[C.pre], &post = layers[C.post];
      int nkx, nky;
      pvhost_num_kernels(&pre.loc, &post.loc, &nkx, &nky);
      int nK = nkx * nky * pre.loc.nf;
      std::string init = C.prm.str("weightInitType");
      if (init != "FileWeight") {
         fatal(C.name + ": weightInitType " + init + " is not supported (FileWeight)");
      }
      C.hostW = readWeights(C.prm.str("initWeightsFile"), C.nxp, C.nyp, C.nfp, nK);
      C.W     = static_cast<float *>(D.alloc(C.hostW.size() * 4));
      D.h2d(C.W, C.hostW.data(), C.hostW.size() * 4);
      std::string norm = C.prm.str("normalizeMethod", "none");
      if (norm == "normalizeL2" && C.prm.num("normalizeOnInitialize", 1) != 0) {
         if (C.prm.num("normalizeArborsIndividually", 0) != 0 || C.prm.num("normalizeFromPostPerspective", 0) != 0
             || C.prm.num("rMinX", 0) != 0 || C.prm.num("rMinY", 0) != 0 || C.prm.num("normalize_cutoff", 0) != 0
             || C.prm.num("nonnegativeConstraintFlag", 0) != 0) {
            fatal(C.name + ": only plain normalizeL2 is supported");
         }
         check(pvhip_normalize_l2(C.W, nK, C.nxp * C.nyp * C.nfp, (float)C.prm.num("strength", 1.0),
                     (float)C.prm.num("minL2NormTolerated", 0.0), D.stream),
               "pvhip_normalize_l2");
         D.d2h(C.hostW.data(), C.W, C.hostW.size() * 4);
      }
      else if (norm != "none" && norm != "") {
         fatal(C.name + ": normalizeMethod " + norm + " is not supported");
      }
   }
   for (auto &C : conns) {
      if (C.mode == C_SCATTER && C.arith != PVHIP_ARITH_EXACT && transposedRoute(C)) {
         C.mode   = C_GATHER;
         C.routed = true;
      }
   }
   for (auto &C : conns) {
      if (C.mode == C_SCATTER && layers[C.pre].sparse) {
         layers[C.pre].needActive = true;
      }
   }
   for (auto &L : layers) {
      if (L.needActive) {
         size_t next = (size_t)nbatch * (L.loc.nx + 2 * L.halo) * (L.loc.ny + 2 * L.halo) * L.loc.nf;
         size_t npos = (size_t)(L.loc.nx + 2 * L.halo) * (L.loc.ny + 2 * L.halo) + 1;
         L.entries   = static_cast<pvhip_sparse_entry *>(D.alloc(next * 8));
         L.numActive = static_cast<int64_t *>(D.alloc(nbatch * 8));
         L.posStart  = static_cast<uint32_t *>(D.alloc(nbatch * npos * 4));
         L.compactScratch = static_cast<uint32_t *>(D.alloc(pvhip_active_compact_scratch_bytes(&L.loc)));
         check(pvhip_active_compact(&L.loc, L.A, L.entries, L.numActive, L.posStart, L.compactScratch, D.stream),
               "pvhip_active_compact");
      }
   }
   for (auto &C : conns) {
      if (C.mode == C_GATHER || C.mode == C_SCATTER) {
         buildConn(C);
      }
   }
   // probes
   for (auto &p : probes) {
      p.values.assign(nbatch, 0.0);
      p.prev.assign(nbatch, 0.0);
      if (p.keyword == "AdaptiveTimeScaleProbe") {
         p.targetName   = p.prm.str("targetName");
         p.triggerLayer = p.prm.str("triggerLayerName");
         p.baseMax      = p.prm.num("baseMax", 1.0);
         p.baseMin      = p.prm.num("baseMin", 1.0);
         p.tauFactor    = p.prm.num("tauFactor", 1.0);
         p.growthFactor = p.prm.num("growthFactor", 1.0);
         p.ts.assign(nbatch, p.baseMin);
         p.tsMax.assign(nbatch, p.baseMax);
         p.tsTrue.assign(nbatch, -1.0);
      }
      else if (p.keyword == "RequireAllZeroActivityProbe") {
         p.target        = layerIndex(p.prm.str("targetLayer"));
         p.nnzThreshold  = (double)(float)p.prm.num("nnzThreshold", 0.0);
         p.exitOnFailure = p.prm.num("exitOnFailure", 1) != 0;
         p.immediateExit = p.exitOnFailure && p.prm.num("immediateExitOnFailure", 0) != 0;
      }
      else if (p.keyword != "ColumnEnergyProbe") {
         p.target      = layerIndex(p.prm.str("targetLayer"));
         p.coefficient = p.prm.num("coefficient", 1.0);
         p.energyProbe = p.prm.str("energyProbe");
         if (p.keyword == "L1NormLCAProbe" || p.keyword == "FirmThresholdCostFnLCAProbe") {
            p.coefficient = (double)(float)layers[p.target].prm.num("VThresh", 0.0);
         }
      }
      if (p.prm.num("textOutputFlag", 1) != 0 && p.prm.has("probeOutputFile") && !p.prm.str("probeOutputFile").empty()) {
         std::string path = outputPath + "/" + p.prm.str("probeOutputFile");
         p.out            = std::fopen(path.c_str(), "w");
      }
   }
   energyDev = static_cast<double *>(D.alloc(nbatch * 8));
   check(pvhip_stream_sync(D.stream), "initialisation");
}

// The power of two s with bound*s in [120, 240): FP8 (e4m3) range with headroom.
float fp8Scale(double bound) {
   if (!(bound > 0.0) || !std::isfinite(bound)) {
      return 1.0f;
   }
   int e;
   std::frexp(240.0 / bound, &e);
   return std::ldexp(1.0f, e - 1);
}

// Bound on |activity| of a gather's source for its FP8 scale: an error layer
// fed by input layers holds errScale*(input - reconstruction), within
// 2*|errScale|*max|input| once the reconstruction approximates the input;
// another source has no such bound and uses 1 (a connection may set
// fp8SourceBound).
double Net::sourceBound(const Layer &pre) {
   if (pre.kind != L_ERROR) {
      return 1.0;
   }
   double m = 0.0;
   for (auto &C : conns) {
      if (&layers[C.post] == &pre && C.mode == C_IDENT && layers[C.pre].kind == L_INPUT) {
         for (auto &fr : layers[C.pre].input.frames) {
            for (float v : fr) {
               m = std::max(m, (double)std::fabs(v));
            }
         }
      }
   }
   return m > 0.0 ? 2.0 * std::fabs(pre.errScale) * m : 1.0;
}

// Pre-perspective weights of C (its own pre -> post) on the host: a HyPerConn's
// or CloneConn's are the original's; a TransposeConn's are the transpose of
// its original's (TransposeWeights::transposeShared).
std::vector<float> prePerspectiveWeights(Net &net, Conn &C, Conn &O) {
   if (C.keyword != "TransposeConn") {
      return O.hostW;
   }
   Layer &opre = net.layers[O.pre], &opost = net.layers[O.post];
   int pkx, pky;
   pvhost_num_kernels(&opost.loc, &opre.loc, &pkx, &pky);
   std::vector<float> wt((size_t)pkx * pky * opost.loc.nf * C.nxp * C.nyp * C.nfp);
   if (pvhost_transpose_weights(&opre.loc, &opost.loc, O.nxp, O.nyp, O.nfp, O.hostW.data(), wt.data()) != 0) {
      fatal(C.name + ": weight transpose failed");
   }
   return wt;
}

// The MFMA route of a GPU scatter: a post-perspective gather with the
// transposed weights, taken when the post-perspective patch exists, the
// pre layer's halo covers it, and the phase x feature channels fit one tile
// (several data patches per feature need pre.nf a multiple of the MFMA K
// step and at most 64 output channels).
bool Net::transposedRoute(Conn &C) {
   Layer &pre = layers[C.pre], &post = layers[C.post];
   int ck  = C.arith == PVHIP_ARITH_BF16 ? 32 : 16;
   int nxq = pvhost_post_patch_size(C.nxp, pre.loc.nx, post.loc.nx);
   int nyq = pvhost_post_patch_size(C.nyp, pre.loc.ny, post.loc.ny);
   int pkx, pky;
   pvhost_num_kernels(&post.loc, &pre.loc, &pkx, &pky);
   int m       = pkx * pky * post.loc.nf;
   bool single = pkx * pky == 1;
   if (nxq <= 0 || nyq <= 0 || (!single && (pre.loc.nf % ck != 0 || m > 64))) {
      return false;
   }
   pvhip_loc src = pre.loc;
   src.halo      = {pre.halo, pre.halo, pre.halo, pre.halo};
   std::vector<int32_t> sx(post.loc.nx), sy(post.loc.ny), kx(post.loc.nx), ky(post.loc.ny);
   int nkx, nky;
   return pvhost_recv_post_tables(&src, &post.loc, nxq, nyq, sx.data(), sy.data(), kx.data(), ky.data(), &nkx,
                &nky)
          == 0;
}

void Net::buildConn(Conn &C) {
   Layer &pre = layers[C.pre], &post = layers[C.post];
   Conn &O    = C.keyword == "HyPerConn" ? C : conns[connIndex(C.original)];
   if (C.mode == C_GATHER) {
      // post-perspective weights [kernels of post -> pre][nyq][nxq][pre.nf]:
      // a TransposeConn gathering reads its original's pre-perspective weights
      // directly (TransposeWeightsPair.cpp:141-147); the others transpose
      // their pre-perspective weights (TransposeWeights::transposeShared).
      if (C.keyword == "TransposeConn" && !C.routed) {
         C.W   = O.W;
         C.nxq = O.nxp;
         C.nyq = O.nyp;
      }
      else {
         std::vector<float> w = prePerspectiveWeights(*this, C, O);
         C.nxq                = pvhost_post_patch_size(C.nxp, pre.loc.nx, post.loc.nx);
         C.nyq                = pvhost_post_patch_size(C.nyp, pre.loc.ny, post.loc.ny);
         int pkx, pky;
         pvhost_num_kernels(&post.loc, &pre.loc, &pkx, &pky);
         std::vector<float> wt((size_t)pkx * pky * post.loc.nf * C.nxq * C.nyq * pre.loc.nf);
         if (pvhost_transpose_weights(&pre.loc, &post.loc, C.nxp, C.nyp, C.nfp, w.data(), wt.data()) != 0) {
            fatal(C.name + ": weight transpose failed");
         }
         C.W = static_cast<float *>(D.alloc(wt.size() * 4));
         D.h2d(C.W, wt.data(), wt.size() * 4);
      }
      std::vector<int32_t> sx(post.loc.nx), sy(post.loc.ny), kx(post.loc.nx), ky(post.loc.ny);
      if (pvhost_recv_post_tables(&pre.loc, &post.loc, C.nxq, C.nyq, sx.data(), sy.data(), kx.data(), ky.data(),
                &C.nkx, &C.nky)
          != 0) {
         fatal(C.name + ": halo too small for the gather");
      }
      C.tab[0]  = D.table(sx);
      C.tab[1]  = D.table(sy);
      C.tab[2]  = D.table(kx);
      C.tab[3]  = D.table(ky);
      int nK    = C.nkx * C.nky * post.loc.nf;
      int psize = C.nxq * C.nyq * pre.loc.nf;
      if (C.nkx * C.nky > 1 && C.arith != PVHIP_ARITH_EXACT) {
         int x0, nx, y0, ny;
         if (pvhost_recv_post_inverse(&pre.loc, &post.loc, C.nxq, C.nyq, nullptr, nullptr, &x0, &nx, &y0, &ny) != 0) {
            fatal(C.name + ": inverse tables failed");
         }
         std::vector<int32_t> ix((size_t)C.nkx * nx), iy((size_t)C.nky * ny);
         pvhost_recv_post_inverse(&pre.loc, &post.loc, C.nxq, C.nyq, ix.data(), iy.data(), &x0, &nx, &y0, &ny);
         C.inv[0] = D.table(ix);
         C.inv[1] = D.table(iy);
         C.win[0] = x0;
         C.win[1] = nx;
         C.win[2] = y0;
         C.win[3] = ny;
      }
      if (C.arith == PVHIP_ARITH_FP8 && C.nkx * C.nky > 1) {
         C.arith = PVHIP_ARITH_BF16; // the channel GEMM has BF16 and FP32 forms
      }
      if (C.arith == PVHIP_ARITH_BF16) {
         C.Wlp = D.alloc(pvhip_recv_post_bf16_weights_bytes(nK, psize));
         check(pvhip_recv_post_bf16_weights(C.W, C.Wlp, nK, psize, D.stream), "bf16 weights");
      }
      if (C.arith == PVHIP_ARITH_FP8) {
         C.Wlp = D.alloc(pvhip_recv_post_fp8_weights_bytes(nK, psize));
         check(pvhip_recv_post_fp8_weights(C.W, C.Wlp, nK, psize, D.stream), "fp8 weights");
         C.srcScale = fp8Scale(C.prm.num("fp8SourceBound", sourceBound(pre)));
      }
      pvhip_recv_post_args a = gatherArgs(C);
      C.scratchBytes         = pvhip_recv_post_scratch_bytes(&a);
      if (C.scratchBytes > 0) {
         C.scratch = D.alloc(C.scratchBytes);
      }
   }
   else {
      if (C.keyword == "TransposeConn") {
         std::vector<float> w = prePerspectiveWeights(*this, C, O);
         C.W                  = static_cast<float *>(D.alloc(w.size() * 4));
         D.h2d(C.W, w.data(), w.size() * 4);
      }
      else {
         C.W = O.W;
      }
      int ncol = 0, nrow = 0;
      if (pvhost_recv_pre_tables(&pre.loc, &post.loc, C.nxp, C.nyp, nullptr, nullptr, nullptr, nullptr, &ncol,
                nullptr, nullptr, nullptr, nullptr, &nrow)
          != 0) {
         fatal(C.name + ": scatter tables failed");
      }
      std::vector<int32_t> cp(post.loc.nx + 1), cx(std::max(ncol, 1)), cw(std::max(ncol, 1)),
            ck(std::max(ncol, 1)), rp((size_t)post.loc.ny * C.nyp + 1), ry(std::max(nrow, 1)),
            rw(std::max(nrow, 1)), rk(std::max(nrow, 1));
      if (pvhost_recv_pre_tables(&pre.loc, &post.loc, C.nxp, C.nyp, cp.data(), cx.data(), cw.data(), ck.data(),
                &ncol, rp.data(), ry.data(), rw.data(), rk.data(), &nrow)
          != 0) {
         fatal(C.name + ": scatter tables failed");
      }
      std::vector<int32_t> *v[8] = {&cp, &cx, &cw, &ck, &rp, &ry, &rw, &rk};
      for (int i = 0; i < 8; i++) {
         C.tab[i] = D.table(*v[i]);
      }
      pvhost_num_kernels(&pre.loc, &post.loc, &C.nkx, &C.nky);
      if (C.arith == PVHIP_ARITH_BF16) {
         C.arith = PVHIP_ARITH_FP32; // the ordered scatter has EXACT and FP32 forms
      }
   }
}

pvhip_recv_post_args Net::gatherArgs(Conn &C) {
   Layer &pre = layers[C.pre], &post = layers[C.post];
   size_t nres = (size_t)nbatch * post.loc.nx * post.loc.ny * post.loc.nf;
   pvhip_recv_post_args a{};
   a.pre           = pre.loc;
   a.post          = post.loc;
   a.nxp           = C.nxq;
   a.nyp           = C.nyq;
   a.nfp           = pre.loc.nf;
   a.num_kernels_x = C.nkx;
   a.num_kernels_y = C.nky;
   a.start_x       = C.tab[0];
   a.start_y       = C.tab[1];
   a.kernel_x      = C.tab[2];
   a.kernel_y      = C.tab[3];
   a.dt_factor     = 1.0f;
   a.pre_data      = pre.A;
   a.weights       = C.W;
   a.post_gsyn     = post.gsyn ? post.gsyn + (size_t)C.channel * nres : nullptr;
   a.accumulate    = 1;
   a.arith         = C.arith;
   a.weights_lp    = C.Wlp;
   a.inverse_x     = C.inv[0];
   a.inverse_y     = C.inv[1];
   a.window_x0     = C.win[0];
   a.window_nx     = C.win[1];
   a.window_y0     = C.win[2];
   a.window_ny     = C.win[3];
   a.scratch       = C.scratch;
   a.scratch_bytes = C.scratchBytes;
   a.cgemm_tile      = C.tile;
   a.cgemm_splits    = C.splits;
   a.lp_source_scale = C.srcScale;
   return a;
}

// --autotune: the channel GEMM's launch shape (window starts per workgroup,
// channel ranges) of every FP32/BF16 gather with several data patches per
// feature, timed with HIP events on the column's own buffers (median of 5
// launches after 2); the fastest is kept.  Deliveries write only GSyn, which
// every step resets.
void Net::autotune() {
   pvhip_event *e0 = nullptr, *e1 = nullptr;
   check(pvhip_event_create(&e0), "pvhip_event_create");
   check(pvhip_event_create(&e1), "pvhip_event_create");
   for (auto &C : conns) {
      if (C.mode != C_GATHER || !C.inv[0] || C.arith == PVHIP_ARITH_EXACT) {
         continue;
      }
      int chunks   = layers[C.pre].loc.nf / (C.arith == PVHIP_ARITH_BF16 ? 32 : 16);
      float bestMs = 0.0f;
      int bestT = 0, bestK = 0;
      for (int tile : {64, 128, 256}) {
         for (int k : {1, 2, 4, 8}) {
            if (k > chunks) {
               continue;
            }
            C.tile                 = tile;
            C.splits               = k;
            pvhip_recv_post_args a = gatherArgs(C);
            size_t need            = pvhip_recv_post_scratch_bytes(&a);
            if (need > C.scratchBytes) {
               C.scratch      = D.alloc(need);
               C.scratchBytes = need;
               a              = gatherArgs(C);
            }
            std::vector<float> ms;
            for (int rep = 0; rep < 7; rep++) {
               check(pvhip_event_record(e0, D.stream), "pvhip_event_record");
               check(pvhip_recv_post(&a, D.stream), "pvhip_recv_post");
               check(pvhip_event_record(e1, D.stream), "pvhip_event_record");
               check(pvhip_stream_sync(D.stream), "pvhip_stream_sync");
               float t = 0.0f;
               check(pvhip_event_elapsed_ms(e0, e1, &t), "pvhip_event_elapsed_ms");
               if (rep >= 2) {
                  ms.push_back(t);
               }
            }
            std::sort(ms.begin(), ms.end());
            float med = ms[ms.size() / 2];
            if (bestT == 0 || med < bestMs) {
               bestMs = med;
               bestT  = tile;
               bestK  = k;
            }
         }
      }
      C.tile   = bestT;
      C.splits = bestK;
      std::printf("pvrun: %s: channel GEMM tile %d, splits %d (%.4f ms)\n", C.name.c_str(), bestT, bestK, bestMs);
   }
   pvhip_event_destroy(e0);
   pvhip_event_destroy(e1);
}

void Net::deliver(Conn &C) {
   Layer &pre = layers[C.pre], &post = layers[C.post];
   size_t nres = (size_t)nbatch * post.loc.nx * post.loc.ny * post.loc.nf;
   float *g    = post.gsyn + (size_t)C.channel * nres;
   if (C.mode == C_IDENT) {
      check(pvhip_ident_deliver(&pre.loc, pre.A, g, 1, D.stream), "pvhip_ident_deliver");
   }
   else if (C.mode == C_GATHER) {
      pvhip_recv_post_args a = gatherArgs(C);
      check(pvhip_recv_post(&a, D.stream), "pvhip_recv_post");
   }
   else if (C.mode == C_SCATTER) {
      pvhip_recv_pre_args a{};
      a.pre = pre.loc;
      a.post = post.loc;
      a.nxp = C.nxp;
      a.nyp = C.nyp;
      a.nfp = C.nfp;
      a.num_kernels_x = C.nkx;
      a.num_kernels_y = C.nky;
      a.col_ptr = C.tab[0];
      a.col_xe = C.tab[1];
      a.col_wx = C.tab[2];
      a.col_kx = C.tab[3];
      a.row_ptr = C.tab[4];
      a.row_ye = C.tab[5];
      a.row_wy = C.tab[6];
      a.row_ky = C.tab[7];
      a.dt_factor = 1.0f;
      if (pre.needActive) {
         a.active = pre.entries;
         a.pos_start = pre.posStart;
      }
      else {
         a.pre_data = pre.A;
      }
      a.weights = C.W;
      a.post_gsyn = g;
      a.accumulate = 1;
      a.arith = C.arith;
      check(pvhip_recv_pre(&a, D.stream), "pvhip_recv_pre");
   }
}

// resetGSynBuffers, then every connection into L adds its channel
void Net::receive(Layer &L) {
   if (L.kind == L_INPUT) {
      return;
   }
   size_t nres = (size_t)nbatch * L.loc.nx * L.loc.ny * L.loc.nf;
   check(pvhip_fill_zero(L.gsyn, (size_t)L.numChannels * nres * 4, D.stream), "reset GSyn");
   for (auto &C : conns) {
      if (&layers[C.post] == &L && C.mode != C_NONE) {
         deliver(C);
      }
   }
}

void Net::updateLayer(Layer &L, double t) {
   if (L.kind == L_INPUT) {
      // LayerUpdateController::needUpdate (LayerUpdateController.cpp:318-335)
      if (L.displayPeriod > 0) {
         double n = (t - L.lastUpdate) / L.displayPeriod;
         if (std::fabs(n - std::nearbyint(n)) * L.displayPeriod < 0.5 * dtCol) {
            for (int b = 0; b < nbatch; b++) {
               const auto &fr = L.input.frames[(L.frameIndex + b) % L.input.frames.size()];
               std::vector<float> ext((size_t)(L.loc.nx + 2 * L.halo) * (L.loc.ny + 2 * L.halo) * L.loc.nf, 0.0f);
               for (int y = 0; y < L.loc.ny; y++) {
                  std::memcpy(&ext[((size_t)(y + L.halo) * (L.loc.nx + 2 * L.halo) + L.halo) * L.loc.nf],
                        &fr[(size_t)y * L.loc.nx * L.loc.nf], (size_t)L.loc.nx * L.loc.nf * 4);
               }
               D.h2d(L.A + (size_t)b * ext.size(), ext.data(), ext.size() * 4);
            }
            L.frameIndex = (L.frameIndex + nbatch) % (int)L.input.frames.size();
            L.lastUpdate = t;
         }
      }
      return;
   }
   if (L.kind == L_LCA) {
      std::vector<float> e(nbatch);
      pvhip_lca_exp_tau(L.dt.data(), nbatch, L.tau, e.data());
      D.h2d(L.expTau, e.data(), nbatch * 4);
      pvhip_lca_update_args u{};
      u.loc = L.loc;
      u.num_channels = L.numChannels;
      u.channel_coeffs[0] = 1.0f;
      u.channel_coeffs[1] = -1.0f;
      u.gsyn = L.gsyn;
      u.V = L.V;
      u.A = L.A;
      u.exp_tau = L.expTau;
      u.self_interact = L.selfInteract;
      u.transfer = L.transfer;
      u.activity_from_v = 1;
      if (L.momentum) {
         check(pvhip_momentum_lca_update(&u, L.momentumRate, L.prevDrive, D.stream),
               "pvhip_momentum_lca_update");
      }
      else {
         check(pvhip_lca_update(&u, D.stream), "pvhip_lca_update");
      }
   }
   else {
      pvhip_layer_update_args u{};
      u.loc = L.loc;
      u.num_channels = L.numChannels;
      u.channel_coeffs[0] = 1.0f;
      u.channel_coeffs[1] = -1.0f;
      u.gsyn = L.gsyn;
      u.V = L.V;
      u.A = L.A;
      u.scale = L.errScale;
      u.apply_scale = L.kind == L_ERROR ? 1 : 0;
      u.transfer = L.transfer;
      check(pvhip_layer_update(&u, D.stream), "pvhip_layer_update");
   }
   if (L.needActive) {
      check(pvhip_active_compact(&L.loc, L.A, L.entries, L.numActive, L.posStart, L.compactScratch, D.stream),
            "pvhip_active_compact");
   }
}

std::vector<double> Net::probeEnergy(Probe &p) {
   if (p.keyword == "ColumnEnergyProbe") {
      std::vector<double> sum(nbatch, 0.0);
      for (auto &q : probes) {
         if (q.energyProbe == p.name) {
            std::vector<double> v = probeEnergy(q);
            for (int b = 0; b < nbatch; b++) {
               sum[b] += q.coefficient * v[b];
            }
         }
      }
      return sum;
   }
   Layer &L = layers[p.target];
   int kind = 0;
   double p0 = 0, p1 = 0;
   if (p.keyword == "L2NormProbe") {
      kind = 0;
   }
   else if (p.keyword == "FirmThresholdCostFnLCAProbe") {
      kind = 2;
      p0   = (double)(float)L.prm.num("VThresh", 0.0);
      p1   = (double)(float)L.prm.num("VWidth", 0.0);
   }
   else {
      kind = 1;
   }
   check(pvhip_cost_sum_f64(&L.loc, L.A, 1, kind, p0, p1, energyDev, D.stream), "pvhip_cost_sum_f64");
   std::vector<double> v(nbatch);
   D.d2h(v.data(), energyDev, nbatch * 8);
   if (p.keyword == "L2NormProbe") {
      double ex = p.prm.num("exponent", 1.0);
      for (auto &x : v) {
         x = ex == 2.0 ? x : std::pow(x, ex / 2.0);
      }
   }
   return v;
}

// RequireAllZeroActivityProbe (RequireAllZeroActivityProbe.cpp, CheckStatsAllZeros.cpp):
// a batch element fails when any restricted activity has |a| > nnzThreshold;
// the first failure is reported, at once with immediateExitOnFailure, else
// at the end, where exitOnFailure makes the run fail.
void Net::checkAllZero(Probe &p, double t) {
   Layer &L = layers[p.target];
   check(pvhip_cost_sum_f64(&L.loc, L.A, 1, 3, p.nnzThreshold, 0.0, energyDev, D.stream), "pvhip_cost_sum_f64");
   std::vector<double> cnt(nbatch);
   D.d2h(cnt.data(), energyDev, nbatch * 8);
   for (int b = 0; b < nbatch; b++) {
      if (p.out) {
         std::fprintf(p.out, "t = %g, b = %d, numNonzero = %.0f\n", t, b, cnt[b]);
      }
      if (cnt[b] != 0.0) {
         std::fprintf(stderr, "Probe %s has nonzero activity at time %g: batch element %d has %.0f values exceeding "
                      "the threshold %g\n", p.name.c_str(), t, b, cnt[b], p.nnzThreshold);
         if (!p.failed) {
            p.failed       = true;
            p.firstFailure = t;
         }
         if (p.immediateExit) {
            fatal("probe " + p.name + " failed");
         }
      }
   }
}

void Net::step(double t) {
   // AdaptTimestep (AdaptiveTimeScaleProbe::calcValues + calcTimesteps)
   for (auto &p : probes) {
      if (p.keyword != "AdaptiveTimeScaleProbe") {
         continue;
      }
      bool trig = false;
      if (!p.triggerLayer.empty()) {
         Layer &T = layers[layerIndex(p.triggerLayer)];
         if (T.kind == L_INPUT && T.displayPeriod > 0) {
            double n = (t - T.lastUpdate) / T.displayPeriod;
            trig     = std::fabs(n - std::nearbyint(n)) * T.displayPeriod < 0.5 * dtCol;
         }
      }
      int ti = probeIndex(p.targetName);
      for (int b = 0; b < nbatch; b++) {
         double Edt = trig || ti < 0 ? -1.0 : probes[ti].prev[b];
         double E0  = p.tsTrue[b];
         p.tsTrue[b] = Edt;
         if (Edt == E0) {
            continue;
         }
         double dE = (E0 - Edt) / p.ts[b];
         if (dE < 0.0 || E0 <= 0.0 || Edt <= 0.0) {
            p.ts[b]    = p.baseMin;
            p.tsMax[b] = p.baseMax;
         }
         else {
            p.ts[b] = p.tauFactor * (E0 / dE);
            if (p.ts[b] >= p.tsMax[b]) {
               p.ts[b] = p.tsMax[b];
               p.tsMax[b] *= (1 + p.growthFactor);
            }
         }
      }
      for (auto &L : layers) {
         if (L.kind == L_LCA && L.adaptiveProbe == p.name) {
            L.dt = p.ts;
         }
      }
      if (p.out) { // AdaptiveTimeScaleProbeOutputter::printTimeScaleData
         for (int b = 0; b < nbatch; b++) {
            if (p.prm.num("writeTimeScaleFieldnames", 1) != 0) {
               std::fprintf(p.out, "sim_time = %f\n\tbatch = %d, timeScale = %10.8f, timeScaleTrue = %10.8f, "
                            "timeScaleMax = %10.8f\n", t, b, p.ts[b], p.tsTrue[b], p.tsMax[b]);
            }
            else {
               std::fprintf(p.out, "%f, %d, %10.8f, %10.8f, %10.8f\n", t, b, p.ts[b], p.tsTrue[b], p.tsMax[b]);
            }
         }
      }
   }
   int maxPhase = 0;
   for (auto &L : layers) {
      maxPhase = std::max(maxPhase, L.phase);
   }
   for (int ph = 0; ph <= maxPhase; ph++) {
      for (auto &L : layers) {
         if (L.phase == ph) {
            receive(L);
         }
      }
      for (auto &L : layers) {
         if (L.phase == ph) {
            updateLayer(L, t);
         }
      }
   }
   for (auto &p : probes) {
      if (p.keyword == "RequireAllZeroActivityProbe") {
         checkAllZero(p, t);
      }
   }
   // ColProbeOutputState: energies of this step
   for (auto &p : probes) {
      if (p.keyword == "ColumnEnergyProbe") {
         p.prev = probeEnergy(p);
         if (p.out) {
            for (int b = 0; b < nbatch; b++) {
               // ColumnEnergyOutputter.cpp:39
               std::fprintf(p.out, "%10f, %d, %10.9f\n", t, b, p.prev[b]);
            }
         }
      }
   }
   output(t);
}

void Net::output(double t) {
   for (auto &L : layers) {
      if (L.writeStep <= 0 || t < L.initialWriteTime) {
         continue;
      }
      double n = (t - L.initialWriteTime) / L.writeStep;
      if (std::fabs(n - std::nearbyint(n)) > 1e-9) {
         continue;
      }
      size_t nxe = L.loc.nx + 2 * L.halo, nye = L.loc.ny + 2 * L.halo;
      std::vector<float> ext((size_t)nbatch * nxe * nye * L.loc.nf);
      D.d2h(ext.data(), L.A, ext.size() * 4);
      for (int b = 0; b < nbatch; b++) {
         std::vector<float> fr((size_t)L.loc.nx * L.loc.ny * L.loc.nf);
         for (int y = 0; y < L.loc.ny; y++) {
            std::memcpy(&fr[(size_t)y * L.loc.nx * L.loc.nf],
                  &ext[(((size_t)b * nye + y + L.halo) * nxe + L.halo) * L.loc.nf], (size_t)L.loc.nx * L.loc.nf * 4);
         }
         L.outFrames.push_back(std::move(fr));
         L.outTimes.push_back(t);
      }
   }
}

int Net::finish() {
   for (auto &L : layers) {
      if (!L.outFrames.empty()) {
         writeActivity(outputPath + "/" + L.name + ".pvp", L.loc.nx, L.loc.ny, L.loc.nf, L.outFrames, L.outTimes);
      }
   }
   int status = 0;
   for (auto &p : probes) {
      if (p.out) {
         std::fclose(p.out);
      }
      if (p.failed) {
         std::fprintf(stderr, "Probe %s has nonzero activity beginning at time %g\n", p.name.c_str(),
               p.firstFailure);
         if (p.exitOnFailure) {
            status = 1;
         }
      }
   }
   return status;
}

} // namespace

int main(int argc, char **argv) {
   std::string paramsPath, outDir;
   long steps = -1;
   bool tune  = false;
   for (int i = 1; i < argc; i++) {
      std::string a = argv[i];
      if (a == "-p" && i + 1 < argc) {
         paramsPath = argv[++i];
      }
      else if (a == "-o" && i + 1 < argc) {
         outDir = argv[++i];
      }
      else if (a == "--steps" && i + 1 < argc) {
         steps = std::atol(argv[++i]);
      }
      else if (a == "--autotune") {
         tune = true;
      }
      else {
         std::fprintf(stderr, "usage: pvrun -p file.params [-o outputPath] [--steps N] [--autotune]\n");
         return 2;
      }
   }
   if (paramsPath.empty()) {
      std::fprintf(stderr, "usage: pvrun -p file.params [-o outputPath] [--steps N] [--autotune]\n");
      return 2;
   }
   pvparams *P = nullptr;
   if (pvparams_load(paramsPath.c_str(), &P) != 0) {
      std::fprintf(stderr, "pvrun: %s: %s\n", paramsPath.c_str(), pvparams_error());
      return 1;
   }
   int rc = 0;
   try {
      Net net;
      net.P          = P;
      net.outputPath = outDir;
      net.build();
      if (tune) {
         net.autotune();
      }
      long n = steps >= 0 ? steps : (long)std::lround(net.stopTime / net.dtCol);
      for (long s = 1; s <= n; s++) {
         net.step((double)s * net.dtCol);
      }
      check(pvhip_stream_sync(net.D.stream), "final sync");
      rc = net.finish();
      std::printf("pvrun: %ld steps done%s\n", n, rc ? ", a probe failed" : "");
   } catch (const std::exception &e) {
      std::fprintf(stderr, "pvrun: %s\n", e.what());
      rc = 1;
   }
   pvparams_free(P);
   return rc;
}
