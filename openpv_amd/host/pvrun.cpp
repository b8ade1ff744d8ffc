// pvrun -- runs the layers, connections and probes of an OpenPV .params file
// on the GPU through libpvhip (the drop-in for OpenPV's run of the same file
// for the delivery + LCA path).
//
//   pvrun -p file.params [-o outputPath] [--steps N] [--autotune]
//
// Supported objects (the LCA column of tests/GPULCATest and LCATest):
//   HyPerCol; PvpLayer, ANNErrorLayer, HyPerLCALayer, MomentumLCALayer, ISTALayer, ANNLayer,
//   HyPerLayer; IdentConn, HyPerConn and MomentumConn (FileWeight or IdentWeight,
//   normalizeL2, plasticity), TransposeConn, CloneConn; connection delays;
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
#include <chrono>
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

// Shared weights (type 5), one arbor, one frame per timestamp: each frame is a
// weight header (BufferUtils::buildSharedWeightHeader, BufferUtilsPvp.cpp:12-98:
// nx, ny, nf = the data patches along x, y, f; minVal / maxVal over the
// weights) followed by every patch as {u16 nxp, u16 nyp, u32 offset 0}
// (unshrunken) and its nxp*nyp*nfp floats (WeightsFileIO::writeSharedWeights,
// WeightsFileIO.cpp:268-298).
void writeWeights(const std::string &path, int nkx, int nky, int nkf, int nxp, int nyp, int nfp,
      const std::vector<std::vector<float>> &frames, const std::vector<double> &times) {
   std::ofstream f(path, std::ios::binary);
   size_t psize = (size_t)nxp * nyp * nfp;
   int32_t npatch = nkx * nky * nkf;
   for (size_t i = 0; i < frames.size(); i++) {
      const std::vector<float> &w = frames[i];
      float mn = w.empty() ? 0.0f : *std::min_element(w.begin(), w.end());
      float mx = w.empty() ? 0.0f : *std::max_element(w.begin(), w.end());
      int32_t h[18] = {104, 26, 5, nkx, nky, nkf, 1, 0, 4, 3, 1, 1, nkx, nky, 0, 0, 1, 1};
      f.write(reinterpret_cast<const char *>(h), sizeof h);
      f.write(reinterpret_cast<const char *>(&times[i]), sizeof(double));
      int32_t pq[3] = {nxp, nyp, nfp};
      f.write(reinterpret_cast<const char *>(pq), sizeof pq);
      f.write(reinterpret_cast<const char *>(&mn), 4);
      f.write(reinterpret_cast<const char *>(&mx), 4);
      f.write(reinterpret_cast<const char *>(&npatch), 4);
      for (int32_t p = 0; p < npatch; p++) {
         uint16_t dims[2] = {(uint16_t)nxp, (uint16_t)nyp};
         uint32_t off     = 0;
         f.write(reinterpret_cast<const char *>(dims), sizeof dims);
         f.write(reinterpret_cast<const char *>(&off), sizeof off);
         f.write(reinterpret_cast<const char *>(w.data() + p * psize), psize * 4);
      }
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
   bool ista     = false;          // ISTALayer (ISTAInternalStateBuffer)
   float momentumRate = 0.0f;      // LCAMomentumRate
   float *prevDrive = nullptr;     // restricted, zero at start
   std::string adaptiveProbe;
   float *expTau = nullptr;
   std::vector<float> expTauHost; // the values expTau holds (uploaded when they change)
   std::vector<double> dt;
   // input
   PvpFrames input;
   int displayPeriod = 0, frameIndex = 0;
   double lastUpdate = 1.0;
   bool mirrorBC = false;
   // DataStore levels beyond 0 (src/columns/DataStore.cpp:18-38): ring[(head + d - 1) % D]
   // holds the activity published d steps before the live buffer A; allocated
   // when a connection reads this layer with a delay
   int levels = 1, head = 0;
   std::vector<float *> ring;
   // chunked BF16 copy of A ([nbatch][nf/32][nyExt][nxExt][32]), kept by the
   // layer's update when a dense BF16 scatter reads this layer
   void *Ab = nullptr;
   // HyPerLCA layer fed by one gather only: that connection's index, the
   // gather then applies the update in its epilogue (GSyn not materialised)
   int fusedConn = -1;
   // ANNErrorLayer fed by two IdentConn only (channels 0 and 1): their
   // indices, the update then runs pvhip_error_update (GSyn not materialised)
   int errIn = -1, errRec = -1;
   // layer fed by one dense scatter only: that connection's index, the
   // scatter then applies the update in its epilogue (GSyn not materialised)
   int denseFusedConn = -1;
   // LCA layer whose update writes only the chunked BF16 copy Ab (nothing in
   // the step reads its FP32 activity): A is derived from V on request (act)
   bool aLazy = false, aStale = false;
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
   bool dense    = false;    // a BF16 or FP32 GPU scatter run as pvhip_recv_pre_dense
   bool denseF32 = false;    // its FP32 form (FP32 activity and weights)
   void *Wdense  = nullptr;  // its BF16 (FP32) weights [nyp*nxp*nfp][pre.nf]
   int dx0 = 0, dy0 = 0;     // its patch start of source extended column / row 0
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
   // a delivery whose weights derive from a plastic original: dev weights =
   // permutation of the original's (index table; NULL when it reads them directly)
   int32_t *perm = nullptr;
   size_t permN  = 0;
   // plasticity of an original (HebbianUpdater / MomentumUpdater)
   bool plastic = false, immediate = true, normalizeDw = true, momentum = false, deprecatedTau = false;
   float dWMax = 0.0f, dWMaxDecayFactor = 0.0f, dwFactor = 0.0f, wFactor = 0.0f;
   int dWMaxDecayInterval = 0, dWMaxDecayTimer = 0;
   std::string triggerLayer;
   double triggerOffset = 0.0, updatePeriod = 0.0, lastUpdateTime = 0.0, weightUpdateTime = 0.0;
   int dwArith = PVHIP_ARITH_EXACT;
   int delay = 0; // delay in timesteps (ArborList::convertDelay)
   pvhip_sparse_entry *dEntries = nullptr; // active list of the delayed level (sparse scatter)
   int64_t *dNumActive = nullptr;
   uint32_t *dPosStart = nullptr, *dScratch = nullptr;
   float *dW = nullptr, *prevDw = nullptr;
   int64_t *counts = nullptr;
   int32_t *hebbTab[6] = {};
   int hebbWin[4] = {};
   void *hebbScratch = nullptr;
   size_t hebbScratchBytes = 0;
   bool updatedNow = false; // weights changed this timestep (NormalizeBase::weightsHaveUpdated)
   int lpKernels = 0, lpPatch = 0; // shape of the low-precision shadow of a gather's weights
   int momentumCode = 0;
   float momentumTau = 0.0f, momentumDecay = 0.0f;
   // NormalizeL2 of an original
   bool normalizeL2 = false, normalizeOnWeightUpdate = true;
   float strength = 1.0f, minL2 = 0.0f;
   // weight output of an original (WeightsPair::outputState): writeStep >= 0
   double writeStep = -1.0, writeTime = 0.0;
   std::vector<std::vector<float>> outFrames;
   std::vector<double> outTimes;
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
   // Checkpointer (src/checkpointing/Checkpointer.cpp): checkpointWrite with
   // checkpointWriteTriggerMode "step", initializeFromCheckpointDir
   bool checkpointWrite = false;
   std::string checkpointWriteDir, initFromCheckpointDir;
   long checkpointInterval = 1;
   int checkpointIndexWidth = -1;
   std::string outputPath, gpuArith = "fp32"; // outputPath: -o, else the HyPerCol's
   std::string gpuDwArith = "exact";          // dW correlation form of plastic connections
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
   bool denseRoute(Conn &c);
   void syncBf16(Layer &L);
   void deriveDense(Conn &c);
   void buildConn(Conn &c);
   void step(double t);
   pvhip_recv_post_args gatherArgs(Conn &c);
   pvhip_recv_pre_dense_args denseArgs(Conn &c);
   double sourceBound(const Layer &pre);
   void autotune();
   void deliver(Conn &c);
   void receive(Layer &L);
   void updateLayer(Layer &L, double t);
   void checkAllZero(Probe &p, double t);
   void loadInputFrames(Layer &L);
   float *levelData(Layer &L, int delay);
   float *act(Layer &L);
   void advanceDataStore(Layer &L);
   void readPlasticity(Conn &C);
   void momentumFactors(Conn &C);
   pvhip_hebbian_args hebbianArgs(Conn &C);
   bool needWeightUpdate(Conn &C, double t);
   void computeDw(Conn &C);
   void applyDw(Conn &C);
   void updateConnections(double t);
   void refreshDerived(Conn &O);
   std::vector<double> probeEnergy(Probe &p);
   void output(double t);
   int finish();
   std::string checkpointDir(long step, long finalStep) const;
   void writeCheckpoint(const std::string &dir, double simTime, long step);
   double readCheckpoint(const std::string &dir);
   std::vector<std::vector<float>> restrictedFrames(Layer &L, const float *ext, bool extended);
   void setRestrictedFrames(Layer &L, float *buf, bool extended, const std::vector<std::vector<float>> &frames,
         size_t first);
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
   if (c.has("dwArithmetic")) {
      gpuDwArith = c.str("dwArithmetic");
   }
   checkpointWrite = c.num("checkpointWrite", 0) != 0;
   if (checkpointWrite) {
      checkpointWriteDir = c.str("checkpointWriteDir");
      if (checkpointWriteDir.empty()) {
         fatal("checkpointWrite is set, so checkpointWriteDir is required");
      }
      std::string mode = c.str("checkpointWriteTriggerMode", "step");
      if (mode != "step" && mode != "Step" && mode != "STEP") {
         fatal("checkpointWriteTriggerMode \"" + mode + "\" is not supported (step)");
      }
      checkpointInterval = (long)c.num("checkpointWriteStepInterval", 1);
      if (checkpointInterval <= 0) {
         fatal("checkpointWriteStepInterval must be positive");
      }
      checkpointIndexWidth = (int)c.num("checkpointIndexWidth", -1);
   }
   initFromCheckpointDir = c.str("initializeFromCheckpointDir");
   if (dtCol != 1.0) {
      fatal("only dt = 1 is supported");
   }
   for (int g = 0; g < pvparams_num_groups(P); g++) {
      std::string kw = pvparams_group_keyword(P, g);
      Params q{P, g};
      if (kw == "PvpLayer" || kw == "ANNErrorLayer" || kw == "HyPerLCALayer" || kw == "MomentumLCALayer"
          || kw == "ISTALayer" || kw == "ANNLayer" || kw == "HyPerLayer") {
         Layer L;
         L.name  = pvparams_group_name(P, g);
         L.prm   = q;
         L.kind  = kw == "PvpLayer" ? L_INPUT : kw == "ANNErrorLayer" ? L_ERROR
                 : kw == "HyPerLCALayer" || kw == "MomentumLCALayer" || kw == "ISTALayer" ? L_LCA : L_ANN;
         L.momentum = kw == "MomentumLCALayer";
         L.ista     = kw == "ISTALayer";
         L.phase = (int)q.num("phase", 0);
         L.loc.nbatch = nbatch;
         L.loc.nx     = (int)std::lround(nxGlobal * q.num("nxScale", 1.0));
         L.loc.ny     = (int)std::lround(nyGlobal * q.num("nyScale", 1.0));
         L.loc.nf     = (int)q.num("nf", 1);
         L.sparse     = q.num("sparseLayer", 0) != 0;
         L.writeStep  = (int)q.num("writeStep", -1);
         L.initialWriteTime = q.num("initialWriteTime", 0);
         L.mirrorBC = q.num("mirrorBCflag", 0) != 0;
         if (q.num("valueBC", 0) != 0) {
            fatal(L.name + ": a nonzero valueBC is not supported");
         }
         layers.push_back(std::move(L));
      }
      else if (kw == "IdentConn" || kw == "HyPerConn" || kw == "MomentumConn" || kw == "TransposeConn"
               || kw == "CloneConn") {
         Conn C;
         C.name     = pvparams_group_name(P, g);
         C.keyword  = kw == "MomentumConn" ? "HyPerConn" : kw; // a HyPerConn with a MomentumUpdater
         C.momentum = kw == "MomentumConn";
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
      // ArborList::initializeDelays / convertDelay (ArborList.cpp:102-124): one
      // delay for the one arbor, in timesteps
      double d[4] = {0};
      int nd      = pvparams_array(P, C.prm.g, "delay", d, 4);
      if (nd > 1) {
         fatal(C.name + ": one delay per connection (one arbor)");
      }
      C.delay = nd > 0 ? (int)std::nearbyint(d[0] / dtCol) : 0;
      if (C.delay < 0) {
         fatal(C.name + ": negative delay");
      }
      Layer &pl = layers[C.pre];
      pl.levels = std::max(pl.levels, C.delay + 1);
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
      if (L.mirrorBC && L.halo > 0) { // a layer without a halo has no border to mirror into
         fatal(L.name + ": mirrorBCflag is supported for layers without a halo only");
      }
      L.loc.halo = {L.halo, L.halo, L.halo, L.halo};
      size_t nres = (size_t)nbatch * L.loc.nx * L.loc.ny * L.loc.nf;
      size_t next = (size_t)nbatch * (L.loc.nx + 2 * L.halo) * (L.loc.ny + 2 * L.halo) * L.loc.nf;
      L.V    = static_cast<float *>(D.alloc(nres * 4));
      L.A    = static_cast<float *>(D.alloc(next * 4));
      L.gsyn = static_cast<float *>(D.alloc((size_t)L.numChannels * nres * 4));
      for (int lv = 1; lv < L.levels; lv++) { // older DataStore levels start at zero (RingBuffer)
         L.ring.push_back(static_cast<float *>(D.alloc(next * 4)));
      }
      bool updGpu = L.prm.num("updateGpu", 0) != 0;
      if (L.kind == L_INPUT) {
         L.input         = readActivity(L.prm.str("inputPath"));
         L.displayPeriod = (int)L.prm.num("displayPeriod", 0);
         if (L.input.nx != L.loc.nx || L.input.ny != L.loc.ny || L.input.nf != L.loc.nf) {
            fatal(L.name + ": input file size differs from the layer (autoResizeFlag is not supported)");
         }
         // InputActivityBuffer::initializeState: the first frames without advancing the index
         loadInputFrames(L);
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
         else if (init != "ZeroV" && init != "ConstantV") {
            fatal(L.name + ": InitVType " + init + " is not supported (ZeroV, ConstantV, InitVFromFile)");
         }
      }
      if (L.kind != L_INPUT && L.prm.str("InitVType", "ZeroV") == "ConstantV") { // InitV.cpp: valueV
         std::vector<float> v(nres, (float)L.prm.num("valueV", 0.0));
         D.h2d(L.V, v.data(), v.size() * 4);
      }
      else if (L.kind != L_INPUT && L.kind != L_LCA && L.prm.str("InitVType", "ZeroV") != "ZeroV") {
         fatal(L.name + ": InitVType " + L.prm.str("InitVType") + " is not supported (ZeroV, ConstantV)");
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
      if (init == "IdentWeight") {
         // InitIdentWeights: a one-to-one connection's weight 1 from feature k to
         // feature k (1x1 patches)
         if (C.nxp != 1 || C.nyp != 1 || C.nfp != pre.loc.nf || nK != pre.loc.nf) {
            fatal(C.name + ": IdentWeight is supported for 1x1 patches between layers of equal nf");
         }
         C.hostW.assign((size_t)nK * C.nfp, 0.0f);
         for (int k = 0; k < nK; k++) {
            C.hostW[(size_t)k * C.nfp + k] = 1.0f;
         }
      }
      else if (init == "FileWeight") {
         C.hostW = readWeights(C.prm.str("initWeightsFile"), C.nxp, C.nyp, C.nfp, nK);
      }
      else {
         fatal(C.name + ": weightInitType " + init + " is not supported (FileWeight, IdentWeight)");
      }
      C.W     = static_cast<float *>(D.alloc(C.hostW.size() * 4));
      D.h2d(C.W, C.hostW.data(), C.hostW.size() * 4);
      std::string norm = C.prm.str("normalizeMethod", "none");
      if (norm == "normalizeL2") {
         if (C.prm.num("normalizeArborsIndividually", 0) != 0 || C.prm.num("normalizeFromPostPerspective", 0) != 0
             || C.prm.num("rMinX", 0) != 0 || C.prm.num("rMinY", 0) != 0 || C.prm.num("normalize_cutoff", 0) != 0
             || C.prm.num("nonnegativeConstraintFlag", 0) != 0) {
            fatal(C.name + ": only plain normalizeL2 is supported");
         }
         C.normalizeL2             = true;
         C.strength                = (float)C.prm.num("strength", 1.0);
         C.minL2                   = (float)C.prm.num("minL2NormTolerated", 0.0);
         C.normalizeOnWeightUpdate = C.prm.num("normalizeOnWeightUpdate", 1) != 0;
         if (C.prm.num("normalizeOnInitialize", 1) != 0) {
            check(pvhip_normalize_l2(C.W, nK, C.nxp * C.nyp * C.nfp, C.strength, C.minL2, D.stream),
                  "pvhip_normalize_l2");
            D.d2h(C.hostW.data(), C.W, C.hostW.size() * 4);
         }
      }
      else if (norm != "none" && norm != "") {
         fatal(C.name + ": normalizeMethod " + norm + " is not supported");
      }
      readPlasticity(C);
      // weights are written only when writeStep is given (WeightsPair.cpp:53-95)
      C.writeStep = C.prm.has("writeStep") ? C.prm.num("writeStep", -1.0) : -1.0;
      C.writeTime = C.prm.num("initialWriteTime", 0.0);
      while (C.writeStep > 0 && C.writeTime < 0.0) {
         C.writeTime += C.writeStep;
      }
   }
   for (auto &C : conns) {
      if (C.mode == C_SCATTER && C.arith == PVHIP_ARITH_FP32 && denseRoute(C)) {
         C.dense    = true; // reads the source's FP32 activity
         C.denseF32 = true;
         std::printf("pvrun: %s: dense FP32 scatter (pvhip_recv_pre_dense)\n", C.name.c_str());
      }
      else if (C.mode == C_SCATTER && (C.arith == PVHIP_ARITH_BF16 || C.arith == PVHIP_ARITH_FP8)
               && denseRoute(C)) {
         C.dense  = true; // the FP8 form's scatter runs in BF16, as its transposed route does
         std::printf("pvrun: %s: dense BF16 scatter (pvhip_recv_pre_dense)\n", C.name.c_str());
         Layer &L = layers[C.pre];
         if (!L.Ab) {
            L.Ab = D.alloc((size_t)nbatch * (L.loc.nx + 2 * L.halo) * (L.loc.ny + 2 * L.halo) * L.loc.nf * 2);
            syncBf16(L);
         }
      }
      else if (C.mode == C_SCATTER && C.arith != PVHIP_ARITH_EXACT && transposedRoute(C)) {
         C.mode   = C_GATHER;
         C.routed = true;
      }
   }
   // A HyPerLCALayer whose one channel is fed by one gather alone (one data
   // patch per feature, no delay, its source updated in an earlier phase) has
   // its update applied in the gather's epilogue: the same arithmetic as
   // receive + pvhip_lca_update (LayerInputBuffer::recvAllSynapticInput, then
   // HyPerLCAInternalStateBuffer::updateBuffer), without GSyn's reset, write
   // and read.
   for (size_t li = 0; li < layers.size(); li++) {
      Layer &L = layers[li];
      if (L.kind != L_LCA || L.ista || L.momentum || L.numChannels != 1) {
         continue;
      }
      int found = -1, n = 0;
      for (size_t ci = 0; ci < conns.size(); ci++) {
         Conn &C = conns[ci];
         if (C.post != (int)li || C.mode == C_NONE) {
            continue;
         }
         n++;
         if (C.mode == C_GATHER && C.channel == 0 && C.delay == 0 && !C.inv[0]
             && layers[C.pre].phase < L.phase) {
            found = (int)ci;
         }
      }
      if (n == 1 && found >= 0) {
         L.fusedConn = found;
         std::printf("pvrun: %s: gather %s with the LCA update fused\n", L.name.c_str(),
               conns[found].name.c_str());
      }
   }
   // The same for two other layers of the column, where the receive and the
   // update see the same source activity (every source in another phase, so
   // none updates between the two): an ANNErrorLayer fed by two IdentConn
   // alone (channels 0 and 1) runs pvhip_error_update (both deliveries and
   // ErrScaleInternalStateBuffer's update in one pass), and a layer fed by
   // one dense scatter alone has its update applied in that scatter's
   // epilogue.  Both give the bits of the separate launches.
   for (size_t li = 0; li < layers.size(); li++) {
      Layer &L = layers[li];
      if (L.kind == L_INPUT || L.kind == L_LCA) {
         continue;
      }
      std::vector<int> in;
      for (size_t ci = 0; ci < conns.size(); ci++) {
         if (conns[ci].post == (int)li && conns[ci].mode != C_NONE) {
            in.push_back((int)ci);
         }
      }
      auto otherPhase = [&](int ci) { return layers[conns[ci].pre].phase != L.phase; };
      auto sameShape  = [&](int ci) {
         const pvhip_loc &s = layers[conns[ci].pre].loc;
         return s.nbatch == L.loc.nbatch && s.nx == L.loc.nx && s.ny == L.loc.ny && s.nf == L.loc.nf;
      };
      if (L.kind == L_ERROR && L.numChannels == 2 && in.size() == 2) {
         int c0 = conns[in[0]].channel == 0 ? in[0] : in[1];
         int c1 = c0 == in[0] ? in[1] : in[0];
         bool ok = conns[c0].channel == 0 && conns[c1].channel == 1;
         for (int ci : {c0, c1}) {
            ok = ok && conns[ci].mode == C_IDENT && conns[ci].delay == 0 && otherPhase(ci) && sameShape(ci);
         }
         if (ok) {
            L.errIn  = c0;
            L.errRec = c1;
            std::printf("pvrun: %s: deliveries %s, %s and the update in one pass\n", L.name.c_str(),
                  conns[c0].name.c_str(), conns[c1].name.c_str());
         }
      }
      else if (L.numChannels == 1 && in.size() == 1) {
         Conn &C = conns[in[0]];
         if (C.mode == C_SCATTER && C.dense && C.channel == 0 && C.delay == 0 && otherPhase(in[0])) {
            L.denseFusedConn = in[0];
            std::printf("pvrun: %s: dense scatter %s with the layer update fused\n", L.name.c_str(),
                  C.name.c_str());
         }
      }
   }
   for (auto &C : conns) {
      Layer &L = layers[C.pre];
      if (C.dense) {
         continue; // reads the dense activity, no active list
      }
      if (C.mode == C_SCATTER && L.sparse && C.delay == 0) {
         L.needActive = true;
      }
      else if (C.mode == C_SCATTER && L.sparse) {
         size_t next = (size_t)nbatch * (L.loc.nx + 2 * L.halo) * (L.loc.ny + 2 * L.halo) * L.loc.nf;
         size_t npos = (size_t)(L.loc.nx + 2 * L.halo) * (L.loc.ny + 2 * L.halo) + 1;
         C.dEntries   = static_cast<pvhip_sparse_entry *>(D.alloc(next * 8));
         C.dNumActive = static_cast<int64_t *>(D.alloc(nbatch * 8));
         C.dPosStart  = static_cast<uint32_t *>(D.alloc(nbatch * npos * 4));
         C.dScratch   = static_cast<uint32_t *>(D.alloc(pvhip_active_compact_scratch_bytes(&L.loc)));
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
   // An LCA layer that keeps the chunked BF16 copy for its dense scatter,
   // publishes no active list, writes no output and is no probe's target:
   // its update writes only the copy, and the occasional readers of its FP32
   // activity (other connections, checkpoints) get A = T(V) from act(), the
   // bits the update would write.  A layer read every step keeps writing A
   // in its update, which costs less than deriving it afterwards.
   for (size_t li = 0; li < layers.size(); li++) {
      Layer &L   = layers[li];
      bool probe = false;
      for (auto &p : probes) {
         probe = probe || p.target == (int)li;
      }
      L.aLazy = L.kind == L_LCA && L.Ab && !L.needActive && !L.ista && !L.momentum && !probe
                && L.writeStep <= 0;
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
// `src` holds the original's pre-perspective weights.
std::vector<float> prePerspectiveWeights(Net &net, Conn &C, Conn &O, const std::vector<float> &src) {
   if (C.keyword != "TransposeConn") {
      return src;
   }
   Layer &opre = net.layers[O.pre], &opost = net.layers[O.post];
   int pkx, pky;
   pvhost_num_kernels(&opost.loc, &opre.loc, &pkx, &pky);
   std::vector<float> wt((size_t)pkx * pky * opost.loc.nf * C.nxp * C.nyp * C.nfp);
   if (pvhost_transpose_weights(&opre.loc, &opost.loc, O.nxp, O.nyp, O.nfp, src.data(), wt.data()) != 0) {
      fatal(C.name + ": weight transpose failed");
   }
   return wt;
}

// The weights a delivery reads when they are not its original's own buffer:
// its pre-perspective weights, transposed again for a gather
// (TransposeWeights::transposeShared), from the original's weights `src`.
std::vector<float> deliveryWeights(Net &net, Conn &C, Conn &O, const std::vector<float> &src) {
   std::vector<float> w = prePerspectiveWeights(net, C, O, src);
   if (C.mode != C_GATHER) {
      return w;
   }
   Layer &pre = net.layers[C.pre], &post = net.layers[C.post];
   int pkx, pky;
   pvhost_num_kernels(&post.loc, &pre.loc, &pkx, &pky);
   std::vector<float> wt((size_t)pkx * pky * post.loc.nf * C.nxq * C.nyq * pre.loc.nf);
   if (pvhost_transpose_weights(&pre.loc, &post.loc, C.nxp, C.nyp, C.nfp, w.data(), wt.data()) != 0) {
      fatal(C.name + ": weight transpose failed");
   }
   return wt;
}

// For a delivery of a plastic original: the same derivation as a table of
// indices into the original's weights (pvhip_permute_weights), so that each
// weight update re-derives the delivery's copy on the device
// (WeightsPair::finalizeUpdate, WeightsPair.cpp:261-280).  Each weight is
// carried through the derivation as its index + 1 (exact in float below
// 2^24); 0 marks a weight no original weight reaches.
void derivationTable(Net &net, Conn &C, Conn &O, size_t n) {
   if (O.hostW.size() >= (1u << 24)) {
      fatal(C.name + ": too many weights for the derivation table");
   }
   std::vector<float> code(O.hostW.size());
   for (size_t i = 0; i < code.size(); i++) {
      code[i] = (float)(i + 1);
   }
   std::vector<float> d = deliveryWeights(net, C, O, code);
   if (d.size() != n) {
      fatal(C.name + ": derivation table size differs from the delivery's weights");
   }
   std::vector<int32_t> idx(n);
   for (size_t j = 0; j < n; j++) {
      idx[j] = (int32_t)d[j] - 1;
   }
   C.perm  = net.D.table(idx);
   C.permN = n;
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

// The dense scatter (pvhip_recv_pre_dense, BF16 or FP32): one data patch per source
// feature into a target 4x finer, 16 offsets per class of 4x4, source
// features 128/256/512, no delay, and a source halo that holds only zeros
// (no mirror boundary).
bool Net::denseRoute(Conn &C) {
   Layer &pre = layers[C.pre], &post = layers[C.post];
   int nkx, nky;
   pvhost_num_kernels(&pre.loc, &post.loc, &nkx, &nky);
   int const f = pre.loc.nf;
   if (C.delay != 0 || pre.mirrorBC || nkx != 1 || nky != 1 || post.loc.nx != 4 * pre.loc.nx
       || post.loc.ny != 4 * pre.loc.ny || C.nxp % 4 || C.nyp % 4 || (C.nxp / 4) * (C.nyp / 4) * C.nfp != 16
       || (f != 128 && f != 256 && f != 512)) {
      return false;
   }
   if (pvhost_recv_pre_dense_start(&pre.loc, &post.loc, C.nxp, C.nyp, nullptr, &C.dx0, &C.dy0) != 0) {
      return false;
   }
   // the workgroup's weights and target tile fit its LDS
   pvhip_recv_pre_dense_args a{};
   a.pre          = pre.loc;
   a.post         = post.loc;
   a.nxp          = C.nxp;
   a.nyp          = C.nyp;
   a.nfp          = C.nfp;
   a.start_x0     = C.dx0;
   a.start_y0     = C.dy0;
   a.halo_is_zero = 1;
   a.fp32         = C.arith == PVHIP_ARITH_FP32;
   return pvhip_recv_pre_dense_fits(&a) != 0;
}

// The dense scatter's transposed weights ([nyp*nxp*nfp][pre.nf], BF16 or FP32)
// from the connection's weights.
void Net::deriveDense(Conn &C) {
   int const nf    = layers[C.pre].loc.nf;
   int const patch = C.nxp * C.nyp * C.nfp;
   if (C.denseF32) {
      check(pvhip_recv_pre_dense_weights_f32(C.W, static_cast<float *>(C.Wdense), nf, patch, D.stream),
            "pvhip_recv_pre_dense_weights_f32");
   }
   else {
      check(pvhip_recv_pre_dense_weights(C.W, C.Wdense, nf, patch, D.stream), "pvhip_recv_pre_dense_weights");
   }
}

// A layer's chunked BF16 activity from its FP32 activity (after A was set
// other than by its update: initial state, checkpoint restore).
void Net::syncBf16(Layer &L) {
   if (L.Ab) {
      check(pvhip_activity_to_bf16c(&L.loc, act(L), L.Ab, D.stream), "pvhip_activity_to_bf16c");
   }
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
         C.nxq                 = pvhost_post_patch_size(C.nxp, pre.loc.nx, post.loc.nx);
         C.nyq                 = pvhost_post_patch_size(C.nyp, pre.loc.ny, post.loc.ny);
         std::vector<float> wt = deliveryWeights(*this, C, O, O.hostW);
         C.W                   = static_cast<float *>(D.alloc(wt.size() * 4));
         D.h2d(C.W, wt.data(), wt.size() * 4);
         if (O.plastic) {
            derivationTable(*this, C, O, wt.size());
         }
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
      C.lpKernels = nK;
      C.lpPatch   = psize;
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
         std::vector<float> w = deliveryWeights(*this, C, O, O.hostW);
         C.W                  = static_cast<float *>(D.alloc(w.size() * 4));
         D.h2d(C.W, w.data(), w.size() * 4);
         if (O.plastic) {
            derivationTable(*this, C, O, w.size());
         }
      }
      else {
         C.W = O.W;
      }
      if (C.dense) {
         int const patch = C.nxp * C.nyp * C.nfp;
         C.Wdense        = D.alloc(C.denseF32 ? (size_t)pre.loc.nf * patch * 4
                                              : pvhip_recv_pre_dense_weights_bytes(pre.loc.nf, patch));
         deriveDense(C);
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
   a.pre_data      = levelData(pre, C.delay);
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

// The dense scatter's arguments for C (no output set).
pvhip_recv_pre_dense_args Net::denseArgs(Conn &C) {
   Layer &pre = layers[C.pre], &post = layers[C.post];
   pvhip_recv_pre_dense_args a{};
   a.pre          = pre.loc;
   a.post         = post.loc;
   a.nxp          = C.nxp;
   a.nyp          = C.nyp;
   a.nfp          = C.nfp;
   a.start_x0     = C.dx0;
   a.start_y0     = C.dy0;
   a.halo_is_zero = 1; // denseRoute: no mirror boundary writes the halo
   if (C.denseF32) {
      a.fp32        = 1;
      a.pre_data    = act(pre);
      a.weights_f32 = static_cast<const float *>(C.Wdense);
   }
   else {
      a.pre_data_bf16c = pre.Ab;
      a.weights_bf16   = C.Wdense;
   }
   a.dt_factor = 1.0f;
   return a;
}

void Net::deliver(Conn &C) {
   Layer &pre = layers[C.pre], &post = layers[C.post];
   size_t nres = (size_t)nbatch * post.loc.nx * post.loc.ny * post.loc.nf;
   float *g    = post.gsyn + (size_t)C.channel * nres;
   if (C.mode == C_IDENT) {
      check(pvhip_ident_deliver(&pre.loc, levelData(pre, C.delay), g, 1, D.stream), "pvhip_ident_deliver");
   }
   else if (C.mode == C_GATHER) {
      pvhip_recv_post_args a = gatherArgs(C);
      check(pvhip_recv_post(&a, D.stream), "pvhip_recv_post");
   }
   else if (C.mode == C_SCATTER && C.dense) {
      pvhip_recv_pre_dense_args a = denseArgs(C);
      a.post_gsyn                 = g;
      a.accumulate                = 1;
      check(pvhip_recv_pre_dense(&a, D.stream), "pvhip_recv_pre_dense");
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
      if (pre.needActive && C.delay == 0) {
         a.active = pre.entries;
         a.pos_start = pre.posStart;
      }
      else if (C.dEntries) { // the delayed level's active list (DataStore::updateActiveIndices)
         check(pvhip_active_compact(&pre.loc, levelData(pre, C.delay), C.dEntries, C.dNumActive, C.dPosStart,
                     C.dScratch, D.stream),
               "pvhip_active_compact");
         a.active    = C.dEntries;
         a.pos_start = C.dPosStart;
      }
      else {
         a.pre_data = levelData(pre, C.delay);
      }
      a.weights = C.W;
      a.post_gsyn = g;
      a.accumulate = 1;
      a.arith = C.arith;
      check(pvhip_recv_pre(&a, D.stream), "pvhip_recv_pre");
   }
}

// The input layer's activity from the frames at its index (one per batch
// element), halo zero (InputActivityBuffer::retrieveInput, valueBC 0).
void Net::loadInputFrames(Layer &L) {
   size_t nxe = L.loc.nx + 2 * L.halo, nye = L.loc.ny + 2 * L.halo;
   for (int b = 0; b < nbatch; b++) {
      const auto &fr = L.input.frames[(L.frameIndex + b) % L.input.frames.size()];
      std::vector<float> ext(nxe * nye * L.loc.nf, 0.0f);
      for (int y = 0; y < L.loc.ny; y++) {
         std::memcpy(&ext[((size_t)(y + L.halo) * nxe + L.halo) * L.loc.nf], &fr[(size_t)y * L.loc.nx * L.loc.nf],
               (size_t)L.loc.nx * L.loc.nf * 4);
      }
      D.h2d(L.A + (size_t)b * ext.size(), ext.data(), ext.size() * 4);
   }
}

// L's FP32 activity.  A layer whose update wrote only its BF16 copy (aLazy)
// has it derived here, A = T(V) (ANNActivityBuffer's transfer of the V the
// update wrote: the same bits), once per update, when a reader asks.
float *Net::act(Layer &L) {
   if (L.aStale) {
      check(pvhip_ann_activity(&L.loc, &L.transfer, L.V, L.A, D.stream), "pvhip_ann_activity");
      L.aStale = false;
   }
   return L.A;
}

// DataStore level `delay` of L: the live activity for 0, else the ring slot
// holding the activity published `delay` steps before it.
float *Net::levelData(Layer &L, int delay) {
   if (delay == 0) {
      return act(L);
   }
   return L.ring[(L.head + delay - 1) % (int)L.ring.size()];
}

// Publisher::increaseTimeLevel + publish / copyForward (Publisher.cpp:104-135,
// BasePublisherComponent.cpp:182-197): before the live buffer takes this
// step's activity, the level it held becomes level 1 (every step, whether or
// not the layer updates).
void Net::advanceDataStore(Layer &L) {
   if (L.ring.empty()) {
      return;
   }
   int nlev = (int)L.ring.size();
   L.head   = (L.head + nlev - 1) % nlev;
   size_t next = (size_t)nbatch * (L.loc.nx + 2 * L.halo) * (L.loc.ny + 2 * L.halo) * L.loc.nf;
   check(pvhip_copy_device(L.ring[L.head], act(L), next * 4, D.stream), "DataStore level copy");
}

// resetGSynBuffers, then every connection into L adds its channel
void Net::receive(Layer &L) {
   if (L.kind == L_INPUT || L.fusedConn >= 0 || L.errIn >= 0 || L.denseFusedConn >= 0) {
      return; // a fused layer's deliveries run in its update
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
            loadInputFrames(L);
            syncBf16(L);
            L.frameIndex = (L.frameIndex + nbatch) % (int)L.input.frames.size();
            L.lastUpdate = t;
         }
      }
      return;
   }
   if (L.kind == L_LCA) {
      std::vector<float> e(nbatch);
      if (L.ista) { // (float)dt/tau per batch element (ISTAInternalStateBuffer.kpp:56)
         pvhip_ista_rate(L.dt.data(), nbatch, L.tau, e.data());
      }
      else {
         pvhip_lca_exp_tau(L.dt.data(), nbatch, L.tau, e.data());
      }
      if (e != L.expTauHost) { // a synchronous copy: only when dt changed
         D.h2d(L.expTau, e.data(), nbatch * 4);
         L.expTauHost = e;
      }
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
      u.A_bf16c = L.Ab;
      if (L.aLazy) {
         u.A      = nullptr; // only the BF16 copy; act() derives A when read
         L.aStale = true;
      }
      if (L.fusedConn >= 0) {
         // the gather with the update in its epilogue (one channel, coefficient 1)
         u.num_channels           = 1;
         u.gsyn                   = nullptr;
         pvhip_recv_post_args a   = gatherArgs(conns[L.fusedConn]);
         a.post_gsyn              = nullptr;
         a.accumulate             = 0;
         a.fused_update           = &u;
         check(pvhip_recv_post(&a, D.stream), "pvhip_recv_post (fused LCA update)");
      }
      else if (L.ista) {
         check(pvhip_ista_update(&u, (float)L.prm.num("VThresh", -FLT_MAX), D.stream), "pvhip_ista_update");
      }
      else if (L.momentum) {
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
      if (L.errIn >= 0) {
         // both IdentConn deliveries and the update in one pass
         Layer &in = layers[conns[L.errIn].pre], &rec = layers[conns[L.errRec].pre];
         pvhip_error_update_args x{};
         x.loc               = L.loc;
         x.in_loc            = in.loc;
         x.in                = act(in);
         x.rec_loc           = rec.loc;
         x.rec               = act(rec);
         x.channel_coeffs[0] = 1.0f;
         x.channel_coeffs[1] = -1.0f;
         x.scale             = u.scale;
         x.apply_scale       = u.apply_scale;
         x.transfer          = u.transfer;
         x.V                 = L.V;
         x.A                 = L.A;
         check(pvhip_error_update(&x, D.stream), "pvhip_error_update");
      }
      else if (L.denseFusedConn >= 0) {
         // the dense scatter with this update in its epilogue (one channel)
         u.num_channels               = 1;
         u.gsyn                       = nullptr;
         pvhip_recv_pre_dense_args a  = denseArgs(conns[L.denseFusedConn]);
         a.post_gsyn                  = nullptr;
         a.accumulate                 = 0;
         a.fused_layer                = &u;
         check(pvhip_recv_pre_dense(&a, D.stream), "pvhip_recv_pre_dense (fused layer update)");
      }
      else {
         check(pvhip_layer_update(&u, D.stream), "pvhip_layer_update");
      }
      syncBf16(L);
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
   check(pvhip_cost_sum_f64(&L.loc, act(L), 1, kind, p0, p1, energyDev, D.stream), "pvhip_cost_sum_f64");
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

// ------------------------------------------------------------------ plasticity
// HebbianUpdater (src/weightupdaters/HebbianUpdater.cpp) and MomentumUpdater
// (src/weightupdaters/MomentumUpdater.cpp) of a HyPerConn / MomentumConn:
// the parameters of ioParamsFillGroup (:32-170; MomentumUpdater.cpp:30-145),
// the device buffers of dW, the activation counts and prev_dW.
void Net::readPlasticity(Conn &C) {
   C.plastic = C.prm.num("plasticityFlag", 0) != 0;
   if (!C.plastic) {
      return;
   }
   Layer &pre = layers[C.pre], &post = layers[C.post];
   if (!C.prm.has("dWMax")) {
      fatal(C.name + ": plasticityFlag is set, so dWMax is required");
   }
   C.dWMax              = (float)C.prm.num("dWMax", 0.0);
   C.dWMaxDecayInterval = (int)C.prm.num("dWMaxDecayInterval", 0);
   C.dWMaxDecayFactor   = (float)C.prm.num("dWMaxDecayFactor", 0.0);
   if (C.dWMaxDecayFactor < 0.0f || C.dWMaxDecayFactor >= 1.0f) {
      fatal(C.name + ": dWMaxDecayFactor must be in the interval [0.0, 1.0)");
   }
   C.normalizeDw = C.prm.num("normalizeDw", 1) != 0;
   C.immediate   = C.prm.num("immediateWeightUpdate", 1) != 0;
   if (C.prm.num("useMask", 0) != 0) {
      fatal(C.name + " has useMask set to true. This parameter is obsolete.");
   }
   if (C.prm.num("combine_dW_with_W_flag", 0) != 0) {
      fatal(C.name + ": combine_dW_with_W_flag is not supported");
   }
   C.triggerLayer = C.prm.str("triggerLayerName");
   if (!C.triggerLayer.empty()) {
      layerIndex(C.triggerLayer);
      C.triggerOffset = C.prm.num("triggerOffset", 0.0);
      if (C.triggerOffset < 0) {
         fatal(C.name + ": triggerOffset must be positive");
      }
      if (C.prm.has("weightUpdatePeriod")) {
         fatal(C.name + " sets both triggerLayerName and weightUpdatePeriod; only one of these can be set.");
      }
   }
   else {
      if (!C.prm.has("weightUpdatePeriod")) {
         fatal(C.name + ": weightUpdatePeriod is required without a triggerLayerName");
      }
      C.updatePeriod = C.prm.num("weightUpdatePeriod", 1.0);
      double t0      = C.prm.num("initialWeightUpdateTime", 0.0);
      C.weightUpdateTime = t0;
      if (C.weightUpdateTime < 0.0) {
         while (C.weightUpdateTime <= 0.0) {
            C.weightUpdateTime += C.updatePeriod;
         }
      }
      C.lastUpdateTime = t0;
   }
   if (C.momentum) {
      std::string m = C.prm.str("momentumMethod");
      int code      = m == "viscosity" ? PVHIP_MOMENTUM_VISCOSITY
                      : m == "simple"  ? PVHIP_MOMENTUM_SIMPLE
                      : m == "alex"    ? PVHIP_MOMENTUM_ALEX
                                       : -1;
      if (code < 0) {
         fatal("MomentumUpdater " + C.name + ": momentumMethod of " + m
               + " is not known. Options are \"viscosity\", \"simple\", and \"alex\".");
      }
      C.momentumCode = code;
      // timeConstantTau wins over the deprecated momentumTau (MomentumUpdater.cpp:114-133)
      C.deprecatedTau = !C.prm.has("timeConstantTau") && C.prm.has("momentumTau");
      double dflt     = code == PVHIP_MOMENTUM_VISCOSITY ? 100.0 : code == PVHIP_MOMENTUM_SIMPLE ? 0.25 : 0.9;
      C.momentumTau   = (float)(C.deprecatedTau ? C.prm.num("momentumTau", 0.0) : C.prm.num("timeConstantTau", dflt));
      C.momentumDecay = (float)C.prm.num("momentumDecay", 0.0);
      if (!C.prm.str("initPrev_dWFile").empty()) {
         fatal(C.name + ": initPrev_dWFile is not supported");
      }
      momentumFactors(C); // the parameter checks, before any device work of the update
   }
   size_t n = C.hostW.size();
   C.dW     = static_cast<float *>(D.alloc(n * 4));
   if (C.normalizeDw) {
      C.counts = static_cast<int64_t *>(D.alloc(n * 8));
   }
   if (C.momentum) {
      C.prevDw = static_cast<float *>(D.alloc(n * 4));
   }
   // PatchGeometry over the pre extended axes: patch sizes, first kept item,
   // post extended start of each patch (aPostOffset)
   std::vector<int32_t> t[6];
   for (int axis = 0; axis < 2; axis++) {
      int ne = axis == 0 ? pre.loc.nx + pre.loc.halo.lt + pre.loc.halo.rt : pre.loc.ny + pre.loc.halo.dn + pre.loc.halo.up;
      std::vector<int32_t> dim(ne), ps(ne), pr(ne), pe(ne), us(ne);
      if (pvhost_patch_axis(&pre.loc, &post.loc, C.nxp, C.nyp, axis, dim.data(), ps.data(), pr.data(), pe.data(),
                us.data())
          != 0) {
         fatal(C.name + ": patch geometry failed");
      }
      t[axis]     = dim;
      t[2 + axis] = ps;
      t[4 + axis] = pe;
   }
   for (int i = 0; i < 6; i++) {
      C.hebbTab[i] = D.table(t[i]);
   }
   std::string ar = gpuDwArith;
   C.dwArith      = ar == "fp32" ? PVHIP_ARITH_FP32 : ar == "bf16" ? PVHIP_ARITH_BF16 : PVHIP_ARITH_EXACT;
   if (ar != "exact" && ar != "fp32" && ar != "bf16") {
      fatal("dwArithmetic \"" + ar + "\" is not exact, fp32 or bf16");
   }
   if (C.dwArith != PVHIP_ARITH_EXACT) {
      int nkx, nky;
      pvhost_num_kernels(&pre.loc, &post.loc, &nkx, &nky);
      if (nkx * nky != 1
          || pvhost_hebbian_windows(&pre.loc, &post.loc, C.nxp, C.nyp, &C.hebbWin[0], &C.hebbWin[1], &C.hebbWin[2],
                   &C.hebbWin[3])
                   != 0) {
         std::printf("pvrun: %s: the dW correlation runs in the EXACT form (the MFMA form needs one data patch per "
                     "feature and clipped windows)\n",
               C.name.c_str());
         C.dwArith = PVHIP_ARITH_EXACT;
      }
      else {
         pvhip_hebbian_args a = hebbianArgs(C);
         C.hebbScratchBytes   = pvhip_hebbian_dw_scratch_bytes(&a);
         if (C.hebbScratchBytes > 0) {
            C.hebbScratch = D.alloc(C.hebbScratchBytes);
         }
      }
   }
   std::printf("pvrun: %s: plastic (%s, dWMax %g, %s, %s)\n", C.name.c_str(),
         C.momentum ? "MomentumUpdater" : "HebbianUpdater", C.dWMax,
         C.triggerLayer.empty() ? "weightUpdatePeriod" : "triggered", C.immediate ? "immediate" : "delayed");
}

// The factors of MomentumUpdater::applyMomentum (or applyMomentumDeprecated)
// for the current dWMax (alex scales the decay by it).
void Net::momentumFactors(Conn &C) {
   int st = C.deprecatedTau
                  ? pvhip_momentum_factors_deprecated(C.momentumCode, C.momentumTau, C.momentumDecay, C.dWMax,
                          &C.dwFactor, &C.wFactor)
                  : pvhip_momentum_factors(C.momentumCode, C.momentumTau, C.momentumDecay, C.dWMax, &C.dwFactor,
                          &C.wFactor);
   if (st != PVHIP_OK) {
      fatal("MomentumUpdater " + C.name + ": " + pvhip_last_error());
   }
}

pvhip_hebbian_args Net::hebbianArgs(Conn &C) {
   Layer &pre = layers[C.pre], &post = layers[C.post];
   pvhip_hebbian_args a{};
   a.pre  = pre.loc;
   a.post = post.loc;
   a.nxp  = C.nxp;
   a.nyp  = C.nyp;
   a.nfp  = C.nfp;
   pvhost_num_kernels(&pre.loc, &post.loc, &a.num_kernels_x, &a.num_kernels_y);
   a.dim_x           = C.hebbTab[0];
   a.dim_y           = C.hebbTab[1];
   a.pstart_x        = C.hebbTab[2];
   a.pstart_y        = C.hebbTab[3];
   a.apost_x         = C.hebbTab[4];
   a.apost_y         = C.hebbTab[5];
   a.dw_max          = C.dWMax;
   a.pre_data        = levelData(pre, C.delay); // HebbianUpdater::update_dW: getLayerData(delay)
   a.post_data       = act(post);
   a.dw              = C.dW;
   a.counts          = C.counts;
   a.arith           = C.dwArith;
   a.window_stride_x = C.hebbWin[0];
   a.window_start_x  = C.hebbWin[1];
   a.window_stride_y = C.hebbWin[2];
   a.window_start_y  = C.hebbWin[3];
   a.scratch         = C.hebbScratch;
   a.scratch_bytes   = C.hebbScratchBytes;
   return a;
}

// HebbianUpdater::needUpdate (HebbianUpdater.cpp:361-374): the trigger layer's
// LayerUpdateController::needUpdate at t + triggerOffset, or every
// weightUpdatePeriod from the last update.
bool Net::needWeightUpdate(Conn &C, double t) {
   if (!C.triggerLayer.empty()) {
      Layer &T  = layers[layerIndex(C.triggerLayer)];
      double tt = t + C.triggerOffset;
      if (T.kind != L_INPUT) {
         return true; // a layer without a trigger of its own updates every timestep
      }
      if (T.displayPeriod <= 0) {
         return false;
      }
      double n = (tt - T.lastUpdate) / T.displayPeriod;
      return std::fabs(n - std::nearbyint(n)) * T.displayPeriod < 0.5 * dtCol;
   }
   double n = (t - C.lastUpdateTime) / C.updatePeriod;
   return std::fabs(n - std::nearbyint(n)) * C.updatePeriod < 0.5 * dtCol;
}

// updateLocal_dW: clear dW (and the activation counts), then update_dW
// (HebbianUpdater.cpp:384-560); reduce_dW is the identity on one rank.
void Net::computeDw(Conn &C) {
   size_t n = C.hostW.size();
   check(pvhip_fill_zero(C.dW, n * 4, D.stream), "clear dW");
   if (C.counts) {
      check(pvhip_fill_zero(C.counts, n * 8, D.stream), "clear activation counts");
   }
   pvhip_hebbian_args a = hebbianArgs(C);
   check(pvhip_hebbian_dw(&a, D.stream), "pvhip_hebbian_dw");
}

// blockingNormalize_dW + updateArbors (HebbianUpdater.cpp:773-833), or
// MomentumUpdater::updateWeights (MomentumUpdater.cpp:276-294).
void Net::applyDw(Conn &C) {
   size_t n = C.hostW.size();
   if (C.momentum) {
      momentumFactors(C);
      check(pvhip_momentum_apply(C.W, C.dW, C.prevDw, C.counts, n, C.normalizeDw, C.dwFactor, C.wFactor,
                  C.deprecatedTau, D.stream),
            "pvhip_momentum_apply");
   }
   else {
      check(pvhip_hebbian_apply(C.W, C.dW, C.counts, n, C.normalizeDw, D.stream), "pvhip_hebbian_apply");
   }
}

// WeightsPair::finalizeUpdate for the deliveries that read O's weights: their
// own copies re-derived on the device, and the low-precision shadows.
void Net::refreshDerived(Conn &O) {
   for (auto &X : conns) {
      if (X.mode != C_GATHER && X.mode != C_SCATTER) {
         continue;
      }
      if (&X != &O && X.original != O.name) {
         continue;
      }
      if (X.perm) {
         check(pvhip_permute_weights(O.W, X.perm, X.W, X.permN, D.stream), "pvhip_permute_weights");
      }
      if (X.dense) {
         deriveDense(X);
      }
      if (X.Wlp && X.arith == PVHIP_ARITH_BF16) {
         check(pvhip_recv_post_bf16_weights(X.W, X.Wlp, X.lpKernels, X.lpPatch, D.stream), "bf16 weights");
      }
      else if (X.Wlp && X.arith == PVHIP_ARITH_FP8) {
         check(pvhip_recv_post_fp8_weights(X.W, X.Wlp, X.lpKernels, X.lpPatch, D.stream), "fp8 weights");
      }
   }
}

// HyPerCol::advanceTime after the layer phases (HyPerCol.cpp:776-779):
// ConnectionUpdate (HebbianUpdater::updateState, :343-359), ConnectionNormalize
// (NormalizeBase::respondConnectionNormalize with normalizeOnWeightUpdate),
// ConnectionFinalizeUpdate.
void Net::updateConnections(double t) {
   for (auto &C : conns) {
      C.updatedNow = false;
      if (!C.plastic || !needWeightUpdate(C, t)) {
         continue;
      }
      if (C.immediate) {
         computeDw(C);
         applyDw(C);
      }
      else {
         applyDw(C);
         computeDw(C);
      }
      if (C.dWMaxDecayInterval > 0 && --C.dWMaxDecayTimer < 0) { // decay_dWMax (:835-844)
         C.dWMaxDecayTimer = C.dWMaxDecayInterval;
         C.dWMax *= 1.0f - C.dWMaxDecayFactor;
      }
      C.lastUpdateTime = t;
      if (C.triggerLayer.empty()) { // computeNewWeightUpdateTime (:846-853)
         while (t >= C.weightUpdateTime) {
            C.weightUpdateTime += C.updatePeriod;
         }
      }
      C.updatedNow = true;
   }
   for (auto &C : conns) {
      if (C.updatedNow && C.normalizeL2 && C.normalizeOnWeightUpdate) {
         int nK = (int)(C.hostW.size() / ((size_t)C.nxp * C.nyp * C.nfp));
         check(pvhip_normalize_l2(C.W, nK, C.nxp * C.nyp * C.nfp, C.strength, C.minL2, D.stream),
               "pvhip_normalize_l2");
      }
   }
   for (auto &C : conns) {
      if (C.updatedNow) {
         refreshDerived(C);
      }
   }
}

// RequireAllZeroActivityProbe (RequireAllZeroActivityProbe.cpp, CheckStatsAllZeros.cpp):
// a batch element fails when any restricted activity has |a| > nnzThreshold;
// the first failure is reported, at once with immediateExitOnFailure, else
// at the end, where exitOnFailure makes the run fail.
void Net::checkAllZero(Probe &p, double t) {
   Layer &L = layers[p.target];
   check(pvhip_cost_sum_f64(&L.loc, act(L), 1, 3, p.nnzThreshold, 0.0, energyDev, D.stream), "pvhip_cost_sum_f64");
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
            advanceDataStore(L);
            updateLayer(L, t);
         }
      }
   }
   updateConnections(t);
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
   for (auto &C : conns) { // WeightsPair::outputState (WeightsPair.cpp:327-337)
      if (C.keyword == "HyPerConn" && C.writeStep >= 0 && t >= C.writeTime) {
         C.writeTime += C.writeStep;
         std::vector<float> w(C.hostW.size());
         D.d2h(w.data(), C.W, w.size() * 4);
         C.outFrames.push_back(std::move(w));
         C.outTimes.push_back(t);
      }
   }
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
      D.d2h(ext.data(), act(L), ext.size() * 4);
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

// ------------------------------------------------------------------ checkpoints
// Checkpointer::makeCheckpointDirectoryFromCurrentStep (Checkpointer.cpp:785-795):
// <checkpointWriteDir>/Checkpoint<step>, zero-padded to checkpointIndexWidth
// or to the width of the final step number.
std::string Net::checkpointDir(long step, long finalStep) const {
   int width = checkpointIndexWidth >= 0 ? checkpointIndexWidth
                                          : (int)std::floor(std::log10((double)std::max(finalStep, 1L))) + 1;
   char num[64];
   std::snprintf(num, sizeof num, "%0*ld", width, step);
   return checkpointWriteDir + "/Checkpoint" + num;
}

// The restricted part of a layer buffer, one frame per batch element (the
// LayerFile of a checkpoint entry writes restricted data, fileExtended false).
std::vector<std::vector<float>> Net::restrictedFrames(Layer &L, const float *buf, bool extended) {
   size_t nxe = L.loc.nx + (extended ? 2 * L.halo : 0), nye = L.loc.ny + (extended ? 2 * L.halo : 0);
   int h      = extended ? L.halo : 0;
   std::vector<float> all((size_t)nbatch * nxe * nye * L.loc.nf);
   D.d2h(all.data(), buf, all.size() * 4);
   std::vector<std::vector<float>> out;
   for (int b = 0; b < nbatch; b++) {
      std::vector<float> fr((size_t)L.loc.nx * L.loc.ny * L.loc.nf);
      for (int y = 0; y < L.loc.ny; y++) {
         std::memcpy(&fr[(size_t)y * L.loc.nx * L.loc.nf], &all[(((size_t)b * nye + y + h) * nxe + h) * L.loc.nf],
               (size_t)L.loc.nx * L.loc.nf * 4);
      }
      out.push_back(std::move(fr));
   }
   return out;
}

// Inverse of restrictedFrames for frames[first .. first + nbatch); the halo
// of an extended buffer is left zero (valueBC 0, no mirror boundary).
void Net::setRestrictedFrames(Layer &L, float *buf, bool extended, const std::vector<std::vector<float>> &frames,
      size_t first) {
   size_t nxe = L.loc.nx + (extended ? 2 * L.halo : 0), nye = L.loc.ny + (extended ? 2 * L.halo : 0);
   int h      = extended ? L.halo : 0;
   if (frames.size() < first + nbatch) {
      fatal(L.name + ": checkpoint holds too few frames");
   }
   std::vector<float> all((size_t)nbatch * nxe * nye * L.loc.nf, 0.0f);
   for (int b = 0; b < nbatch; b++) {
      const auto &fr = frames[first + b];
      if (fr.size() != (size_t)L.loc.nx * L.loc.ny * L.loc.nf) {
         fatal(L.name + ": checkpoint frame size differs from the layer");
      }
      for (int y = 0; y < L.loc.ny; y++) {
         std::memcpy(&all[(((size_t)b * nye + y + h) * nxe + h) * L.loc.nf], &fr[(size_t)y * L.loc.nx * L.loc.nf],
               (size_t)L.loc.nx * L.loc.nf * 4);
      }
   }
   D.h2d(buf, all.data(), all.size() * 4);
}

namespace {
template <typename T>
void writeBin(const std::string &dir, const std::string &name, const T *v, size_t n) {
   // CheckpointEntryData::write (CheckpointEntryData.tpp:14-27): raw values and a text copy
   std::ofstream b(dir + "/" + name + ".bin", std::ios::binary);
   b.write(reinterpret_cast<const char *>(v), sizeof(T) * n);
   std::ofstream t(dir + "/" + name + ".txt");
   for (size_t i = 0; i < n; i++) {
      t << v[i] << "\n";
   }
}
template <typename T>
std::vector<T> readBin(const std::string &dir, const std::string &name, size_t n) {
   std::ifstream f(dir + "/" + name + ".bin", std::ios::binary);
   std::vector<T> v(n);
   if (!f || !f.read(reinterpret_cast<char *>(v.data()), sizeof(T) * n)) {
      fatal("checkpoint " + dir + ": cannot read " + name + ".bin");
   }
   return v;
}
} // namespace

// Checkpointer::checkpointToDirectory (Checkpointer.cpp:817-852) for the
// objects pvrun builds, in the reference's file names and formats: per layer
// <name>_V.pvp, <name>_A.pvp (restricted, one frame per batch element),
// <name>_Delays.pvp (every DataStore level), <name>_prevDrive.pvp
// (MomentumLCALayer), <name>_lastUpdateTime.bin and <name>_FrameNumbers.bin
// (input layers); per original connection <name>_W.pvp, and when plastic
// <name>_lastUpdateTime.bin, <name>_weightUpdateTime.bin (no trigger),
// <name>_dW.pvp (delayed updates), <name>_prev_dW.pvp (MomentumUpdater); per
// adaptive time-scale probe <name>_timescaleinfo.bin; timeinfo.bin last.
void Net::writeCheckpoint(const std::string &dir, double simTime, long step) {
   std::error_code ec;
   std::filesystem::create_directories(dir, ec);
   if (ec) {
      fatal("cannot create the checkpoint directory " + dir);
   }
   for (auto &L : layers) {
      auto put = [&](const std::string &label, const std::vector<std::vector<float>> &frames) {
         std::vector<double> times(frames.size(), simTime);
         writeActivity(dir + "/" + L.name + "_" + label + ".pvp", L.loc.nx, L.loc.ny, L.loc.nf, frames, times);
      };
      if (L.kind != L_INPUT) {
         put("V", restrictedFrames(L, L.V, false));
      }
      put("A", restrictedFrames(L, act(L), true));
      std::vector<std::vector<float>> levels;
      for (int lv = 0; lv < L.levels; lv++) {
         auto f = restrictedFrames(L, levelData(L, lv), true);
         levels.insert(levels.end(), f.begin(), f.end());
      }
      put("Delays", levels);
      if (L.momentum) {
         put("prevDrive", restrictedFrames(L, L.prevDrive, false));
      }
      if (L.kind == L_INPUT) {
         std::vector<int> idx(nbatch);
         for (int b = 0; b < nbatch; b++) {
            idx[b] = (L.frameIndex + b) % (int)L.input.frames.size();
         }
         writeBin(dir, L.name + "_FrameNumbers", idx.data(), idx.size());
         writeBin(dir, L.name + "_lastUpdateTime", &L.lastUpdate, 1);
      }
   }
   for (auto &C : conns) {
      if (C.keyword != "HyPerConn") {
         continue;
      }
      int nkx, nky;
      pvhost_num_kernels(&layers[C.pre].loc, &layers[C.post].loc, &nkx, &nky);
      auto putW = [&](const std::string &label, const float *d) {
         std::vector<float> w(C.hostW.size());
         D.d2h(w.data(), d, w.size() * 4);
         writeWeights(dir + "/" + C.name + "_" + label + ".pvp", nkx, nky, layers[C.pre].loc.nf, C.nxp, C.nyp,
               C.nfp, {w}, {simTime});
      };
      putW("W", C.W);
      if (C.plastic) {
         writeBin(dir, C.name + "_lastUpdateTime", &C.lastUpdateTime, 1);
         if (C.triggerLayer.empty()) {
            writeBin(dir, C.name + "_weightUpdateTime", &C.weightUpdateTime, 1);
         }
         if (!C.immediate) {
            putW("dW", C.dW);
         }
         if (C.momentum) {
            putW("prev_dW", C.prevDw);
         }
      }
   }
   for (auto &p : probes) {
      if (p.keyword == "AdaptiveTimeScaleProbe") { // CheckpointEntryTimeScaleInfo::write
         std::vector<double> v;
         for (int b = 0; b < nbatch; b++) {
            v.push_back(p.ts[b]);
            v.push_back(p.tsTrue[b]);
            v.push_back(p.tsMax[b]);
         }
         writeBin(dir, p.name + "_timescaleinfo", v.data(), v.size());
      }
   }
   struct {
      double simTime;
      long step;
   } timeinfo{simTime, step};
   std::ofstream b(dir + "/timeinfo.bin", std::ios::binary);
   b.write(reinterpret_cast<const char *>(&timeinfo), sizeof timeinfo);
   std::ofstream t(dir + "/timeinfo.txt");
   t << "time = " << simTime << "\nnextCheckpointStep = " << step << "\n";
}

// initializeFromCheckpointDir: the state writeCheckpoint saved; returns its
// simulation time.  Derived weights and low-precision shadows are re-derived
// and every active list rebuilt from the restored activity.
double Net::readCheckpoint(const std::string &dir) {
   struct {
      double simTime;
      long step;
   } timeinfo{};
   {
      std::ifstream b(dir + "/timeinfo.bin", std::ios::binary);
      if (!b || !b.read(reinterpret_cast<char *>(&timeinfo), sizeof timeinfo)) {
         fatal("checkpoint " + dir + ": cannot read timeinfo.bin");
      }
   }
   for (auto &L : layers) {
      auto get = [&](const std::string &label) {
         PvpFrames f = readActivity(dir + "/" + L.name + "_" + label + ".pvp");
         if (f.nx != L.loc.nx || f.ny != L.loc.ny || f.nf != L.loc.nf) {
            fatal("checkpoint " + dir + ": " + L.name + "_" + label + ".pvp has another size");
         }
         return f.frames;
      };
      if (L.kind != L_INPUT) {
         setRestrictedFrames(L, L.V, false, get("V"), 0);
      }
      L.aStale = false; // A is read from the checkpoint
      setRestrictedFrames(L, L.A, true, get("A"), 0);
      syncBf16(L);
      auto levels = get("Delays");
      for (int lv = 1; lv < L.levels; lv++) {
         setRestrictedFrames(L, levelData(L, lv), true, levels, (size_t)lv * nbatch);
      }
      if (L.momentum) {
         setRestrictedFrames(L, L.prevDrive, false, get("prevDrive"), 0);
      }
      if (L.kind == L_INPUT) {
         L.frameIndex = readBin<int>(dir, L.name + "_FrameNumbers", nbatch)[0];
         L.lastUpdate = readBin<double>(dir, L.name + "_lastUpdateTime", 1)[0];
      }
   }
   for (auto &C : conns) {
      if (C.keyword != "HyPerConn") {
         continue;
      }
      int nK = (int)(C.hostW.size() / ((size_t)C.nxp * C.nyp * C.nfp));
      auto getW = [&](const std::string &label, float *d) {
         std::vector<float> w = readWeights(dir + "/" + C.name + "_" + label + ".pvp", C.nxp, C.nyp, C.nfp, nK);
         D.h2d(d, w.data(), w.size() * 4);
      };
      getW("W", C.W);
      if (C.plastic) {
         C.lastUpdateTime = readBin<double>(dir, C.name + "_lastUpdateTime", 1)[0];
         if (C.triggerLayer.empty()) {
            C.weightUpdateTime = readBin<double>(dir, C.name + "_weightUpdateTime", 1)[0];
         }
         if (!C.immediate) {
            getW("dW", C.dW);
         }
         if (C.momentum) {
            getW("prev_dW", C.prevDw);
         }
      }
      refreshDerived(C);
   }
   for (auto &p : probes) {
      if (p.keyword == "AdaptiveTimeScaleProbe") {
         auto v = readBin<double>(dir, p.name + "_timescaleinfo", (size_t)3 * nbatch);
         for (int b = 0; b < nbatch; b++) {
            p.ts[b]     = v[3 * b];
            p.tsTrue[b] = v[3 * b + 1];
            p.tsMax[b]  = v[3 * b + 2];
         }
      }
   }
   for (auto &L : layers) {
      if (L.needActive) {
         check(pvhip_active_compact(&L.loc, L.A, L.entries, L.numActive, L.posStart, L.compactScratch, D.stream),
               "pvhip_active_compact");
      }
   }
   for (auto &p : probes) { // the energies of the restored state, as the next AdaptTimestep reads them
      if (p.keyword == "ColumnEnergyProbe") {
         p.prev = probeEnergy(p);
      }
   }
   check(pvhip_stream_sync(D.stream), "checkpoint read");
   return timeinfo.simTime;
}

int Net::finish() {
   for (auto &L : layers) {
      if (!L.outFrames.empty()) {
         writeActivity(outputPath + "/" + L.name + ".pvp", L.loc.nx, L.loc.ny, L.loc.nf, L.outFrames, L.outTimes);
      }
   }
   for (auto &C : conns) {
      if (!C.outFrames.empty()) {
         int nkx, nky;
         pvhost_num_kernels(&layers[C.pre].loc, &layers[C.post].loc, &nkx, &nky);
         writeWeights(outputPath + "/" + C.name + ".pvp", nkx, nky, layers[C.pre].loc.nf, C.nxp, C.nyp, C.nfp,
               C.outFrames, C.outTimes);
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
      long n     = steps >= 0 ? steps : (long)std::lround(net.stopTime / net.dtCol);
      long first = 1;
      if (!net.initFromCheckpointDir.empty()) {
         first = std::lround(net.readCheckpoint(net.initFromCheckpointDir) / net.dtCol) + 1;
      }
      // HyPerCol::advanceTimeLoop: checkpointWrite before each step (checkpoint k
      // holds the state after k steps), then a final checkpoint
      check(pvhip_stream_sync(net.D.stream), "sync before the run");
      auto const t0 = std::chrono::steady_clock::now();
      for (long s = first; s <= n; s++) {
         bool restartStep = s == first && first > 1; // the checkpoint just read
         if (net.checkpointWrite && (s - 1) % net.checkpointInterval == 0 && !restartStep) {
            net.writeCheckpoint(net.checkpointDir(s - 1, n), (double)(s - 1) * net.dtCol, s - 1);
         }
         net.step((double)s * net.dtCol);
      }
      if (net.checkpointWrite) {
         check(pvhip_stream_sync(net.D.stream), "sync before the final checkpoint");
         net.writeCheckpoint(net.checkpointDir(n, n), (double)n * net.dtCol, n);
      }
      check(pvhip_stream_sync(net.D.stream), "final sync");
      double const secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      rc = net.finish();
      std::printf("pvrun: %ld steps done%s (%.3f s, %.4f ms per step with the host's own work)\n", n - first + 1,
            rc ? ", a probe failed" : "", secs, n >= first ? 1e3 * secs / (double)(n - first + 1) : 0.0);
   } catch (const std::exception &e) {
      std::fprintf(stderr, "pvrun: %s\n", e.what());
      rc = 1;
   }
   pvparams_free(P);
   return rc;
}
