// Internal helpers shared by the libpvhip translation units (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "pvhip.h"

namespace pvhip {

// Records the message for pvhip_last_error() and returns `code`.
int set_error(int code, const char *fmt, ...);
void clear_error();
// Reads and clears the tiles the gather kernels refused since the last call
// (bit 1: a source window outside its bounds, bit 2: off the alignment class).
int gather_faults(unsigned *mask);
// Turns refused tiles into an error (after the caller has synchronised).
int check_faults();

#define PVHIP_TRY(call)                                                                     \
   do {                                                                                    \
      hipError_t e_ = (call);                                                              \
      if (e_ != hipSuccess) {                                                              \
         return ::pvhip::set_error(                                                        \
               PVHIP_ERR_RUNTIME, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),   \
               __FILE__, __LINE__);                                                        \
      }                                                                                    \
   } while (0)

#define PVHIP_LAUNCHED(name)                                                                \
   do {                                                                                    \
      hipError_t e_ = hipGetLastError();                                                   \
      if (e_ != hipSuccess) {                                                              \
         return ::pvhip::set_error(                                                        \
               PVHIP_ERR_RUNTIME, "%s launch failed: %s", name, hipGetErrorString(e_));    \
      }                                                                                    \
   } while (0)

#define PVHIP_REQUIRE(cond, ...)                                                            \
   do {                                                                                    \
      if (!(cond)) {                                                                       \
         return ::pvhip::set_error(PVHIP_ERR_INVALID, __VA_ARGS__);                        \
      }                                                                                    \
   } while (0)

// Geometry of a layer buffer, usable on host and device.
struct Geo {
   int nbatch, nx, ny, nf, lt, rt, dn, up;
   __host__ __device__ int nxExt() const { return nx + lt + rt; }
   __host__ __device__ int nyExt() const { return ny + dn + up; }
   __host__ __device__ int numRes() const { return nx * ny * nf; }
   __host__ __device__ int numExt() const { return nxExt() * nyExt() * nf; }
   // kIndexExtended (src/utils/conversions.hpp:401-407): x offset lt, y offset up.
   __host__ __device__ int kex(int k) const {
      int kf = k % nf;
      int kx = (k / nf) % nx;
      int ky = k / (nx * nf);
      return kf + ((kx + lt) + (ky + up) * nxExt()) * nf;
   }
};

inline Geo geo(const pvhip_loc &l) {
   return Geo{l.nbatch, l.nx, l.ny, l.nf, l.halo.lt, l.halo.rt, l.halo.dn, l.halo.up};
}

inline bool loc_valid(const pvhip_loc &l) {
   return l.nbatch > 0 && l.nx > 0 && l.ny > 0 && l.nf > 0 && l.halo.lt >= 0 && l.halo.rt >= 0
          && l.halo.dn >= 0 && l.halo.up >= 0
          && (double)l.nbatch * (l.nx + l.halo.lt + l.halo.rt) * (l.ny + l.halo.dn + l.halo.up)
                         * l.nf
                   < 2147483647.0;
}

// Transfer function parameters passed by value to kernels.
struct Transfer {
   int mode;
   float VThresh, AMin, AMax, AShift, VWidth;
   int numVertices;
   float vV[PVHIP_MAX_VERTICES];
   float vA[PVHIP_MAX_VERTICES];
   float slopes[PVHIP_MAX_VERTICES + 1];
};

inline int make_transfer(const pvhip_transfer &t, Transfer *out) {
   out->mode    = t.mode;
   out->VThresh = t.VThresh;
   out->AMin    = t.AMin;
   out->AMax    = t.AMax;
   out->AShift  = t.AShift;
   out->VWidth  = t.VWidth;
   out->numVertices = t.num_vertices;
   if (t.mode == PVHIP_TRANSFER_VERTICES) {
      PVHIP_REQUIRE(
            t.num_vertices >= 1 && t.num_vertices <= PVHIP_MAX_VERTICES,
            "transfer: num_vertices %d outside [1, %d]", t.num_vertices, PVHIP_MAX_VERTICES);
   }
   else {
      PVHIP_REQUIRE(t.mode == PVHIP_TRANSFER_VTHRESH, "transfer: unknown mode %d", t.mode);
   }
   for (int i = 0; i < PVHIP_MAX_VERTICES; i++) {
      out->vV[i] = t.vertices_v[i];
      out->vA[i] = t.vertices_a[i];
   }
   for (int i = 0; i <= PVHIP_MAX_VERTICES; i++) {
      out->slopes[i] = t.slopes[i];
   }
   return PVHIP_OK;
}

// A = transfer(V), VTHRESH form: HyPerActivityBuffer copy, applyVThresh,
// applyAMax (src/components/ANNActivityBuffer.cpp:490-556).
__device__ __forceinline__ float apply_vthresh(const Transfer &t, float V) {
   float a = V;
   if (t.VThresh > -FLT_MAX) {
      if (V < t.VThresh) {
         a = t.AMin;
      }
      else if (V < t.VThresh + t.VWidth) {
         float span = t.VThresh + t.VWidth;
         span       = span - t.AShift;
         span       = span - t.AMin;
         float num  = span * (V - t.VThresh);
         a          = t.AMin + num / t.VWidth;
      }
      else {
         a = a - t.AShift;
      }
   }
   if (t.AMax < FLT_MAX) {
      if (a > t.AMax) {
         a = t.AMax;
      }
   }
   return a;
}

// A = transfer(V) with the reference's exact float operation order.
//  VTHRESH: apply_vthresh.
//  VERTICES: applyVerticesANNActivityBufferOnGPU (src/components/ANNActivityBuffer.kpp:45-76).
__device__ __forceinline__ float apply_transfer(const Transfer &t, float V) {
   if (t.mode == PVHIP_TRANSFER_VTHRESH) {
      return apply_vthresh(t, V);
   }
   // Vertex arrays are read with compile-time indices only (unrolled over
   // the maximum), so the descriptor stays in registers / kernel arguments.
   int const last = t.numVertices - 1;
   float vLast = t.vV[0], aLast = t.vA[0], sLast = t.slopes[1];
#pragma unroll
   for (int i = 0; i < PVHIP_MAX_VERTICES; i++) {
      if (i == last) {
         vLast = t.vV[i];
         aLast = t.vA[i];
         sLast = t.slopes[i + 1];
      }
   }
   float activity = 0.0f;
   if (V < t.vV[0]) {
      activity = t.vA[0] + t.slopes[0] * (V - t.vV[0]);
   }
   else if (V >= vLast) {
      activity = aLast + sLast * (V - vLast);
   }
   else {
      bool stop = false;
#pragma unroll
      for (int v = 0; v < PVHIP_MAX_VERTICES - 1; v++) {
         if (!stop && v < last) {
            if (V < t.vV[v]) {
               stop = true; // jumps are continuous from the right
            }
            else if (V == t.vV[v]) {
               activity = t.vA[v];
            }
            else if (V > t.vV[v] && V < t.vV[v + 1]) {
               activity = t.vA[v] + t.slopes[v + 1] * (V - t.vV[v]);
            }
         }
      }
   }
   return activity;
}

// Parameters of the LCA update passed by value (see pvhip_lca_update_args).
struct LCAParams {
   Geo loc;
   int numChannels;
   float coeffs[PVHIP_MAX_CHANNELS];
   const float *gsyn;
   float *V;
   float *A;
   const float *expTau;
   float selfInteract;
   Transfer transfer;
   int aFromV; // A(t-1) = transfer(V(t-1)) is recomputed instead of read
   __bf16 *Ab; // optional chunked BF16 copy of A (pvhip_activity_to_bf16c layout);
               // A may then be NULL (with aFromV): only the copy is written
};

// Element of extended neuron kex (batch b) in the chunked BF16 layout:
// [nbatch][nf/32][nyExt][nxExt][32].
__device__ __forceinline__ size_t bf16c_index(const Geo &g, int b, int kex) {
   int const f   = kex % g.nf;
   int const pos = kex / g.nf;
   size_t const plane = (size_t)g.nxExt() * g.nyExt();
   return (((size_t)b * (g.nf / 32) + f / 32) * plane + pos) * 32 + f % 32;
}

int make_lca_params(const pvhip_lca_update_args &a, LCAParams *p);

// Transfer selected at compile time: MODE 1 is the VTHRESH form only,
// MODE 0 dispatches on t.mode.
template <int MODE>
__device__ __forceinline__ float transfer_as(const Transfer &t, float V) {
   return MODE == 1 ? apply_vthresh(t, V) : apply_transfer(t, V);
}

// One neuron of the fused K6+K7 update given its accumulated input G
// (HyPerLCAInternalStateBuffer.kpp:51-59, then the transfer).
template <int MODE = 0>
__device__ __forceinline__ void lca_update_one(const LCAParams &p, int b, int k, float G) {
   float const e   = p.expTau[b];
   size_t const vi = (size_t)b * p.loc.numRes() + k;
   size_t const ai = (size_t)b * p.loc.numExt() + p.loc.kex(k);
   float const Vold = p.V[vi];
   float const Aold = p.aFromV ? transfer_as<MODE>(p.transfer, Vold) : p.A[ai];
   float t1    = e * Vold;
   float t2    = 1.0f - e;
   float si    = p.selfInteract * Aold;
   float t3    = G + si;
   float t4    = t2 * t3;
   float Vnew  = t1 + t4;
   float Anew  = transfer_as<MODE>(p.transfer, Vnew);
   p.V[vi]     = Vnew;
   if (p.A) {
      p.A[ai] = Anew;
   }
   if (p.Ab) {
      p.Ab[bf16c_index(p.loc, b, p.loc.kex(k))] = (__bf16)Anew;
   }
}

constexpr int kWave = 64;

// Scratch of the ordered compaction (chunk counts) for a layer.
size_t compact_scratch_bytes(const Geo &g);

// Ordered active list of layer g from per-segment staging: segCnt holds the
// entry count of each (restricted position, feature block) segment in raster
// order (segsPerPos blocks of segW features per position) and is scanned in
// place; stg holds each segment's entries in slots of segW.  Requires an
// activity halo of zeros.
int active_from_segments(
      const Geo &g, uint32_t *segCnt, const pvhip_sparse_entry *stg, int segsPerPos, int segW,
      pvhip_sparse_entry *entries, int64_t *numActive, uint32_t *posStart, hipStream_t s);

} // namespace pvhip
