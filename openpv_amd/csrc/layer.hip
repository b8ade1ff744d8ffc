// Layer-update kernels of libpvhip: the LCA V/A update (K6+K7+K8 fused), the
// ANN/ErrScale updates, IdentDelivery, GSyn accumulation and the ordered
// active-index compaction (DataStore::updateActiveIndices on the device).
//
// Every kernel here is elementwise or a scan and HBM-bound.  All float
// arithmetic follows the reference's operation order with one rounding per
// operation (the library is built with -ffp-contract=off).
#include "pvhip_internal.hpp"

#include <mutex>

namespace pvhip {

int make_lca_params(const pvhip_lca_update_args &a, LCAParams *p) {
   PVHIP_REQUIRE(loc_valid(a.loc), "lca_update: invalid layer loc");
   PVHIP_REQUIRE(
         a.num_channels >= 0 && a.num_channels <= PVHIP_MAX_CHANNELS,
         "lca_update: num_channels %d outside [0,%d]", a.num_channels, PVHIP_MAX_CHANNELS);
   PVHIP_REQUIRE(a.V && a.exp_tau, "lca_update: NULL V or exp_tau");
   PVHIP_REQUIRE(a.A || (a.A_bf16c && a.activity_from_v),
         "lca_update: NULL A (allowed only with A_bf16c and activity_from_v)");
   PVHIP_REQUIRE(a.num_channels == 0 || a.gsyn, "lca_update: NULL gsyn");
   p->loc         = geo(a.loc);
   p->numChannels = a.num_channels;
   for (int i = 0; i < PVHIP_MAX_CHANNELS; i++) {
      p->coeffs[i] = a.channel_coeffs[i];
   }
   p->gsyn         = a.gsyn;
   p->V            = a.V;
   p->A            = a.A;
   p->expTau       = a.exp_tau;
   p->selfInteract = a.self_interact ? 1.0f : 0.0f;
   p->aFromV       = a.activity_from_v ? 1 : 0;
   p->Ab           = static_cast<__bf16 *>(a.A_bf16c);
   PVHIP_REQUIRE(!a.A_bf16c || (a.loc.nf % 32 == 0 && reinterpret_cast<uintptr_t>(a.A_bf16c) % 16 == 0),
         "lca_update: the chunked BF16 activity needs nf %% 32 == 0 (nf %d) and 16-byte alignment",
         a.loc.nf);
   return make_transfer(a.transfer, &p->transfer);
}

namespace {

// G = sum_ch coeff[ch]*GSyn[ch] starting from 0.0f (GSynAccumulator.kpp:36-40).
__device__ __forceinline__ float accumulate_channels(
      int numChannels, const float *coeffs, const float *gsyn, size_t n, size_t index) {
   float value = 0.0f;
#pragma unroll
   for (int ch = 0; ch < PVHIP_MAX_CHANNELS; ch++) {
      if (ch < numChannels) {
         float prod = coeffs[ch] * gsyn[(size_t)ch * n + index];
         value      = value + prod;
      }
   }
   return value;
}

// Elementwise kernels: grid (ceil(nx*ny*nf / kRowSpan), nbatch); a workgroup
// covers kRowSpan consecutive restricted neurons of one batch element, each
// thread kRowIter of them kRowThreads apart (unit-stride accesses across a
// wave; one 32-bit division per neuron for its row, no 64-bit division).
constexpr int kRowThreads = 256;
constexpr int kRowIter    = 4;
constexpr int kRowSpan    = kRowThreads * kRowIter;

inline dim3 row_grid(const Geo &g) {
   return dim3((unsigned)((g.numRes() + kRowSpan - 1) / kRowSpan), (unsigned)g.nbatch);
}

inline bool row_grid_ok(const Geo &g) { return g.nbatch <= 65535; }

// Offset of row y of batch b in an extended buffer, at the first restricted column.
__device__ __forceinline__ size_t ext_row(const Geo &g, int b, int y) {
   return (size_t)b * g.numExt() + ((size_t)(y + g.up) * g.nxExt() + g.lt) * g.nf;
}

__global__ void __launch_bounds__(kRowThreads) lca_update_kernel(LCAParams p) {
   int const numRes = p.loc.numRes();
   int const b      = blockIdx.y;
   size_t const n   = (size_t)p.loc.nbatch * numRes;
#pragma unroll
   for (int u = 0; u < kRowIter; u++) {
      int const k = blockIdx.x * kRowSpan + u * kRowThreads + threadIdx.x;
      if (k < numRes) {
         float G = accumulate_channels(p.numChannels, p.coeffs, p.gsyn, n, (size_t)b * numRes + k);
         lca_update_one(p, b, k, G);
      }
   }
}

// MomentumLCA V update (MomentumLCAInternalStateBuffer.kpp:44-63) with the
// GSyn accumulation and the transfer fused: the HyPerLCA operations up to the
// drive, then V = (e*V + drive) + rate*prevDrive in the reference's order.
__global__ void __launch_bounds__(kRowThreads) momentum_lca_update_kernel(
      LCAParams p, float rate, float *prevDrive) {
   int const numRes = p.loc.numRes();
   int const b      = blockIdx.y;
   size_t const n   = (size_t)p.loc.nbatch * numRes;
   float const e    = p.expTau[b];
#pragma unroll
   for (int u = 0; u < kRowIter; u++) {
      int const k = blockIdx.x * kRowSpan + u * kRowThreads + threadIdx.x;
      if (k < numRes) {
         size_t const vi = (size_t)b * numRes + k;
         size_t const ai = (size_t)b * p.loc.numExt() + p.loc.kex(k);
         float const G    = accumulate_channels(p.numChannels, p.coeffs, p.gsyn, n, vi);
         float const Vold = p.V[vi];
         float const Aold = p.aFromV ? apply_transfer(p.transfer, Vold) : p.A[ai];
         float t2         = 1.0f - e;
         float si         = p.selfInteract * Aold;
         float t3         = G + si;
         float drive      = t2 * t3;
         float t1         = e * Vold;
         float t5         = t1 + drive;
         float mom        = rate * prevDrive[vi];
         float Vnew       = t5 + mom;
         float Anew       = apply_transfer(p.transfer, Vnew);
         prevDrive[vi]    = drive;
         p.V[vi]          = Vnew;
         if (p.A) {
            p.A[ai] = Anew;
         }
         if (p.Ab) {
            p.Ab[bf16c_index(p.loc, b, p.loc.kex(k))] = (__bf16)Anew;
         }
      }
   }
}

// ISTA V update (ISTAInternalStateBuffer.kpp:34-57, the reference's K9) with
// the GSyn accumulation and the transfer fused as in lca_update_kernel:
//   V += rate[b]*(G - VThresh*sign(A(t-1)[kex])),  rate[b] = (float)dt[b]/tau,
// in the reference's operation order; A(t-1) recomputed from V when aFromV.
__global__ void __launch_bounds__(kRowThreads) ista_update_kernel(LCAParams p, float vthresh) {
   int const numRes = p.loc.numRes();
   int const b      = blockIdx.y;
   size_t const n   = (size_t)p.loc.nbatch * numRes;
   float const rate = p.expTau[b];
#pragma unroll
   for (int u = 0; u < kRowIter; u++) {
      int const k = blockIdx.x * kRowSpan + u * kRowThreads + threadIdx.x;
      if (k < numRes) {
         size_t const vi = (size_t)b * numRes + k;
         size_t const ai = (size_t)b * p.loc.numExt() + p.loc.kex(k);
         float const G    = accumulate_channels(p.numChannels, p.coeffs, p.gsyn, n, vi);
         float const Vold = p.V[vi];
         float const Aold = p.aFromV ? apply_transfer(p.transfer, Vold) : p.A[ai];
         float sign       = 0.0f;
         if (Aold != 0.0f) {
            sign = Aold / fabsf(Aold);
         }
         float thr   = vthresh * sign;
         float drive = G - thr;
         float step  = rate * drive;
         float Vnew  = Vold + step;
         float Anew  = apply_transfer(p.transfer, Vnew);
         p.V[vi]     = Vnew;
         if (p.A) {
            p.A[ai] = Anew;
         }
         if (p.Ab) {
            p.Ab[bf16c_index(p.loc, b, p.loc.kex(k))] = (__bf16)Anew;
         }
      }
   }
}

struct LayerParams {
   Geo loc;
   int numChannels;
   float coeffs[PVHIP_MAX_CHANNELS];
   const float *gsyn;
   float *V;
   float *A;
   float scale;
   int applyScale;
   Transfer transfer;
};

__global__ void __launch_bounds__(kRowThreads) layer_update_kernel(LayerParams p) {
   int const numRes = p.loc.numRes();
   int const rowLen = p.loc.nx * p.loc.nf;
   int const b      = blockIdx.y;
   size_t const n   = (size_t)p.loc.nbatch * numRes;
#pragma unroll
   for (int u = 0; u < kRowIter; u++) {
      int const k = blockIdx.x * kRowSpan + u * kRowThreads + threadIdx.x;
      if (k < numRes) {
         size_t const index = (size_t)b * numRes + k;
         float V            = accumulate_channels(p.numChannels, p.coeffs, p.gsyn, n, index);
         if (p.applyScale) {
            V = V * p.scale; // ErrScaleInternalStateBuffer.cpp:51
         }
         p.V[index] = V;
         if (p.A) {
            int const y = k / rowLen;
            p.A[ext_row(p.loc, b, y) + (k - y * rowLen)] = apply_transfer(p.transfer, V);
         }
      }
   }
}

__global__ void __launch_bounds__(kRowThreads)
      ann_activity_kernel(Geo loc, Transfer t, const float *V, float *A) {
   int const numRes = loc.numRes();
   int const rowLen = loc.nx * loc.nf;
   int const b      = blockIdx.y;
#pragma unroll
   for (int u = 0; u < kRowIter; u++) {
      int const k = blockIdx.x * kRowSpan + u * kRowThreads + threadIdx.x;
      if (k < numRes) {
         int const y = k / rowLen;
         A[ext_row(loc, b, y) + (k - y * rowLen)] = apply_transfer(t, V[(size_t)b * numRes + k]);
      }
   }
}

__global__ void __launch_bounds__(256) gsyn_accumulate_kernel(
      int n, int numChannels, Transfer coeffHolder, const float *gsyn, float *out) {
   int const index = blockIdx.x * blockDim.x + threadIdx.x;
   if (index >= n) {
      return;
   }
   out[index] = accumulate_channels(numChannels, coeffHolder.vV, gsyn, (size_t)n, (size_t)index);
}

// IdentDelivery dense branch (IdentDelivery.cpp:146-162): GSyn += A restricted.
__global__ void __launch_bounds__(kRowThreads)
      ident_deliver_kernel(Geo loc, const float *pre, float *gsyn, int accumulate) {
   int const numRes = loc.numRes();
   int const rowLen = loc.nx * loc.nf;
   int const b      = blockIdx.y;
#pragma unroll
   for (int u = 0; u < kRowIter; u++) {
      int const k = blockIdx.x * kRowSpan + u * kRowThreads + threadIdx.x;
      if (k < numRes) {
         int const y        = k / rowLen;
         size_t const index = (size_t)b * numRes + k;
         float a            = pre[ext_row(loc, b, y) + (k - y * rowLen)];
         float base         = accumulate ? gsyn[index] : 0.0f;
         gsyn[index]        = base + a;
      }
   }
}

// An error layer fed by two identity connections (channel 0 from `in`,
// channel 1 from `rec`), the residual of an LCA column: the two
// IdentDelivery passes onto reset GSyn, GSynAccumulator, ErrScale and the
// transfer in one pass, with the same operations in the same order as the
// separate kernels (so the same bits, signed zeros included).
__global__ void __launch_bounds__(kRowThreads) error_update_kernel(
      Geo loc, Geo inLoc, const float *in, Geo recLoc, const float *rec, float c0, float c1,
      int applyScale, float scale, Transfer t, float *V, float *A) {
   int const numRes = loc.numRes();
   int const rowLen = loc.nx * loc.nf;
   int const b      = blockIdx.y;
   // every element's two reads first (clamped in bounds), so they are in
   // flight together rather than one round trip per element
   float in0[kRowIter], rec0[kRowIter];
#pragma unroll
   for (int u = 0; u < kRowIter; u++) {
      int const k = min(blockIdx.x * kRowSpan + u * kRowThreads + threadIdx.x, numRes - 1);
      int const y = k / rowLen;
      int const j = k - y * rowLen;
      in0[u]      = in[ext_row(inLoc, b, y) + j];
      rec0[u]     = rec[ext_row(recLoc, b, y) + j];
   }
#pragma unroll
   for (int u = 0; u < kRowIter; u++) {
      int const k = blockIdx.x * kRowSpan + u * kRowThreads + threadIdx.x;
      if (k < numRes) {
         int const y    = k / rowLen;
         int const j    = k - y * rowLen;
         float const g0 = 0.0f + in0[u]; // IdentDelivery, reset GSyn
         float const g1 = 0.0f + rec0[u];
         float value    = 0.0f;          // GSynAccumulator
         float prod     = c0 * g0;
         value          = value + prod;
         prod           = c1 * g1;
         value          = value + prod;
         if (applyScale) {
            value = value * scale;
         }
         V[(size_t)b * numRes + k]  = value;
         A[ext_row(loc, b, y) + j] = apply_transfer(t, value);
      }
   }
}

// The same update with 16-byte accesses: each thread takes four consecutive
// elements of one row (rows of a multiple of 4 values), reads them with one
// load per input, writes V with one store and A with AW-wide stores (A's rows
// start on multiples of AW values).  Per element the operations are those of
// error_update_kernel.
template <int AW>
__global__ void __launch_bounds__(kRowThreads) error_update_vec_kernel(
      Geo loc, Geo inLoc, const float *in, Geo recLoc, const float *rec, float c0, float c1,
      int applyScale, float scale, Transfer t, float *V, float *A) {
   int const numRes = loc.numRes();
   int const rowLen = loc.nx * loc.nf;
   int const b      = blockIdx.y;
   int const k      = (blockIdx.x * kRowThreads + threadIdx.x) * 4;
   if (k >= numRes) {
      return;
   }
   int const y = k / rowLen;
   int const j = k - y * rowLen;
   float4 const i4 = *reinterpret_cast<const float4 *>(in + ext_row(inLoc, b, y) + j);
   float4 const r4 = *reinterpret_cast<const float4 *>(rec + ext_row(recLoc, b, y) + j);
   float const iv[4] = {i4.x, i4.y, i4.z, i4.w};
   float const rv[4] = {r4.x, r4.y, r4.z, r4.w};
   float v[4], a[4];
#pragma unroll
   for (int u = 0; u < 4; u++) {
      float const g0 = 0.0f + iv[u]; // IdentDelivery, reset GSyn
      float const g1 = 0.0f + rv[u];
      float value    = 0.0f;         // GSynAccumulator
      float prod     = c0 * g0;
      value          = value + prod;
      prod           = c1 * g1;
      value          = value + prod;
      if (applyScale) {
         value = value * scale;
      }
      v[u] = value;
      a[u] = apply_transfer(t, value);
   }
   *reinterpret_cast<float4 *>(V + (size_t)b * numRes + k) = make_float4(v[0], v[1], v[2], v[3]);
   float *Ar = A + ext_row(loc, b, y) + j;
   if (AW == 4) {
      *reinterpret_cast<float4 *>(Ar) = make_float4(a[0], a[1], a[2], a[3]);
   }
   else {
      *reinterpret_cast<float2 *>(Ar)     = make_float2(a[0], a[1]);
      *reinterpret_cast<float2 *>(Ar + 2) = make_float2(a[2], a[3]);
   }
}

// ---------------------------------------------------------------- compaction
//
// Ordered stream compaction of the extended buffer of each batch element
// (DataStore.cpp:55-62).  Chunks of kChunk consecutive elements per
// workgroup; pass 1 counts, pass 2 scans the chunk counts of each batch
// element, pass 3 writes entries and the per-position segment starts.

constexpr int kCompactThreads = 256;
constexpr int kCompactIters   = 16;
constexpr int kCompactVec     = 4; // elements per thread per iteration
constexpr int kChunk          = kCompactThreads * kCompactIters * kCompactVec;

// Loads elements [e, e+VEC) of one batch element's extended buffer (zeros past
// the end); VEC == 4 requires numExt % 4 == 0 (16-byte aligned rows).
template <int VEC>
__device__ __forceinline__ void load_vec(const float *Ab, int e, int numExt, float (&a)[kCompactVec]) {
   if (VEC == 4) {
      float4 v = e < numExt ? *reinterpret_cast<const float4 *>(Ab + e) : make_float4(0.f, 0.f, 0.f, 0.f);
      a[0] = v.x;
      a[1] = v.y;
      a[2] = v.z;
      a[3] = v.w;
   }
   else {
#pragma unroll
      for (int j = 0; j < kCompactVec; j++) {
         a[j] = e + j < numExt ? Ab[e + j] : 0.0f;
      }
   }
}

template <int VEC>
__global__ void __launch_bounds__(kCompactThreads)
      compact_count_kernel(const float *A, int numExt, uint32_t *chunkCount, int chunksPerBatch) {
   int const b         = blockIdx.y;
   int const chunk     = blockIdx.x;
   float const *Ab     = A + (size_t)b * numExt;
   int const base      = chunk * kChunk;
   int count           = 0;
#pragma unroll 4
   for (int it = 0; it < kCompactIters; it++) {
      int e = base + (it * kCompactThreads + threadIdx.x) * kCompactVec;
      float a[kCompactVec];
      load_vec<VEC>(Ab, e, numExt, a);
#pragma unroll
      for (int j = 0; j < kCompactVec; j++) {
         count += a[j] != 0.0f ? 1 : 0;
      }
   }
   __shared__ int red[kCompactThreads / kWave];
   for (int off = kWave / 2; off > 0; off >>= 1) {
      count += __shfl_down(count, off, kWave);
   }
   int lane = threadIdx.x % kWave;
   int wave = threadIdx.x / kWave;
   if (lane == 0) {
      red[wave] = count;
   }
   __syncthreads();
   if (threadIdx.x == 0) {
      int total = 0;
      for (int w = 0; w < kCompactThreads / kWave; w++) {
         total += red[w];
      }
      chunkCount[(size_t)b * chunksPerBatch + chunk] = (uint32_t)total;
   }
}

// One workgroup per batch element: exclusive scan of its chunk counts, in place.
__global__ void __launch_bounds__(1024)
      compact_scan_kernel(uint32_t *chunkCount, int chunksPerBatch, int64_t *numActive) {
   int const b        = blockIdx.x;
   uint32_t *c        = chunkCount + (size_t)b * chunksPerBatch;
   __shared__ uint32_t partial[1024];
   int const per      = (chunksPerBatch + blockDim.x - 1) / blockDim.x;
   int const begin    = threadIdx.x * per;
   int const end      = min(begin + per, chunksPerBatch);
   uint32_t local     = 0;
   for (int i = begin; i < end; i++) {
      local += c[i];
   }
   partial[threadIdx.x] = local;
   __syncthreads();
   // Hillis-Steele inclusive scan over the per-thread sums.
   for (int off = 1; off < (int)blockDim.x; off <<= 1) {
      uint32_t v = threadIdx.x >= (unsigned)off ? partial[threadIdx.x - off] : 0u;
      __syncthreads();
      partial[threadIdx.x] += v;
      __syncthreads();
   }
   uint32_t run = threadIdx.x > 0 ? partial[threadIdx.x - 1] : 0u;
   for (int i = begin; i < end; i++) {
      uint32_t v = c[i];
      c[i]       = run;
      run += v;
   }
   if (threadIdx.x == blockDim.x - 1) {
      numActive[b] = (int64_t)partial[blockDim.x - 1];
   }
}

template <int VEC>
__global__ void __launch_bounds__(kCompactThreads) compact_write_kernel(
      const float *A, int numExt, int nf, const uint32_t *chunkOffset, int chunksPerBatch,
      pvhip_sparse_entry *entries, uint32_t *posStart, int numPos, const int64_t *numActive) {
   int const b      = blockIdx.y;
   int const chunk  = blockIdx.x;
   float const *Ab  = A + (size_t)b * numExt;
   pvhip_sparse_entry *Eb = entries + (size_t)b * numExt;
   uint32_t *Pb     = posStart ? posStart + (size_t)b * (numPos + 1) : nullptr;
   int const base   = chunk * kChunk;
   uint32_t run     = chunkOffset[(size_t)b * chunksPerBatch + chunk];
   int const lane   = threadIdx.x % kWave;
   int const wave   = threadIdx.x / kWave;
   __shared__ uint32_t waveCount[kCompactThreads / kWave];
   for (int it = 0; it < kCompactIters; it++) {
      int e = base + (it * kCompactThreads + threadIdx.x) * kCompactVec;
      float a[kCompactVec];
      load_vec<VEC>(Ab, e, numExt, a);
      uint32_t mine = 0;
#pragma unroll
      for (int j = 0; j < kCompactVec; j++) {
         mine += (e + j < numExt && a[j] != 0.0f) ? 1u : 0u;
      }
      // inclusive prefix of the per-thread counts across the wave
      uint32_t incl = mine;
#pragma unroll
      for (int off = 1; off < kWave; off <<= 1) {
         uint32_t v = __shfl_up(incl, off, kWave);
         if (lane >= off) {
            incl += v;
         }
      }
      if (lane == kWave - 1) {
         waveCount[wave] = incl;
      }
      __syncthreads();
      uint32_t waveBase = 0, total = 0;
#pragma unroll
      for (int w = 0; w < kCompactThreads / kWave; w++) {
         uint32_t c = waveCount[w];
         waveBase += w < wave ? c : 0u;
         total += c;
      }
      uint32_t pos = run + waveBase + incl - mine; // entries before element e
#pragma unroll
      for (int j = 0; j < kCompactVec; j++) {
         int ej = e + j;
         if (ej < numExt) {
            if (Pb && (ej % nf) == 0) {
               Pb[ej / nf] = pos;
            }
            if (a[j] != 0.0f) {
               Eb[pos].index = (uint32_t)ej;
               Eb[pos].value = a[j];
               pos++;
            }
         }
      }
      run += total;
      __syncthreads();
   }
   if (Pb && chunk == 0 && threadIdx.x == 0) {
      Pb[numPos] = (uint32_t)numActive[b];
   }
}

// ---------------------------------------------------------------- FP64 probes

__global__ void __launch_bounds__(1024) cost_sum_kernel(
      Geo loc, const float *buf, int extended, int kind, double p0, double p1, double *out) {
   int const b           = blockIdx.x;
   int const numNeurons  = loc.numRes();
   size_t const stride   = extended ? (size_t)loc.numExt() : (size_t)numNeurons;
   float const *B        = buf + (size_t)b * stride;
   double const tpw      = p0 + p1;
   double const maxCost  = 0.5 * tpw;
   double const quad     = 0.5 / tpw;
   double sum            = 0.0;
   for (int k = threadIdx.x; k < numNeurons; k += blockDim.x) {
      double v = (double)B[extended ? loc.kex(k) : k];
      double c;
      if (kind == 0) {
         c = v * v;
      }
      else if (kind == 1) {
         c = fabs(v);
      }
      else if (kind == 2) {
         double ax = fabs(v);
         c         = ax < tpw ? ax * (1.0 - quad * ax) : maxCost;
      }
      else {
         c = fabs(v) > p0 ? 1.0 : 0.0;
      }
      sum += c;
   }
   __shared__ double red[1024 / kWave];
   for (int off = kWave / 2; off > 0; off >>= 1) {
      sum += __shfl_down(sum, off, kWave);
   }
   if (threadIdx.x % kWave == 0) {
      red[threadIdx.x / kWave] = sum;
   }
   __syncthreads();
   if (threadIdx.x == 0) {
      double total = 0.0;
      for (int w = 0; w < (int)(blockDim.x / kWave); w++) {
         total += red[w];
      }
      out[b] = total;
   }
}

// NormalizeL2 in the reference's float arithmetic (NormalizeL2.cpp:93-125,
// NormalizeBase.cpp:184-193, NormalizeMultiply.cpp:251-254): per data patch
// sumsq += w*w in order, l2 = sqrtf(sumsq), w *= strength / l2.  A patch
// whose sumsq is within minL2 of zero ends the loop (the reference breaks out
// of its patch loop), so it and every later patch stay unchanged.  One
// workgroup; thread p owns patch p (strided).
__global__ void __launch_bounds__(1024) normalize_l2_exact_kernel(
      float *w, int numPatches, int patchSize, float strength, float minL2) {
   __shared__ int firstBad;
   if (threadIdx.x == 0) {
      firstBad = numPatches;
   }
   __syncthreads();
   for (int p = threadIdx.x; p < numPatches; p += blockDim.x) {
      float const *d = w + (size_t)p * patchSize;
      float sumsq    = 0.0f;
      for (int k = 0; k < patchSize; k++) {
         float const x  = d[k];
         float const sq = x * x;
         sumsq          = sumsq + sq;
      }
      if (fabsf(sumsq) <= minL2) {
         atomicMin(&firstBad, p);
      }
   }
   __syncthreads();
   int const stop = firstBad;
   for (int p = threadIdx.x; p < stop; p += blockDim.x) {
      float *d    = w + (size_t)p * patchSize;
      float sumsq = 0.0f;
      for (int k = 0; k < patchSize; k++) {
         float const x  = d[k];
         float const sq = x * x;
         sumsq          = sumsq + sq;
      }
      float const l2   = sqrtf(sumsq);
      float const mult = strength / l2;
      for (int k = 0; k < patchSize; k++) {
         d[k] = d[k] * mult;
      }
   }
}

__global__ void __launch_bounds__(256) normalize_l2_f64_kernel(
      float *w, int patchSize, double strength, double minL2) {
   float *data = w + (size_t)blockIdx.x * patchSize;
   double s    = 0.0;
   for (int k = threadIdx.x; k < patchSize; k += blockDim.x) {
      double x = (double)data[k];
      s += x * x;
   }
   __shared__ double red[256 / kWave];
   __shared__ double mult;
   for (int off = kWave / 2; off > 0; off >>= 1) {
      s += __shfl_down(s, off, kWave);
   }
   if (threadIdx.x % kWave == 0) {
      red[threadIdx.x / kWave] = s;
   }
   __syncthreads();
   if (threadIdx.x == 0) {
      double total = 0.0;
      for (int i = 0; i < (int)(blockDim.x / kWave); i++) {
         total += red[i];
      }
      mult = total <= minL2 ? 1.0 : strength / sqrt(total);
   }
   __syncthreads();
   double m = mult;
   for (int k = threadIdx.x; k < patchSize; k += blockDim.x) {
      data[k] = (float)((double)data[k] * m);
   }
}

__global__ void __launch_bounds__(256) hebbian_apply_kernel(
      float *w, float *dw, const int64_t *counts, size_t n, int normalizeDw) {
   size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (i >= n) {
      return;
   }
   float d = dw[i];
   if (normalizeDw) { // HebbianUpdater.cpp:794-800
      long divisor = (long)counts[i];
      if (divisor != 0) {
         d = d / (float)divisor;
      }
      else {
         d = 0.0f;
      }
      dw[i] = d;
   }
   w[i] = w[i] + d; // HebbianUpdater.cpp:828-830
}

// MomentumUpdater::updateWeights per weight (MomentumUpdater.cpp:276-294):
// normalize_dW as hebbian_apply_kernel, applyMomentum (:309-329, or the
// deprecated momentumTau form :346-360), prev_dW = dW, w += dW.
__global__ void __launch_bounds__(256) momentum_apply_kernel(
      float *w, float *dw, float *prev, const int64_t *counts, size_t n, int normalizeDw,
      float dwFactor, float wFactor, int deprecatedTau) {
   size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (i >= n) {
      return;
   }
   float d = dw[i];
   if (normalizeDw) {
      long divisor = (long)counts[i];
      d            = divisor != 0 ? d / (float)divisor : 0.0f;
   }
   float const wo = w[i];
   float const pd = prev[i];
   if (deprecatedTau) {
      float a = dwFactor * pd;
      float c = wFactor * wo;
      d       = d + (a - c);
   }
   else {
      d = d * (1.0f - dwFactor);
      d = d + dwFactor * pd;
      d = d - wFactor * wo;
   }
   dw[i]   = d;
   prev[i] = d;
   w[i]    = wo + d;
}

// dst[i] = src[index[i]] (0 where index[i] < 0): the shared-weight
// transposition of TransposeWeights::transposeShared as a table
// (pvhost_transpose_weights_index), so a weight refresh stays on the device.
__global__ void __launch_bounds__(256) permute_weights_kernel(
      const float *src, const int32_t *index, float *dst, size_t n) {
   size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (i >= n) {
      return;
   }
   int32_t const j = index[i];
   dst[i]          = j >= 0 ? src[j] : 0.0f;
}

inline unsigned blocks_for(size_t n, unsigned threads) { return (unsigned)((n + threads - 1) / threads); }

// ------------------------------------------- active list from epilogue segments
//
// A delivery epilogue that applies the update can also emit the active
// entries of each (restricted position, feature block of segW features)
// segment into a staging slot of segW entries, in ascending feature order,
// and the segment's count.  Segments in (position, feature block) order are
// in ascending extended index order when the halo holds no actives, so an
// exclusive scan of the counts (compact_scan_kernel) and a copy give the
// ordered list of DataStore::updateActiveIndices.

// Eight lanes per segment (a segment holds a few entries on average):
// segment s = 32 * blockIdx.x + threadIdx.x / 8.
constexpr int kSegLanes = 8;
__global__ void __launch_bounds__(256) segment_copy_kernel(
      const pvhip_sparse_entry *stg, const uint32_t *segOff, int segsPerBatch, int segW,
      const int64_t *numActive, pvhip_sparse_entry *entries, int numExt) {
   int const b = blockIdx.y;
   int const s = blockIdx.x * (256 / kSegLanes) + threadIdx.x / kSegLanes;
   if (s >= segsPerBatch) {
      return;
   }
   uint32_t const *Ob = segOff + (size_t)b * segsPerBatch;
   uint32_t const off = Ob[s];
   uint32_t const n   = (s + 1 < segsPerBatch ? Ob[s + 1] : (uint32_t)numActive[b]) - off;
   uint2 const *src   = reinterpret_cast<const uint2 *>(stg + ((size_t)b * segsPerBatch + s) * segW);
   uint2 *dst         = reinterpret_cast<uint2 *>(entries + (size_t)b * numExt + off);
   for (uint32_t j = threadIdx.x % kSegLanes; j < n; j += kSegLanes) {
      dst[j] = src[j];
   }
}

// pos_start[b][p] for every extended position p: the actives before the
// first restricted position at or after p in raster order.
__global__ void __launch_bounds__(256) segment_pos_kernel(
      Geo g, const uint32_t *segOff, int segsPerPos, const int64_t *numActive, uint32_t *posStart) {
   int const b      = blockIdx.y;
   int const numPos = g.nxExt() * g.nyExt();
   int const p      = blockIdx.x * blockDim.x + threadIdx.x;
   if (p > numPos) {
      return;
   }
   int const numPosR = g.nx * g.ny;
   int r             = numPosR;
   if (p < numPos) {
      int const xe = p % g.nxExt();
      int const ye = p / g.nxExt();
      if (ye < g.up) {
         r = 0;
      }
      else if (ye < g.up + g.ny) {
         int const y = ye - g.up;
         r           = xe < g.lt ? y * g.nx : (xe < g.lt + g.nx ? y * g.nx + xe - g.lt : (y + 1) * g.nx);
      }
   }
   uint32_t const v = r < numPosR ? segOff[((size_t)b * numPosR + r) * segsPerPos] : (uint32_t)numActive[b];
   posStart[(size_t)b * (numPos + 1) + p] = v;
}

// Chunked BF16 copy of an extended activity buffer: one thread per 8 output
// values (16 bytes), read as two float4 of the PV layout.
__global__ void __launch_bounds__(256) activity_to_bf16c_kernel(Geo g, const float *A, __bf16 *out) {
   size_t const plane = (size_t)g.nxExt() * g.nyExt();
   int const chunks   = g.nf / 32;
   size_t const n8    = (size_t)g.nbatch * chunks * plane * 4;
   size_t const i     = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (i >= n8) {
      return;
   }
   int const q        = (int)(i % 4);
   size_t const pos   = (i / 4) % plane;
   size_t const bc    = i / (4 * plane); // b * chunks + chunk
   int const chunk    = (int)(bc % chunks);
   size_t const b     = bc / chunks;
   float const *src   = A + b * (size_t)g.numExt() + pos * g.nf + chunk * 32 + q * 8;
   float4 const lo    = *reinterpret_cast<const float4 *>(src);
   float4 const hi    = *reinterpret_cast<const float4 *>(src + 4);
   typedef __bf16 bf16x8v __attribute__((ext_vector_type(8)));
   bf16x8v v{(__bf16)lo.x, (__bf16)lo.y, (__bf16)lo.z, (__bf16)lo.w,
             (__bf16)hi.x, (__bf16)hi.y, (__bf16)hi.z, (__bf16)hi.w};
   *reinterpret_cast<bf16x8v *>(out + i * 8) = v;
}

} // namespace

size_t compact_scratch_bytes(const Geo &g) {
   return (size_t)g.nbatch * (size_t)((g.numExt() + kChunk - 1) / kChunk) * sizeof(uint32_t);
}

int active_from_segments(
      const Geo &g, uint32_t *segCnt, const pvhip_sparse_entry *stg, int segsPerPos, int segW,
      pvhip_sparse_entry *entries, int64_t *numActive, uint32_t *posStart, hipStream_t s) {
   PVHIP_REQUIRE(segW >= 1, "active list: segment width %d", segW);
   int const segsPerBatch = g.nx * g.ny * segsPerPos;
   hipLaunchKernelGGL(compact_scan_kernel, dim3(g.nbatch), dim3(1024), 0, s, segCnt, segsPerBatch,
         numActive);
   PVHIP_LAUNCHED("compact_scan");
   int const segsPerBlock = 256 / kSegLanes;
   hipLaunchKernelGGL(segment_copy_kernel, dim3((segsPerBatch + segsPerBlock - 1) / segsPerBlock, g.nbatch),
         dim3(256), 0, s,
         stg, segCnt, segsPerBatch, segW, numActive, entries, g.numExt());
   PVHIP_LAUNCHED("segment_copy");
   if (posStart) {
      int const numPos = g.nxExt() * g.nyExt();
      hipLaunchKernelGGL(segment_pos_kernel, dim3((numPos + 1 + 255) / 256, g.nbatch), dim3(256), 0,
            s, g, segCnt, segsPerPos, numActive, posStart);
      PVHIP_LAUNCHED("segment_pos");
   }
   return PVHIP_OK;
}

int compact_launch(const pvhip_loc &l, const float *A, pvhip_sparse_entry *entries, int64_t *num_active,
      uint32_t *pos_start, uint32_t *chunk_scratch, pvhip_stream stream);

// Workspace of pvhip_active_compact_noalloc (the compaction's per-chunk
// counts) in the library's own device memory.  In host code `scratch` names a
// host-side stand-in, not a device address: the address of each GPU's copy
// comes from hipGetSymbolAddress with that GPU current, once per device, and
// only that address is passed to kernels.
constexpr size_t kNoAllocScratchBytes = 256 * 1024;
constexpr int kMaxDevices             = 64;

} // namespace pvhip

static __device__ __align__(256) uint32_t scratch[pvhip::kNoAllocScratchBytes / sizeof(uint32_t)];

namespace pvhip {
namespace {
std::mutex g_workspace_mu;
uint32_t *g_workspace[kMaxDevices] = {};
} // namespace

// The current device's copy of `scratch` (looked up on its first use).
int noalloc_workspace(uint32_t **out) {
   int dev = -1;
   PVHIP_TRY(hipGetDevice(&dev));
   PVHIP_REQUIRE(dev >= 0 && dev < kMaxDevices, "noalloc workspace: device %d beyond %d", dev, kMaxDevices);
   std::lock_guard<std::mutex> lock(g_workspace_mu);
   if (g_workspace[dev] == nullptr) {
      void *ptr = nullptr;
      PVHIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
      PVHIP_REQUIRE(ptr != nullptr, "noalloc workspace: no device address for device %d", dev);
      g_workspace[dev] = static_cast<uint32_t *>(ptr);
   }
   *out = g_workspace[dev];
   return PVHIP_OK;
}

} // namespace pvhip

using namespace pvhip;

extern "C" {

int pvhip_lca_update(const pvhip_lca_update_args *args, pvhip_stream stream) {
   PVHIP_REQUIRE(args != nullptr, "pvhip_lca_update: NULL args");
   LCAParams p;
   int st = make_lca_params(*args, &p);
   if (st) {
      return st;
   }
   PVHIP_REQUIRE(row_grid_ok(p.loc), "lca_update: more than 65535 rows or batch elements");
   hipLaunchKernelGGL(lca_update_kernel, row_grid(p.loc), dim3(kRowThreads), 0, (hipStream_t)stream, p);
   PVHIP_LAUNCHED("lca_update");
   return PVHIP_OK;
}

int pvhip_momentum_lca_update(
      const pvhip_lca_update_args *args, float momentum_rate, float *prev_drive,
      pvhip_stream stream) {
   PVHIP_REQUIRE(args != nullptr, "pvhip_momentum_lca_update: NULL args");
   PVHIP_REQUIRE(prev_drive != nullptr, "pvhip_momentum_lca_update: NULL prev_drive");
   LCAParams p;
   int st = make_lca_params(*args, &p);
   if (st) {
      return st;
   }
   PVHIP_REQUIRE(row_grid_ok(p.loc), "momentum_lca_update: more than 65535 batch elements");
   hipLaunchKernelGGL(momentum_lca_update_kernel, row_grid(p.loc), dim3(kRowThreads), 0,
         (hipStream_t)stream, p, momentum_rate, prev_drive);
   PVHIP_LAUNCHED("momentum_lca_update");
   return PVHIP_OK;
}

int pvhip_ista_update(const pvhip_lca_update_args *args, float vthresh, pvhip_stream stream) {
   PVHIP_REQUIRE(args != nullptr, "ista_update: NULL args");
   LCAParams p;
   int st = make_lca_params(*args, &p);
   if (st) {
      return st;
   }
   PVHIP_REQUIRE(row_grid_ok(p.loc), "ista_update: nbatch too large");
   hipLaunchKernelGGL(ista_update_kernel, row_grid(p.loc), dim3(kRowThreads), 0, (hipStream_t)stream, p,
         vthresh);
   PVHIP_LAUNCHED("ista_update");
   return PVHIP_OK;
}

void pvhip_ista_rate(const double *dt, int nbatch, float tau, float *rate_out) {
   for (int b = 0; b < nbatch; b++) {
      rate_out[b] = (float)dt[b] / tau; // ISTAInternalStateBuffer.kpp:56
   }
}

int pvhip_activity_to_bf16c(const pvhip_loc *loc, const float *A, void *out, pvhip_stream stream) {
   PVHIP_REQUIRE(loc != nullptr && loc_valid(*loc), "activity_to_bf16c: invalid loc");
   PVHIP_REQUIRE(loc->nf % 32 == 0, "activity_to_bf16c: nf %d is not a multiple of 32", loc->nf);
   PVHIP_REQUIRE(A != nullptr && out != nullptr, "activity_to_bf16c: NULL buffer");
   PVHIP_REQUIRE(((uintptr_t)A % 16) == 0 && ((uintptr_t)out % 16) == 0,
         "activity_to_bf16c: buffers must be 16-byte aligned");
   Geo const g    = geo(*loc);
   size_t const n = (size_t)g.nbatch * g.numExt() / 8;
   hipLaunchKernelGGL(activity_to_bf16c_kernel, dim3(blocks_for(n, 256)), dim3(256), 0,
         (hipStream_t)stream, g, A, static_cast<__bf16 *>(out));
   PVHIP_LAUNCHED("activity_to_bf16c");
   return PVHIP_OK;
}

int pvhip_gsyn_accumulate(
      int n, int num_channels, const float *coeffs_host, const float *gsyn, float *out,
      pvhip_stream stream) {
   PVHIP_REQUIRE(n >= 0, "gsyn_accumulate: negative n");
   PVHIP_REQUIRE(
         num_channels >= 0 && num_channels <= PVHIP_MAX_CHANNELS, "gsyn_accumulate: bad channel count");
   PVHIP_REQUIRE(out && (gsyn || num_channels == 0), "gsyn_accumulate: NULL buffer");
   if (n == 0) {
      return PVHIP_OK;
   }
   Transfer holder{};
   for (int i = 0; i < num_channels; i++) {
      holder.vV[i] = coeffs_host[i];
   }
   hipLaunchKernelGGL(gsyn_accumulate_kernel, dim3(blocks_for(n, 256)), dim3(256), 0,
         (hipStream_t)stream, n, num_channels, holder, gsyn, out);
   PVHIP_LAUNCHED("gsyn_accumulate");
   return PVHIP_OK;
}

int pvhip_layer_update(const pvhip_layer_update_args *a, pvhip_stream stream) {
   PVHIP_REQUIRE(a != nullptr, "pvhip_layer_update: NULL args");
   PVHIP_REQUIRE(loc_valid(a->loc), "layer_update: invalid loc");
   PVHIP_REQUIRE(a->num_channels >= 0 && a->num_channels <= PVHIP_MAX_CHANNELS, "layer_update: bad channel count");
   PVHIP_REQUIRE(a->V != nullptr, "layer_update: NULL V");
   PVHIP_REQUIRE(a->num_channels == 0 || a->gsyn, "layer_update: NULL gsyn");
   LayerParams p;
   p.loc         = geo(a->loc);
   p.numChannels = a->num_channels;
   for (int i = 0; i < PVHIP_MAX_CHANNELS; i++) {
      p.coeffs[i] = a->channel_coeffs[i];
   }
   p.gsyn       = a->gsyn;
   p.V          = a->V;
   p.A          = a->A;
   p.scale      = a->scale;
   p.applyScale = a->apply_scale;
   int st       = make_transfer(a->transfer, &p.transfer);
   if (st) {
      return st;
   }
   PVHIP_REQUIRE(row_grid_ok(p.loc), "layer_update: more than 65535 rows or batch elements");
   hipLaunchKernelGGL(layer_update_kernel, row_grid(p.loc), dim3(kRowThreads), 0, (hipStream_t)stream, p);
   PVHIP_LAUNCHED("layer_update");
   return PVHIP_OK;
}

int pvhip_ann_activity(
      const pvhip_loc *loc, const pvhip_transfer *transfer, const float *V, float *A,
      pvhip_stream stream) {
   PVHIP_REQUIRE(loc && transfer && V && A, "ann_activity: NULL argument");
   PVHIP_REQUIRE(loc_valid(*loc), "ann_activity: invalid loc");
   Transfer t;
   int st = make_transfer(*transfer, &t);
   if (st) {
      return st;
   }
   Geo g = geo(*loc);
   PVHIP_REQUIRE(row_grid_ok(g), "ann_activity: more than 65535 rows or batch elements");
   hipLaunchKernelGGL(ann_activity_kernel, row_grid(g), dim3(kRowThreads), 0, (hipStream_t)stream, g, t, V, A);
   PVHIP_LAUNCHED("ann_activity");
   return PVHIP_OK;
}

int pvhip_error_update(const pvhip_error_update_args *a, pvhip_stream stream) {
   PVHIP_REQUIRE(a && a->in && a->rec && a->V && a->A, "error_update: NULL argument");
   PVHIP_REQUIRE(loc_valid(a->loc) && loc_valid(a->in_loc) && loc_valid(a->rec_loc), "error_update: invalid loc");
   PVHIP_REQUIRE(a->in_loc.nbatch == a->loc.nbatch && a->in_loc.nx == a->loc.nx && a->in_loc.ny == a->loc.ny
                       && a->in_loc.nf == a->loc.nf && a->rec_loc.nbatch == a->loc.nbatch
                       && a->rec_loc.nx == a->loc.nx && a->rec_loc.ny == a->loc.ny && a->rec_loc.nf == a->loc.nf,
         "error_update: the identity connections need equal sizes");
   Transfer t;
   int st = make_transfer(a->transfer, &t);
   if (st) {
      return st;
   }
   Geo g = geo(a->loc);
   PVHIP_REQUIRE(row_grid_ok(g), "error_update: more than 65535 rows or batch elements");
   // 16-byte form when every row segment of four values is aligned for it:
   // rows of a multiple of 4 values, the inputs' and V's rows on 16-byte
   // boundaries, A's rows on multiples of 4 (else 2) values
   auto rows_aligned = [](const Geo &x, const void *p, int m) {
      return ((uintptr_t)p % (4 * m)) == 0 && (x.numExt() % m) == 0 && ((x.nxExt() * x.nf) % m) == 0
             && (((x.up * x.nxExt() + x.lt) * x.nf) % m) == 0;
   };
   Geo const gi = geo(a->in_loc), gr = geo(a->rec_loc);
   bool const vec = (g.nx * g.nf) % 4 == 0 && rows_aligned(gi, a->in, 4) && rows_aligned(gr, a->rec, 4)
                    && (uintptr_t)a->V % 16 == 0 && rows_aligned(g, a->A, 2);
   if (vec) {
      dim3 const grid((unsigned)((g.numRes() / 4 + kRowThreads - 1) / kRowThreads), (unsigned)g.nbatch);
      if (rows_aligned(g, a->A, 4)) {
         hipLaunchKernelGGL(error_update_vec_kernel<4>, grid, dim3(kRowThreads), 0, (hipStream_t)stream, g, gi,
               a->in, gr, a->rec, a->channel_coeffs[0], a->channel_coeffs[1], a->apply_scale, a->scale, t,
               a->V, a->A);
      }
      else {
         hipLaunchKernelGGL(error_update_vec_kernel<2>, grid, dim3(kRowThreads), 0, (hipStream_t)stream, g, gi,
               a->in, gr, a->rec, a->channel_coeffs[0], a->channel_coeffs[1], a->apply_scale, a->scale, t,
               a->V, a->A);
      }
      PVHIP_LAUNCHED("error_update");
      return PVHIP_OK;
   }
   hipLaunchKernelGGL(error_update_kernel, row_grid(g), dim3(kRowThreads), 0, (hipStream_t)stream, g, gi, a->in,
         gr, a->rec, a->channel_coeffs[0], a->channel_coeffs[1], a->apply_scale, a->scale, t, a->V, a->A);
   PVHIP_LAUNCHED("error_update");
   return PVHIP_OK;
}

int pvhip_ident_deliver(
      const pvhip_loc *loc, const float *pre_data, float *post_gsyn, int accumulate,
      pvhip_stream stream) {
   PVHIP_REQUIRE(loc && pre_data && post_gsyn, "ident_deliver: NULL argument");
   PVHIP_REQUIRE(loc_valid(*loc), "ident_deliver: invalid loc");
   Geo g = geo(*loc);
   PVHIP_REQUIRE(row_grid_ok(g), "ident_deliver: more than 65535 rows or batch elements");
   hipLaunchKernelGGL(ident_deliver_kernel, row_grid(g), dim3(kRowThreads), 0, (hipStream_t)stream, g,
         pre_data, post_gsyn, accumulate);
   PVHIP_LAUNCHED("ident_deliver");
   return PVHIP_OK;
}

int pvhip_active_compact(
      const pvhip_loc *loc, const float *A, pvhip_sparse_entry *entries, int64_t *num_active,
      uint32_t *pos_start, uint32_t *chunk_scratch, pvhip_stream stream) {
   PVHIP_REQUIRE(loc && A && entries && num_active, "active_compact: NULL argument");
   PVHIP_REQUIRE(loc_valid(*loc), "active_compact: invalid loc");
   return compact_launch(*loc, A, entries, num_active, pos_start, chunk_scratch, stream);
}

int pvhip_active_compact_noalloc(
      const pvhip_loc *loc, const float *A, pvhip_sparse_entry *entries, int64_t *num_active,
      uint32_t *pos_start, pvhip_stream stream) {
   PVHIP_REQUIRE(loc && A && entries && num_active, "active_compact_noalloc: NULL argument");
   PVHIP_REQUIRE(loc_valid(*loc), "active_compact_noalloc: invalid loc");
   size_t const need = compact_scratch_bytes(geo(*loc));
   PVHIP_REQUIRE(need <= kNoAllocScratchBytes,
         "active_compact_noalloc: %d batch elements of %d extended neurons need %zu bytes of chunk counts, "
         "beyond the library's %zu-byte workspace (use pvhip_active_compact with a caller workspace)",
         loc->nbatch, geo(*loc).numExt(), need, kNoAllocScratchBytes);
   uint32_t *ws = nullptr;
   int st       = noalloc_workspace(&ws);
   if (st) {
      return st;
   }
   // every chunk count is written by compact_count_kernel before the scan
   // reads it, so nothing left by an earlier call needs clearing
   return compact_launch(*loc, A, entries, num_active, pos_start, ws, stream);
}

} // extern "C"

namespace pvhip {

int compact_launch(const pvhip_loc &l, const float *A, pvhip_sparse_entry *entries, int64_t *num_active,
      uint32_t *pos_start, uint32_t *chunk_scratch, pvhip_stream stream) {
   const pvhip_loc *loc = &l;
   Geo g              = geo(*loc);
   int const numExt   = g.numExt();
   int const chunks   = (numExt + kChunk - 1) / kChunk;
   int const numPos   = g.nxExt() * g.nyExt();
   PVHIP_REQUIRE(g.nbatch <= 65535, "active_compact: nbatch %d too large", g.nbatch);
   PVHIP_REQUIRE(chunk_scratch != nullptr, "active_compact: NULL chunk_scratch");
   uint32_t *scratch = chunk_scratch;
   dim3 grid(chunks, g.nbatch);
   bool const vec4 = numExt % 4 == 0 && ((uintptr_t)A & 15) == 0;
   if (vec4) {
      hipLaunchKernelGGL(compact_count_kernel<4>, grid, dim3(kCompactThreads), 0, (hipStream_t)stream,
            A, numExt, scratch, chunks);
   }
   else {
      hipLaunchKernelGGL(compact_count_kernel<1>, grid, dim3(kCompactThreads), 0, (hipStream_t)stream,
            A, numExt, scratch, chunks);
   }
   PVHIP_LAUNCHED("compact_count");
   hipLaunchKernelGGL(compact_scan_kernel, dim3(g.nbatch), dim3(1024), 0, (hipStream_t)stream,
         scratch, chunks, num_active);
   PVHIP_LAUNCHED("compact_scan");
   if (vec4) {
      hipLaunchKernelGGL(compact_write_kernel<4>, grid, dim3(kCompactThreads), 0, (hipStream_t)stream,
            A, numExt, g.nf, scratch, chunks, entries, pos_start, numPos, num_active);
   }
   else {
      hipLaunchKernelGGL(compact_write_kernel<1>, grid, dim3(kCompactThreads), 0, (hipStream_t)stream,
            A, numExt, g.nf, scratch, chunks, entries, pos_start, numPos, num_active);
   }
   PVHIP_LAUNCHED("compact_write");
   return PVHIP_OK;
}

} // namespace pvhip

extern "C" {

int pvhip_cost_sum_f64(
      const pvhip_loc *loc, const float *buffer, int extended, int kind, double p0, double p1,
      double *out, pvhip_stream stream) {
   PVHIP_REQUIRE(loc && buffer && out, "cost_sum_f64: NULL argument");
   PVHIP_REQUIRE(loc_valid(*loc), "cost_sum_f64: invalid loc");
   PVHIP_REQUIRE(kind >= 0 && kind <= 3, "cost_sum_f64: unknown kind %d", kind);
   Geo g = geo(*loc);
   hipLaunchKernelGGL(cost_sum_kernel, dim3(g.nbatch), dim3(1024), 0, (hipStream_t)stream, g,
         buffer, extended, kind, p0, p1, out);
   PVHIP_LAUNCHED("cost_sum_f64");
   return PVHIP_OK;
}

int pvhip_normalize_l2_f64(
      float *weights, int num_patches, int patch_size, double strength, double min_l2,
      pvhip_stream stream) {
   PVHIP_REQUIRE(weights && num_patches > 0 && patch_size > 0, "normalize_l2_f64: bad argument");
   hipLaunchKernelGGL(normalize_l2_f64_kernel, dim3(num_patches), dim3(256), 0,
         (hipStream_t)stream, weights, patch_size, strength, min_l2);
   PVHIP_LAUNCHED("normalize_l2_f64");
   return PVHIP_OK;
}

int pvhip_normalize_l2(
      float *weights, int num_patches, int patch_size, float strength, float min_l2,
      pvhip_stream stream) {
   PVHIP_REQUIRE(weights && num_patches > 0 && patch_size > 0, "normalize_l2: bad argument");
   hipLaunchKernelGGL(normalize_l2_exact_kernel, dim3(1), dim3(1024), 0, (hipStream_t)stream,
         weights, num_patches, patch_size, strength, min_l2);
   PVHIP_LAUNCHED("normalize_l2");
   return PVHIP_OK;
}

int pvhip_hebbian_apply(
      float *weights, float *dw, const int64_t *counts, size_t n, int normalize_dw,
      pvhip_stream stream) {
   PVHIP_REQUIRE(weights && dw, "hebbian_apply: NULL buffer");
   PVHIP_REQUIRE(!normalize_dw || counts, "hebbian_apply: normalize_dw needs counts");
   if (n == 0) {
      return PVHIP_OK;
   }
   hipLaunchKernelGGL(hebbian_apply_kernel, dim3(blocks_for(n, 256)), dim3(256), 0,
         (hipStream_t)stream, weights, dw, counts, n, normalize_dw);
   PVHIP_LAUNCHED("hebbian_apply");
   return PVHIP_OK;
}

int pvhip_momentum_factors(
      int method, float tau, float decay, float dw_max, float *dw_factor, float *w_factor) {
   PVHIP_REQUIRE(dw_factor && w_factor, "momentum_factors: NULL output");
   PVHIP_REQUIRE(decay >= 0.0f && decay <= 1.0f,
         "momentum_factors: momentumDecay must be between 0 and 1 inclusive (%g)", (double)decay);
   switch (method) {
      case PVHIP_MOMENTUM_VISCOSITY:
         PVHIP_REQUIRE(tau >= 0.0f, "momentum_factors: viscosity needs timeConstantTau >= 0 (%g)",
               (double)tau);
         *dw_factor = tau != 0.0f ? std::exp(-1.0f / tau) : 0.0f;
         *w_factor  = decay;
         return PVHIP_OK;
      case PVHIP_MOMENTUM_SIMPLE:
      case PVHIP_MOMENTUM_ALEX:
         PVHIP_REQUIRE(tau >= 0.0f && tau < 1.0f,
               "momentum_factors: simple and alex need 0 <= timeConstantTau < 1 (%g)", (double)tau);
         *dw_factor = tau;
         *w_factor  = method == PVHIP_MOMENTUM_ALEX ? decay * dw_max : decay;
         return PVHIP_OK;
      default: PVHIP_REQUIRE(false, "momentum_factors: unknown momentumMethod %d", method);
   }
}

int pvhip_momentum_factors_deprecated(
      int method, float momentum_tau, float decay, float dw_max, float *dw_factor, float *w_factor) {
   PVHIP_REQUIRE(dw_factor && w_factor, "momentum_factors: NULL output");
   PVHIP_REQUIRE(decay >= 0.0f && decay <= 1.0f,
         "momentum_factors: momentumDecay must be between 0 and 1 inclusive (%g)", (double)decay);
   switch (method) { // applyMomentumDeprecated (MomentumUpdater.cpp:333-343): no range check
      case PVHIP_MOMENTUM_VISCOSITY:
         *dw_factor = std::exp(-1.0f / momentum_tau);
         *w_factor  = decay;
         return PVHIP_OK;
      case PVHIP_MOMENTUM_SIMPLE:
      case PVHIP_MOMENTUM_ALEX:
         *dw_factor = momentum_tau;
         *w_factor  = method == PVHIP_MOMENTUM_ALEX ? decay * dw_max : decay;
         return PVHIP_OK;
      default: PVHIP_REQUIRE(false, "momentum_factors: unknown momentumMethod %d", method);
   }
}

int pvhip_permute_weights(
      const float *src, const int32_t *index, float *dst, size_t n, pvhip_stream stream) {
   PVHIP_REQUIRE(src && index && dst, "permute_weights: NULL buffer");
   PVHIP_REQUIRE(src != dst, "permute_weights: src and dst must differ");
   if (n == 0) {
      return PVHIP_OK;
   }
   hipLaunchKernelGGL(permute_weights_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, (hipStream_t)stream, src,
         index, dst, n);
   PVHIP_LAUNCHED("permute_weights");
   return PVHIP_OK;
}

int pvhip_momentum_apply(
      float *weights, float *dw, float *prev_dw, const int64_t *counts, size_t n,
      int normalize_dw, float dw_factor, float w_factor, int deprecated_tau,
      pvhip_stream stream) {
   PVHIP_REQUIRE(weights && dw && prev_dw, "momentum_apply: NULL buffer");
   PVHIP_REQUIRE(!normalize_dw || counts, "momentum_apply: normalize_dw needs counts");
   if (n == 0) {
      return PVHIP_OK;
   }
   hipLaunchKernelGGL(momentum_apply_kernel, dim3(blocks_for(n, 256)), dim3(256), 0,
         (hipStream_t)stream, weights, dw, prev_dw, counts, n, normalize_dw, dw_factor, w_factor,
         deprecated_tau);
   PVHIP_LAUNCHED("momentum_apply");
   return PVHIP_OK;
}

size_t pvhip_active_compact_scratch_bytes(const pvhip_loc *loc) {
   if (!loc || !loc_valid(*loc)) {
      return 0;
   }
   return compact_scratch_bytes(geo(*loc));
}

} // extern "C"
