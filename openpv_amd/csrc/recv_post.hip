// Post-perspective gather (PostsynapticPerspectiveConvolveDelivery /
// CudaRecvPost replacement).  Three kernels:
//   direct   -- one thread per target neuron, any geometry, EXACT order;
//   tiled    -- LDS-staged source tile and per-row weight staging, shared
//               weights with one data patch per feature (the LCA geometries),
//               EXACT (mul, add rounded separately, reference order) or FP32
//               (FMA) arithmetic;
//   mfma     -- implicit GEMM on v_mfma_f32_16x16x32_bf16: weights as the
//               A operand (features x patch), the im2col of an LDS source
//               tile as the B operand (patch x positions), FP32 accumulate.
// Each kernel can apply the HyPerLCA update in its epilogue (fused_update).
//
// Reference order (PostsynapticPerspectiveConvolveDelivery.cpp:107-146):
//   for ky: dv = 0; for k < nxp*nfp: dv += a[k]*w[k];  GSyn += dtFactor*dv
#include "pvhip_internal.hpp"

#include <algorithm>
#include <type_traits>

namespace pvhip {
namespace {

struct PostParams {
   Geo pre, post;
   int nxp, nyp, nfp;
   int nKx, nKy;
   const int32_t *startX, *startY, *kerX, *kerY;
   float dtf;
   const float *pre_data;
   const float *W;
   float *gsyn;
   int accumulate;
   int fused;
   LCAParams lca;
   // Active-list emission of the fused epilogue (full tiles only): entries of
   // each (restricted position, MF-feature block) segment into a slot of MF
   // entries of stg, and the segment's count into segCnt.
   int emit;
   pvhip_sparse_entry *stg;
   uint32_t *segCnt;
   // FP8 form: per-kernel weight scales (after the FP8 shadow) and the source scale
   const float *lpWScale;
   float lpSrcScale;
   // BF16 channel GEMM: pre_data in the chunked BF16 layout, or NULL
   const __bf16 *preB16;
   // Channel GEMM with the target layer's update fused (pvhip_layer_update
   // with this delivery as channel 0): V and A written instead of GSyn
   int layer;
   Geo layerLoc; // the target layer (restricted size of post; its own halo)
   float layerCoeff, layerScale;
   int layerApplyScale;
   float *layerV, *layerA;
   Transfer layerT;
};

// Set by a gather kernel that refuses a tile whose source window breaks the
// bounds the host sized the launch for (geometry tables that do not belong to
// the locs of the call, or window starts off the alignment class the host
// assumed); the tile is then left unwritten.  Read and cleared by
// gather_faults(), which pvhip_stream_sync / pvhip_device_sync /
// pvhip_check_faults turn into an error.
enum : unsigned { FAULT_WINDOW = 1u, FAULT_ALIGN = 2u };
__device__ unsigned int g_gather_fault;

__device__ __forceinline__ void report_fault(unsigned bit) { atomicOr(&g_gather_fault, bit); }

// Target-layer update of one delivered value G (restricted index gi, extended
// index ai), the operations of layer_update_kernel with one channel.
__device__ __forceinline__ void fused_layer_store(const PostParams &p, size_t gi, size_t ai, float G) {
   float V    = 0.0f;
   float prod = p.layerCoeff * G;
   V          = V + prod;
   if (p.layerApplyScale) {
      V = V * p.layerScale;
   }
   p.layerV[gi] = V;
   if (p.layerA) {
      p.layerA[ai] = apply_transfer(p.layerT, V);
   }
}

__device__ __forceinline__ void post_epilogue(const PostParams &p, int b, int k, float G) {
   if (p.fused) {
      lca_update_one(p.lca, b, k, G);
   }
   else {
      p.gsyn[(size_t)b * p.post.numRes() + k] = G;
   }
}

// ------------------------------------------------------------------ direct
__global__ void __launch_bounds__(256) recv_post_direct_kernel(PostParams p) {
   size_t const n     = (size_t)p.post.nbatch * p.post.numRes();
   size_t const index = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (index >= n) {
      return;
   }
   int const numRes = p.post.numRes();
   int const b      = (int)(index / numRes);
   int const k      = (int)(index % numRes);
   int const f      = k % p.post.nf;
   int const x      = (k / p.post.nf) % p.post.nx;
   int const y      = k / (p.post.nf * p.post.nx);
   int const ker    = f + (p.kerX[x] + p.kerY[y] * p.nKx) * p.post.nf;
   int const K      = p.nxp * p.nfp;
   int const sy     = p.pre.nxExt() * p.pre.nf;
   float const *w   = p.W + (size_t)ker * p.nyp * K;
   float const *a   = p.pre_data + (size_t)b * p.pre.numExt()
                    + ((size_t)p.startY[y] * p.pre.nxExt() + p.startX[x]) * p.pre.nf;
   float g = p.accumulate ? p.gsyn[index] : 0.0f;
   for (int ky = 0; ky < p.nyp; ky++) {
      float dv = 0.0f;
      for (int kk = 0; kk < K; kk++) {
         float prod = a[ky * sy + kk] * w[ky * K + kk];
         dv         = dv + prod;
      }
      float scaled = p.dtf * dv;
      g            = g + scaled;
   }
   post_epilogue(p, b, k, g);
}

// ------------------------------------------------------------------ tiled
// Workgroup: TX x TY target positions x FB features of one batch element;
// 256 threads, each TM features x TN positions.  Source tile rows/cols are
// bounded by the host (dynamic LDS).
constexpr int TX = 8, TY = 8, TN = 4;

template <int FB, bool EXACT>
__global__ void __launch_bounds__(256) recv_post_tiled_kernel(PostParams p, int srcRowsMax, int srcColsMax) {
   extern __shared__ float smem[];
   constexpr int TM      = FB / 16;                // features per thread
   constexpr int FG      = FB / TM;                // feature groups
   static_assert(FG * (TX * TY / TN) == 256, "tile shape must use 256 threads");
   int const K           = p.nxp * p.nfp;
   int const pitch       = srcColsMax * p.nfp;      // LDS row pitch (floats)
   float *src            = smem;
   float *wl             = smem + srcRowsMax * pitch; // [K][FB]
   int const tilesX      = (p.post.nx + TX - 1) / TX;
   int const tile        = blockIdx.x;
   int const x0          = (tile % tilesX) * TX;
   int const y0          = (tile / tilesX) * TY;
   int const f0          = blockIdx.y * FB;
   int const b           = blockIdx.z;
   int const xl          = min(x0 + TX, p.post.nx) - 1;
   int const yl          = min(y0 + TY, p.post.ny) - 1;
   int const srcX0       = p.startX[x0];
   int const srcY0       = p.startY[y0];
   int const srcCols     = p.startX[xl] - srcX0 + p.nxp;
   int const srcRows     = p.startY[yl] - srcY0 + p.nyp;
   int const tid         = threadIdx.x;
   // Tables outside the bound the host sized LDS for, or patches reaching
   // outside the extended source buffer: the tile is refused and reported.
   if (srcCols > srcColsMax || srcRows > srcRowsMax || srcX0 < 0 || srcY0 < 0
       || srcX0 + srcCols > p.pre.nxExt() || srcY0 + srcRows > p.pre.nyExt()) {
      if (tid == 0) {
         report_fault(FAULT_WINDOW);
      }
      return;
   }
   int const preRow      = p.pre.nxExt() * p.pre.nf;
   float const *preB     = p.pre_data + (size_t)b * p.pre.numExt();

   // Source tile -> LDS.
   int const rowElems = srcCols * p.nfp;
   for (int i = tid; i < srcRows * rowElems; i += 256) {
      int r = i / rowElems;
      int c = i % rowElems;
      src[r * pitch + c] = preB[(size_t)(srcY0 + r) * preRow + (size_t)srcX0 * p.nfp + c];
   }

   int const fg = tid % FG;
   int const pg = tid / FG;
   int aoff[TN];
   bool valid[TN];
   int kOut[TN];
#pragma unroll
   for (int j = 0; j < TN; j++) {
      int pj   = pg * TN + j;
      int x    = x0 + pj % TX;
      int y    = y0 + pj / TX;
      valid[j] = x < p.post.nx && y < p.post.ny;
      int xc   = valid[j] ? x : x0;
      int yc   = valid[j] ? y : y0;
      aoff[j]  = (p.startY[yc] - srcY0) * pitch + (p.startX[xc] - srcX0) * p.nfp;
      kOut[j]  = (yc * p.post.nx + xc) * p.post.nf;
   }
   float acc[TM][TN];
#pragma unroll
   for (int i = 0; i < TM; i++) {
#pragma unroll
      for (int j = 0; j < TN; j++) {
         int f     = f0 + fg * TM + i;
         acc[i][j] = (p.accumulate && valid[j] && f < p.post.nf)
                           ? p.gsyn[(size_t)b * p.post.numRes() + kOut[j] + f]
                           : 0.0f;
      }
   }
   int const patchSize = p.nyp * K;
   for (int ky = 0; ky < p.nyp; ky++) {
      __syncthreads();
      for (int i = tid; i < FB * K; i += 256) {
         int fi = i / K;
         int kk = i % K;
         int f  = f0 + fi;
         wl[kk * FB + fi] = f < p.post.nf ? p.W[(size_t)f * patchSize + ky * K + kk] : 0.0f;
      }
      __syncthreads();
      float dv[TM][TN];
#pragma unroll
      for (int i = 0; i < TM; i++) {
#pragma unroll
         for (int j = 0; j < TN; j++) {
            dv[i][j] = 0.0f;
         }
      }
      float const *srow = src + ky * pitch;
      for (int kk = 0; kk < K; kk++) {
         float w[TM];
#pragma unroll
         for (int i = 0; i < TM; i++) {
            w[i] = wl[kk * FB + fg * TM + i];
         }
         float a[TN];
#pragma unroll
         for (int j = 0; j < TN; j++) {
            a[j] = srow[aoff[j] + kk];
         }
#pragma unroll
         for (int i = 0; i < TM; i++) {
#pragma unroll
            for (int j = 0; j < TN; j++) {
               if (EXACT) {
                  float prod = a[j] * w[i];
                  dv[i][j]   = dv[i][j] + prod;
               }
               else {
                  dv[i][j] = __builtin_fmaf(a[j], w[i], dv[i][j]);
               }
            }
         }
      }
#pragma unroll
      for (int i = 0; i < TM; i++) {
#pragma unroll
         for (int j = 0; j < TN; j++) {
            float scaled = p.dtf * dv[i][j];
            acc[i][j]    = acc[i][j] + scaled;
         }
      }
   }
#pragma unroll
   for (int j = 0; j < TN; j++) {
      if (!valid[j]) {
         continue;
      }
#pragma unroll
      for (int i = 0; i < TM; i++) {
         int f = f0 + fg * TM + i;
         if (f < p.post.nf) {
            post_epilogue(p, b, kOut[j] + f, acc[i][j]);
         }
      }
   }
}

// ------------------------------------------------------------------ mfma (bf16)
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x8 __attribute__((ext_vector_type(8)));

// Small-channel implicit GEMM (sources whose patch rows hold nxp*nfp values,
// a multiple of 8): M = features (one data patch per feature), N = target
// positions of a 16x8 tile, K = nyp*nxp*nfp in the weights' own order.
// Workgroup: 4 waves; wave w owns MW features for all 128 positions.
// Fragment layouts of v_mfma_f32_16x16x32_bf16 (lane l, r = l%16, g = l/16):
//   A[r][8g..8g+7], B[8g..8g+7][r], D[4g..4g+3][r].
// The source tile is staged once in LDS as bf16; a B fragment is 8
// consecutive k of one patch row, i.e. 8 consecutive LDS values, read with
// the widest load the tile's alignment class ALIGN (elements) allows.
constexpr int MTX = 16, MTY = 8;        // position tile
constexpr int MPB = MTX * MTY / 16;     // 16-position blocks per wave
constexpr int MF  = 64;                 // features per workgroup
constexpr int MFB = MF / 4 / 16;        // 16-feature blocks per wave

// Element types of the gather's low-precision forms.  A fragment is 8
// consecutive k (v_mfma_f32_16x16x32_bf16 and v_mfma_f32_16x16x32_fp8_fp8
// share the lane layout: A[r][8g..8g+7], B[8g..8g+7][r]).  FP8 values are
// scaled (weights per kernel, source per call, powers of two) so that they
// sit in the e4m3 range; the accumulators are scaled back before the
// epilogue.
struct GatherBF16 {
   using T    = __bf16;
   using Frag = bf16x8;
   using RawA = uint4; // 8 elements as raw 128-bit registers
   static constexpr bool SCALED = false;
   // four waves per SIMD (128 registers, the fused update spills only on the
   // padded tail's path) with two k-steps of weights in flight: measured
   // faster than three waves with four k-steps
   static constexpr int WAVES = 4; // waves per SIMD the gather is compiled for
   static constexpr int KPRE  = 2; // k-steps whose weight fragments are fetched together
   __device__ static T cvt(float x) { return (__bf16)x; }
   __device__ static f32x4 mfma(RawA a, Frag b, f32x4 c) {
      return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), b, c, 0, 0, 0);
   }
};
struct GatherFP8 {
   using T    = uint8_t;
   using Frag = long;
   using RawA = uint2;
   static constexpr bool SCALED = true;
   static constexpr int WAVES   = 1; // no bound (four waves measured slower, with two or four k-steps)
   static constexpr int KPRE    = 4;
   __device__ static T cvt(float x) {
      return (T)(__builtin_amdgcn_cvt_pk_fp8_f32(x, 0.0f, 0, false) & 0xff);
   }
   __device__ static f32x4 mfma(RawA a, Frag b, f32x4 c) {
      return __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(__builtin_bit_cast(long, a), b, c, 0, 0, 0);
   }
};

// The FP32 form on the matrix cores: a fragment is the same 8 consecutive k
// in FP32, consumed by eight v_mfma_f32_16x16x4f32 (MFMA j sums k = 8g + j
// of lane group g, the same permutation of k for A and B).
struct GatherF32 {
   using T    = float;
   using Frag = f32x8;
   using RawA = f32x8;
   static constexpr bool SCALED = false;
   static constexpr int WAVES   = 1; // no bound: four waves would spill the FP32 fragments
   static constexpr int KPRE    = 4;
   __device__ static T cvt(float x) { return x; }
   __device__ static f32x4 mfma(RawA a, Frag b, f32x4 c) {
#pragma unroll
      for (int j = 0; j < 8; j++) {
         c = __builtin_amdgcn_mfma_f32_16x16x4f32(a[j], b[j], c, 0, 0, 0);
      }
      return c;
   }
};

// 8 consecutive FP32 values at p, read with the widest load ALIGN allows.
template <int ALIGN>
__device__ __forceinline__ f32x8 lds_frag(const float *p) {
   f32x8 v;
   if (ALIGN >= 4) {
      f32x4 a = *reinterpret_cast<const f32x4 *>(p);
      f32x4 b = *reinterpret_cast<const f32x4 *>(p + 4);
      v       = f32x8{a[0], a[1], a[2], a[3], b[0], b[1], b[2], b[3]};
   }
   else if (ALIGN >= 2) {
#pragma unroll
      for (int e = 0; e < 8; e += 2) {
         float2 a = *reinterpret_cast<const float2 *>(p + e);
         v[e]     = a.x;
         v[e + 1] = a.y;
      }
   }
   else {
#pragma unroll
      for (int e = 0; e < 8; e++) {
         v[e] = p[e];
      }
   }
   return v;
}

// 8 consecutive FP8 values at p, read with the widest load ALIGN allows.
template <int ALIGN>
__device__ __forceinline__ long lds_frag8(const uint8_t *p) {
   if (ALIGN >= 8) {
      return *reinterpret_cast<const long *>(p);
   }
   uint32_t lo, hi;
   if (ALIGN >= 4) {
      lo = *reinterpret_cast<const uint32_t *>(p);
      hi = *reinterpret_cast<const uint32_t *>(p + 4);
   }
   else if (ALIGN >= 2) {
      const uint16_t *q = reinterpret_cast<const uint16_t *>(p);
      lo = (uint32_t)q[0] | ((uint32_t)q[1] << 16);
      hi = (uint32_t)q[2] | ((uint32_t)q[3] << 16);
   }
   else {
      lo = (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
      hi = (uint32_t)p[4] | ((uint32_t)p[5] << 8) | ((uint32_t)p[6] << 16) | ((uint32_t)p[7] << 24);
   }
   return (long)(((uint64_t)hi << 32) | lo);
}

template <int ALIGN>
__device__ __forceinline__ bf16x8 lds_frag(const __bf16 *p) {
   bf16x8 v;
   if (ALIGN >= 8) {
      v = *reinterpret_cast<const bf16x8 *>(p);
   }
   else if (ALIGN >= 4) {
      bf16x4 a = *reinterpret_cast<const bf16x4 *>(p);
      bf16x4 b = *reinterpret_cast<const bf16x4 *>(p + 4);
      v        = bf16x8{a[0], a[1], a[2], a[3], b[0], b[1], b[2], b[3]};
   }
   else if (ALIGN >= 2) {
      bf16x2 a = *reinterpret_cast<const bf16x2 *>(p);
      bf16x2 b = *reinterpret_cast<const bf16x2 *>(p + 2);
      bf16x2 c = *reinterpret_cast<const bf16x2 *>(p + 4);
      bf16x2 d = *reinterpret_cast<const bf16x2 *>(p + 6);
      v        = bf16x8{a[0], a[1], b[0], b[1], c[0], c[1], d[0], d[1]};
   }
   else {
#pragma unroll
      for (int e = 0; e < 8; e++) {
         v[e] = p[e];
      }
   }
   return v;
}

// Output tile of the BF16 gather staged in LDS for the epilogue: one row
// of MF features per position, padded by 4 floats (conflict-free float4
// writes of the MFMA D fragments).
constexpr int MOP = MF + 4;

// Four consecutive features (k .. k+3) of one neuron position with the
// partial sum acc (not yet scaled by dtf): GSyn store or the fused LCA
// update with 16-byte loads and non-temporal 16-byte stores of V and A
// (neither is read again by this step's gather; the stores stream past the
// caches).  EPI: 0 GSyn only, 1 fused update with the VTHRESH transfer,
// 2 fused update with either transfer.  Vo: the old V when EPI == 1.
template <int EPI>
__device__ __forceinline__ void post_epilogue4(
      const PostParams &p, int b, int k, f32x4 acc, f32x4 Vo) {
   size_t const gi = (size_t)b * p.post.numRes() + k;
   float G[4];
   if (p.accumulate) {
      float4 base = *reinterpret_cast<const float4 *>(p.gsyn + gi);
      G[0] = base.x + p.dtf * acc[0];
      G[1] = base.y + p.dtf * acc[1];
      G[2] = base.z + p.dtf * acc[2];
      G[3] = base.w + p.dtf * acc[3];
   }
   else {
#pragma unroll
      for (int i = 0; i < 4; i++) {
         G[i] = 0.0f + p.dtf * acc[i];
      }
   }
   if (EPI == 0) {
      *reinterpret_cast<float4 *>(p.gsyn + gi) = make_float4(G[0], G[1], G[2], G[3]);
      return;
   }
   const LCAParams &q = p.lca;
   float const e      = q.expTau[b];
   size_t const ai    = (size_t)b * q.loc.numExt() + q.loc.kex(k);
   if (EPI == 2) {
      Vo = *reinterpret_cast<const f32x4 *>(q.V + gi); // not prefetched: see the kernel
   }
   float ao[4];
   if (q.aFromV) {
#pragma unroll
      for (int i = 0; i < 4; i++) {
         ao[i] = transfer_as<EPI == 1 ? 1 : 0>(q.transfer, Vo[i]);
      }
   }
   else {
      float4 Ao = *reinterpret_cast<const float4 *>(q.A + ai);
      ao[0] = Ao.x;
      ao[1] = Ao.y;
      ao[2] = Ao.z;
      ao[3] = Ao.w;
   }
   f32x4 vn, an;
#pragma unroll
   for (int i = 0; i < 4; i++) {
      float t1 = e * Vo[i];
      float t2 = 1.0f - e;
      float si = q.selfInteract * ao[i];
      float t3 = G[i] + si;
      float t4 = t2 * t3;
      vn[i]    = t1 + t4;
      an[i]    = transfer_as<EPI == 1 ? 1 : 0>(q.transfer, vn[i]);
   }
   *reinterpret_cast<f32x4 *>(q.V + gi) = vn;   // read again by the next step
   if (q.A) {
      __builtin_nontemporal_store(an, reinterpret_cast<f32x4 *>(q.A + ai));
   }
   if (q.Ab) {
      // features k..k+3 sit in one 32-feature chunk (k % 4 == 0)
      *reinterpret_cast<bf16x4 *>(q.Ab + bf16c_index(q.loc, b, (int)(ai - (size_t)b * q.loc.numExt()))) =
            bf16x4{(__bf16)an[0], (__bf16)an[1], (__bf16)an[2], (__bf16)an[3]};
   }
}

// Epilogue row layout: MEQ threads cover the MF features of one position
// (4 each); a pass covers MEPP positions; MEP passes cover the tile.
constexpr int MEQ  = MF / 4;
constexpr int MEPP = 256 / MEQ;
constexpr int MEP  = MTX * MTY / MEPP;
static_assert(MTX == 16 && 16 % MEPP == 0, "row-layout epilogue assumes 16-wide position tiles");

// Fused VTHRESH update of a full tile in the row layout (the common case):
// thread (epr, eq) updates features ef .. ef+3 of positions epr + MEPP*it,
// it < MEP, whose partial sums sit at out[it * MEPP * MOP].  Same arithmetic,
// bit for bit, as post_epilogue4<1>; index math without divisions and the
// threshold as selects (the ramp of a firm threshold only when it exists).

__device__ __forceinline__ float vthresh_select(
      float v, bool thr, bool ramp, bool clamp, float VT, float AMin, float AShift, float AMax,
      float VW) {
   float a = v;
   if (thr) {
      a = v < VT ? AMin : v - AShift;
      if (ramp && !(v < VT) && v < VT + VW) {
         float span = VT + VW;
         span       = span - AShift;
         span       = span - AMin;
         float num  = span * (v - VT);
         a          = AMin + num / VW;
      }
   }
   if (clamp) {
      a = a > AMax ? AMax : a;
   }
   return a;
}

// The plain threshold (no ramp, no AMax), dtFactor 1, A(t-1) from V, only the
// chunked BF16 copy of A written: the batched LCA's BF16 / FP8 step.
__device__ __forceinline__ bool lca_plain_bf16(const PostParams &p) {
   const LCAParams &q = p.lca;
   Transfer const &t  = q.transfer;
   return t.VThresh > -FLT_MAX && !(t.VThresh + t.VWidth > t.VThresh) && !(t.AMax < FLT_MAX) && q.aFromV
          && p.dtf == 1.0f && q.A == nullptr && q.Ab != nullptr;
}

// PLAIN: the caller has checked lca_plain_bf16(p) (the plain threshold, no
// FP32 activity, the chunked BF16 copy written): straight-line code for it.
template <bool EMIT, bool PLAIN = false>
__device__ __forceinline__ void lca_epilogue_vthresh(
      const PostParams &p, int b, int x0, int y0, int epr, int eq, int ef, const float *out,
      const f32x4 *vpre) {
   const LCAParams &q = p.lca;
   float const e      = q.expTau[b];
   float const ome    = 1.0f - e;
   float const si     = q.selfInteract;
   Transfer const &t  = q.transfer;
   bool const thr     = t.VThresh > -FLT_MAX;
   bool const ramp    = t.VThresh + t.VWidth > t.VThresh;
   bool const clamp   = t.AMax < FLT_MAX;
   bool const simple  = PLAIN || (thr && !ramp && !clamp && q.aFromV && p.dtf == 1.0f); // uniform
   int const nf       = p.post.nf;
   int const rowV     = p.post.nx * nf;
   int const rowA     = q.loc.nxExt() * nf;
   float *Vp = q.V + (size_t)b * p.post.numRes() + (size_t)(y0 * p.post.nx + x0 + epr) * nf + ef;
   float *Ap = q.A ? q.A + (size_t)b * q.loc.numExt()
                           + (size_t)((y0 + q.loc.up) * q.loc.nxExt() + x0 + epr + q.loc.lt) * nf + ef
                   : nullptr; // NULL: only the chunked BF16 copy is written (A(t-1) from V)
   // chunked BF16 copy of A (optional): 4 features of one 32-feature chunk
   __bf16 *Abp = q.Ab ? q.Ab + ((((size_t)b * (nf / 32) + ef / 32) * q.loc.nyExt() + y0 + q.loc.up)
                                        * q.loc.nxExt() + x0 + epr + q.loc.lt) * 32 + ef % 32
                      : nullptr;
#pragma unroll
   for (int it = 0; it < MEP; it++) {
      int const yl    = it * MEPP / MTX;
      int const xo    = it * MEPP % MTX * nf;
      f32x4 const acc = *reinterpret_cast<const f32x4 *>(out + it * MEPP * MOP);
      f32x4 const vo  = vpre[it];
      f32x4 vn, an;
      if (simple) {
         // the plain threshold (no ramp, no AMax), dtFactor 1, A(t-1) from V:
         // the same operations as the general path with those terms dropped
#pragma unroll
         for (int i = 0; i < 4; i++) {
            float const ao = vo[i] < t.VThresh ? t.AMin : vo[i] - t.AShift;
            float G        = 0.0f + acc[i];
            float t1       = e * vo[i];
            float sa       = si * ao;
            float t3       = G + sa;
            float t4       = ome * t3;
            vn[i]          = t1 + t4;
            an[i]          = vn[i] < t.VThresh ? t.AMin : vn[i] - t.AShift;
         }
      }
      else {
         f32x4 ao;
         if (q.aFromV) {
#pragma unroll
            for (int i = 0; i < 4; i++) {
               ao[i] = vthresh_select(vo[i], thr, ramp, clamp, t.VThresh, t.AMin, t.AShift, t.AMax, t.VWidth);
            }
         }
         else {
            ao = *reinterpret_cast<const f32x4 *>(Ap + yl * rowA + xo);
         }
#pragma unroll
         for (int i = 0; i < 4; i++) {
            float G  = 0.0f + p.dtf * acc[i];
            float t1 = e * vo[i];
            float sa = si * ao[i];
            float t3 = G + sa;
            float t4 = ome * t3;
            vn[i]    = t1 + t4;
            an[i]    = vthresh_select(vn[i], thr, ramp, clamp, t.VThresh, t.AMin, t.AShift, t.AMax, t.VWidth);
         }
      }
      // V and the BF16 activity go through the caches (the next step reads
      // both; a plain store measured faster than a streaming one here).
      *reinterpret_cast<f32x4 *>(Vp + yl * rowV + xo) = vn;
      if (!PLAIN && Ap) {
         __builtin_nontemporal_store(an, reinterpret_cast<f32x4 *>(Ap + yl * rowA + xo));
      }
      if (PLAIN || Abp) {
         *reinterpret_cast<bf16x4 *>(Abp + (yl * q.loc.nxExt() + it * MEPP % MTX) * 32) =
               bf16x4{(__bf16)an[0], (__bf16)an[1], (__bf16)an[2], (__bf16)an[3]};
      }
      if (EMIT) {
         // Rank of this thread's actives within its position's segment: the
         // actives of lower lanes of the same MEQ-lane group, from ballots.
         uint64_t const m0 = __ballot(an[0] != 0.0f);
         uint64_t const m1 = __ballot(an[1] != 0.0f);
         uint64_t const m2 = __ballot(an[2] != 0.0f);
         uint64_t const m3 = __ballot(an[3] != 0.0f);
         int const lane      = threadIdx.x % kWave;
         uint64_t const grp  = ((MEQ == 64 ? ~0ull : ((1ull << MEQ) - 1)) << (lane - eq));
         uint64_t const lowr = grp & ((1ull << lane) - 1);
         int rank = __popcll(m0 & lowr) + __popcll(m1 & lowr) + __popcll(m2 & lowr) + __popcll(m3 & lowr);
         int const x      = x0 + epr + it * MEPP % MTX;
         int const y      = y0 + yl;
         size_t const seg = ((size_t)b * p.post.ny * p.post.nx + (size_t)y * p.post.nx + x)
                                  * (nf / MF) + ef / MF;
         uint2 *e         = reinterpret_cast<uint2 *>(p.stg + seg * MF);
         uint32_t kex     = (uint32_t)(((y + q.loc.up) * q.loc.nxExt() + x + q.loc.lt) * nf + ef);
#pragma unroll
         for (int i = 0; i < 4; i++) {
            if (an[i] != 0.0f) {
               e[rank++] = make_uint2(kex + i, __float_as_uint(an[i]));
            }
         }
         if (eq == MEQ - 1) {
            p.segCnt[seg] = (uint32_t)(__popcll(m0 & grp) + __popcll(m1 & grp) + __popcll(m2 & grp)
                                       + __popcll(m3 & grp));
         }
      }
   }
}

// EPI: 0 GSyn only, 1 fused VTHRESH update, 2 fused update with either
// transfer, 3 as 1 plus the active-list emission of full tiles.
template <int ALIGN, int EPI, typename E = GatherBF16>
__global__ void __launch_bounds__(256, E::WAVES) recv_post_mfma_bf16_kernel(
      PostParams p, const typename E::T *Wb, int srcRowsMax, int srcColsMax, int Kpad) {
   using T = typename E::T;
   extern __shared__ float smem[];
   int const K       = p.nxp * p.nfp * p.nyp;
   int const rowK    = p.nxp * p.nfp;
   int const pitch   = (srcColsMax * p.nfp + 7) / 8 * 8; // elements per LDS row
   T *src            = reinterpret_cast<T *>(smem);
   int *koff         = reinterpret_cast<int *>(smem + (srcRowsMax * pitch * sizeof(T) + 3) / 4); // [Kpad]
   int const tilesX  = (p.post.nx + MTX - 1) / MTX;
   int const tile    = blockIdx.x;
   int const x0      = (tile % tilesX) * MTX;
   int const y0      = (tile / tilesX) * MTY;
   int const fBase   = blockIdx.y * MF;
   int const b       = blockIdx.z;
   int const xl      = min(x0 + MTX, p.post.nx) - 1;
   int const yl      = min(y0 + MTY, p.post.ny) - 1;
   int const srcX0   = p.startX[x0];
   int const srcY0   = p.startY[y0];
   int const srcCols = p.startX[xl] - srcX0 + p.nxp;
   int const srcRows = p.startY[yl] - srcY0 + p.nyp;
   int const tid     = threadIdx.x;
   if (srcCols > srcColsMax || srcRows > srcRowsMax || srcX0 < 0 || srcY0 < 0
       || srcX0 + srcCols > p.pre.nxExt() || srcY0 + srcRows > p.pre.nyExt()) {
      if (tid == 0) {
         report_fault(FAULT_WINDOW); // tile refused: tables do not match the locs
      }
      return;
   }
   int const lane = tid % 64;
   int const wave = tid / 64;
   int const r    = lane % 16;
   int const g    = lane / 16;
   int const fWave = fBase + wave * (MF / 4);
   bool const vec4 = (p.post.nf & 3) == 0;
   // Epilogue layout (row layout): thread t owns features fBase + 4*(t % MEQ) ..
   // +3 of positions (t / MEQ) + MEPP * it, it < MEP.  The old V of those neurons
   // (fused VTHRESH update) is fetched first so that its latency overlaps the
   // staging and the MFMA loop; the vertex transfer needs those registers itself.
   int const eq  = tid % MEQ;
   int const epr = tid / MEQ;
   int const ef  = fBase + 4 * eq;
   f32x4 vpre[MEP];
#pragma unroll
   for (int it = 0; it < MEP; it++) {
      vpre[it] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
      if ((EPI == 1 || EPI == 3) && vec4) {
         int pj = it * MEPP + epr;
         int x  = x0 + pj % MTX;
         int y  = y0 + pj / MTX;
         if (x < p.post.nx && y < p.post.ny && ef < p.post.nf) {
            vpre[it] = *reinterpret_cast<const f32x4 *>(
                  p.lca.V + (size_t)b * p.post.numRes() + (y * p.post.nx + x) * p.post.nf + ef);
         }
      }
   }
   // Weight (A) fragments: KPRE k-steps per group, rows clamped to the last
   // feature (rows past it give outputs the epilogue never stores) and k
   // clamped inside Kpad, so every load is unconditional.  The first group is
   // issued here, before the staging; later groups are prefetched during the
   // MFMAs of the one before.
   constexpr int KPRE = E::KPRE; // k-steps whose A fragments are fetched together
   using RawA         = typename E::RawA;
   const T *rowA[MFB];
#pragma unroll
   for (int fb = 0; fb < MFB; fb++) {
      rowA[fb] = Wb + (size_t)min(fWave + fb * 16 + r, p.post.nf - 1) * Kpad;
   }
   auto loadA = [&](RawA (&dst)[KPRE][MFB], int kb) {
#pragma unroll
      for (int st = 0; st < KPRE; st++) {
         int const kk = min(kb + 32 * st, Kpad - 32) + 8 * g;
#pragma unroll
         for (int fb = 0; fb < MFB; fb++) {
            dst[st][fb] = *reinterpret_cast<const RawA *>(rowA[fb] + kk);
         }
      }
   };
   RawA aC[KPRE][MFB], aN[KPRE][MFB];
   loadA(aC, 0);
   int const preRow  = p.pre.nxExt() * p.pre.nf;
   float const *preB = p.pre_data + (size_t)b * p.pre.numExt();
   int const rowElems = srcCols * p.nfp;
   int const total    = srcRows * rowElems;
   // Element i = r * rowElems + c of the window, stepped by 256 with a carry
   // instead of a division per element.
   int const dR = 256 / rowElems;
   int const dC = 256 - dR * rowElems;
   int sr       = tid / rowElems;
   int sc       = tid - sr * rowElems;
   float const *winB = preB + (size_t)srcY0 * preRow + (size_t)srcX0 * p.nfp;
   for (int base = tid; base < total; base += 256 * 8) {
      // unconditional loads (clamped position) so all eight are in flight together
      float v[8];
      int rr[8], cc[8];
#pragma unroll
      for (int u = 0; u < 8; u++) {
         rr[u] = sr;
         cc[u] = sc;
         bool const ok = sr < srcRows;
         v[u] = winB[(size_t)(ok ? sr : srcRows - 1) * preRow + (ok ? sc : rowElems - 1)];
         sc += dC;
         sr += dR;
         if (sc >= rowElems) {
            sc -= rowElems;
            sr++;
         }
      }
#pragma unroll
      for (int u = 0; u < 8; u++) {
         if (rr[u] < srcRows) {
            src[rr[u] * pitch + cc[u]] = E::cvt(E::SCALED ? v[u] * p.lpSrcScale : v[u]);
         }
      }
   }
   for (int kk = tid; kk < Kpad; kk += 256) {
      koff[kk] = kk < K ? (kk / rowK) * pitch + (kk % rowK) : -1;
   }
   __syncthreads();

   int pbase[MPB];
#pragma unroll
   for (int pb = 0; pb < MPB; pb++) {
      int pj    = pb * 16 + r;
      int x     = x0 + pj % MTX;
      int y     = y0 + pj / MTX;
      bool ok   = x < p.post.nx && y < p.post.ny;
      pbase[pb] = ok ? (p.startY[y] - srcY0) * pitch + (p.startX[x] - srcX0) * p.nfp : 0;
   }
   if (ALIGN > 1) {
      // vector LDS reads need every window start on the alignment class
      bool bad = false;
#pragma unroll
      for (int pb = 0; pb < MPB; pb++) {
         bad = bad || (pbase[pb] % ALIGN) != 0;
      }
      if (__syncthreads_or(bad)) {
         if (tid == 0) {
            report_fault(FAULT_ALIGN); // tile refused: a window start off the alignment class
         }
         return;
      }
   }
   f32x4 acc[MFB][MPB];
#pragma unroll
   for (int fb = 0; fb < MFB; fb++) {
#pragma unroll
      for (int pb = 0; pb < MPB; pb++) {
         acc[fb][pb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
      }
   }
   // B fragment of k-step k0 at position block pb, the k values masked by koff
   // (per-element reads: patch rows not a multiple of 8, or zero padding)
   auto frag_masked = [&](int pb, const int *kos) {
      typename E::Frag bfrag;
      if constexpr (E::SCALED) {
         uint64_t v = 0;
#pragma unroll
         for (int e = 0; e < 8; e++) {
            v |= (uint64_t)(kos[e] >= 0 ? src[pbase[pb] + kos[e]] : (T)0) << (8 * e);
         }
         bfrag = (long)v;
      }
      else {
#pragma unroll
         for (int e = 0; e < 8; e++) {
            bfrag[e] = kos[e] >= 0 ? src[pbase[pb] + kos[e]] : (T)0.0f;
         }
      }
      return bfrag;
   };
   // Complete groups of KPRE k-steps whose 32 k are all real (no zero
   // padding): one vector LDS read per fragment, no branch and no condition
   // inside the group, so the accumulators stay in the MFMA registers.
   int const Kfull  = ALIGN != 0 ? K / 32 * 32 : 0;
   int const KfullG = Kfull / (32 * KPRE) * (32 * KPRE);
   for (int kb = 0; kb < KfullG; kb += 32 * KPRE) {
      if (kb + 32 * KPRE < KfullG) {
         loadA(aN, kb + 32 * KPRE); // in flight during this group's MFMAs
      }
#pragma unroll
      for (int st = 0; st < KPRE; st++) {
         int const ko = koff[kb + 32 * st + 8 * g];
#pragma unroll
         for (int pb = 0; pb < MPB; pb++) {
            typename E::Frag bfrag;
            if constexpr (E::SCALED) {
               bfrag = lds_frag8<ALIGN>(src + pbase[pb] + ko);
            }
            else {
               bfrag = lds_frag<ALIGN>(src + pbase[pb] + ko);
            }
#pragma unroll
            for (int fb = 0; fb < MFB; fb++) {
               acc[fb][pb] = E::mfma(aC[st][fb], bfrag, acc[fb][pb]);
            }
         }
      }
#pragma unroll
      for (int st = 0; st < KPRE; st++) {
#pragma unroll
         for (int fb = 0; fb < MFB; fb++) {
            aC[st][fb] = aN[st][fb];
         }
      }
   }
   // the remaining k-steps one at a time (all of them for ALIGN == 0): vector
   // reads while the step is all real, masked reads for the padded one
   for (int k0 = KfullG; k0 < Kpad; k0 += 32) {
      int const kk = k0 + 8 * g;
      RawA a1[MFB];
#pragma unroll
      for (int fb = 0; fb < MFB; fb++) {
         a1[fb] = *reinterpret_cast<const RawA *>(rowA[fb] + kk);
      }
      if (ALIGN != 0 && k0 + 32 <= K) {
         int const ko = koff[kk];
#pragma unroll
         for (int pb = 0; pb < MPB; pb++) {
            typename E::Frag bfrag;
            if constexpr (E::SCALED) {
               bfrag = lds_frag8<ALIGN>(src + pbase[pb] + ko);
            }
            else {
               bfrag = lds_frag<ALIGN>(src + pbase[pb] + ko);
            }
#pragma unroll
            for (int fb = 0; fb < MFB; fb++) {
               acc[fb][pb] = E::mfma(a1[fb], bfrag, acc[fb][pb]);
            }
         }
      }
      else {
         int kos[8];
#pragma unroll
         for (int e = 0; e < 8; e++) {
            kos[e] = koff[kk + e];
         }
#pragma unroll
         for (int pb = 0; pb < MPB; pb++) {
            typename E::Frag bfrag = frag_masked(pb, kos);
#pragma unroll
            for (int fb = 0; fb < MFB; fb++) {
               acc[fb][pb] = E::mfma(a1[fb], bfrag, acc[fb][pb]);
            }
         }
      }
   }
   if constexpr (E::SCALED) {
      // back from the scaled operands: D[f][pos] / (weight scale of f * source scale)
#pragma unroll
      for (int fb = 0; fb < MFB; fb++) {
#pragma unroll
         for (int i = 0; i < 4; i++) {
            int const f       = fWave + fb * 16 + 4 * g + i;
            float const unscl = f < p.post.nf ? 1.0f / (p.lpWScale[f] * p.lpSrcScale) : 0.0f;
#pragma unroll
            for (int pb = 0; pb < MPB; pb++) {
               acc[fb][pb][i] *= unscl;
            }
         }
      }
   }
   __syncthreads();
   float *outl = smem; // [MTX*MTY][MOP]
#pragma unroll
   for (int pb = 0; pb < MPB; pb++) {
#pragma unroll
      for (int fb = 0; fb < MFB; fb++) {
         *reinterpret_cast<f32x4 *>(outl + (pb * 16 + r) * MOP + wave * (MF / 4) + fb * 16 + 4 * g) =
               acc[fb][pb];
      }
   }
   __syncthreads();
   bool const fullTile = vec4 && fBase + MF <= p.post.nf && x0 + MTX <= p.post.nx
                         && y0 + MTY <= p.post.ny;
   if ((EPI == 1 || EPI == 3) && fullTile) {
      lca_epilogue_vthresh<EPI == 3>(p, b, x0, y0, epr, eq, ef, outl + epr * MOP + 4 * eq, vpre);
      return;
   }
#pragma unroll
   for (int it = 0; it < MEP; it++) {
      int const pj   = it * MEPP + epr;
      int const x    = x0 + pj % MTX;
      int const y    = y0 + pj / MTX;
      int const kpos = (y * p.post.nx + x) * p.post.nf;
      f32x4 const a4 = *reinterpret_cast<const f32x4 *>(outl + pj * MOP + 4 * eq);
      if (!(x < p.post.nx && y < p.post.ny)) {
         // position outside the layer: nothing to store
      }
      else if (ef + 3 < p.post.nf && vec4) {
         post_epilogue4<EPI == 3 ? 1 : EPI>(p, b, kpos + ef, a4, vpre[it]);
      }
      else {
#pragma unroll
         for (int i = 0; i < 4; i++) {
            if (ef + i < p.post.nf) {
               size_t const gi = (size_t)b * p.post.numRes() + kpos + ef + i;
               float const G   = (p.accumulate ? p.gsyn[gi] : 0.0f) + p.dtf * a4[i];
               if (EPI == 0) {
                  p.gsyn[gi] = G;
               }
               else {
                  lca_update_one<EPI == 1 || EPI == 3 ? 1 : 0>(p.lca, b, kpos + ef + i, G);
               }
            }
         }
      }
   }
}

// Persistent form of the gather + fused LCA update for the common case (full
// 16x8x64 tiles, VTHRESH transfer, patches of KS*32 values read with vector
// LDS loads): a workgroup keeps its 64 features' weight fragments in
// registers and walks `ipw` tiles (over positions and batch), fetching the
// next tile's old V and source window into registers before the current tile
// runs its MFMAs and epilogue, so the loads of one tile overlap the work and
// the stores of the one before.  Launched for the FP32 and BF16 forms whose
// weight rows are unpadded (K a multiple of 32).  Per neuron the arithmetic is that of
// recv_post_mfma_bf16_kernel<ALIGN, 1, E> (same MFMA order, same epilogue):
// the same bits.
constexpr int PSREG = 16; // source window values per thread held in registers (window <= 4096 values)

template <int ALIGN, typename E, int KS>
__global__ void __launch_bounds__(256) recv_post_mfma_persist_kernel(
      PostParams p, const typename E::T *Wb, int srcRowsMax, int srcColsMax, int items, int ipw) {
   using T    = typename E::T;
   using RawA = typename E::RawA;
   extern __shared__ float smem[];
   constexpr int K  = KS * 32;
   int const rowK   = p.nxp * p.nfp;
   int const pitch  = (srcColsMax * p.nfp + 7) / 8 * 8;
   T *src           = reinterpret_cast<T *>(smem);
   float *outl      = smem; // [MTX*MTY][MOP], after the MFMAs of a tile
   int const tilesX = p.post.nx / MTX;
   int const tilesXY = tilesX * (p.post.ny / MTY);
   int const fBase  = blockIdx.y * MF;
   int const tid    = threadIdx.x;
   int const lane   = tid % 64;
   int const wave   = tid / 64;
   int const r      = lane % 16;
   int const g      = lane / 16;
   int const fWave  = fBase + wave * (MF / 4);
   int const eq     = tid % MEQ;
   int const epr    = tid / MEQ;
   int const ef     = fBase + 4 * eq;
   int const preRow = p.pre.nxExt() * p.pre.nf;
   // weight fragments of every k-step, once
   RawA aAll[KS][MFB];
#pragma unroll
   for (int fb = 0; fb < MFB; fb++) {
      const T *rowA = Wb + (size_t)(fWave + fb * 16 + r) * K;
#pragma unroll
      for (int st = 0; st < KS; st++) {
         aAll[st][fb] = *reinterpret_cast<const RawA *>(rowA + 32 * st + 8 * g);
      }
   }
   int koffR[KS]; // LDS offset of this lane group's 8 k of each k-step (no fragment straddles a patch row)
#pragma unroll
   for (int st = 0; st < KS; st++) {
      int const kk = 32 * st + 8 * g;
      koffR[st]    = (kk / rowK) * pitch + (kk % rowK);
   }
   int const first = blockIdx.x * ipw;
   int const last  = min(first + ipw, items);
   if (first >= last) {
      return; // whole workgroup
   }
   // geometry of one tile: batch, corner, window origin, window size
   struct Tile {
      int b, x0, y0, srcX0, srcY0, rowElems, total;
   };
   auto tile_of = [&](int item) {
      Tile t;
      t.b        = item / tilesXY;
      int const tl = item % tilesXY;
      t.x0       = (tl % tilesX) * MTX;
      t.y0       = (tl / tilesX) * MTY;
      t.srcX0    = p.startX[t.x0];
      t.srcY0    = p.startY[t.y0];
      // the host's tables keep every window inside the extended source and
      // within the staged size; tables that do not are reported, and the
      // clamps keep the loads in bounds
      int const cols = p.startX[t.x0 + MTX - 1] - t.srcX0 + p.nxp;
      int const rows = p.startY[t.y0 + MTY - 1] - t.srcY0 + p.nyp;
      if (cols > srcColsMax || rows > srcRowsMax || t.srcX0 < 0 || t.srcY0 < 0
          || t.srcX0 + cols > p.pre.nxExt() || t.srcY0 + rows > p.pre.nyExt()) {
         if (tid == 0) {
            report_fault(FAULT_WINDOW);
         }
      }
      int const srcCols = min(cols, srcColsMax);
      int const srcRows = min(rows, srcRowsMax);
      t.srcX0    = max(0, min(t.srcX0, p.pre.nxExt() - srcCols));
      t.srcY0    = max(0, min(t.srcY0, p.pre.nyExt() - srcRows));
      t.rowElems = srcCols * p.nfp;
      t.total    = srcRows * t.rowElems;
      return t;
   };
   // source window of a tile -> registers (element i = tid + 256*u, clamped)
   auto fetch_src = [&](const Tile &t, float (&sv)[PSREG]) {
      float const *winB = p.pre_data + (size_t)t.b * p.pre.numExt() + (size_t)t.srcY0 * preRow
                          + (size_t)t.srcX0 * p.nfp;
      int sr = tid / t.rowElems;
      int sc = tid - sr * t.rowElems;
      int const dR = 256 / t.rowElems;
      int const dC = 256 - dR * t.rowElems;
      int const rows = t.total / t.rowElems;
#pragma unroll
      for (int u = 0; u < PSREG; u++) {
         bool const ok = sr < rows;
         sv[u] = winB[(size_t)(ok ? sr : rows - 1) * preRow + (ok ? sc : t.rowElems - 1)];
         sc += dC;
         sr += dR;
         if (sc >= t.rowElems) {
            sc -= t.rowElems;
            sr++;
         }
      }
   };
   auto store_src = [&](const Tile &t, const float (&sv)[PSREG]) {
      int sr = tid / t.rowElems;
      int sc = tid - sr * t.rowElems;
      int const dR = 256 / t.rowElems;
      int const dC = 256 - dR * t.rowElems;
      int const rows = t.total / t.rowElems;
#pragma unroll
      for (int u = 0; u < PSREG; u++) {
         if (sr < rows) {
            src[sr * pitch + sc] = E::cvt(E::SCALED ? sv[u] * p.lpSrcScale : sv[u]);
         }
         sc += dC;
         sr += dR;
         if (sc >= t.rowElems) {
            sc -= t.rowElems;
            sr++;
         }
      }
   };
   auto fetch_v = [&](const Tile &t, f32x4 (&v)[MEP]) {
#pragma unroll
      for (int it = 0; it < MEP; it++) {
         int const pj = it * MEPP + epr;
         int const x  = t.x0 + pj % MTX;
         int const y  = t.y0 + pj / MTX;
         v[it] = *reinterpret_cast<const f32x4 *>(
               p.lca.V + (size_t)t.b * p.post.numRes() + (y * p.post.nx + x) * p.post.nf + ef);
      }
   };
   auto bases = [&](const Tile &t, int (&pbase)[MPB]) {
#pragma unroll
      for (int pb = 0; pb < MPB; pb++) {
         int const pj = pb * 16 + r;
         int const x  = t.x0 + pj % MTX;
         int const y  = t.y0 + pj / MTX;
         pbase[pb]    = (p.startY[y] - t.srcY0) * pitch + (p.startX[x] - t.srcX0) * p.nfp;
      }
   };

   bool const plain = lca_plain_bf16(p); // the same for every tile
   Tile cur = tile_of(first);
   f32x4 vpre[MEP], vpreN[MEP];
   float sv[PSREG];
   int pbase[MPB], pbaseN[MPB];
   fetch_v(cur, vpre);
   fetch_src(cur, sv);
   bases(cur, pbase);
   store_src(cur, sv);
   __syncthreads();
   for (int item = first; item < last; item++) {
      f32x4 acc[MFB][MPB];
#pragma unroll
      for (int fb = 0; fb < MFB; fb++) {
#pragma unroll
         for (int pb = 0; pb < MPB; pb++) {
            acc[fb][pb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
         }
      }
      // the next tile's loads, in flight during this tile's MFMAs and epilogue
      bool const more = item + 1 < last; // the same for the whole workgroup
      Tile nxt        = cur;
      // Wave priority by phase: the waves issuing the next tile's loads and
      // running the MFMAs go before the waves of the other workgroup on the
      // CU that are in their epilogue (VALU and stores).  The workgroups then
      // drift out of step, so that loads, MFMAs and stores overlap; measured
      // 4-7 % faster than one priority throughout (no effect on results).
      __builtin_amdgcn_s_setprio(2);
      if (more) {
         nxt = tile_of(item + 1);
         fetch_v(nxt, vpreN);
         fetch_src(nxt, sv);
         bases(nxt, pbaseN);
      }
#pragma unroll
      for (int st = 0; st < KS; st++) {
#pragma unroll
         for (int pb = 0; pb < MPB; pb++) {
            typename E::Frag bfrag;
            if constexpr (E::SCALED) {
               bfrag = lds_frag8<ALIGN>(src + pbase[pb] + koffR[st]);
            }
            else {
               bfrag = lds_frag<ALIGN>(src + pbase[pb] + koffR[st]);
            }
#pragma unroll
            for (int fb = 0; fb < MFB; fb++) {
               acc[fb][pb] = E::mfma(aAll[st][fb], bfrag, acc[fb][pb]);
            }
         }
      }
      if constexpr (E::SCALED) {
#pragma unroll
         for (int fb = 0; fb < MFB; fb++) {
#pragma unroll
            for (int i = 0; i < 4; i++) {
               float const unscl = 1.0f / (p.lpWScale[fWave + fb * 16 + 4 * g + i] * p.lpSrcScale);
#pragma unroll
               for (int pb = 0; pb < MPB; pb++) {
                  acc[fb][pb][i] *= unscl;
               }
            }
         }
      }
      __builtin_amdgcn_s_setprio(1);
      __syncthreads(); // every wave is done reading the source tile
#pragma unroll
      for (int pb = 0; pb < MPB; pb++) {
#pragma unroll
         for (int fb = 0; fb < MFB; fb++) {
            *reinterpret_cast<f32x4 *>(outl + (pb * 16 + r) * MOP + wave * (MF / 4) + fb * 16 + 4 * g) =
                  acc[fb][pb];
         }
      }
      __syncthreads();
      if (plain) {
         lca_epilogue_vthresh<false, true>(p, cur.b, cur.x0, cur.y0, epr, eq, ef, outl + epr * MOP + 4 * eq, vpre);
      }
      else {
         lca_epilogue_vthresh<false>(p, cur.b, cur.x0, cur.y0, epr, eq, ef, outl + epr * MOP + 4 * eq, vpre);
      }
      if (more) {
         __syncthreads(); // every wave is done reading the output tile
         store_src(nxt, sv);
#pragma unroll
         for (int it = 0; it < MEP; it++) {
            vpre[it] = vpreN[it];
         }
#pragma unroll
         for (int pb = 0; pb < MPB; pb++) {
            pbase[pb] = pbaseN[pb];
         }
         cur = nxt;
         __syncthreads();
      }
   }
}

// ------------------------------------------------------------------ mfma (channel GEMM)
// Implicit GEMM for sources with nfp % CK == 0 channels and any number of
// data patches (targets larger than the source, e.g. the transposed form of
// a one-to-many scatter): M = every kernel (nKx*nKy*post.nf, padded to 16),
// N = distinct window starts (qx, qy), K = nyp*nxp*nfp taken as
// (channel chunk, ky, kx, channel).  Output (kernel m, window q) belongs to
// the target neuron whose kernel index and window start are (m, q): the
// inverse tables map (kernel x-index, qx) -> target x.
//
// Tiling: the window-start grid is cut into strips of QW columns; a
// workgroup takes CNT consecutive window starts of one strip in row-major
// order, so a tile wastes at most one partial tile per strip, and stages
// the source rows those starts touch (full strip width plus the patch
// halo).  Per channel chunk the source (converted to T) and the weights of
// every tap are staged in LDS with 16 bytes of padding per row of CK
// values (conflict-free 16-byte fragment reads); the next chunk's global
// loads are issued before the MFMAs of the current one.
// T = __bf16 uses v_mfma_f32_16x16x32_bf16 (CK = 32); T = float uses
// v_mfma_f32_16x16x4f32 (CK = 16, lane group g holds channels 4g..4g+3 and
// MFMA j of a chunk sums channel 4g+j, the same permutation for A and B).
// Window starts per workgroup: CNT = 4 waves * NBW blocks of 16 (NBW = 1, 2,
// 4: 64, 128 or 256 starts), chosen per launch (default 128).
constexpr int CSV  = 12;  // source float4 per thread carried across a chunk in registers
constexpr int CAV  = 4;   // weight 16-byte pieces per thread carried in registers

template <typename T>
struct CGemm;
template <>
struct CGemm<__bf16> {
   static constexpr int CK = 32;
   static constexpr int PP = 40; // LDS pitch (elements) of one row of CK values
   static constexpr int KL = 8;  // k values of one lane group in a fragment
   using Frag              = bf16x8;
};
template <>
struct CGemm<float> {
   static constexpr int CK = 16;
   static constexpr int PP = 20;
   static constexpr int KL = 4;
   using Frag              = float4;
};

struct CStrip {
   int qx0, qy0, nqx, nqy; // window-start grid in extended source coordinates
   int M;                  // kernels (GEMM rows) before padding
   int QW, nStrips, tilesPerStrip;
   int rowsMax, twMax;     // staged source tile bound (rows, columns)
   int remap;              // 1: workgroups of one batch share an XCD (grid % 8 == 0)
   int splits;             // channel ranges (gridDim.y); > 1: partial sums to part
   float *part;            // [splits][nOut] FP32 partials, or NULL
   size_t nOut;            // nbatch * post restricted size
};

__device__ __forceinline__ void store4(__bf16 *d, float4 v) {
   bf16x4 h = bf16x4{(__bf16)v.x, (__bf16)v.y, (__bf16)v.z, (__bf16)v.w};
   *reinterpret_cast<bf16x4 *>(d) = h;
}
__device__ __forceinline__ void store4(float *d, float4 v) { *reinterpret_cast<float4 *>(d) = v; }

// SRCB (T = __bf16 only): the source is p.preB16, the chunked BF16 copy of
// pre_data ([nbatch][C/32][nyExt][nxExt][32]), staged without conversion in
// 16-byte items of 8 channels; otherwise FP32 pre_data converted to T.
template <typename T, int MB, int CNBW, bool SRCB = false>
__global__ void __launch_bounds__(256) recv_post_cgemm_kernel(
      PostParams p, const T *Wq, const int32_t *invX, const int32_t *invY, CStrip cs) {
   static_assert(!SRCB || sizeof(T) == 2, "the chunked BF16 source feeds the BF16 GEMM only");
   constexpr int CNT = 4 * CNBW * 16;
   // Wq: weights in T, layout [M][nyp][nxp][C] (the post weights' own layout;
   // the BF16 shadow of pvhip_recv_post_bf16_weights or the FP32 weights).
   constexpr int CK = CGemm<T>::CK;
   constexpr int PP = CGemm<T>::PP;
   constexpr int Mp = MB * 16;
   constexpr int EP = 16 / sizeof(T); // elements per 16-byte piece
   extern __shared__ float smem[];
   T *srcl = reinterpret_cast<T *>(smem);           // [rowsMax][twMax][PP]
   T *wl   = srcl + (size_t)cs.rowsMax * cs.twMax * PP; // [taps + 1][Mp][PP], the last all zero

   int L = blockIdx.x;
   if (cs.remap) {
      L = (L % 8) * (gridDim.x / 8) + L / 8; // hardware places workgroup L on XCD L % 8
   }
   int const perBatch = cs.nStrips * cs.tilesPerStrip;
   int const b        = L / perBatch;
   int const strip    = (L % perBatch) / cs.tilesPerStrip;
   int const tile     = L % cs.tilesPerStrip;
   int const qxS      = strip * cs.QW;
   int const QWs      = min(cs.QW, cs.nqx - qxS);
   int const nS       = QWs * cs.nqy;
   int const n0       = tile * CNT;
   if (b >= p.post.nbatch || n0 >= nS) {
      return; // whole workgroup: no barrier is skipped by part of it
   }
   int const qyA  = n0 / QWs;
   int const qyB  = (min(n0 + CNT, nS) - 1) / QWs;
   int const rows = qyB - qyA + p.nyp;
   int const tw   = QWs + p.nxp - 1;
   int const tid  = threadIdx.x;
   int const lane = tid % 64;
   int const wave = tid / 64;
   int const r    = lane % 16;
   int const g    = lane / 16;
   int const C    = p.nfp;
   int const taps = p.nyp * p.nxp;
   int const patch = taps * C;
   int const nxExt = p.pre.nxExt();
   int const nyExt = p.pre.nyExt();
   float const *preB = p.pre_data + (size_t)b * p.pre.numExt();
   // chunked source: this batch element's planes of nyExt*nxExt*32 values
   size_t const plane32 = (size_t)nyExt * nxExt * 32;
   __bf16 const *preH   = SRCB ? p.preB16 + (size_t)b * (C / 32) * plane32 : nullptr;
   constexpr int SI     = SRCB ? 8 : 4; // channels per source item
   int const srcItems   = rows * tw * (CK / SI);
   int const wItems   = taps * Mp * (CK / EP);
   // the zero weight tap read by the tap loop past the last tap (ordered
   // before any read by the first chunk's barriers)
   for (int i = tid; i < Mp * PP; i += 256) {
      wl[(size_t)taps * Mp * PP + i] = T(0.0f);
   }
   // this workgroup's channel range: chunks [ks*chunks/splits, (ks+1)*chunks/splits)
   int const chunks = C / CK;
   int const ks     = blockIdx.y;
   int const cBeg   = (ks * chunks / cs.splits) * CK;
   int const cEnd   = ((ks + 1) * chunks / cs.splits) * CK;

   // LDS offset of each MFMA column (window start) of this lane at tap 0
   int nbase[CNBW];
#pragma unroll
   for (int nb = 0; nb < CNBW; nb++) {
      int n     = min(n0 + (wave * CNBW + nb) * 16 + r, nS - 1);
      nbase[nb] = ((n / QWs - qyA) * cs.twMax + n % QWs) * PP;
   }

   // Item offsets do not depend on the chunk: computed once.  A source item
   // is 4 channels of one staged position; a weight item 16 bytes of one
   // (tap, kernel) row of CK channels.
   // (FP32 source: offset in the batch element's PV-layout activity, to which
   // the chunk's first channel is added; chunked source: offset within a
   // chunk plane.)
   auto src_goff = [&](int i) {
      int q4 = i % (CK / SI);
      int cc = (i / (CK / SI)) % tw;
      int rr = i / ((CK / SI) * tw);
      int ye = min(cs.qy0 + qyA + rr, nyExt - 1);
      int xe = min(cs.qx0 + qxS + cc, nxExt - 1);
      return SRCB ? (ye * nxExt + xe) * 32 + SI * q4 : (ye * nxExt + xe) * C + SI * q4;
   };
   auto src_loff = [&](int i) {
      int q4 = i % (CK / SI);
      int cc = (i / (CK / SI)) % tw;
      int rr = i / ((CK / SI) * tw);
      return (rr * cs.twMax + cc) * PP + SI * q4;
   };
   // source item: 4 FP32 channels, or 8 BF16 channels of the chunked copy
   using SV   = std::conditional_t<SRCB, uint4, float4>;
   auto sload = [&](int goff, int c0) -> SV {
      if constexpr (SRCB) {
         return *reinterpret_cast<const uint4 *>(preH + (size_t)(c0 / 32) * plane32 + goff);
      }
      else {
         return *reinterpret_cast<const float4 *>(preB + goff + c0);
      }
   };
   auto sstore = [&](int loff, SV v) {
      if constexpr (SRCB) {
         *reinterpret_cast<uint4 *>(srcl + loff) = v;
      }
      else {
         store4(srcl + loff, v);
      }
   };
   auto w_goff = [&](int j) { // -1: padding kernel row (zeros)
      int pc = j % (CK / EP);
      int m  = (j / (CK / EP)) % Mp;
      int t  = j / ((CK / EP) * Mp);
      return m < cs.M ? m * patch + t * C + pc * EP : -1;
   };
   auto w_loff = [&](int j) {
      int pc = j % (CK / EP);
      int m  = (j / (CK / EP)) % Mp;
      int t  = j / ((CK / EP) * Mp);
      return (t * Mp + m) * PP + pc * EP;
   };
   // Channel chunks in flight in registers: the chunked BF16 source and the
   // FP32 GEMM (16-channel chunks) need few registers per chunk, so two chunks
   // are fetched ahead; the BF16 GEMM converting an FP32 source one.
   constexpr int DEPTH = (SRCB || sizeof(T) == 4) ? 2 : 1;
   constexpr int NSV   = DEPTH == 2 ? 6 : CSV; // source items per thread and chunk held in registers
   int sG[NSV], sL[NSV], wG[CAV], wL[CAV];
#pragma unroll
   for (int u = 0; u < NSV; u++) {
      int i = tid + u * 256;
      sG[u] = src_goff(min(i, srcItems - 1));
      sL[u] = i < srcItems ? src_loff(i) : -1;
   }
#pragma unroll
   for (int u = 0; u < CAV; u++) {
      int j = tid + u * 256;
      wG[u] = w_goff(min(j, wItems - 1));
      wL[u] = j < wItems ? w_loff(j) : -1;
   }
   SV sv[DEPTH][NSV];
   uint4 wv[DEPTH][CAV];
   // global -> register set d for chunk c0 (the first NSV / CAV items per thread)
   auto fetch = [&](int d, int c0) {
#pragma unroll
      for (int u = 0; u < NSV; u++) {
         sv[d][u] = sload(sG[u], c0);
      }
#pragma unroll
      for (int u = 0; u < CAV; u++) {
         wv[d][u] = wG[u] >= 0 ? *reinterpret_cast<const uint4 *>(Wq + (size_t)wG[u] + c0)
                               : make_uint4(0u, 0u, 0u, 0u);
      }
   };
   // register set d (and any items beyond it, loaded now) -> LDS for chunk c0
   auto stage = [&](int d, int c0) {
#pragma unroll
      for (int u = 0; u < NSV; u++) {
         if (sL[u] >= 0) {
            sstore(sL[u], sv[d][u]);
         }
      }
      for (int i = tid + NSV * 256; i < srcItems; i += 256) {
         sstore(src_loff(i), sload(src_goff(i), c0));
      }
#pragma unroll
      for (int u = 0; u < CAV; u++) {
         if (wL[u] >= 0) {
            *reinterpret_cast<uint4 *>(wl + wL[u]) = wv[d][u];
         }
      }
      for (int j = tid + CAV * 256; j < wItems; j += 256) {
         int go  = w_goff(j);
         uint4 v = go >= 0 ? *reinterpret_cast<const uint4 *>(Wq + (size_t)go + c0) : make_uint4(0u, 0u, 0u, 0u);
         *reinterpret_cast<uint4 *>(wl + w_loff(j)) = v;
      }
   };

   f32x4 acc[MB][CNBW];
#pragma unroll
   for (int mb = 0; mb < MB; mb++) {
#pragma unroll
      for (int nb = 0; nb < CNBW; nb++) {
         acc[mb][nb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
      }
   }
   // Software-pipelined over the taps, two register sets in turn: the
   // fragments of the next tap are read from LDS while the MFMAs of the
   // current one run.  Loads are issued unconditionally (no branch, so each
   // MFMA waits only for its own set): a tap index past the last one reads
   // the zero weight tap at index `taps` (adds nothing).  Tap offsets step
   // (ky, kx) without a division.
   using Frag = typename CGemm<T>::Frag;
   auto load_frags = [&](Frag *af, Frag *bfr, int u, int ky, int kx) {
      bool const real  = u < taps;
      T const *wt      = wl + (real ? u : taps) * Mp * PP;
      int const tapOff = real ? (ky * cs.twMax + kx) * PP : 0;
#pragma unroll
      for (int mb = 0; mb < MB; mb++) {
         af[mb] = *reinterpret_cast<const Frag *>(wt + (mb * 16 + r) * PP + CGemm<T>::KL * g);
      }
#pragma unroll
      for (int nb = 0; nb < CNBW; nb++) {
         bfr[nb] = *reinterpret_cast<const Frag *>(srcl + nbase[nb] + tapOff + CGemm<T>::KL * g);
      }
   };
   auto mfma_tap = [&](const Frag *af, const Frag *bfr) {
#pragma unroll
      for (int nb = 0; nb < CNBW; nb++) {
#pragma unroll
         for (int mb = 0; mb < MB; mb++) {
            if constexpr (CK == 32) {
               acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[mb], bfr[nb], acc[mb][nb], 0, 0, 0);
            }
            else {
               acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[mb].x, bfr[nb].x, acc[mb][nb], 0, 0, 0);
               acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[mb].y, bfr[nb].y, acc[mb][nb], 0, 0, 0);
               acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[mb].z, bfr[nb].z, acc[mb][nb], 0, 0, 0);
               acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[mb].w, bfr[nb].w, acc[mb][nb], 0, 0, 0);
            }
         }
      }
   };
   auto compute = [&]() { // every tap of the staged chunk
      Frag afC[MB], bfC[CNBW], afN[MB], bfN[CNBW];
      int ky = 0, kx = 0; // tap of the next load
      auto advance = [&]() {
         if (++kx == p.nxp) {
            kx = 0;
            ky++;
         }
      };
      load_frags(afC, bfC, 0, 0, 0);
      advance();
      for (int t = 0; t < taps; t += 2) {
         load_frags(afN, bfN, t + 1, ky, kx);
         advance();
         mfma_tap(afC, bfC);
         load_frags(afC, bfC, t + 2, ky, kx);
         advance();
         mfma_tap(afN, bfN);
      }
   };
   int const nCh = (cEnd - cBeg) / CK;
#pragma unroll
   for (int d = 0; d < DEPTH; d++) {
      if (d < nCh) {
         fetch(d, cBeg + d * CK);
      }
   }
   for (int i0 = 0; i0 < nCh; i0 += DEPTH) {
#pragma unroll
      for (int d = 0; d < DEPTH; d++) {
         int const i = i0 + d;
         if (i < nCh) { // the same for the whole workgroup: no barrier is skipped by part of it
            int const c0 = cBeg + i * CK;
            __syncthreads(); // every wave is done reading the previous chunk
            stage(d, c0);
            __syncthreads();
            if (i + DEPTH < nCh) {
               fetch(d, c0 + DEPTH * CK); // in flight during the MFMAs of the next DEPTH chunks
            }
            compute();
         }
      }
   }
   // Lane holds D[4g+i][r]: kernel mb*16+4g+i at window start n of block nb.
   // The target coordinates of every element are looked up first (clamped
   // indices, so the loads are unconditional and in flight together), then
   // the elements are stored.
   int const nf = p.post.nf;
   int tx[CNBW][MB][4], ty[CNBW][MB][4];
#pragma unroll
   for (int nb = 0; nb < CNBW; nb++) {
      int const n  = min(n0 + (wave * CNBW + nb) * 16 + r, nS - 1);
      int const qx = qxS + n % QWs;
      int const qy = n / QWs;
#pragma unroll
      for (int mb = 0; mb < MB; mb++) {
#pragma unroll
         for (int i = 0; i < 4; i++) {
            int const m  = min(mb * 16 + 4 * g + i, cs.M - 1);
            int const kx = (m / nf) % p.nKx;
            int const ky = m / (nf * p.nKx);
            tx[nb][mb][i] = invX[kx * cs.nqx + qx];
            ty[nb][mb][i] = invY[ky * cs.nqy + qy];
         }
      }
   }
#pragma unroll
   for (int nb = 0; nb < CNBW; nb++) {
      bool const inside = n0 + (wave * CNBW + nb) * 16 + r < nS;
#pragma unroll
      for (int mb = 0; mb < MB; mb++) {
#pragma unroll
         for (int i = 0; i < 4; i++) {
            int const m = mb * 16 + 4 * g + i;
            int const x = tx[nb][mb][i];
            int const y = ty[nb][mb][i];
            if (inside && m < cs.M && x >= 0 && y >= 0) {
               size_t gi = (size_t)b * p.post.numRes() + ((size_t)y * p.post.nx + x) * nf + m % nf;
               if (cs.part) {
                  cs.part[(size_t)ks * cs.nOut + gi] = acc[mb][nb][i];
               }
               else if (p.layer) {
                  Geo const &L = p.layerLoc;
                  size_t ai    = (size_t)b * L.numExt() + ((size_t)(y + L.up) * L.nxExt() + x + L.lt) * nf + m % nf;
                  fused_layer_store(p, gi, ai, 0.0f + p.dtf * acc[mb][nb][i]);
               }
               else {
                  float base = p.accumulate ? p.gsyn[gi] : 0.0f;
                  p.gsyn[gi] = base + p.dtf * acc[mb][nb][i];
               }
            }
         }
      }
   }
}

// Split channel GEMM with the target layer's update fused: the partials of
// each target neuron added in split order, then fused_layer_store.  Grid
// (ceil(numRes / 1024), nbatch), four neurons per thread, unit stride.
__global__ void __launch_bounds__(256) cgemm_reduce_layer_kernel(PostParams p, const float *part, int splits, size_t n) {
   int const numRes = p.post.numRes();
   int const rowLen = p.post.nx * p.post.nf;
   int const b      = blockIdx.y;
#pragma unroll
   for (int u = 0; u < 4; u++) {
      int const k = blockIdx.x * 1024 + u * 256 + threadIdx.x;
      if (k < numRes) {
         size_t const i = (size_t)b * numRes + k;
         float acc      = part[i];
         for (int s = 1; s < splits; s++) {
            acc += part[(size_t)s * n + i];
         }
         int const y     = k / rowLen;
         Geo const &L    = p.layerLoc;
         size_t const ai = (size_t)b * L.numExt() + ((size_t)(y + L.up) * L.nxExt() + L.lt) * L.nf + (k - y * rowLen);
         fused_layer_store(p, i, ai, 0.0f + p.dtf * acc);
      }
   }
}

// Split channel GEMM: GSyn (+)= dtf * sum of the split partials, added in
// split order (deterministic).  VEC: n % 4 == 0, one float4 per thread.
template <bool VEC>
__global__ void __launch_bounds__(256) cgemm_reduce_kernel(
      const float *part, int splits, size_t n, float *gsyn, int accumulate, float dtf) {
   size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
   if constexpr (VEC) {
      if (i >= n / 4) {
         return;
      }
      float4 acc = reinterpret_cast<const float4 *>(part)[i];
      for (int k = 1; k < splits; k++) {
         float4 v = reinterpret_cast<const float4 *>(part + (size_t)k * n)[i];
         acc.x += v.x;
         acc.y += v.y;
         acc.z += v.z;
         acc.w += v.w;
      }
      float4 base = accumulate ? reinterpret_cast<const float4 *>(gsyn)[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      float4 out;
      out.x = base.x + dtf * acc.x;
      out.y = base.y + dtf * acc.y;
      out.z = base.z + dtf * acc.z;
      out.w = base.w + dtf * acc.w;
      reinterpret_cast<float4 *>(gsyn)[i] = out;
   }
   else {
      if (i >= n) {
         return;
      }
      float acc = part[i];
      for (int k = 1; k < splits; k++) {
         acc += part[(size_t)k * n + i];
      }
      float base = accumulate ? gsyn[i] : 0.0f;
      gsyn[i]    = base + dtf * acc;
   }
}

// FP8 weight shadow: one workgroup per kernel (data patch); its scale is the
// power of two that brings max |w| into [120, 240) (the e4m3 range with
// headroom), stored after the [kernels][Kpad] bytes.
__global__ void __launch_bounds__(256) weights_to_fp8_kernel(
      const float *W, uint8_t *Wq, float *scales, int K, int Kpad) {
   int const f = blockIdx.x;
   float m     = 0.0f;
   for (int k = threadIdx.x; k < K; k += 256) {
      m = fmaxf(m, fabsf(W[(size_t)f * K + k]));
   }
   for (int off = kWave / 2; off > 0; off >>= 1) {
      m = fmaxf(m, __shfl_down(m, off, kWave));
   }
   __shared__ float red[256 / kWave];
   if (threadIdx.x % kWave == 0) {
      red[threadIdx.x / kWave] = m;
   }
   __syncthreads();
   m = red[0];
   for (int w = 1; w < 256 / kWave; w++) {
      m = fmaxf(m, red[w]);
   }
   float sc = 1.0f;
   if (m > 0.0f) {
      int e;
      frexpf(240.0f / m, &e); // 240/m = q * 2^e, q in [0.5, 1)
      sc = ldexpf(1.0f, e - 1);
   }
   if (threadIdx.x == 0) {
      scales[f] = sc;
   }
   for (int k = threadIdx.x; k < Kpad; k += 256) {
      float v               = k < K ? W[(size_t)f * K + k] * sc : 0.0f;
      Wq[(size_t)f * Kpad + k] = GatherFP8::cvt(v);
   }
}

__global__ void weights_to_bf16_kernel(
      const float *W, __bf16 *Wb, int numKernels, int K, int Kpad) {
   size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
   size_t n = (size_t)numKernels * Kpad;
   if (i >= n) {
      return;
   }
   int f  = (int)(i / Kpad);
   int kk = (int)(i % Kpad);
   Wb[i]  = kk < K ? (__bf16)W[(size_t)f * K + kk] : (__bf16)0.0f;
}

// Halo a source of n_pre needs on each side for a post-perspective patch of
// `patch` values into a target of n_post: requiredConvolveMargin
// (src/utils/requiredConvolveMargin.hpp:14-40) of the connection's own
// pre-perspective patch, rewritten in terms of the post-perspective patch of
// PatchSize::calcPostPatchSize (src/components/PatchSize.cpp:98-130).
// -1: the patch does not fit the scale.
int gather_margin(int n_pre, int n_post, int patch) {
   if (n_pre <= 0 || n_post <= 0 || patch <= 0) {
      return -1;
   }
   if (n_pre == n_post) {
      return patch % 2 == 1 ? (patch - 1) / 2 : -1; // one to one: odd patches
   }
   if (n_pre > n_post) { // many to one, ratio r: pre-perspective patch patch/r
      int const r = n_pre / n_post;
      if (n_pre % n_post || patch % r) {
         return -1;
      }
      return (patch / r - 1) * r / 2;
   }
   // one to many: the pre-perspective patch is patch * (n_post / n_pre)
   return n_post % n_pre ? -1 : patch / 2;
}

int build_params(const pvhip_recv_post_args &a, PostParams *p) {
   PVHIP_REQUIRE(loc_valid(a.pre) && loc_valid(a.post), "recv_post: invalid loc");
   PVHIP_REQUIRE(a.pre.nbatch == a.post.nbatch, "recv_post: batch sizes differ");
   PVHIP_REQUIRE(a.nxp > 0 && a.nyp > 0 && a.nfp == a.pre.nf, "recv_post: patch nfp must equal pre nf");
   PVHIP_REQUIRE(a.num_kernels_x > 0 && a.num_kernels_y > 0, "recv_post: bad kernel counts");
   PVHIP_REQUIRE(a.start_x && a.start_y && a.kernel_x && a.kernel_y, "recv_post: NULL geometry table");
   PVHIP_REQUIRE(a.pre_data && a.weights, "recv_post: NULL pre_data or weights");
   {
      // every side of the source halo must cover the reach of the patches
      // (the margin the connection requires of its pre layer,
      // requiredConvolveMargin in terms of the post-perspective patch)
      int const mx = gather_margin(a.pre.nx, a.post.nx, a.nxp);
      int const my = gather_margin(a.pre.ny, a.post.ny, a.nyp);
      PVHIP_REQUIRE(mx >= 0 && my >= 0, "recv_post: patch %dx%d does not fit the scale %d->%d x %d->%d", a.nxp,
            a.nyp, a.pre.nx, a.post.nx, a.pre.ny, a.post.ny);
      PVHIP_REQUIRE(a.pre.halo.lt >= mx && a.pre.halo.rt >= mx && a.pre.halo.dn >= my && a.pre.halo.up >= my,
            "recv_post: source halo (lt %d, rt %d, dn %d, up %d) smaller than the margin %d x %d the %dx%d "
            "post-perspective patch needs",
            a.pre.halo.lt, a.pre.halo.rt, a.pre.halo.dn, a.pre.halo.up, mx, my, a.nxp, a.nyp);
   }
   p->pre      = geo(a.pre);
   p->post     = geo(a.post);
   p->nxp      = a.nxp;
   p->nyp      = a.nyp;
   p->nfp      = a.nfp;
   p->nKx      = a.num_kernels_x;
   p->nKy      = a.num_kernels_y;
   p->startX   = a.start_x;
   p->startY   = a.start_y;
   p->kerX     = a.kernel_x;
   p->kerY     = a.kernel_y;
   p->dtf      = a.dt_factor;
   p->pre_data = a.pre_data;
   p->W        = a.weights;
   p->gsyn     = a.post_gsyn;
   p->accumulate = a.accumulate;
   p->preB16     = static_cast<const __bf16 *>(a.pre_data_bf16c);
   p->emit       = 0;
   p->stg        = nullptr;
   p->segCnt     = nullptr;
   p->layer      = a.fused_layer != nullptr;
   if (p->layer) {
      const pvhip_layer_update_args &l = *a.fused_layer;
      PVHIP_REQUIRE(!a.fused_update && !a.accumulate,
            "recv_post: fused_layer needs accumulate == 0 and no fused_update");
      PVHIP_REQUIRE(loc_valid(l.loc) && l.loc.nbatch == a.post.nbatch && l.loc.nx == a.post.nx
                          && l.loc.ny == a.post.ny && l.loc.nf == a.post.nf,
            "recv_post: fused_layer loc differs from post loc");
      PVHIP_REQUIRE(l.V != nullptr, "recv_post: fused_layer needs V");
      p->layerLoc        = geo(l.loc);
      p->layerCoeff      = l.channel_coeffs[0];
      p->layerScale      = l.scale;
      p->layerApplyScale = l.apply_scale;
      p->layerV          = l.V;
      p->layerA          = l.A;
      int st             = make_transfer(l.transfer, &p->layerT);
      if (st) {
         return st;
      }
   }
   p->fused      = a.fused_update != nullptr;
   if (p->fused) {
      PVHIP_REQUIRE(!a.accumulate, "recv_post: fused update requires accumulate == 0");
      const pvhip_lca_update_args &u = *a.fused_update;
      PVHIP_REQUIRE(
            u.loc.nbatch == a.post.nbatch && u.loc.nx == a.post.nx && u.loc.ny == a.post.ny
                  && u.loc.nf == a.post.nf,
            "recv_post: fused update loc differs from post loc");
      PVHIP_REQUIRE(
            u.num_channels == 1 && u.channel_coeffs[0] == 1.0f,
            "recv_post: fused update needs one channel with coefficient 1");
      pvhip_lca_update_args u2 = u;
      u2.gsyn                  = nullptr;
      u2.num_channels          = 0;
      int st                   = make_lca_params(u2, &p->lca);
      if (st) {
         return st;
      }
   }
   else {
      PVHIP_REQUIRE(a.post_gsyn != nullptr || p->layer, "recv_post: NULL post_gsyn");
   }
   return PVHIP_OK;
}

// Staging of the epilogue-emitted active list: MF-entry slots for every
// restricted neuron, then one count per segment (256-byte aligned).
size_t segment_stage_bytes(const pvhip_loc &l) {
   return ((size_t)l.nbatch * l.nx * l.ny * l.nf * sizeof(pvhip_sparse_entry) + 255) / 256 * 256;
}
size_t segment_scratch_bytes(const pvhip_loc &l) {
   return segment_stage_bytes(l) + (size_t)l.nbatch * l.nx * l.ny * ((l.nf + MF - 1) / MF) * sizeof(uint32_t);
}

int recv_post_launch(const pvhip_recv_post_args *args, PostParams &p, hipStream_t s);

// Channel GEMM launch shape: strips, LDS, grid, and the channel split that
// gives the launch about CGEMM_WG_TARGET workgroups (a workgroup runs its
// channel chunks one after another, each staged through LDS, so a short
// grid leaves the chip waiting on those loads).
constexpr long long CGEMM_WG_TARGET = 2048;

struct CGemmPlan {
   CStrip cs;
   size_t lds;
   long long blocks; // workgroups per channel range
   int MB;
   int NBW; // 16-start MFMA blocks per wave: CNT = 64 * NBW window starts per workgroup
   bool bf;
};

int cgemm_plan(const pvhip_recv_post_args *args, const PostParams &p, CGemmPlan *pl) {
   PVHIP_REQUIRE(args->inverse_x && args->inverse_y && args->window_nx > 0 && args->window_ny > 0,
         "recv_post: targets larger than the source need inverse tables (pvhost_recv_post_inverse)");
   PVHIP_REQUIRE(!p.fused, "recv_post: fused update needs one data patch per feature");
   bool bf = args->arith == PVHIP_ARITH_BF16;
   int CK  = bf ? 32 : 16;
   PVHIP_REQUIRE(p.nfp % CK == 0, "recv_post: MFMA path with %d data patches needs nfp %% %d == 0",
         p.nKx * p.nKy, CK);
   int M  = p.nKx * p.nKy * p.post.nf;
   int MB = (M + 15) / 16;
   PVHIP_REQUIRE(MB <= 4, "recv_post: %d kernels exceed the MFMA tile (64)", M);
   MB = MB <= 1 ? 1 : (MB <= 2 ? 2 : 4);
   if (bf) {
      PVHIP_REQUIRE(args->weights_lp != nullptr,
            "recv_post: BF16 channel GEMM needs weights_lp (pvhip_recv_post_bf16_weights)");
   }
   int const tile = args->cgemm_tile;
   PVHIP_REQUIRE(tile == 0 || tile == 64 || tile == 128 || tile == 256,
         "recv_post: cgemm_tile %d is not 0, 64, 128 or 256", tile);
   int const NBW = tile == 0 ? 2 : tile / 64;
   int const CNT = 64 * NBW;
   CStrip cs;
   cs.qx0 = args->window_x0;
   cs.qy0 = args->window_y0;
   cs.nqx = args->window_nx;
   cs.nqy = args->window_ny;
   cs.M   = M;
   size_t const esz  = bf ? 2 : 4;
   int const PP      = bf ? CGemm<__bf16>::PP : CGemm<float>::PP;
   size_t const wlds = ((size_t)p.nyp * p.nxp + 1) * MB * 16 * PP * esz; // + the zero tap
   size_t lds        = 0;
   // widest strip (at most 64 starts) whose staged tile fits the LDS budget
   for (int strips = (cs.nqx + 63) / 64;; strips++) {
      cs.QW      = (cs.nqx + strips - 1) / strips;
      cs.nStrips = (cs.nqx + cs.QW - 1) / cs.QW;
      // a tile of CNT starts in a strip of width w spans <= (w + CNT - 2) / w + 1
      // start rows; the narrowest strip (the last, width >= 1) gives the bound
      int const wmin = cs.nqx - (cs.nStrips - 1) * cs.QW;
      cs.rowsMax = (wmin + CNT - 2) / wmin + p.nyp;
      cs.twMax   = cs.QW + p.nxp - 1;
      lds        = (size_t)cs.rowsMax * cs.twMax * PP * esz + wlds;
      if (lds <= 96 * 1024 || cs.QW == 1) {
         break;
      }
   }
   PVHIP_REQUIRE(lds <= 160 * 1024, "recv_post: channel GEMM tile needs %zu bytes of LDS", lds);
   cs.tilesPerStrip = (cs.QW * cs.nqy + CNT - 1) / CNT;
   long long blocks = (long long)p.post.nbatch * cs.nStrips * cs.tilesPerStrip;
   PVHIP_REQUIRE(blocks < (1LL << 31), "recv_post: channel GEMM grid too large");
   cs.remap     = blocks % 8 == 0 ? 1 : 0;
   int const chunks = p.nfp / CK;
   PVHIP_REQUIRE(args->cgemm_splits >= 0, "recv_post: negative cgemm_splits");
   if (args->cgemm_splits > 0) {
      cs.splits = std::min(args->cgemm_splits, std::min(chunks, 65535));
   }
   else if (bf) {
      cs.splits = 1; // measured: the BF16 chunks do not gain from a split
   }
   else {
      // ranges of at least 8 chunks, enough of them for about CGEMM_WG_TARGET workgroups
      long long sp = (CGEMM_WG_TARGET + blocks - 1) / blocks;
      cs.splits    = (int)std::min<long long>(std::max<long long>(sp, 1), std::max(chunks / 8, 1));
   }
   cs.part      = nullptr;
   cs.nOut      = 0;
   pl->cs       = cs;
   pl->lds      = lds;
   pl->blocks   = blocks;
   pl->MB       = MB;
   pl->NBW      = NBW;
   pl->bf       = bf;
   return PVHIP_OK;
}

size_t cgemm_scratch_bytes(const PostParams &p, const CGemmPlan &pl) {
   return pl.cs.splits > 1 ? (size_t)pl.cs.splits * p.post.nbatch * p.post.numRes() * sizeof(float) : 0;
}

} // namespace

int gather_faults(unsigned *mask) {
   unsigned v = 0, zero = 0;
   PVHIP_TRY(hipMemcpyFromSymbol(&v, HIP_SYMBOL(g_gather_fault), sizeof v, 0, hipMemcpyDeviceToHost));
   if (v != 0) {
      PVHIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(g_gather_fault), &zero, sizeof zero, 0, hipMemcpyHostToDevice));
   }
   *mask = v;
   return PVHIP_OK;
}

} // namespace pvhip

using namespace pvhip;

extern "C" {

size_t pvhip_recv_post_scratch_bytes(const pvhip_recv_post_args *args) {
   if (!args || (args->arith != PVHIP_ARITH_BF16 && args->arith != PVHIP_ARITH_FP32)) {
      return 0;
   }
   // the size depends on the geometry and the form only: data pointers may be
   // NULL here, and stand-ins (never dereferenced) let the checks pass
   static float unused;
   pvhip_recv_post_args a = *args;
   a.pre_data             = a.pre_data ? a.pre_data : &unused;
   a.weights              = a.weights ? a.weights : &unused;
   a.weights_lp           = a.weights_lp ? a.weights_lp : &unused;
   a.post_gsyn            = a.post_gsyn ? a.post_gsyn : &unused;
   a.fused_layer          = nullptr; // writes V/A instead of GSyn: the same scratch
   PostParams p;
   if (build_params(a, &p) != PVHIP_OK || (p.nKx == 1 && p.nKy == 1)) {
      return 0;
   }
   CGemmPlan pl;
   if (cgemm_plan(&a, p, &pl) != PVHIP_OK) {
      return 0;
   }
   return cgemm_scratch_bytes(p, pl);
}

size_t pvhip_recv_post_active_scratch_bytes(const pvhip_recv_post_args *args) {
   if (!args || !args->fused_update || !loc_valid(args->fused_update->loc)) {
      return 0;
   }
   const pvhip_loc &l = args->fused_update->loc;
   return std::max(segment_scratch_bytes(l), compact_scratch_bytes(geo(l)));
}

int pvhip_recv_post_bf16_weights(
      const float *weights, void *weights_bf16, int num_kernels, int patch_size, pvhip_stream stream) {
   PVHIP_REQUIRE(weights && weights_bf16 && num_kernels > 0 && patch_size > 0, "bf16 weights: bad argument");
   int Kpad = (patch_size + 31) / 32 * 32;
   size_t n = (size_t)num_kernels * Kpad;
   hipLaunchKernelGGL(weights_to_bf16_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
         (hipStream_t)stream, weights, (__bf16 *)weights_bf16, num_kernels, patch_size, Kpad);
   PVHIP_LAUNCHED("weights_to_bf16");
   return PVHIP_OK;
}

size_t pvhip_recv_post_bf16_weights_bytes(int num_kernels, int patch_size) {
   return (size_t)num_kernels * ((patch_size + 31) / 32 * 32) * 2;
}

int pvhip_recv_post_fp8_weights(
      const float *weights, void *weights_fp8, int num_kernels, int patch_size, pvhip_stream stream) {
   PVHIP_REQUIRE(weights && weights_fp8 && num_kernels > 0 && patch_size > 0, "fp8 weights: bad argument");
   int Kpad     = (patch_size + 31) / 32 * 32;
   uint8_t *q   = static_cast<uint8_t *>(weights_fp8);
   float *scale = reinterpret_cast<float *>(q + (size_t)num_kernels * Kpad);
   hipLaunchKernelGGL(weights_to_fp8_kernel, dim3((unsigned)num_kernels), dim3(256), 0, (hipStream_t)stream,
         weights, q, scale, patch_size, Kpad);
   PVHIP_LAUNCHED("weights_to_fp8");
   return PVHIP_OK;
}

size_t pvhip_recv_post_fp8_weights_bytes(int num_kernels, int patch_size) {
   // bytes [kernels][Kpad] (Kpad a multiple of 32, so the scales stay aligned), then float scales
   return (size_t)num_kernels * ((patch_size + 31) / 32 * 32) + (size_t)num_kernels * sizeof(float);
}

int pvhip_recv_post(const pvhip_recv_post_args *args, pvhip_stream stream) {
   PVHIP_REQUIRE(args != nullptr, "pvhip_recv_post: NULL args");
   PostParams p;
   int st = build_params(*args, &p);
   if (st) {
      return st;
   }
   hipStream_t s                = (hipStream_t)stream;
   const pvhip_active_out *aout = args->active_out;
   if (aout) {
      PVHIP_REQUIRE(p.fused, "recv_post: active_out needs fused_update");
      PVHIP_REQUIRE(args->fused_update->A != nullptr, "recv_post: active_out needs the fused update's A");
      PVHIP_REQUIRE(aout->entries && aout->num_active && aout->scratch,
            "recv_post: active_out needs entries, num_active and scratch");
      PVHIP_REQUIRE(aout->scratch_bytes >= pvhip_recv_post_active_scratch_bytes(args),
            "recv_post: active_out scratch of %zu bytes, %zu needed", aout->scratch_bytes,
            pvhip_recv_post_active_scratch_bytes(args));
   }
   st = recv_post_launch(args, p, s);
   if (st || !aout) {
      return st;
   }
   const pvhip_lca_update_args &u = *args->fused_update;
   if (p.emit) {
      return active_from_segments(geo(u.loc), p.segCnt, p.stg, u.loc.nf / MF, MF, aout->entries,
            aout->num_active, aout->pos_start, s);
   }
   return pvhip_active_compact(&u.loc, u.A, aout->entries, aout->num_active, aout->pos_start,
         static_cast<uint32_t *>(aout->scratch), stream);
}

} // extern "C"

namespace pvhip {
namespace {

int recv_post_launch(const pvhip_recv_post_args *args, PostParams &p, hipStream_t s) {
   int const arith   = args->arith;
   bool const single = p.nKx == 1 && p.nKy == 1; // one data patch per feature
   PVHIP_REQUIRE(p.post.nbatch <= 65535, "recv_post: nbatch too large");
   // Target <= source along both axes: the window starts advance by the
   // source/target ratio, so a tile of T positions spans (T-1)*ratio + patch.
   int const sx = p.pre.nx > p.post.nx ? p.pre.nx / p.post.nx : 1;
   int const sy = p.pre.ny > p.post.ny ? p.pre.ny / p.post.ny : 1;

   if ((arith == PVHIP_ARITH_BF16 || arith == PVHIP_ARITH_FP32) && !single) {
      // Several data patches per feature: channel GEMM with inverse tables.
      CGemmPlan pl;
      int st = cgemm_plan(args, p, &pl);
      if (st) {
         return st;
      }
      if (pl.cs.splits > 1 && (!args->scratch || args->scratch_bytes < cgemm_scratch_bytes(p, pl))) {
         pl.cs.splits = 1; // no room for the partials: one range of channels
      }
      pl.cs.part = pl.cs.splits > 1 ? static_cast<float *>(args->scratch) : nullptr;
      pl.cs.nOut = (size_t)p.post.nbatch * p.post.numRes();
      const void *Wq = pl.bf ? args->weights_lp : (const void *)args->weights;
      dim3 grid((unsigned)pl.blocks, (unsigned)pl.cs.splits);
      bool const srcb = pl.bf && p.preB16 != nullptr;
      if (srcb) {
         PVHIP_REQUIRE(p.pre.nf % 32 == 0 && reinterpret_cast<uintptr_t>(p.preB16) % 16 == 0,
               "recv_post: pre_data_bf16c needs pre nf %% 32 == 0 and a 16-byte aligned buffer");
      }
#define PVHIP_CGEMM3(T, MBV, NB, SB)                                                              \
   hipLaunchKernelGGL((recv_post_cgemm_kernel<T, MBV, NB, SB>), grid, dim3(256), pl.lds, s, p,    \
         reinterpret_cast<const T *>(Wq), args->inverse_x, args->inverse_y, pl.cs)
#define PVHIP_CGEMM(T, MBV, SB)                                                                   \
   do {                                                                                          \
      if (pl.NBW == 1) PVHIP_CGEMM3(T, MBV, 1, SB);                                              \
      else if (pl.NBW == 2) PVHIP_CGEMM3(T, MBV, 2, SB);                                         \
      else PVHIP_CGEMM3(T, MBV, 4, SB);                                                          \
   } while (0)
      if (srcb) {
         if (pl.MB == 1) PVHIP_CGEMM(__bf16, 1, true);
         else if (pl.MB == 2) PVHIP_CGEMM(__bf16, 2, true);
         else PVHIP_CGEMM(__bf16, 4, true);
      }
      else if (pl.bf) {
         if (pl.MB == 1) PVHIP_CGEMM(__bf16, 1, false);
         else if (pl.MB == 2) PVHIP_CGEMM(__bf16, 2, false);
         else PVHIP_CGEMM(__bf16, 4, false);
      }
      else {
         if (pl.MB == 1) PVHIP_CGEMM(float, 1, false);
         else if (pl.MB == 2) PVHIP_CGEMM(float, 2, false);
         else PVHIP_CGEMM(float, 4, false);
      }
#undef PVHIP_CGEMM3
#undef PVHIP_CGEMM
      PVHIP_LAUNCHED("recv_post_cgemm");
      if (pl.cs.splits > 1 && p.layer) {
         dim3 rg((unsigned)((p.post.numRes() + 1023) / 1024), (unsigned)p.post.nbatch);
         hipLaunchKernelGGL(cgemm_reduce_layer_kernel, rg, dim3(256), 0, s, p, pl.cs.part, pl.cs.splits, pl.cs.nOut);
         PVHIP_LAUNCHED("recv_post_cgemm_reduce_layer");
      }
      else if (pl.cs.splits > 1) {
         size_t n     = pl.cs.nOut;
         bool aligned = (reinterpret_cast<uintptr_t>(p.gsyn) % 16 == 0)
                        && (reinterpret_cast<uintptr_t>(pl.cs.part) % 16 == 0);
         if (n % 4 == 0 && aligned) {
            hipLaunchKernelGGL(cgemm_reduce_kernel<true>, dim3((unsigned)((n / 4 + 255) / 256)), dim3(256), 0, s,
                  pl.cs.part, pl.cs.splits, n, p.gsyn, p.accumulate, p.dtf);
         }
         else {
            hipLaunchKernelGGL(cgemm_reduce_kernel<false>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                  pl.cs.part, pl.cs.splits, n, p.gsyn, p.accumulate, p.dtf);
         }
         PVHIP_LAUNCHED("recv_post_cgemm_reduce");
      }
      return PVHIP_OK;
   }

   PVHIP_REQUIRE(arith != PVHIP_ARITH_FP8 || single,
         "recv_post: the FP8 form needs one data patch per feature (the channel GEMM has BF16 and FP32 forms)");
   PVHIP_REQUIRE(!p.layer, "recv_post: fused_layer is supported by the channel GEMM (BF16 and FP32 forms with "
                           "several data patches per feature) only");
   // FP32 on the matrix cores: the gather kernel of the low-precision forms
   // with FP32 operands, for patches of a multiple of 32 values (the FP32
   // weights are then their own padded copy) whose tile fits the LDS.
   bool mfma32 = false;
   if (arith == PVHIP_ARITH_FP32) {
      int const K          = p.nxp * p.nfp * p.nyp;
      int const srcColsMax = (MTX - 1) * sx + p.nxp;
      int const srcRowsMax = (MTY - 1) * sy + p.nyp;
      int const pitch      = (srcColsMax * p.nfp + 7) / 8 * 8;
      size_t const lds     = std::max(((size_t)srcRowsMax * pitch * 4 + 3) / 4 * 4 + (size_t)K * sizeof(int),
                                      (size_t)MTX * MTY * MOP * sizeof(float));
      mfma32 = K % 32 == 0 && lds <= 80 * 1024; // two workgroups per CU at least
   }
   if (arith == PVHIP_ARITH_BF16 || arith == PVHIP_ARITH_FP8 || mfma32) {
      bool const fp8 = arith == PVHIP_ARITH_FP8;
      PVHIP_REQUIRE(mfma32 || args->weights_lp != nullptr,
            "recv_post: %s needs weights_lp (pvhip_recv_post_%s_weights)", fp8 ? "FP8" : "BF16",
            fp8 ? "fp8" : "bf16");
      int const rowK = p.nxp * p.nfp;
      int const K    = rowK * p.nyp;
      int const Kpad = (K + 31) / 32 * 32;
      int const srcColsMax = (MTX - 1) * sx + p.nxp;
      int const srcRowsMax = (MTY - 1) * sy + p.nyp;
      int const pitch      = (srcColsMax * p.nfp + 7) / 8 * 8;
      size_t const esz = fp8 ? 1 : (mfma32 ? 4 : 2);
      size_t lds = ((size_t)srcRowsMax * pitch * esz + 3) / 4 * 4 + (size_t)Kpad * sizeof(int);
      lds        = std::max(lds, (size_t)MTX * MTY * MOP * sizeof(float)); // epilogue tile
      PVHIP_REQUIRE(lds <= 160 * 1024, "recv_post: MFMA tile needs %zu bytes of LDS", lds);
      // alignment class of a fragment's LDS address: rows are 8-aligned, window
      // starts differ by multiples of sx*nfp, fragments start at multiples of 8
      int step  = sx * p.nfp;
      int align = step % 8 == 0 ? 8 : (step % 4 == 0 ? 4 : (step % 2 == 0 ? 2 : 1));
      if (rowK % 8 != 0) {
         align = 0; // a fragment may straddle patch rows: per-element offsets
      }
      int const tilesX = (p.post.nx + MTX - 1) / MTX;
      int const tilesY = (p.post.ny + MTY - 1) / MTY;
      dim3 grid(tilesX * tilesY, (p.post.nf + MF - 1) / MF, p.post.nbatch);
      if (fp8) {
         float const ss = args->lp_source_scale;
         PVHIP_REQUIRE(ss >= 0.0f && ss == ss && ss < 3.0e38f, "recv_post: bad lp_source_scale %g", ss);
         p.lpSrcScale = ss > 0.0f ? ss : 1.0f;
         p.lpWScale   = reinterpret_cast<const float *>(static_cast<const uint8_t *>(args->weights_lp)
                                                      + (size_t)(p.nKx * p.nKy * p.post.nf) * Kpad);
      }
      int epi = !p.fused ? 0 : (p.lca.transfer.mode == PVHIP_TRANSFER_VTHRESH ? 1 : 2);
      const pvhip_active_out *aout = args->active_out;
      if (aout && aout->halo_is_zero && epi == 1 && p.post.nf % MF == 0 && p.post.nx % MTX == 0
          && p.post.ny % MTY == 0) {
         // every tile is full: the epilogue emits the active list
         p.emit   = 1;
         p.stg    = static_cast<pvhip_sparse_entry *>(aout->scratch);
         p.segCnt = reinterpret_cast<uint32_t *>(
               static_cast<char *>(aout->scratch) + segment_stage_bytes(args->fused_update->loc));
         epi      = 3;
      }
      // the persistent kernel for full tiles of the fused VTHRESH update
      // (measured faster than one tile per workgroup: FP32 by keeping its
      // weight fragments in registers across tiles, BF16 and FP8 by
      // overlapping the next tile's loads with the current tile's MFMAs and
      // stores)
      int const KS       = K / 32;
      bool const persist = epi == 1 && align >= 2 && K % 32 == 0 && K == Kpad && (KS == 6 || KS == 8)
                           && p.post.nx % MTX == 0 && p.post.ny % MTY == 0 && p.post.nf % MF == 0
                           && (long long)srcRowsMax * srcColsMax * p.nfp <= 256LL * PSREG;
      if (persist) {
         int const items = tilesX * tilesY * p.post.nbatch;
         int const fblk  = p.post.nf / MF;
         // about 512 workgroups (two per CU), each walking ipw tiles
         int const ipw  = std::max(1, (int)(((long long)items * fblk + 511) / 512));
         dim3 pgrid((unsigned)((items + ipw - 1) / ipw), (unsigned)fblk);
         size_t const plds = std::max(((size_t)srcRowsMax * pitch * esz + 3) / 4 * 4,
                                      (size_t)MTX * MTY * MOP * sizeof(float));
#define PVHIP_PERSIST(AL, KSV)                                                                   \
   do {                                                                                          \
      if (mfma32) {                                                                              \
         hipLaunchKernelGGL((recv_post_mfma_persist_kernel<AL, GatherF32, KSV>), pgrid, dim3(256), plds, s, \
               p, args->weights, srcRowsMax, srcColsMax, items, ipw);                            \
      }                                                                                          \
      else if (fp8) {                                                                            \
         hipLaunchKernelGGL((recv_post_mfma_persist_kernel<AL, GatherFP8, KSV>), pgrid, dim3(256), plds, s, \
               p, static_cast<const uint8_t *>(args->weights_lp), srcRowsMax, srcColsMax, items, ipw); \
      }                                                                                          \
      else {                                                                                     \
         hipLaunchKernelGGL((recv_post_mfma_persist_kernel<AL, GatherBF16, KSV>), pgrid, dim3(256), plds, s, \
               p, static_cast<const __bf16 *>(args->weights_lp), srcRowsMax, srcColsMax, items, ipw); \
      }                                                                                          \
   } while (0)
#define PVHIP_PERSIST_KS(AL)                                                                     \
   do {                                                                                          \
      if (KS == 8) PVHIP_PERSIST(AL, 8);                                                         \
      else PVHIP_PERSIST(AL, 6);                                                                 \
   } while (0)
         if (align == 8) PVHIP_PERSIST_KS(8);
         else if (align == 4) PVHIP_PERSIST_KS(4);
         else PVHIP_PERSIST_KS(2);
#undef PVHIP_PERSIST_KS
#undef PVHIP_PERSIST
         PVHIP_LAUNCHED("recv_post_mfma_persist");
         return PVHIP_OK;
      }
#define PVHIP_MFMA(AL, EP)                                                                       \
   do {                                                                                          \
      if (mfma32) {                                                                              \
         hipLaunchKernelGGL((recv_post_mfma_bf16_kernel<AL, EP, GatherF32>), grid, dim3(256), lds, \
               s, p, args->weights, srcRowsMax, srcColsMax, Kpad);                               \
      }                                                                                          \
      else if (fp8) {                                                                            \
         hipLaunchKernelGGL((recv_post_mfma_bf16_kernel<AL, EP, GatherFP8>), grid, dim3(256), lds, \
               s, p, static_cast<const uint8_t *>(args->weights_lp), srcRowsMax, srcColsMax, Kpad); \
      }                                                                                          \
      else {                                                                                     \
         hipLaunchKernelGGL((recv_post_mfma_bf16_kernel<AL, EP, GatherBF16>), grid, dim3(256), lds, \
               s, p, static_cast<const __bf16 *>(args->weights_lp), srcRowsMax, srcColsMax, Kpad); \
      }                                                                                          \
   } while (0)
#define PVHIP_MFMA_EPI(AL)                                                                       \
   do {                                                                                          \
      if (epi == 0) PVHIP_MFMA(AL, 0);                                                           \
      else if (epi == 1) PVHIP_MFMA(AL, 1);                                                      \
      else if (epi == 3) PVHIP_MFMA(AL, 3);                                                      \
      else PVHIP_MFMA(AL, 2);                                                                    \
   } while (0)
      if (align == 8) PVHIP_MFMA_EPI(8);
      else if (align == 4) PVHIP_MFMA_EPI(4);
      else if (align == 2) PVHIP_MFMA_EPI(2);
      else if (align == 1) PVHIP_MFMA_EPI(1);
      else PVHIP_MFMA_EPI(0);
#undef PVHIP_MFMA_EPI
#undef PVHIP_MFMA
      PVHIP_LAUNCHED(fp8 ? "recv_post_mfma_fp8" : (mfma32 ? "recv_post_mfma_f32" : "recv_post_mfma_bf16"));
      return PVHIP_OK;
   }
   if ((arith == PVHIP_ARITH_EXACT || arith == PVHIP_ARITH_FP32) && single) {
      int const K          = p.nxp * p.nfp;
      int const srcColsMax = (TX - 1) * sx + p.nxp;
      int const srcRowsMax = (TY - 1) * sy + p.nyp;
      int const tilesX     = (p.post.nx + TX - 1) / TX;
      int const tilesY     = (p.post.ny + TY - 1) / TY;
      for (int FB : {64, 32}) {
         size_t lds = ((size_t)srcRowsMax * srcColsMax * p.nfp + (size_t)K * FB) * sizeof(float);
         if (lds > 160 * 1024) {
            continue;
         }
         dim3 grid(tilesX * tilesY, (p.post.nf + FB - 1) / FB, p.post.nbatch);
         if (FB == 64) {
            if (arith == PVHIP_ARITH_EXACT) {
               hipLaunchKernelGGL((recv_post_tiled_kernel<64, true>), grid, dim3(256), lds, s, p, srcRowsMax, srcColsMax);
            }
            else {
               hipLaunchKernelGGL((recv_post_tiled_kernel<64, false>), grid, dim3(256), lds, s, p, srcRowsMax, srcColsMax);
            }
         }
         else {
            if (arith == PVHIP_ARITH_EXACT) {
               hipLaunchKernelGGL((recv_post_tiled_kernel<32, true>), grid, dim3(256), lds, s, p, srcRowsMax, srcColsMax);
            }
            else {
               hipLaunchKernelGGL((recv_post_tiled_kernel<32, false>), grid, dim3(256), lds, s, p, srcRowsMax, srcColsMax);
            }
         }
         PVHIP_LAUNCHED("recv_post_tiled");
         return PVHIP_OK;
      }
   }
   if (arith == PVHIP_ARITH_EXACT || arith == PVHIP_ARITH_FP32) {
      size_t n = (size_t)p.post.nbatch * p.post.numRes();
      hipLaunchKernelGGL(recv_post_direct_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p);
      PVHIP_LAUNCHED("recv_post_direct");
      return PVHIP_OK;
   }
   return set_error(PVHIP_ERR_UNSUPPORTED, "recv_post: arith %d not supported for this geometry", arith);
}

} // namespace
} // namespace pvhip
