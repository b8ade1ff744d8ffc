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
};

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
   // outside the extended source buffer, leave this tile unwritten.
   if (srcCols > srcColsMax || srcRows > srcRowsMax || srcX0 < 0 || srcY0 < 0
       || srcX0 + srcCols > p.pre.nxExt() || srcY0 + srcRows > p.pre.nyExt()) {
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
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Workgroup: 64 target positions (8x8) x MF features of one batch element,
// 4 waves; wave w owns features [w*MF/4, (w+1)*MF/4) for all 64 positions.
// Fragment layouts of v_mfma_f32_16x16x32_bf16 (lane l, r = l%16, g = l/16):
//   A[r][8g..8g+7], B[8g..8g+7][r], D[4g..4g+3][r].
constexpr int MF       = 128;          // features per workgroup
constexpr int WAVE_F   = MF / 4;       // features per wave (2 blocks of 16)
constexpr int FBLK     = WAVE_F / 16;  // 16-row blocks per wave
constexpr int PBLK     = 4;            // 16-position blocks (64 positions)

__device__ __forceinline__ __bf16 to_bf16(float x) { return (__bf16)x; }

__global__ void __launch_bounds__(256) recv_post_mfma_bf16_kernel(
      PostParams p, const __bf16 *Wb, int srcRowsMax, int srcColsMax, int Kpad) {
   extern __shared__ float smem[];
   int const K       = p.nxp * p.nfp * p.nyp; // full patch
   int const rowK    = p.nxp * p.nfp;
   int const pitch   = srcColsMax * p.nfp;
   float *src        = smem;                                        // fp32 source tile
   int *koff         = reinterpret_cast<int *>(smem + srcRowsMax * pitch); // [Kpad]
   int const tilesX  = (p.post.nx + TX - 1) / TX;
   int const tile    = blockIdx.x;
   int const x0      = (tile % tilesX) * TX;
   int const y0      = (tile / tilesX) * TY;
   int const fBase   = blockIdx.y * MF;
   int const b       = blockIdx.z;
   int const xl      = min(x0 + TX, p.post.nx) - 1;
   int const yl      = min(y0 + TY, p.post.ny) - 1;
   int const srcX0   = p.startX[x0];
   int const srcY0   = p.startY[y0];
   int const srcCols = p.startX[xl] - srcX0 + p.nxp;
   int const srcRows = p.startY[yl] - srcY0 + p.nyp;
   int const tid     = threadIdx.x;
   if (srcCols > srcColsMax || srcRows > srcRowsMax || srcX0 < 0 || srcY0 < 0
       || srcX0 + srcCols > p.pre.nxExt() || srcY0 + srcRows > p.pre.nyExt()) {
      return;
   }
   int const preRow  = p.pre.nxExt() * p.pre.nf;
   float const *preB = p.pre_data + (size_t)b * p.pre.numExt();

   int const rowElems = srcCols * p.nfp;
   for (int i = tid; i < srcRows * rowElems; i += 256) {
      int r = i / rowElems;
      int c = i % rowElems;
      src[r * pitch + c] = preB[(size_t)(srcY0 + r) * preRow + (size_t)srcX0 * p.nfp + c];
   }
   for (int k = tid; k < Kpad; k += 256) {
      koff[k] = k < K ? (k / rowK) * pitch + (k % rowK) : -1;
   }
   __syncthreads();

   int const lane = tid % 64;
   int const wave = tid / 64;
   int const r    = lane % 16;
   int const g    = lane / 16;

   // Per-lane position base offsets for the 4 position blocks (column r).
   int pbase[PBLK];
#pragma unroll
   for (int pb = 0; pb < PBLK; pb++) {
      int pj    = pb * 16 + r;
      int x     = x0 + pj % TX;
      int y     = y0 + pj / TX;
      bool ok   = x < p.post.nx && y < p.post.ny;
      pbase[pb] = ok ? (p.startY[y] - srcY0) * pitch + (p.startX[x] - srcX0) * p.nfp : 0;
   }

   f32x4 acc[FBLK][PBLK];
#pragma unroll
   for (int fb = 0; fb < FBLK; fb++) {
#pragma unroll
      for (int pb = 0; pb < PBLK; pb++) {
         acc[fb][pb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
      }
   }
   int const fWave = fBase + wave * WAVE_F;
   for (int k0 = 0; k0 < Kpad; k0 += 32) {
      int const kk = k0 + 8 * g;
      bf16x8 afrag[FBLK];
#pragma unroll
      for (int fb = 0; fb < FBLK; fb++) {
         int f = fWave + fb * 16 + r;
         if (f < p.post.nf) {
            afrag[fb] = *reinterpret_cast<const bf16x8 *>(Wb + (size_t)f * Kpad + kk);
         }
         else {
            afrag[fb] = bf16x8{};
         }
      }
      int ko[8];
#pragma unroll
      for (int e = 0; e < 8; e++) {
         ko[e] = koff[kk + e];
      }
#pragma unroll
      for (int pb = 0; pb < PBLK; pb++) {
         bf16x8 bfrag;
#pragma unroll
         for (int e = 0; e < 8; e++) {
            bfrag[e] = ko[e] >= 0 ? to_bf16(src[pbase[pb] + ko[e]]) : to_bf16(0.0f);
         }
#pragma unroll
         for (int fb = 0; fb < FBLK; fb++) {
            acc[fb][pb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afrag[fb], bfrag, acc[fb][pb], 0, 0, 0);
         }
      }
   }
   // Epilogue: lane holds D[4g+i][r] = feature fWave+fb*16+4g+i at position pb*16+r.
#pragma unroll
   for (int pb = 0; pb < PBLK; pb++) {
      int pj = pb * 16 + r;
      int x  = x0 + pj % TX;
      int y  = y0 + pj / TX;
      if (x >= p.post.nx || y >= p.post.ny) {
         continue;
      }
      int kpos = (y * p.post.nx + x) * p.post.nf;
#pragma unroll
      for (int fb = 0; fb < FBLK; fb++) {
#pragma unroll
         for (int i = 0; i < 4; i++) {
            int f = fWave + fb * 16 + 4 * g + i;
            if (f < p.post.nf) {
               float base = p.accumulate ? p.gsyn[(size_t)b * p.post.numRes() + kpos + f] : 0.0f;
               float G    = base + p.dtf * acc[fb][pb][i];
               post_epilogue(p, b, kpos + f, G);
            }
         }
      }
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

int build_params(const pvhip_recv_post_args &a, PostParams *p) {
   PVHIP_REQUIRE(loc_valid(a.pre) && loc_valid(a.post), "recv_post: invalid loc");
   PVHIP_REQUIRE(a.pre.nbatch == a.post.nbatch, "recv_post: batch sizes differ");
   PVHIP_REQUIRE(a.nxp > 0 && a.nyp > 0 && a.nfp == a.pre.nf, "recv_post: patch nfp must equal pre nf");
   PVHIP_REQUIRE(a.num_kernels_x > 0 && a.num_kernels_y > 0, "recv_post: bad kernel counts");
   PVHIP_REQUIRE(a.start_x && a.start_y && a.kernel_x && a.kernel_y, "recv_post: NULL geometry table");
   PVHIP_REQUIRE(a.pre_data && a.weights, "recv_post: NULL pre_data or weights");
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
      PVHIP_REQUIRE(a.post_gsyn != nullptr, "recv_post: NULL post_gsyn");
   }
   return PVHIP_OK;
}

} // namespace
} // namespace pvhip

using namespace pvhip;

extern "C" {

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

int pvhip_recv_post(const pvhip_recv_post_args *args, pvhip_stream stream) {
   PVHIP_REQUIRE(args != nullptr, "pvhip_recv_post: NULL args");
   PostParams p;
   int st = build_params(*args, &p);
   if (st) {
      return st;
   }
   hipStream_t s = (hipStream_t)stream;
   int const arith = args->arith;
   bool const oneKernelPerFeature = p.nKx == 1 && p.nKy == 1;
   // Source tile bound: the tables are monotone non-decreasing (calcPatchStartInPost),
   // so a tile of TX columns spans at most the largest start difference + nxp.
   int maxDX = 0, maxDY = 0;
   if (oneKernelPerFeature) {
      // The tables live on the device; the host side of the delivery passes
      // their spans through the stride rule: span <= (T-1)*ceil(stride) + nxp.
      int sx = p.pre.nx > p.post.nx ? p.pre.nx / p.post.nx : 1;
      int sy = p.pre.ny > p.post.ny ? p.pre.ny / p.post.ny : 1;
      maxDX  = (TX - 1) * sx;
      maxDY  = (TY - 1) * sy;
   }
   int const srcColsMax = maxDX + p.nxp;
   int const srcRowsMax = maxDY + p.nyp;
   int const tilesX     = (p.post.nx + TX - 1) / TX;
   int const tilesY     = (p.post.ny + TY - 1) / TY;
   PVHIP_REQUIRE(p.post.nbatch <= 65535, "recv_post: nbatch too large");

   if (arith == PVHIP_ARITH_BF16 && oneKernelPerFeature) {
      PVHIP_REQUIRE(args->weights_lp != nullptr, "recv_post: BF16 needs weights_lp (pvhip_recv_post_bf16_weights)");
      int const K    = p.nxp * p.nyp * p.nfp;
      int const Kpad = (K + 31) / 32 * 32;
      size_t lds = (size_t)srcRowsMax * srcColsMax * p.nfp * sizeof(float) + (size_t)Kpad * sizeof(int);
      PVHIP_REQUIRE(lds <= 160 * 1024, "recv_post: BF16 tile needs %zu bytes of LDS", lds);
      dim3 grid(tilesX * tilesY, (p.post.nf + MF - 1) / MF, p.post.nbatch);
      hipLaunchKernelGGL(recv_post_mfma_bf16_kernel, grid, dim3(256), lds, s, p,
            (const __bf16 *)args->weights_lp, srcRowsMax, srcColsMax, Kpad);
      PVHIP_LAUNCHED("recv_post_mfma_bf16");
      return PVHIP_OK;
   }
   if ((arith == PVHIP_ARITH_EXACT || arith == PVHIP_ARITH_FP32) && oneKernelPerFeature) {
      int const K = p.nxp * p.nfp;
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

} // extern "C"
