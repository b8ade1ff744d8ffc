// HebbianUpdater dW for shared weights (HebbianUpdater.cpp:476-627).
//
// EXACT form: one thread per dW element; a wave covers 64 consecutive items
// of one data patch, so the presynaptic activity it walks is wave-uniform.
// For a dW element the reference accumulates (HebbianUpdater.cpp:501-528):
//   batch ascending, then the presynaptic extended rows and columns of that
//   data patch ascending, skipping a_pre == 0, adding (dWMax*a_pre)*a_post.
// Counts (normalizeDw) increment once per contributing presynaptic neuron.
#include "pvhip_internal.hpp"

#include <algorithm>

typedef __bf16 bf16x8v __attribute__((ext_vector_type(8)));
typedef float f32x8v __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace pvhip {
namespace {

struct HebbParams {
   Geo pre, post;
   int nxp, nyp, nfp;
   int nKx, nKy;
   const int32_t *dimX, *dimY, *pstartX, *pstartY, *apostX, *apostY;
   float dwMax;
   const float *preData, *postData;
   float *dw;
   int64_t *counts;
};

__global__ void __launch_bounds__(256) hebbian_exact_kernel(HebbParams p) {
   int const patchSize  = p.nxp * p.nyp * p.nfp;
   int const numKernels = p.nKx * p.nKy * p.pre.nf;
   size_t const n       = (size_t)numKernels * patchSize;
   size_t const index   = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (index >= n) {
      return;
   }
   int const d    = (int)(index / patchSize);
   int const item = (int)(index % patchSize);
   int const f    = d % p.pre.nf;
   int const dx   = (d / p.pre.nf) % p.nKx;
   int const dy   = d / (p.pre.nf * p.nKx);
   int const F    = item % p.nfp;
   int const wx   = (item / p.nfp) % p.nxp;
   int const wy   = item / (p.nfp * p.nxp);
   int const nxExt = p.pre.nxExt();
   int const nyExt = p.pre.nyExt();
   int const nxExtPost = p.post.nxExt();
   // First extended column/row whose data index is (dx, dy):
   // (xe - lt) mod nKx == dx  (Weights.cpp:212-216).
   int x0 = (dx + p.pre.lt) % p.nKx;
   int y0 = (dy + p.pre.up) % p.nKy;
   float acc   = p.dw[index];
   int64_t cnt = p.counts ? p.counts[index] : 0;
   for (int b = 0; b < p.pre.nbatch; b++) {
      float const *preB  = p.preData + (size_t)b * p.pre.numExt();
      float const *postB = p.postData + (size_t)b * p.post.numExt();
      for (int ye = y0; ye < nyExt; ye += p.nKy) {
         int const ry = wy - p.pstartY[ye];
         bool const rowOk = ry >= 0 && ry < p.dimY[ye];
         for (int xe = x0; xe < nxExt; xe += p.nKx) {
            float const preact = preB[((size_t)ye * nxExt + xe) * p.pre.nf + f];
            if (preact == 0.0f) {
               continue;
            }
            int const rx = wx - p.pstartX[xe];
            if (!rowOk || rx < 0 || rx >= p.dimX[xe]) {
               continue;
            }
            float const apost =
                  postB[((size_t)(p.apostY[ye] + ry) * nxExtPost + p.apostX[xe] + rx) * p.post.nf + F];
            cnt++;
            float t = p.dwMax * preact;
            t       = t * apost;
            acc     = acc + t;
         }
      }
   }
   p.dw[index] = acc;
   if (p.counts) {
      p.counts[index] = cnt;
   }
}

// ------------------------------------------------------------ MFMA correlation
//
// dW[f][item] = dWMax * sum_{b, ye, xe} a_pre[b][ye][xe][f] * P[b](window(ye, xe) + item)
// for connections with one data patch per feature: a GEMM with M = features,
// N = patch items (ky, kx, f'), K = batch x pre extended positions.  P is the
// post activity with every value outside the post restricted region read as
// 0, which is the reference's shrunken patch (pvhost_hebbian_windows checks
// the geometry).  Counts (normalizeDw) are the same correlation of the
// indicators a_pre != 0 and "inside the restricted region".
//
// Workgroup: 64 features x 256 items, 4 waves; wave w owns items
// 64w .. 64w+63 (4 MFMA column blocks) for all 64 features (4 row blocks).
// K runs over pre extended rows (b, ye); per row the 64-feature activity
// slab is staged transposed in LDS ([f][xe], T) and the post rows of the
// windows are staged phase-split ([ky][f'][phase][xq], T) so that the values
// for 8 consecutive pre columns xe of one item are 8 consecutive LDS
// elements.  A lane's 8 K values are pre columns 8g..8g+7 of a 32-column
// step; the FP32 form runs eight v_mfma_f32_16x16x4f32, MFMA j taking column
// 8g+j (the same permutation of K for both operands).  BF16 rows are stored
// twice, the second copy shifted by one element, so a fragment starting at
// any element is read as four aligned 32-bit words.
constexpr int HMT = 64;  // features per workgroup
constexpr int HNT = 256; // items per workgroup
constexpr int HNB = HNT / 4 / 16; // 16-item blocks per wave

struct HebbMfma {
   Geo pre, post;
   int nxp, nyp, nfp, N, nK;
   int sx, cx0, qx0, sy, cy; // window: post ext x = sx*(xq + qx0) + phase, rows sy*ye + cy + ky
   int kxPad, lq, apitch;     // 32-padded pre ext width, B row length, A row pitch (elements)
   int rowsPerChunk;
   const float *preData, *postData;
   float *part; // [chunks][nK][N]
};

template <typename T>
struct HebbFrag;
template <>
struct HebbFrag<__bf16> {
   typedef bf16x8v type;
};
template <>
struct HebbFrag<float> {
   typedef f32x8v type;
};

template <typename T>
__device__ __forceinline__ T hcvt(float v) {
   return (T)v;
}

// B fragment: 8 consecutive elements of a B row starting at element e.
template <typename T>
__device__ __forceinline__ typename HebbFrag<T>::type hload_b(const T *bl, int copyStride, int e);
template <>
__device__ __forceinline__ bf16x8v hload_b<__bf16>(const __bf16 *bl, int copyStride, int e) {
   // copy 1 holds row[j + 1] at j: an odd start is an even start there
   const __bf16 *p = (e & 1) ? bl + copyStride + e - 1 : bl + e;
   const uint32_t *w = reinterpret_cast<const uint32_t *>(p);
   uint32_t v[4] = {w[0], w[1], w[2], w[3]};
   return __builtin_bit_cast(bf16x8v, v);
}
template <>
__device__ __forceinline__ f32x8v hload_b<float>(const float *bl, int, int e) {
   f32x8v v;
#pragma unroll
   for (int j = 0; j < 8; j++) {
      v[j] = bl[e + j];
   }
   return v;
}

// Staging of a pre row, software-pipelined: the next row's activity values
// and window values are loaded into registers (HAREG and HBR x HBX per
// thread, any rest loaded when stored) while the MFMAs of the current row run.  The window's post rows sit in a
// ring of nyp rows (slot = post row mod nyp), so a pre row following the one
// before in the same image stages only the post rows the window advanced by
// (sy of them), not all nyp.
constexpr int HAREG = 32;
// window values held per lane: HBR B rows per wave (B row = post row, feature,
// phase) x HBX values 64 apart along the row
constexpr int HBR = 4;
constexpr int HBX = 3;

template <typename T, bool COUNT>
__global__ void __launch_bounds__(256) hebbian_mfma_kernel(HebbMfma h) {
   extern __shared__ float smem[];
   int const copies = sizeof(T) == 2 ? 2 : 1;
   int const nbrow  = h.nyp * h.nfp * h.sx; // B rows (ring slot, feature, phase)
   int const cstr   = (nbrow * h.lq + 7) / 8 * 8; // elements per B copy (16-byte multiple)
   T *bl            = reinterpret_cast<T *>(smem);
   T *al            = bl + (size_t)copies * cstr; // [HMT][apitch]
   int const f0     = blockIdx.x * HMT;
   int const n0     = blockIdx.y * HNT;
   int const chunk  = blockIdx.z;
   int const nyExt  = h.pre.nyExt();
   int const nxExt  = h.pre.nxExt();
   int const rows   = h.pre.nbatch * nyExt;
   int const r0     = chunk * h.rowsPerChunk;
   int const r1     = min(r0 + h.rowsPerChunk, rows);
   int const tid    = threadIdx.x;
   int const lane   = tid % 64;
   int const wave   = tid / 64;
   int const r      = lane % 16;
   int const g      = lane / 16;
   int const nxExtP = h.post.nxExt();
   int const rowB   = h.nfp * h.sx * h.lq; // B elements of one post row
   int const aItems = HMT * h.kxPad;

   // this lane's item of each item block: window row ky, feature fp, and the
   // B element offset within a ring slot
   int nky[HNB], nsub[HNB];
#pragma unroll
   for (int nb = 0; nb < HNB; nb++) {
      int n     = min(n0 + (wave * HNB + nb) * 16 + r, h.N - 1);
      int fp    = n % h.nfp;
      int kx    = (n / h.nfp) % h.nxp;
      nky[nb]   = n / (h.nfp * h.nxp);
      int ph    = (h.cx0 + kx) % h.sx;
      nsub[nb]  = (fp * h.sx + ph) * h.lq + (h.cx0 + kx) / h.sx + 8 * g;
   }
   f32x4 acc[4][HNB];
#pragma unroll
   for (int mb = 0; mb < 4; mb++) {
#pragma unroll
      for (int nb = 0; nb < HNB; nb++) {
         acc[mb][nb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
      }
   }
   auto ring = [&](int py) { int m = py % h.nyp; return m < 0 ? m + h.nyp : m; };
   // activity element i (feature m = i % HMT, pre column xe = i / HMT) of pre row `row`
   auto a_val = [&](int row, int i) -> float {
      int const b  = row / nyExt;
      int const ye = row % nyExt;
      int const m  = i % HMT;
      int const xe = i / HMT;
      float v      = 0.0f;
      if (xe < nxExt && f0 + m < h.nK) {
         v = h.preData[(((size_t)b * nyExt + ye) * nxExt + xe) * h.pre.nf + f0 + m];
         if (COUNT) {
            v = v != 0.0f ? 1.0f : 0.0f;
         }
      }
      return v;
   };
   // post rows a pre row stages: [pyA, pyB) (all nyp rows, or the sy new ones
   // when the row before was the previous pre row of the same image)
   auto new_rows = [&](int row, int &pyA, int &pyB) {
      int const ye = row % nyExt;
      int const wy = h.sy * ye + h.cy;
      pyA          = wy;
      pyB          = wy + h.nyp;
      if (row > r0 && ye > 0 && h.sy < h.nyp) {
         pyA = wy + h.nyp - h.sy;
      }
   };
   // element tid + 256*u of the activity slab is feature tid % HMT of pre
   // column tid / HMT + 4*u (256 is a multiple of HMT)
   static_assert(256 % HMT == 0, "activity slab walk assumes HMT divides 256");
   int const am  = tid % HMT;
   int const ax0 = tid / HMT;
   // window value (B row br of the staged post rows [pyA, ..), position xq),
   // and its ring slot; B rows are walked one per wave, lanes along xq, so the
   // row's decomposition is wave-uniform
   auto b_row = [&](int row, int pyA, int br, int &py, int &fp, int &ph) {
      int const pr = h.nfp * h.sx;
      py           = pyA + br / pr;
      fp           = (br % pr) / h.sx;
      ph           = br % h.sx;
   };
   auto b_elem = [&](int b, int py, int fp, int ph, int xq) -> float {
      int const px = h.sx * (xq + h.qx0) + ph;
      bool const in = py >= h.post.up && py < h.post.up + h.post.ny && px >= h.post.lt
                      && px < h.post.lt + h.post.nx;
      return in ? (COUNT ? 1.0f : h.postData[(size_t)b * h.post.numExt() + ((size_t)py * nxExtP + px) * h.post.nf + fp])
                : 0.0f;
   };
   auto b_put = [&](int py, int fp, int ph, int xq, float v) {
      int const idx = ring(py) * rowB + (fp * h.sx + ph) * h.lq + xq;
      T const t     = hcvt<T>(v);
      bl[idx]       = t;
      if (copies == 2 && xq > 0) {
         bl[cstr + idx - 1] = t; // copy 1: row shifted left by one element
      }
   };
   float va[HAREG], vb[HBR][HBX];
   auto fetch = [&](int row) {
      int pyA, pyB;
      new_rows(row, pyA, pyB);
      int const nbr = (pyB - pyA) * h.nfp * h.sx; // B rows to stage
      int const b   = row / nyExt;
#pragma unroll
      for (int j = 0; j < HBR; j++) {
         int const br = wave + 4 * j;
         int py = 0, fp = 0, ph = 0;
         b_row(row, pyA, min(br, max(nbr - 1, 0)), py, fp, ph);
#pragma unroll
         for (int x = 0; x < HBX; x++) {
            int const xq = lane + 64 * x;
            vb[j][x]     = (br < nbr && xq < h.lq) ? b_elem(b, py, fp, ph, xq) : 0.0f;
         }
      }
      float const *base = h.preData + ((size_t)row * nxExt + ax0) * h.pre.nf + f0 + am; // row = b*nyExt + ye
      bool const fok    = f0 + am < h.nK;
#pragma unroll
      for (int u = 0; u < HAREG; u++) {
         int const xe = ax0 + (256 / HMT) * u;
         float v      = 0.0f;
         if (fok && xe < nxExt) {
            v = base[(size_t)(256 / HMT) * u * h.pre.nf];
            if (COUNT) {
               v = v != 0.0f ? 1.0f : 0.0f;
            }
         }
         va[u] = v;
      }
   };
   auto stage = [&](int row) {
#pragma unroll
      for (int u = 0; u < HAREG; u++) {
         int const i = tid + 256 * u;
         if (i < aItems) {
            al[(i % HMT) * h.apitch + i / HMT] = hcvt<T>(va[u]);
         }
      }
      for (int i = tid + 256 * HAREG; i < aItems; i += 256) {
         al[(i % HMT) * h.apitch + i / HMT] = hcvt<T>(a_val(row, i));
      }
      int pyA, pyB;
      new_rows(row, pyA, pyB);
      int const nbr = (pyB - pyA) * h.nfp * h.sx;
      int const b   = row / nyExt;
      for (int br = wave; br < nbr; br += 4) {
         int py, fp, ph;
         b_row(row, pyA, br, py, fp, ph);
         int const j = br / 4; // register slot of this B row, if any
         for (int xq = lane; xq < h.lq; xq += 64) {
            int const x = xq / 64;
            float v;
            if (j < HBR && x < HBX) {
               // the prefetched value (the register index is a compile-time
               // choice: select among the slots)
               v = 0.0f;
#pragma unroll
               for (int jj = 0; jj < HBR; jj++) {
#pragma unroll
                  for (int xx = 0; xx < HBX; xx++) {
                     v = (jj == j && xx == x) ? vb[jj][xx] : v;
                  }
               }
            }
            else {
               v = b_elem(b, py, fp, ph, xq);
            }
            b_put(py, fp, ph, xq, v);
         }
      }
   };
   if (r0 < r1) {
      fetch(r0);
   }
   for (int row = r0; row < r1; row++) {
      __syncthreads(); // the previous row's fragments are consumed
      stage(row);
      __syncthreads();
      if (row + 1 < r1) {
         fetch(row + 1); // in flight during this row's MFMAs
      }
      int const wy = h.sy * (row % nyExt) + h.cy;
      int bbase[HNB];
#pragma unroll
      for (int nb = 0; nb < HNB; nb++) {
         bbase[nb] = ring(wy + nky[nb]) * rowB + nsub[nb];
      }
      for (int x0 = 0; x0 < h.kxPad; x0 += 32) {
         typename HebbFrag<T>::type af[4], bf[HNB];
#pragma unroll
         for (int mb = 0; mb < 4; mb++) {
            T const *ap = al + (mb * 16 + r) * h.apitch + x0 + 8 * g;
            if constexpr (sizeof(T) == 2) {
               af[mb] = *reinterpret_cast<const bf16x8v *>(ap);
            }
            else {
               f32x4 lo = *reinterpret_cast<const f32x4 *>(ap);
               f32x4 hi = *reinterpret_cast<const f32x4 *>(ap + 4);
               af[mb]   = f32x8v{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
            }
         }
#pragma unroll
         for (int nb = 0; nb < HNB; nb++) {
            bf[nb] = hload_b<T>(bl, cstr, bbase[nb] + x0);
         }
#pragma unroll
         for (int mb = 0; mb < 4; mb++) {
#pragma unroll
            for (int nb = 0; nb < HNB; nb++) {
               if constexpr (sizeof(T) == 2) {
                  acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[mb], bf[nb], acc[mb][nb], 0, 0, 0);
               }
               else {
#pragma unroll
                  for (int j = 0; j < 8; j++) {
                     acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[mb][j], bf[nb][j], acc[mb][nb], 0, 0, 0);
                  }
               }
            }
         }
      }
   }
   // D[4g+i][r]: feature f0 + mb*16 + 4g + i, item of lane r in block nb
   float *out = h.part + (size_t)chunk * h.nK * h.N;
#pragma unroll
   for (int mb = 0; mb < 4; mb++) {
#pragma unroll
      for (int nb = 0; nb < HNB; nb++) {
         int const n = n0 + (wave * HNB + nb) * 16 + r;
#pragma unroll
         for (int i = 0; i < 4; i++) {
            int const f = f0 + mb * 16 + 4 * g + i;
            if (f < h.nK && n < h.N) {
               out[(size_t)f * h.N + n] = acc[mb][nb][i];
            }
         }
      }
   }
}

// dw[i] += dWMax * sum_c part[c][i] (chunks in order); counts[i] += sum_c cpart[c][i].
__global__ void __launch_bounds__(256) hebbian_reduce_kernel(
      const float *part, const float *cpart, int chunks, size_t n, float dwMax, float *dw,
      int64_t *counts) {
   size_t const i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (i >= n) {
      return;
   }
   float s = 0.0f;
   for (int c = 0; c < chunks; c++) {
      s = s + part[(size_t)c * n + i];
   }
   dw[i] = dw[i] + dwMax * s;
   if (counts) {
      float cs = 0.0f;
      for (int c = 0; c < chunks; c++) {
         cs = cs + cpart[(size_t)c * n + i];
      }
      counts[i] += (int64_t)cs;
   }
}

struct HebbPlan {
   int chunks, rowsPerChunk, kxPad, lq, apitch;
   size_t lds, partBytes;
};

HebbPlan hebb_plan(const pvhip_hebbian_args &a, bool bf) {
   HebbPlan pl;
   Geo pre    = geo(a.pre);
   int nK     = a.pre.nf;
   int N      = a.nxp * a.nyp * a.nfp;
   int mt     = (nK + HMT - 1) / HMT;
   int nt     = (N + HNT - 1) / HNT;
   int rows   = pre.nbatch * pre.nyExt();
   // enough workgroups to fill the chip twice, at least 4 rows per chunk
   // (or the caller's choice)
   int want   = std::max(1, 512 / std::max(1, mt * nt));
   pl.rowsPerChunk = a.rows_per_chunk > 0 ? std::min(a.rows_per_chunk, rows)
                                          : std::max(4, (rows + want - 1) / want);
   pl.chunks       = (rows + pl.rowsPerChunk - 1) / pl.rowsPerChunk;
   pl.kxPad        = (pre.nxExt() + 31) / 32 * 32;
   int sx          = std::max(1, a.window_stride_x);
   int cx          = a.window_start_x;
   int qx0         = cx >= 0 ? cx / sx : -((-cx + sx - 1) / sx);
   int cx0         = cx - sx * qx0;
   pl.lq           = (pl.kxPad + (cx0 + a.nxp - 1) / sx + 2 + 1) / 2 * 2;
   pl.apitch       = pl.kxPad + 8;
   size_t esz      = bf ? 2 : 4;
   int copies      = bf ? 2 : 1;
   size_t cstr     = ((size_t)a.nyp * a.nfp * sx * pl.lq + 7) / 8 * 8;
   pl.lds          = ((size_t)copies * cstr + (size_t)HMT * pl.apitch) * esz;
   pl.partBytes    = (size_t)pl.chunks * nK * N * sizeof(float);
   return pl;
}

} // namespace
} // namespace pvhip

using namespace pvhip;

extern "C" {

int pvhip_hebbian_dw(const pvhip_hebbian_args *a, pvhip_stream stream) {
   PVHIP_REQUIRE(a != nullptr, "pvhip_hebbian_dw: NULL args");
   PVHIP_REQUIRE(loc_valid(a->pre) && loc_valid(a->post), "hebbian_dw: invalid loc");
   PVHIP_REQUIRE(a->pre.nbatch == a->post.nbatch, "hebbian_dw: batch sizes differ");
   PVHIP_REQUIRE(a->nfp == a->post.nf, "hebbian_dw: nfp must equal post nf");
   PVHIP_REQUIRE(a->num_kernels_x > 0 && a->num_kernels_y > 0, "hebbian_dw: bad kernel counts");
   PVHIP_REQUIRE(
         a->dim_x && a->dim_y && a->pstart_x && a->pstart_y && a->apost_x && a->apost_y,
         "hebbian_dw: NULL geometry table");
   PVHIP_REQUIRE(a->pre_data && a->post_data && a->dw, "hebbian_dw: NULL buffer");
   PVHIP_REQUIRE(a->rows_per_chunk >= 0, "hebbian_dw: negative rows_per_chunk");
   if (a->arith == PVHIP_ARITH_FP32 || a->arith == PVHIP_ARITH_BF16) {
      bool const bf = a->arith == PVHIP_ARITH_BF16;
      PVHIP_REQUIRE(a->num_kernels_x == 1 && a->num_kernels_y == 1,
            "hebbian_dw: the MFMA forms need one data patch per feature");
      PVHIP_REQUIRE(a->window_stride_x > 0 && a->window_stride_y > 0,
            "hebbian_dw: the MFMA forms need the window geometry (pvhost_hebbian_windows)");
      HebbPlan pl = hebb_plan(*a, bf);
      size_t need = pl.partBytes * (a->counts ? 2 : 1);
      PVHIP_REQUIRE(a->scratch && a->scratch_bytes >= need,
            "hebbian_dw: scratch of %zu bytes, %zu needed", a->scratch_bytes, need);
      PVHIP_REQUIRE(pl.lds <= 160 * 1024, "hebbian_dw: tile needs %zu bytes of LDS", pl.lds);
      HebbMfma h;
      h.pre     = geo(a->pre);
      h.post    = geo(a->post);
      h.nxp     = a->nxp;
      h.nyp     = a->nyp;
      h.nfp     = a->nfp;
      h.N       = a->nxp * a->nyp * a->nfp;
      h.nK      = a->pre.nf;
      h.sx      = a->window_stride_x;
      h.qx0     = a->window_start_x >= 0 ? a->window_start_x / h.sx
                                         : -((-a->window_start_x + h.sx - 1) / h.sx);
      h.cx0     = a->window_start_x - h.sx * h.qx0;
      h.sy      = a->window_stride_y;
      h.cy      = a->window_start_y;
      h.kxPad   = pl.kxPad;
      h.lq      = pl.lq;
      h.apitch  = pl.apitch;
      h.rowsPerChunk = pl.rowsPerChunk;
      h.preData  = a->pre_data;
      h.postData = a->post_data;
      float *part  = static_cast<float *>(a->scratch);
      float *cpart = a->counts ? part + pl.partBytes / sizeof(float) : nullptr;
      dim3 grid((h.nK + HMT - 1) / HMT, (h.N + HNT - 1) / HNT, pl.chunks);
      hipStream_t s = (hipStream_t)stream;
      h.part        = part;
      if (bf) {
         hipLaunchKernelGGL((hebbian_mfma_kernel<__bf16, false>), grid, dim3(256), pl.lds, s, h);
      }
      else {
         hipLaunchKernelGGL((hebbian_mfma_kernel<float, false>), grid, dim3(256), pl.lds, s, h);
      }
      PVHIP_LAUNCHED("hebbian_mfma");
      if (cpart) {
         h.part = cpart;
         if (bf) {
            hipLaunchKernelGGL((hebbian_mfma_kernel<__bf16, true>), grid, dim3(256), pl.lds, s, h);
         }
         else {
            hipLaunchKernelGGL((hebbian_mfma_kernel<float, true>), grid, dim3(256), pl.lds, s, h);
         }
         PVHIP_LAUNCHED("hebbian_mfma_counts");
      }
      size_t n = (size_t)h.nK * h.N;
      hipLaunchKernelGGL(hebbian_reduce_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
            part, cpart, pl.chunks, n, a->dw_max, a->dw, a->counts);
      PVHIP_LAUNCHED("hebbian_reduce");
      return PVHIP_OK;
   }
   if (a->arith != PVHIP_ARITH_EXACT) {
      return set_error(PVHIP_ERR_UNSUPPORTED, "hebbian_dw: arith %d not built", a->arith);
   }
   HebbParams p;
   p.pre      = geo(a->pre);
   p.post     = geo(a->post);
   p.nxp      = a->nxp;
   p.nyp      = a->nyp;
   p.nfp      = a->nfp;
   p.nKx      = a->num_kernels_x;
   p.nKy      = a->num_kernels_y;
   p.dimX     = a->dim_x;
   p.dimY     = a->dim_y;
   p.pstartX  = a->pstart_x;
   p.pstartY  = a->pstart_y;
   p.apostX   = a->apost_x;
   p.apostY   = a->apost_y;
   p.dwMax    = a->dw_max;
   p.preData  = a->pre_data;
   p.postData = a->post_data;
   p.dw       = a->dw;
   p.counts   = a->counts;
   size_t n   = (size_t)p.nKx * p.nKy * p.pre.nf * p.nxp * p.nyp * p.nfp;
   hipLaunchKernelGGL(hebbian_exact_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
         (hipStream_t)stream, p);
   PVHIP_LAUNCHED("hebbian_exact");
   return PVHIP_OK;
}

size_t pvhip_hebbian_dw_scratch_bytes(const pvhip_hebbian_args *a) {
   if (!a || !loc_valid(a->pre) || a->window_stride_x <= 0 || a->nxp <= 0 || a->nyp <= 0 || a->nfp <= 0) {
      return 0;
   }
   HebbPlan pl = hebb_plan(*a, a->arith == PVHIP_ARITH_BF16);
   return pl.partBytes * (a->counts ? 2 : 1);
}

} // extern "C"
