// Pre-perspective delivery (PresynapticPerspectiveConvolveDelivery /
// CudaRecvPre replacement) as an ordered gather over the active list.
//
// The reference scatters every active presynaptic patch into GSyn
// (PresynapticPerspectiveConvolveDelivery.cpp:150-188); with one thread the
// contributions to a postsynaptic neuron arrive in the order
//   patch row y ascending, then active entry (extended index kex) ascending.
// Here every postsynaptic neuron sums its own contributions in exactly that
// order: for patch row r, for the presynaptic rows whose row r lands on it
// (ascending), for the presynaptic columns whose patch covers it
// (ascending), for the active features of that position (the segment
// [pos_start[p], pos_start[p+1]) of the ordered active list).  No atomics,
// each GSyn element written once, results independent of launch shape.
// Threads of a wave cover consecutive postsynaptic features, so for nf >= 64
// the position walk and the active entries are wave-uniform (scalar loads)
// and the weight reads are coalesced along the feature axis.
#include "pvhip_internal.hpp"

#include <algorithm>

namespace pvhip {
namespace {

struct PreParams {
   Geo pre, post;
   int nxp, nyp, nfp;
   int nKx, nKy;
   const int32_t *colPtr, *colXe, *colWx, *colKx;
   const int32_t *rowPtr, *rowYe, *rowWy, *rowKy;
   float dtf;
   const pvhip_sparse_entry *active;
   const uint32_t *posStart;
   const float *pre_data;
   const float *W;
   float *gsyn;
   int accumulate;
};

template <bool EXACT>
__global__ void __launch_bounds__(256) recv_pre_gather_kernel(PreParams p) {
   size_t const n     = (size_t)p.post.nbatch * p.post.numRes();
   size_t const index = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (index >= n) {
      return;
   }
   int const numRes   = p.post.numRes();
   int const b        = (int)(index / numRes);
   int const k        = (int)(index % numRes);
   int const F        = k % p.post.nf;
   int const X        = (k / p.post.nf) % p.post.nx;
   int const Y        = k / (p.post.nf * p.post.nx);
   int const nfPre    = p.pre.nf;
   int const nxExt    = p.pre.nxExt();
   int const numPos   = nxExt * p.pre.nyExt();
   int const patchSize = p.nxp * p.nyp * p.nfp;
   size_t const preOff = (size_t)b * p.pre.numExt();
   const uint32_t *ps  = p.posStart ? p.posStart + (size_t)b * (numPos + 1) : nullptr;
   const pvhip_sparse_entry *ent = p.active ? p.active + preOff : nullptr;
   int const c0 = p.colPtr[X];
   int const c1 = p.colPtr[X + 1];
   float g      = p.accumulate ? p.gsyn[index] : 0.0f;
   for (int r = 0; r < p.nyp; r++) {
      int const r0 = p.rowPtr[Y * p.nyp + r];
      int const r1 = p.rowPtr[Y * p.nyp + r + 1];
      for (int ri = r0; ri < r1; ri++) {
         int const ye   = p.rowYe[ri];
         int const wy   = p.rowWy[ri];
         int const kyK  = p.rowKy[ri];
         for (int ci = c0; ci < c1; ci++) {
            int const xe   = p.colXe[ci];
            int const wrow = (wy * p.nxp + p.colWx[ci]) * p.nfp + F;
            int const kerBase = (p.colKx[ci] + kyK * p.nKx) * nfPre;
            int const pos  = ye * nxExt + xe;
            if (ps) {
               uint32_t const s    = ps[pos];
               uint32_t const e    = ps[pos + 1];
               uint32_t const kex0 = (uint32_t)pos * (uint32_t)nfPre; // kex = pos*nf + f
               // four entries (and their weights) loaded together, then added
               // in entry order
               constexpr int U = 8;
               uint32_t i      = s;
               for (; i + U <= e; i += U) {
                  pvhip_sparse_entry en[U];
                  float w[U];
#pragma unroll
                  for (int u = 0; u < U; u++) {
                     en[u] = ent[i + u];
                  }
#pragma unroll
                  for (int u = 0; u < U; u++) {
                     w[u] = p.W[(size_t)(kerBase + (int)(en[u].index - kex0)) * patchSize + wrow];
                  }
#pragma unroll
                  for (int u = 0; u < U; u++) {
                     float a = en[u].value;
                     if (a != 0.0f) {
                        a = a * p.dtf;
                        if (EXACT) {
                           float prod = a * w[u];
                           g          = g + prod;
                        }
                        else {
                           g = __builtin_fmaf(a, w[u], g);
                        }
                     }
                  }
               }
               for (; i < e; i++) {
                  pvhip_sparse_entry const en = ent[i];
                  float a = en.value;
                  if (a == 0.0f) {
                     continue;
                  }
                  a           = a * p.dtf;
                  int const f = (int)(en.index - kex0);
                  float w     = p.W[(size_t)(kerBase + f) * patchSize + wrow];
                  if (EXACT) {
                     float prod = a * w;
                     g          = g + prod;
                  }
                  else {
                     g = __builtin_fmaf(a, w, g);
                  }
               }
            }
            else {
               float const *apos = p.pre_data + preOff + (size_t)pos * nfPre;
               for (int f = 0; f < nfPre; f++) {
                  float a = apos[f];
                  if (a == 0.0f) {
                     continue;
                  }
                  a       = a * p.dtf;
                  float w = p.W[(size_t)(kerBase + f) * patchSize + wrow];
                  if (EXACT) {
                     float prod = a * w;
                     g          = g + prod;
                  }
                  else {
                     g = __builtin_fmaf(a, w, g);
                  }
               }
            }
         }
      }
   }
   p.gsyn[index] = g;
}

// ------------------------------------------------------------------ LDS-staged
//
// Same per-neuron order, organised so that the data a workgroup needs sits in
// LDS: a workgroup owns R postsynaptic rows that share their sequence of
// weight rows (rows rowStep apart), TX columns x all features, and G batch
// elements.  For each patch row r (ascending) and each of its rows, the
// presynaptic row ye that lands there is handled by
//   1. staging the weight row wy of every data patch with that kernel row
//      (only when it differs from the row already staged), and
//   2. staging, column chunk by column chunk, the run of the ordered active
//      list that belongs to pre row ye and the tile's pre columns -- one
//      contiguous range of the list, since kex is position-major;
// then every thread walks its own columns (ascending) and their active
// features (ascending) out of LDS.  Work per thread and order per neuron are
// those of recv_pre_gather_kernel; only where the operands come from changes.

struct PreTile {
   int TX;             // post columns per workgroup
   int G;              // batch elements per workgroup
   int rowStep;        // distance between the rows of a workgroup
   int groupsPerRes;   // row groups per residue class
   int cap;            // staged entries per batch element
   int maxCols;        // widest pre-column union of a tile
   int wlds;           // weight row staged in LDS
   int sliceFloats;    // floats of one staged weight row
};

template <int R, bool EXACT>
__global__ void __launch_bounds__(256) recv_pre_lds_kernel(PreParams p, PreTile t) {
   extern __shared__ float smem[];
   int const nfPost  = p.post.nf;
   int const nfPre   = p.pre.nf;
   int const nxExt   = p.pre.nxExt();
   int const numPos  = nxExt * p.pre.nyExt();
   int const nxpnfp  = p.nxp * p.nfp;
   int const patchSize = nxpnfp * p.nyp;
   float *Wl   = smem;
   float *Ev   = smem + (t.wlds ? t.sliceFloats : 0);
   int *Ef     = reinterpret_cast<int *>(Ev + t.G * t.cap);
   int *cs     = Ef + t.G * t.cap;
   int *shared = cs + t.G * (t.maxCols + 1); // [0] chunk end column, [1..2] column range

   int const tid    = threadIdx.x;
   int const perImg = t.TX * nfPost;
   int const g      = tid / perImg;
   int const rem    = tid % perImg;
   int const X0     = blockIdx.x * t.TX;
   int const X      = X0 + rem / nfPost;
   int const F      = rem % nfPost;
   int const b      = blockIdx.z * t.G + g;
   bool const lane  = g < t.G && X < p.post.nx && b < p.post.nbatch;
   int const residue = blockIdx.y / t.groupsPerRes;
   int const gi      = blockIdx.y % t.groupsPerRes;
   int Yj[R];
#pragma unroll
   for (int j = 0; j < R; j++) {
      Yj[j] = residue + t.rowStep * (gi * R + j);
   }
   int const Xlast = min(X0 + t.TX, p.post.nx) - 1;
   if (tid == 0) {
      int lo = nxExt, hi = -1;
      for (int ci = p.colPtr[X0]; ci < p.colPtr[Xlast + 1]; ci++) {
         lo = min(lo, p.colXe[ci]);
         hi = max(hi, p.colXe[ci]);
      }
      shared[1] = lo;
      shared[2] = hi;
   }
   __syncthreads();
   int const xeLo  = shared[1];
   int const xeHi  = shared[2];
   int const c0    = lane ? p.colPtr[X] : 0;
   int const c1    = lane ? p.colPtr[X + 1] : 0;

   float acc[R];
#pragma unroll
   for (int j = 0; j < R; j++) {
      acc[j] = (lane && p.accumulate && Yj[j] < p.post.ny)
                     ? p.gsyn[(size_t)b * p.post.numRes() + ((size_t)Yj[j] * p.post.nx + X) * nfPost + F]
                     : 0.0f;
   }
   int curWy = -1, curKy = -1;
   for (int r = 0; r < p.nyp; r++) {
#pragma unroll
      for (int j = 0; j < R; j++) {
         int const Y = Yj[j];
         if (Y >= p.post.ny) {
            continue; // uniform over the workgroup
         }
         int const r0 = p.rowPtr[Y * p.nyp + r];
         int const r1 = p.rowPtr[Y * p.nyp + r + 1];
         for (int ri = r0; ri < r1; ri++) {
            int const ye  = p.rowYe[ri];
            int const wy  = p.rowWy[ri];
            int const kyK = p.rowKy[ri];
            if (t.wlds && (wy != curWy || kyK != curKy)) {
               __syncthreads();
               int const nk = p.nKx * nfPre;
               for (int i = tid; i < nk * nxpnfp; i += blockDim.x) {
                  int kk = i / nxpnfp;
                  int o  = i % nxpnfp;
                  int kx = kk / nfPre;
                  int f  = kk % nfPre;
                  Wl[i]  = p.W[(size_t)((kx + kyK * p.nKx) * nfPre + f) * patchSize + wy * nxpnfp + o];
               }
               curWy = wy;
               curKy = kyK;
            }
            int cA = xeLo;
            while (cA <= xeHi) {
               __syncthreads();
               // Chunk end: the longest run of columns from cA whose entries fit
               // the staging capacity for every batch element of the group
               // (run length is monotone in the column count).
               int width = min(min(xeHi - cA + 1, t.maxCols), (int)blockDim.x);
               bool fits = false;
               if (tid < width) {
                  fits = true;
                  for (int gg = 0; gg < t.G; gg++) {
                     int bb = blockIdx.z * t.G + gg;
                     if (bb >= p.post.nbatch) {
                        break;
                     }
                     const uint32_t *ps = p.posStart + (size_t)bb * (numPos + 1) + ye * nxExt;
                     if ((int)(ps[cA + tid + 1] - ps[cA]) > t.cap) {
                        fits = false;
                     }
                  }
               }
               int nfit = __syncthreads_count(fits);
               if (tid == 0) {
                  shared[0] = cA + max(nfit, 1) - 1;
               }
               __syncthreads();
               int const cB    = shared[0];
               int const ncols = cB - cA + 1;
               // stage the active-list run of pre row ye, columns [cA, cB]
               for (int gg = 0; gg < t.G; gg++) {
                  int bb = blockIdx.z * t.G + gg;
                  if (bb >= p.post.nbatch) {
                     break;
                  }
                  const uint32_t *ps = p.posStart + (size_t)bb * (numPos + 1) + ye * nxExt;
                  uint32_t const e0  = ps[cA];
                  int const cnt      = (int)(ps[cB + 1] - e0);
                  const pvhip_sparse_entry *src = p.active + (size_t)bb * p.pre.numExt() + e0;
                  for (int i = tid; i < cnt; i += blockDim.x) {
                     pvhip_sparse_entry en   = src[i];
                     Ev[gg * t.cap + i]      = en.value;
                     Ef[gg * t.cap + i]      = (int)(en.index % (uint32_t)nfPre);
                  }
                  for (int c = tid; c <= ncols; c += blockDim.x) {
                     cs[gg * (t.maxCols + 1) + c] = (int)(ps[cA + c] - e0);
                  }
               }
               __syncthreads();
               if (lane) {
                  const int *csg  = cs + g * (t.maxCols + 1);
                  const float *Eg = Ev + g * t.cap;
                  const int *Fg   = Ef + g * t.cap;
                  float a0        = acc[j];
                  for (int ci = c0; ci < c1; ci++) {
                     int const xe = p.colXe[ci];
                     if (xe < cA) {
                        continue;
                     }
                     if (xe > cB) {
                        break;
                     }
                     int const wx   = p.colWx[ci];
                     int const kx   = p.colKx[ci];
                     int const s    = csg[xe - cA];
                     int const e    = csg[xe - cA + 1];
                     int const wofs = wx * p.nfp + F;
                     for (int i = s; i < e; i++) {
                        float a    = Eg[i] * p.dtf;
                        int const f = Fg[i];
                        float w = t.wlds ? Wl[(kx * nfPre + f) * nxpnfp + wofs]
                                         : p.W[(size_t)((kx + kyK * p.nKx) * nfPre + f) * patchSize
                                               + wy * nxpnfp + wofs];
                        if (EXACT) {
                           float prod = a * w;
                           a0         = a0 + prod;
                        }
                        else {
                           a0 = __builtin_fmaf(a, w, a0);
                        }
                     }
                  }
                  acc[j] = a0;
               }
               cA = cB + 1;
            }
         }
      }
   }
   if (lane) {
#pragma unroll
      for (int j = 0; j < R; j++) {
         if (Yj[j] < p.post.ny) {
            p.gsyn[(size_t)b * p.post.numRes() + ((size_t)Yj[j] * p.post.nx + X) * nfPost + F] = acc[j];
         }
      }
   }
}

// ------------------------------------------------------------------ wave form
//
// For postsynaptic layers with many features (nf % 4 == 0, nf >= 64): one
// wave per postsynaptic position, lane l summing features 4l..4l+3 of each
// 256-feature chunk.  The walk over patch rows, presynaptic rows, columns
// and active features is the same for every feature, so it is wave-uniform:
// when the covering presynaptic columns of the position are consecutive
// (at most 64 of them), the entries of one presynaptic row form a single
// run of the ordered active list, read 64 at a time by the lanes and then
// broadcast one by one in list order; otherwise the run of each column is
// read separately.  Each lane keeps the one-thread reference order for its
// features; the weight rows of an entry are 16-byte loads along the feature
// axis.
__device__ __forceinline__ int bcast_i(int v, int j) { return __builtin_amdgcn_readlane(v, j); }
__device__ __forceinline__ float bcast_f(float v, int j) {
   return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j));
}

template <bool EXACT>
__device__ __forceinline__ void acc4(float4 &acc, float a, float4 w) {
   if (EXACT) {
      float p0 = a * w.x, p1 = a * w.y, p2 = a * w.z, p3 = a * w.w;
      acc.x    = acc.x + p0;
      acc.y    = acc.y + p1;
      acc.z    = acc.z + p2;
      acc.w    = acc.w + p3;
   }
   else {
      acc.x = __builtin_fmaf(a, w.x, acc.x);
      acc.y = __builtin_fmaf(a, w.y, acc.y);
      acc.z = __builtin_fmaf(a, w.z, acc.z);
      acc.w = __builtin_fmaf(a, w.w, acc.w);
   }
}

// Wide targets (post.nf >= 64, a multiple of 4), sparse source: one wave per
// target position, lanes over 4 features each.  The terms of the position are
// the active entries of its coverage segments -- (patch row, covering pre
// row, column run) in the reference's order -- each in active-list order.
// Per chunk of 64 segments: lane i loads segment i's pos_start range, a wave
// prefix sum of the lengths numbers every entry, lane j loads entry j (a
// binary search over the prefix finds its segment) and computes its weight
// row offset; the wave then adds the entries in that order, 8 weight loads in
// flight at a time.  Same terms and order as the one-thread delivery.
__device__ __forceinline__ int wave_incl_scan(int v) {
#pragma unroll
   for (int off = 1; off < kWave; off <<= 1) {
      int u = __shfl_up(v, off, kWave);
      if ((int)(threadIdx.x % kWave) >= off) {
         v += u;
      }
   }
   return v;
}

template <bool EXACT>
__global__ void __launch_bounds__(256) recv_pre_wave_kernel(PreParams p) {
   int const lane    = threadIdx.x % kWave;
   size_t const wpos = (size_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
   size_t const nPos = (size_t)p.post.nbatch * p.post.ny * p.post.nx;
   if (wpos >= nPos) {
      return; // the whole wave
   }
   int const perImg  = p.post.ny * p.post.nx;
   int const b       = (int)(wpos / perImg);
   int const Y       = (int)(wpos % perImg) / p.post.nx;
   int const X       = (int)(wpos % perImg) % p.post.nx;
   int const nfPost  = p.post.nf;
   int const nfPre   = p.pre.nf;
   int const nxExt   = p.pre.nxExt();
   int const numPos  = nxExt * p.pre.nyExt();
   int const patchSize = p.nxp * p.nyp * p.nfp;
   const uint32_t *ps  = p.posStart + (size_t)b * (numPos + 1);
   const pvhip_sparse_entry *ent = p.active + (size_t)b * p.pre.numExt();
   int const cA     = p.colPtr[X];
   int const cB     = p.colPtr[X + 1];
   int const ncol   = cB - cA;
   int const xeA    = ncol > 0 ? p.colXe[cA] : 0;
   bool const contig = ncol > 0 && ncol <= kWave && p.colXe[cB - 1] - xeA == ncol - 1;
   int const colwx  = lane < ncol ? p.colWx[cA + lane] : 0;
   int const colkx  = lane < ncol ? p.colKx[cA + lane] : 0;
   int const nruns  = contig ? 1 : ncol;
   int const rA     = p.rowPtr[Y * p.nyp];
   int const rB     = p.rowPtr[(Y + 1) * p.nyp];
   int const nseg   = (rB - rA) * nruns;
   float *gout      = p.gsyn + (size_t)b * p.post.numRes() + ((size_t)Y * p.post.nx + X) * nfPost;
   for (int fc = 0; fc < nfPost; fc += 4 * kWave) {
      int const F    = fc + 4 * lane;
      bool const fOK = F < nfPost;
      int const Fc   = fOK ? F : 0;
      float4 acc     = (p.accumulate && fOK) ? *reinterpret_cast<const float4 *>(gout + F)
                                             : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      for (int s0 = 0; s0 < nseg; s0 += kWave) {
         // lane's segment: covering row ri, column run
         int const sid  = s0 + lane;
         bool const sOK = sid < nseg;
         int const ri   = rA + (sOK ? sid / nruns : 0);
         int const run  = sOK ? sid % nruns : 0;
         int const ye   = p.rowYe[ri];
         int const xe0  = contig ? xeA : p.colXe[cA + run];
         uint32_t const sst = sOK ? ps[ye * nxExt + xe0] : 0u;
         uint32_t const sen = sOK ? ps[ye * nxExt + xe0 + (contig ? ncol : 1)] : 0u;
         int const len      = (int)(sen - sst);
         // weight offset of the segment's patch row (kernel y index, weight y)
         int const rowOff = (p.rowKy[ri] * p.nKx) * nfPre * patchSize + p.rowWy[ri] * p.nxp * p.nfp;
         int const runWx  = contig ? 0 : p.colWx[cA + run];
         int const runKx  = contig ? 0 : p.colKx[cA + run];
         int const incl   = wave_incl_scan(len);
         int const excl   = incl - len;
         int const total  = __builtin_amdgcn_readlane(incl, kWave - 1);
         int const nsegC  = min(nseg - s0, kWave);
         for (int k0 = 0; k0 < total; k0 += kWave) {
            int const k  = k0 + lane;
            bool const kOK = k < total;
            // segment t of entry k: the last segment whose exclusive prefix is <=
            // k (excl[0] = 0 <= k; that segment is never empty).  A branchless
            // search of fixed length: every lane takes every shuffle, because a
            // shuffle reads zero from a lane that is not active.
            int t = 0;
#pragma unroll
            for (int step = kWave / 2; step > 0; step >>= 1) {
               int const mid = t + step;
               int const em  = __shfl(excl, min(mid, nsegC - 1), kWave);
               if (mid < nsegC && em <= k) {
                  t = mid;
               }
            }
            int const segS   = __shfl((int)sst, t, kWave);
            int const segE   = __shfl((int)sen, t, kWave);
            uint32_t const e = (uint32_t)segS + (uint32_t)(k - __shfl(excl, t, kWave));
            // an entry must lie in its segment's range (true by construction;
            // anything else is dropped rather than read out of bounds)
            bool const eOK = kOK && e >= (uint32_t)segS && e < (uint32_t)segE;
            pvhip_sparse_entry en = eOK ? ent[e] : pvhip_sparse_entry{0u, 0.0f};
            int const fl   = (int)(en.index % (uint32_t)nfPre);
            int const cl   = (int)((en.index / (uint32_t)nfPre) % (uint32_t)nxExt) - xeA;
            int const wx   = contig ? __shfl(colwx, cl, kWave) : __shfl(runWx, t, kWave);
            int const kx   = contig ? __shfl(colkx, cl, kWave) : __shfl(runKx, t, kWave);
            int const rowT = __shfl(rowOff, t, kWave);
            int const woff = eOK ? rowT + (kx * nfPre + fl) * patchSize + wx * p.nfp : 0;
            float const av = eOK ? en.value : 0.0f;
            int const n    = min(kWave, total - k0);
            for (int j0 = 0; j0 < n; j0 += 8) {
               float4 w[8];
               float a[8];
#pragma unroll
               for (int u = 0; u < 8; u++) {
                  int const j  = min(j0 + u, n - 1);
                  a[u]         = (j0 + u < n) ? bcast_f(av, j) : 0.0f;
                  int const wo = bcast_i(woff, j);
                  w[u]         = *reinterpret_cast<const float4 *>(p.W + (size_t)wo + Fc);
               }
#pragma unroll
               for (int u = 0; u < 8; u++) {
                  if (j0 + u < n && a[u] != 0.0f) {
                     acc4<EXACT>(acc, a[u] * p.dtf, w[u]);
                  }
               }
            }
         }
      }
      if (fOK) {
         *reinterpret_cast<float4 *>(gout + F) = acc;
      }
   }
}

} // namespace
} // namespace pvhip

using namespace pvhip;

extern "C" {

int pvhip_recv_pre(const pvhip_recv_pre_args *a, pvhip_stream stream) {
   PVHIP_REQUIRE(a != nullptr, "pvhip_recv_pre: NULL args");
   PVHIP_REQUIRE(loc_valid(a->pre) && loc_valid(a->post), "recv_pre: invalid loc");
   PVHIP_REQUIRE(a->pre.nbatch == a->post.nbatch, "recv_pre: batch sizes differ");
   PVHIP_REQUIRE(a->nxp > 0 && a->nyp > 0 && a->nfp == a->post.nf, "recv_pre: patch nfp must equal post nf");
   PVHIP_REQUIRE(a->num_kernels_x > 0 && a->num_kernels_y > 0, "recv_pre: bad kernel counts");
   PVHIP_REQUIRE(
         a->col_ptr && a->col_xe && a->col_wx && a->col_kx && a->row_ptr && a->row_ye && a->row_wy
               && a->row_ky,
         "recv_pre: NULL coverage table");
   PVHIP_REQUIRE(a->weights && a->post_gsyn, "recv_pre: NULL weights or gsyn");
   PVHIP_REQUIRE(
         (a->pos_start && a->active) || (!a->pos_start && a->pre_data),
         "recv_pre: need (active, pos_start) for a sparse pre layer or pre_data for a dense one");
   PVHIP_REQUIRE(
         a->arith == PVHIP_ARITH_EXACT || a->arith == PVHIP_ARITH_FP32,
         "recv_pre: arith %d not supported", a->arith);
   PreParams p;
   p.pre        = geo(a->pre);
   p.post       = geo(a->post);
   p.nxp        = a->nxp;
   p.nyp        = a->nyp;
   p.nfp        = a->nfp;
   p.nKx        = a->num_kernels_x;
   p.nKy        = a->num_kernels_y;
   p.colPtr     = a->col_ptr;
   p.colXe      = a->col_xe;
   p.colWx      = a->col_wx;
   p.colKx      = a->col_kx;
   p.rowPtr     = a->row_ptr;
   p.rowYe      = a->row_ye;
   p.rowWy      = a->row_wy;
   p.rowKy      = a->row_ky;
   p.dtf        = a->dt_factor;
   p.active     = a->active;
   p.posStart   = a->pos_start;
   p.pre_data   = a->pre_data;
   p.W          = a->weights;
   p.gsyn       = a->post_gsyn;
   p.accumulate = a->accumulate;
   hipStream_t s = (hipStream_t)stream;
   bool const exact = a->arith == PVHIP_ARITH_EXACT;

   // Wave form for sparse pre layers into many postsynaptic features.
   if (p.posStart && p.post.nf % 4 == 0 && p.post.nf >= 64) {
      size_t const nPos = (size_t)p.post.nbatch * p.post.ny * p.post.nx;
      dim3 grid((unsigned)((nPos + 3) / 4));
      if (exact) {
         hipLaunchKernelGGL(recv_pre_wave_kernel<true>, grid, dim3(256), 0, s, p);
      }
      else {
         hipLaunchKernelGGL(recv_pre_wave_kernel<false>, grid, dim3(256), 0, s, p);
      }
      PVHIP_LAUNCHED("recv_pre_wave");
      return PVHIP_OK;
   }
   // LDS-staged form for sparse pre layers whose post features fit one workgroup.
   if (p.posStart && p.post.nf <= 256 && p.pre.nf <= 4096) {
      PreTile t;
      t.TX      = std::max(1, std::min(p.post.nx, 256 / p.post.nf));
      t.G       = std::max(1, std::min(p.post.nbatch, 256 / (t.TX * p.post.nf)));
      t.rowStep = p.post.ny > p.pre.ny ? p.post.ny / p.pre.ny : 1;
      int const R = 4;
      int rowsPerRes = (p.post.ny + t.rowStep - 1) / t.rowStep;
      t.groupsPerRes = (rowsPerRes + R - 1) / R;
      // widest union of covering pre columns over the column tiles: bounded by
      // the columns a tile of TX post columns can see (the patch reach in pre
      // columns plus the tile span), computed from the stride rule
      int sPre  = p.pre.nx > p.post.nx ? p.pre.nx / p.post.nx : 1;
      int tPost = p.post.nx > p.pre.nx ? p.post.nx / p.pre.nx : 1;
      int reach = (p.nxp + tPost - 1) / tPost * sPre + sPre;
      t.maxCols = std::min(p.pre.nxExt(), (t.TX + tPost - 1) / tPost * sPre + reach + 2);
      t.sliceFloats = p.nKx * p.pre.nf * p.nxp * p.nfp;
      size_t const budget = 72 * 1024; // two workgroups per CU
      t.wlds = (size_t)t.sliceFloats * 4 <= budget / 2;
      size_t fixed = (t.wlds ? (size_t)t.sliceFloats * 4 : 0) + (size_t)t.G * (t.maxCols + 1) * 4 + 64;
      size_t room  = budget > fixed ? budget - fixed : 0;
      t.cap        = (int)std::min<size_t>(room / (8 * (size_t)t.G), 1u << 20);
      if (t.cap >= p.pre.nf) {
         size_t lds = fixed + (size_t)t.G * t.cap * 8;
         int threads = t.TX * p.post.nf * t.G;
         threads     = (threads + 63) / 64 * 64;
         dim3 grid((p.post.nx + t.TX - 1) / t.TX, t.rowStep * t.groupsPerRes,
               (p.post.nbatch + t.G - 1) / t.G);
         if (exact) {
            hipLaunchKernelGGL((recv_pre_lds_kernel<4, true>), grid, dim3(threads), lds, s, p, t);
         }
         else {
            hipLaunchKernelGGL((recv_pre_lds_kernel<4, false>), grid, dim3(threads), lds, s, p, t);
         }
         PVHIP_LAUNCHED("recv_pre_lds");
         return PVHIP_OK;
      }
   }
   size_t n     = (size_t)p.post.nbatch * p.post.numRes();
   dim3 grid((unsigned)((n + 255) / 256));
   if (exact) {
      hipLaunchKernelGGL(recv_pre_gather_kernel<true>, grid, dim3(256), 0, s, p);
   }
   else {
      hipLaunchKernelGGL(recv_pre_gather_kernel<false>, grid, dim3(256), 0, s, p);
   }
   PVHIP_LAUNCHED("recv_pre_gather");
   return PVHIP_OK;
}

} // extern "C"
