// Pre-perspective delivery of a dense source as a GEMM on the matrix cores
// (PresynapticPerspectiveGPUDelivery / CudaRecvPre replacement for the
// BF16 form; the reference order is PresynapticPerspectiveConvolveDelivery.cpp:
// 61-192, for each source neuron: GSyn[patch] += a * w[patch]).
//
// For a source whose neurons each send one shared patch (one data patch per
// source feature) into a target finer by a factor of 4 along x and y, every
// source position p (all its C channels) and patch offset o give one product
//     P[p][o] = sum_c a[p][c] * w[c][o],
// added to the target neuron at start(p) + o.  That is a GEMM with M = source
// positions, N = patch offsets, K = channels (v_mfma_f32_16x16x32_bf16 on the
// BF16 copy of the activity and a BF16 copy of the weights; in the FP32 form
// v_mfma_f32_16x16x4f32 on the FP32 activity and an FP32 copy of the weights),
// followed by the overlap-add of the products into the target.
//
// Work split.  The target neurons whose row is q modulo 4 receive only the
// offsets whose row is q - start_y modulo 4: one workgroup per (batch
// element, q) owns those target rows outright, keeps them in LDS, and needs
// only that quarter of the weights (staged in LDS once).  Within a workgroup
// the 8 waves take the source positions in groups of DMG blocks of 16 and
// compute all four column classes (target column modulo 4) for them, so each
// source fragment feeds four MFMAs and each weight fragment DMG.  The
// overlap-add runs in rounds separated by barriers: in round rr wave w adds
// its products of column class (w + rr) mod 4 (four waves per round, or all
// eight when the two waves of a class are far enough apart), so no two waves
// touch a target neuron at the same time and every neuron receives its
// products in a fixed order: the result does not depend on scheduling.
//
// Numerics: BF16 (FP32 form: FP32) products, FP32 sums (MFMA order over the
// channels, then the fixed overlap-add order), the form's tolerance, not the
// EXACT order.
#include "pvhip_internal.hpp"

#include <algorithm>
#include <type_traits>

namespace pvhip {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int DST = 4; // target cells per source cell along x and y
constexpr int DMG = 4; // position blocks (16 source positions) per wave and group

// Division by a divisor fixed per launch: n / d = (umulhi(n, mul) + n) >> shift
// for 0 <= n < 2^31.
struct FastDiv {
   uint32_t mul;
   int shift;
};

inline FastDiv make_fastdiv(int d) {
   int l = 0;
   while ((1u << l) < (uint32_t)d) {
      l++;
   }
   FastDiv f;
   f.shift = l;
   f.mul   = (uint32_t)(((uint64_t)1 << 32) * (((uint64_t)1 << l) - (uint64_t)d) / (uint64_t)d + 1);
   return f;
}

struct DenseParams {
   Geo pre, post;
   int nxp, nyp, nfp;
   int x0, y0;             // target x (y) of the unshrunken patch start of source extended column (row) 0
   int px0, py0, pnx, pny; // source positions multiplied (extended coordinates)
   int C;                  // source features (channels)
   const __bf16 *A;        // BF16: chunked activity [nbatch][C/32][nyExt][nxExt][32]
   const __bf16 *Wt;       // BF16: weights [nyp*nxp*nfp][C]
   const float *Af;        // FP32: activity [nbatch][nyExt][nxExt][C]
   const float *Wf;        // FP32: weights [nyp*nxp*nfp][C]
   float dtf;
   float *gsyn;
   int accumulate;
   int layer; // the target layer's update applied instead of writing GSyn
   Geo layerLoc;
   float layerCoeff, layerScale;
   int layerApplyScale;
   float *layerV, *layerA;
   Transfer layerT;
   // LDS tile of a workgroup: target columns Xlo .. Xlo + pitch/nfp - 1 (the
   // target plus the reach of the patches of the multiplied positions past
   // its edges), rows for target rows Ymin .. Ymax (likewise), then 5 rows
   // that absorb the products of the padding positions of the last group.
   int pitch;
   int Xlo, Ymin, Ymax;
   int rows; // rows of every workgroup's tile before the absorbing rows
   FastDiv fdx, fdxf; // division by pnx and by post.nx*nfp
};

__device__ __forceinline__ int pmod4(int a) { return a & 3; } // two's complement: a mod 4 in [0, 4)

// n / d by a multiply for 0 <= n < 2^31 (Granlund-Montgomery, l = ceil(log2 d)).
__device__ __forceinline__ int fdiv(int n, FastDiv f) {
   uint32_t const t = __umulhi((uint32_t)n, f.mul);
   return (int)((t + (uint32_t)n) >> f.shift);
}

// KS = C / KC k-steps.  The source fragments of one k-step are one register
// stage; NB stages are in flight (the loads of k-step s + NB - 1 are issued
// before the MFMAs of k-step s, running on into the next group's first
// k-steps).  NW = 8 waves per workgroup: two per SIMD hide each other's
// latencies.  Wave w keeps the products of class (w + k) mod 4 in its
// accumulator slot k, so that the overlap-add reads static registers.
// ADD selects the overlap-add:
//   ADD_ATOMIC: LDS atomic adds, for patches wider than 16 or position grids
//     narrower than 12 + nxp/4 (the four positions of one add, 4 apart in
//     row-major order, may then share a target neuron); otherwise the adds of
//     one instruction never share one and a plain read-add-write is exact;
//   ADD_QUAD: plain adds, four waves at a time (one per class);
//   ADD_ALL: plain adds, all eight waves at a time.  Waves w and w + 4 hold
//     the same class for source positions 4*16*DMG apart, at least
//     48*DMG + 1 apart position by position; when that spans nyp/4 source
//     rows their target rows differ.
constexpr int ADD_QUAD = 0, ADD_ATOMIC = 1, ADD_ALL = 2;
// F32: FP32 operands on v_mfma_f32_16x16x4f32, KC = 16 channels per k-step;
// otherwise BF16 operands on v_mfma_f32_16x16x32_bf16, KC = 32.  Either way a
// lane's fragment of one k-step is 16 bytes: BF16 channels 8g..8g+7 of its
// row, FP32 channels 4g..4g+3, which the FP32 form spreads over four MFMAs
// (MFMA j takes channel 4g + j from lane group g, for A and B alike; the sum
// over the channels is the same, in another order).
template <bool F32, int KS, int NW, int NB, int ADD>
__global__ void __launch_bounds__(64 * NW) recv_pre_dense_kernel(DenseParams p) {
   constexpr int NT  = 64 * NW;
   constexpr int KC  = F32 ? 16 : 32;        // channels per k-step
   constexpr int EPP = F32 ? 4 : 8;          // channels per 16-byte piece
   using Elt         = typename std::conditional<F32, float, __bf16>::type;
   using Frag        = typename std::conditional<F32, f32x4, bf16x8>::type;
   static_assert(KS % NB == 0 && NB % 2 == 0, "k-step s of every group lives in buffer s % NB");
   static_assert(NW % 4 == 0, "whole rounds of four class owners");
   static_assert(ADD != ADD_ALL || NW == 8, "two waves per class");
   extern __shared__ float smem[];
   Elt *wl   = reinterpret_cast<Elt *>(smem);    // [4 classes][KS][16 columns][KC channels]
   float *ol = smem + 4 * KS * 16 * KC * sizeof(Elt) / 4; // tile [rows + 5][pitch]

   // the four workgroups of one batch element on one XCD (workgroup L runs on
   // XCD L % 8), so that they share the element's activity in that XCD's L2
   int const L = blockIdx.x;
   int const t = L / 8;
   int const q = t % 4;
   int const b = (t / 4) * 8 + L % 8;
   if (b >= p.post.nbatch) {
      return; // whole workgroup
   }
   constexpr int C = KS * KC;
   int const tid   = threadIdx.x;
   int const lane  = tid % 64;
   int const wave  = __builtin_amdgcn_readfirstlane(tid / 64);
   int const r     = lane % 16;
   int const g     = lane / 16;
   int const NXQ   = p.nxp / DST;
   int const oyq   = pmod4(q - p.y0); // offset rows of this workgroup: oyq + 4*jy
   int const jlo   = (p.Ymin - q) >> 2; // tile row 0 = target row q + 4*jlo (floor division)

   // Weights of this workgroup's offsets: per class c (offset column
   // (c - x0) mod 4 + 4*ix) and k-step, 16 columns (iy, ix, f) x KC channels,
   // four 16-byte pieces per column.  Item it = tid + u*NT: its column and
   // piece depend on tid only.
   {
      constexpr int WIT = 4 * KS * 64 / NT;
      int const gg = tid % 4;
      int const n  = (tid / 4) % 16;
      int const f  = n % p.nfp;
      int const ix = (n / p.nfp) % NXQ;
      int const oy = oyq + DST * (n / (p.nfp * NXQ));
      int const jw = tid / 64;
      // element offset of (column, piece) in the [offset][C] weights, per class
      int wcol[4];
#pragma unroll
      for (int c = 0; c < 4; c++) {
         wcol[c] = ((oy * p.nxp + pmod4(c - p.x0) + DST * ix) * p.nfp + f) * C + EPP * gg;
      }
#pragma unroll
      for (int u = 0; u < WIT; u++) {
         int const j = jw + u * NW; // class * KS + k-step
         int const c = j / KS;
         int wo      = wcol[0];
#pragma unroll
         for (int cc = 1; cc < 4; cc++) {
            wo = c == cc ? wcol[cc] : wo;
         }
         const Elt *wsrc;
         if constexpr (F32) {
            wsrc = p.Wf;
         }
         else {
            wsrc = p.Wt;
         }
         uint4 const v = *reinterpret_cast<const uint4 *>(wsrc + wo + (j % KS) * KC);
         *reinterpret_cast<uint4 *>(wl + (j * 16 + n) * KC + EPP * gg) = v;
      }
   }
   for (int i = tid; i < (p.rows + 5) * p.pitch; i += NT) {
      ol[i] = 0.0f;
   }

   int const nxExt      = p.pre.nxExt();
   size_t const plane32 = (size_t)p.pre.nyExt() * nxExt * 32;
   const __bf16 *Ab     = nullptr; // BF16 form: batch element b's chunked activity
   const float *Afb     = nullptr; // FP32 form: its activity
   if constexpr (F32) {
      Afb = p.Af + (size_t)b * p.pre.numExt();
   }
   else {
      Ab = p.A + (size_t)b * (C / 32) * plane32;
   }
   int const nPos       = p.pnx * p.pny;
   int const nMB        = (nPos + 15) / 16;
   int const rounds     = (nMB + NW * DMG - 1) / (NW * DMG);

   // source element offsets (BF16: in chunk plane 0) of this lane's A-fragment rows.
   // A row past the last position reads position 0 instead: its products
   // land in output rows that the overlap-add skips (a row of the product
   // depends on its own A row only), so the loads need no mask.
   int aoff[DMG];
   auto set_rows = [&](int mb0) {
#pragma unroll
      for (int m = 0; m < DMG; m++) {
         int const pos = (mb0 + m) * 16 + r;
         int const pc  = pos < nPos ? pos : 0;
         int const yr  = fdiv(pc, p.fdx);
         aoff[m]       = ((p.py0 + yr) * nxExt + p.px0 + pc - yr * p.pnx) * (F32 ? C : 32) + EPP * g;
      }
   };
   Frag av[NB][DMG];
   auto fetch = [&](int buf, int ks) {
#pragma unroll
      for (int m = 0; m < DMG; m++) {
         if constexpr (F32) {
            av[buf][m] = *reinterpret_cast<const f32x4 *>(Afb + aoff[m] + ks * KC);
         }
         else {
            av[buf][m] = *reinterpret_cast<const bf16x8 *>(Ab + (size_t)ks * plane32 + aoff[m]);
         }
      }
   };
   // weight fragments of k-step ks for the wave's four slots
   int wofs[4];
#pragma unroll
   for (int k = 0; k < 4; k++) {
      wofs[k] = (((wave + k) & 3) * KS * 16 + r) * KC + EPP * g;
   }
   Frag bw[2][4];
   auto loadB = [&](int buf, int ks) {
#pragma unroll
      for (int k = 0; k < 4; k++) {
         bw[buf][k] = *reinterpret_cast<const Frag *>(wl + wofs[k] + ks * 16 * KC);
      }
   };
   // tile offset of this lane's column (jy, jx, f)
   int const fo      = r % p.nfp;
   int const jx      = (r / p.nfp) % NXQ;
   int const jy      = r / (p.nfp * NXQ);
   int const laneOff = jy * p.pitch + DST * jx * p.nfp + fo;
   int const rowOff  = ((p.y0 + oyq - q) >> 2) - jlo; // (y0 + oyq - q) is a multiple of 4

   // One k-step: issue the source loads of k-step fks (buffer (bb + NB - 1)
   // mod NB) and the weight fragments of k-step bks, then the MFMAs of the
   // k-step in buffer bb.  Every k-step issues the same loads, so the wait
   // before each MFMA is for the oldest stage only (a k-step that issued
   // none would make the compiler's wait count cover the youngest stages
   // too, draining the pipeline).
   f32x4 acc[DMG][4];
   auto kstep = [&](int bb, int fks, int bks) {
      fetch((bb + NB - 1) % NB, fks);
      loadB((bb + 1) & 1, bks);
      // keep the loads ahead of this k-step's MFMAs (the scheduler would
      // otherwise sink them next to their use)
      __builtin_amdgcn_sched_barrier(0);
      if constexpr (F32) {
         // channel 4g + j of the k-step: sixteen independent MFMAs between
         // two uses of one accumulator
#pragma unroll
         for (int j = 0; j < 4; j++) {
#pragma unroll
            for (int k = 0; k < 4; k++) {
#pragma unroll
               for (int m = 0; m < DMG; m++) {
                  acc[m][k] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[bb][m][j], bw[bb & 1][k][j], acc[m][k], 0, 0, 0);
               }
            }
         }
      }
      else {
#pragma unroll
         for (int k = 0; k < 4; k++) {
#pragma unroll
            for (int m = 0; m < DMG; m++) {
               acc[m][k] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[bb][m], bw[bb & 1][k], acc[m][k], 0, 0, 0);
            }
         }
      }
   };

   set_rows(wave * DMG);
#pragma unroll
   for (int st = 0; st < NB - 1; st++) {
      fetch(st, st); // in flight during the barrier below
   }
   __syncthreads(); // the staged weights and the cleared tile
   loadB(0, 0);
   for (int rd = 0; rd < rounds; rd++) {
      int const mb0 = (rd * NW + wave) * DMG;
#pragma unroll
      for (int m = 0; m < DMG; m++) {
#pragma unroll
         for (int k = 0; k < 4; k++) {
            acc[m][k] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
         }
      }
#pragma unroll 1
      for (int sg = 0; sg < KS / NB - 1; sg++) {
#pragma unroll
         for (int bb = 0; bb < NB; bb++) {
            int const ks = sg * NB + bb;
            kstep(bb, ks + NB - 1, ks + 1);
         }
      }
      // The group's last NB k-steps issue the loads of the next group's first
      // NB - 1 k-steps (in flight during the overlap-add) and its first weight
      // fragments.  After the last group the next group's positions are all
      // past the end, so those loads read position 0: in bounds, unused.
#pragma unroll
      for (int bb = 0; bb < NB; bb++) {
         if (bb == 1) {
            set_rows(((rd + 1) * NW + wave) * DMG);
         }
         kstep(bb, bb == 0 ? KS - 1 : bb - 1, bb + 1 < NB ? KS - NB + bb + 1 : 0);
      }
      // Overlap-add: lane (r, g), register i of block m, slot k holds the
      // product of source position (mb0 + m)*16 + 4g + i and column r of class
      // c = (w + k) mod 4, i.e. of target row 4*ye + y0 + oyq + 4*jy and column
      // 4*xe + x0 + (c - x0) mod 4 + 4*jx.  Tile address = position part +
      // lane part + class part; a padding position adds into the absorbing
      // rows.
      int pb[DMG][4];
#pragma unroll
      for (int m = 0; m < DMG; m++) {
#pragma unroll
         for (int i = 0; i < 4; i++) {
            int const pos = (mb0 + m) * 16 + 4 * g + i;
            int const yr  = fdiv(pos, p.fdx);
            int const xe  = p.px0 + pos - yr * p.pnx;
            pb[m][i]      = pos < nPos ? (p.py0 + yr + rowOff) * p.pitch + (DST * xe + p.x0 - p.Xlo) * p.nfp
                                       : p.rows * p.pitch;
         }
      }
      // Rounds: in round rr the four waves 4*(rr/4) .. +3 (ADD_ALL: every
      // wave) add, each its slot rr mod 4 (class (w + rr) mod 4), so no wave
      // touches a tile element another wave touches in the same round; LDS
      // keeps a wave's accesses in order, so when the 64 lanes of one add
      // have distinct elements a plain read-add-write is exact (and far
      // cheaper than LDS atomics).
      constexpr int ROUNDS = ADD == ADD_ALL ? 4 : NW;
#pragma unroll
      for (int rr = 0; rr < ROUNDS; rr++) {
         __syncthreads();
         if (ADD == ADD_ALL || wave / 4 == rr / 4) { // the same for the whole wave
            int const k   = rr & 3;
            int const off = laneOff + pmod4(((wave + k) & 3) - p.x0) * p.nfp;
#pragma unroll
            for (int m = 0; m < DMG; m++) {
#pragma unroll
               for (int i = 0; i < 4; i++) {
                  if constexpr (ADD == ADD_ATOMIC) {
                     atomicAdd(&ol[pb[m][i] + off], acc[m][rr & 3][i]);
                  }
                  else {
                     ol[pb[m][i] + off] += acc[m][rr & 3][i];
                  }
               }
            }
         }
      }
   }
   __syncthreads();

   // Epilogue: the rows q + 4j of this batch element, written once.
   int const nxf = p.post.nx * p.nfp;
   int const nJ  = (p.post.ny - q + DST - 1) / DST; // target rows q + 4j < ny
   for (int i = tid; i < nJ * nxf; i += NT) {
      int const j     = fdiv(i, p.fdxf);
      int const xf    = i - j * nxf;
      int const Y     = q + DST * j;
      float const s   = ol[(j - jlo) * p.pitch + xf - p.Xlo * p.nfp];
      size_t const gi = (size_t)b * p.post.numRes() + (size_t)Y * nxf + xf;
      if (p.layer) {
         Geo const &Ll   = p.layerLoc;
         int const X     = xf / p.nfp;
         size_t const ai = (size_t)b * Ll.numExt() + ((size_t)(Y + Ll.up) * Ll.nxExt() + X + Ll.lt) * p.nfp + xf % p.nfp;
         float const G   = 0.0f + p.dtf * s;
         float V         = 0.0f;
         float prod      = p.layerCoeff * G;
         V               = V + prod;
         if (p.layerApplyScale) {
            V = V * p.layerScale;
         }
         p.layerV[gi] = V;
         if (p.layerA) {
            p.layerA[ai] = apply_transfer(p.layerT, V);
         }
      }
      else {
         float const base = p.accumulate ? p.gsyn[gi] : 0.0f;
         p.gsyn[gi]       = base + p.dtf * s;
      }
   }
}

// weights [C][nyp][nxp][nfp] (pre-perspective, one data patch per source
// feature) -> BF16 [nyp*nxp*nfp][C], rounded to nearest even
__global__ void __launch_bounds__(256) recv_pre_dense_weights_kernel(
      const float *w, __bf16 *out, int C, int patch) {
   size_t const i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (i >= (size_t)C * patch) {
      return;
   }
   int const o = (int)(i / C);
   int const c = (int)(i % C);
   out[i]      = (__bf16)w[(size_t)c * patch + o];
}

// FP32 weights [C][patch] -> [patch][C]
__global__ void __launch_bounds__(256) recv_pre_dense_weights_f32_kernel(
      const float *w, float *out, int C, int patch) {
   size_t const i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (i >= (size_t)C * patch) {
      return;
   }
   int const o = (int)(i / C);
   int const c = (int)(i % C);
   out[i]      = w[(size_t)c * patch + o];
}

// weights of four classes x C channels x 16 columns, then the tile
size_t dense_lds_bytes(bool f32, int C, int nJ, int pitch) {
   return (size_t)4 * C * 16 * (f32 ? 4 : 2) + (size_t)nJ * pitch * 4;
}

constexpr size_t kDenseLdsMax = 160 * 1024;

// The positions multiplied (all extended ones, or the restricted ones when
// the halo holds only zeros) and the LDS tile of a workgroup: the target,
// widened to every column / row a multiplied position's patch reaches, then 5
// absorbing rows.  Returns the workgroup's LDS bytes.  Needs p.x0, p.y0 and
// p.pre set.
size_t dense_tile(const pvhip_recv_pre_dense_args &a, DenseParams &p) {
   if (a.halo_is_zero) {
      p.px0 = a.pre.halo.lt;
      p.py0 = a.pre.halo.up;
      p.pnx = a.pre.nx;
      p.pny = a.pre.ny;
   }
   else {
      p.px0 = 0;
      p.py0 = 0;
      p.pnx = p.pre.nxExt();
      p.pny = p.pre.nyExt();
   }
   int const Xmin = std::min(0, DST * p.px0 + p.x0);
   int const Xmax = std::max(a.post.nx - 1, DST * (p.px0 + p.pnx - 1) + p.x0 + a.nxp - 1);
   p.Ymin         = std::min(0, DST * p.py0 + p.y0);
   p.Ymax         = std::max(a.post.ny - 1, DST * (p.py0 + p.pny - 1) + p.y0 + a.nyp - 1);
   p.Xlo          = Xmin;
   p.pitch        = (Xmax - Xmin + 1) * a.nfp;
   p.pitch += (p.pitch % 2 == 0); // odd: consecutive rows start in different banks
   // rows (Y - q)/4 - floor((Ymin - q)/4) for Y <= Ymax, over every q
   p.rows = (p.Ymax - p.Ymin) / DST + 2;
   return dense_lds_bytes(a.fp32 != 0, a.pre.nf, p.rows + 5, p.pitch);
}

} // namespace
} // namespace pvhip

using namespace pvhip;

extern "C" {

int pvhip_recv_pre_dense(const pvhip_recv_pre_dense_args *a, pvhip_stream stream) {
   PVHIP_REQUIRE(a != nullptr, "recv_pre_dense: NULL args");
   PVHIP_REQUIRE(loc_valid(a->pre) && loc_valid(a->post), "recv_pre_dense: invalid loc");
   PVHIP_REQUIRE(a->pre.nbatch == a->post.nbatch, "recv_pre_dense: batch sizes differ");
   PVHIP_REQUIRE(a->post.nx == DST * a->pre.nx && a->post.ny == DST * a->pre.ny,
         "recv_pre_dense: the target must be 4 times the source along x and y (%dx%d -> %dx%d)", a->pre.nx,
         a->pre.ny, a->post.nx, a->post.ny);
   PVHIP_REQUIRE(a->nfp == a->post.nf && a->nxp % DST == 0 && a->nyp % DST == 0
                       && (a->nxp / DST) * (a->nyp / DST) * a->nfp == 16,
         "recv_pre_dense: patch %dx%dx%d is not 16 offsets per class of 4x4", a->nxp, a->nyp, a->nfp);
   int const C = a->pre.nf;
   PVHIP_REQUIRE(C == 128 || C == 256 || C == 512, "recv_pre_dense: %d source features (128, 256 or 512)", C);
   bool const f32 = a->fp32 != 0;
   if (f32) {
      PVHIP_REQUIRE(a->pre_data && a->weights_f32, "recv_pre_dense: FP32 form needs pre_data and weights_f32");
   }
   else {
      PVHIP_REQUIRE(a->pre_data_bf16c && a->weights_bf16, "recv_pre_dense: NULL activity or weights");
   }
   DenseParams p{};
   p.pre   = geo(a->pre);
   p.post  = geo(a->post);
   p.nxp   = a->nxp;
   p.nyp   = a->nyp;
   p.nfp   = a->nfp;
   p.x0    = a->start_x0;
   p.y0    = a->start_y0;
   p.C     = C;
   p.A     = f32 ? nullptr : static_cast<const __bf16 *>(a->pre_data_bf16c);
   p.Wt    = f32 ? nullptr : static_cast<const __bf16 *>(a->weights_bf16);
   p.Af    = f32 ? a->pre_data : nullptr;
   p.Wf    = f32 ? a->weights_f32 : nullptr;
   p.dtf   = a->dt_factor;
   p.gsyn  = a->post_gsyn;
   p.accumulate = a->accumulate;
   p.layer = a->fused_layer != nullptr;
   if (p.layer) {
      const pvhip_layer_update_args &l = *a->fused_layer;
      PVHIP_REQUIRE(!a->accumulate, "recv_pre_dense: fused_layer needs accumulate == 0");
      PVHIP_REQUIRE(loc_valid(l.loc) && l.loc.nbatch == a->post.nbatch && l.loc.nx == a->post.nx
                          && l.loc.ny == a->post.ny && l.loc.nf == a->post.nf,
            "recv_pre_dense: fused_layer loc differs from post loc");
      PVHIP_REQUIRE(l.V != nullptr, "recv_pre_dense: fused_layer needs V");
      p.layerLoc        = geo(l.loc);
      p.layerCoeff      = l.channel_coeffs[0];
      p.layerScale      = l.scale;
      p.layerApplyScale = l.apply_scale;
      p.layerV          = l.V;
      p.layerA          = l.A;
      int st            = make_transfer(l.transfer, &p.layerT);
      if (st) {
         return st;
      }
   }
   else {
      PVHIP_REQUIRE(a->post_gsyn != nullptr, "recv_pre_dense: NULL post_gsyn");
   }
   size_t const lds = dense_tile(*a, p);
   p.fdx         = make_fastdiv(p.pnx);
   p.fdxf        = make_fastdiv(a->post.nx * a->nfp);
   PVHIP_REQUIRE(lds <= kDenseLdsMax, "recv_pre_dense: %zu bytes of LDS (tile rows of %d values)", lds, p.pitch);
   unsigned const groups = (unsigned)((a->post.nbatch + 7) / 8);
   dim3 grid(groups * 8 * 4);
   hipStream_t s = (hipStream_t)stream;
   constexpr int NW = 8, NB = 4;
   // the four positions of one add lie 4 apart in row-major order: in one row
   // their patches (nxp/4 source columns wide) are disjoint when nxp <= 16,
   // across a row wrap when the grid is at least 12 + nxp/4 wide
   bool const atomic = a->nxp > 16 || p.pnx < 12 + a->nxp / DST;
   // waves w and w + 4 add together when positions 48*DMG + 1 apart lie
   // nyp/4 source rows apart
   bool const all = !atomic && (a->nyp / DST) * p.pnx <= 48 * DMG + 1;
#define RPD_LAUNCH(F, KSV)                                                                                \
   if (atomic) {                                                                                         \
      hipLaunchKernelGGL((recv_pre_dense_kernel<F, KSV, NW, NB, ADD_ATOMIC>), grid, dim3(64 * NW), lds, s, p); \
   }                                                                                                     \
   else if (all) {                                                                                       \
      hipLaunchKernelGGL((recv_pre_dense_kernel<F, KSV, NW, NB, ADD_ALL>), grid, dim3(64 * NW), lds, s, p); \
   }                                                                                                     \
   else {                                                                                                \
      hipLaunchKernelGGL((recv_pre_dense_kernel<F, KSV, NW, NB, ADD_QUAD>), grid, dim3(64 * NW), lds, s, p); \
   }
   // k-steps of 32 channels (BF16) or 16 (FP32)
   if (!f32 && C == 128) {
      RPD_LAUNCH(false, 4)
   }
   else if (!f32 && C == 256) {
      RPD_LAUNCH(false, 8)
   }
   else if (!f32) {
      RPD_LAUNCH(false, 16)
   }
   else if (C == 128) {
      RPD_LAUNCH(true, 8)
   }
   else if (C == 256) {
      RPD_LAUNCH(true, 16)
   }
   else {
      RPD_LAUNCH(true, 32)
   }
#undef RPD_LAUNCH
   PVHIP_LAUNCHED("recv_pre_dense");
   return PVHIP_OK;
}

int pvhip_recv_pre_dense_fits(const pvhip_recv_pre_dense_args *a) {
   if (a == nullptr || !loc_valid(a->pre) || !loc_valid(a->post) || a->post.nx != DST * a->pre.nx
       || a->post.ny != DST * a->pre.ny || a->nfp != a->post.nf || a->nxp % DST || a->nyp % DST
       || (a->nxp / DST) * (a->nyp / DST) * a->nfp != 16
       || (a->pre.nf != 128 && a->pre.nf != 256 && a->pre.nf != 512)) {
      return 0;
   }
   DenseParams p{};
   p.pre = geo(a->pre);
   p.x0  = a->start_x0;
   p.y0  = a->start_y0;
   return dense_tile(*a, p) <= kDenseLdsMax;
}

int pvhip_recv_pre_dense_weights(
      const float *weights, void *weights_bf16, int num_features, int patch_size, pvhip_stream stream) {
   PVHIP_REQUIRE(weights && weights_bf16 && num_features > 0 && patch_size > 0,
         "recv_pre_dense_weights: bad argument");
   size_t const n = (size_t)num_features * patch_size;
   hipLaunchKernelGGL(recv_pre_dense_weights_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
         (hipStream_t)stream, weights, static_cast<__bf16 *>(weights_bf16), num_features, patch_size);
   PVHIP_LAUNCHED("recv_pre_dense_weights");
   return PVHIP_OK;
}

int pvhip_recv_pre_dense_weights_f32(
      const float *weights, float *weights_f32, int num_features, int patch_size, pvhip_stream stream) {
   PVHIP_REQUIRE(weights && weights_f32 && num_features > 0 && patch_size > 0,
         "recv_pre_dense_weights_f32: bad argument");
   size_t const n = (size_t)num_features * patch_size;
   hipLaunchKernelGGL(recv_pre_dense_weights_f32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
         (hipStream_t)stream, weights, weights_f32, num_features, patch_size);
   PVHIP_LAUNCHED("recv_pre_dense_weights_f32");
   return PVHIP_OK;
}

size_t pvhip_recv_pre_dense_weights_bytes(int num_features, int patch_size) {
   return (size_t)num_features * patch_size * 2;
}

} // extern "C"
