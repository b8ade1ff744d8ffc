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
               uint32_t const s = ps[pos];
               uint32_t const e = ps[pos + 1];
               for (uint32_t i = s; i < e; i++) {
                  pvhip_sparse_entry const en = ent[i];
                  float a = en.value;
                  if (a == 0.0f) {
                     continue;
                  }
                  a           = a * p.dtf;
                  int const f = (int)(en.index % (uint32_t)nfPre);
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
   size_t n     = (size_t)p.post.nbatch * p.post.numRes();
   dim3 grid((unsigned)((n + 255) / 256));
   if (a->arith == PVHIP_ARITH_EXACT) {
      hipLaunchKernelGGL(recv_pre_gather_kernel<true>, grid, dim3(256), 0, (hipStream_t)stream, p);
   }
   else {
      hipLaunchKernelGGL(recv_pre_gather_kernel<false>, grid, dim3(256), 0, (hipStream_t)stream, p);
   }
   PVHIP_LAUNCHED("recv_pre_gather");
   return PVHIP_OK;
}

} // extern "C"
