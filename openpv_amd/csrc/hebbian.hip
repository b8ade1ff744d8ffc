// HebbianUpdater dW for shared weights (HebbianUpdater.cpp:476-627).
//
// EXACT form: one thread per dW element; a wave covers 64 consecutive items
// of one data patch, so the presynaptic activity it walks is wave-uniform.
// For a dW element the reference accumulates (HebbianUpdater.cpp:501-528):
//   batch ascending, then the presynaptic extended rows and columns of that
//   data patch ascending, skipping a_pre == 0, adding (dWMax*a_pre)*a_post.
// Counts (normalizeDw) increment once per contributing presynaptic neuron.
#include "pvhip_internal.hpp"

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
   if (a->arith != PVHIP_ARITH_EXACT) {
      return set_error(PVHIP_ERR_UNSUPPORTED, "hebbian_dw: only the EXACT form is built");
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

} // extern "C"
