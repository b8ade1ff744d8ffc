// Backward (adjoint) kernels of the unrolled LCA loop.  The reference has no
// adjoint; these are the derivatives of its forward operations:
//   V_t = e*V_{t-1} + (1-e)*(G_t + si*A_{t-1})   (HyPerLCAInternalStateBuffer.kpp:48-59)
//   A_t = T(V_t)                                  (ANNActivityBuffer transfer)
// The deliveries' adjoints are the deliveries themselves (the scatter is the
// transpose of the gather) and the weight gradients are the Hebbian
// correlation (pvhip_hebbian_dw with dWMax = 1).
#include "pvhip_internal.hpp"

namespace pvhip {
namespace {

// dA/dV of the transfer at V (almost everywhere; a jump counts as slope 0).
__device__ __forceinline__ float transfer_slope(const Transfer &t, float V) {
   float s = 1.0f;
   if (t.mode == PVHIP_TRANSFER_VTHRESH) {
      if (t.VThresh > -FLT_MAX) {
         if (V < t.VThresh) {
            s = 0.0f;
         }
         else if (V < t.VThresh + t.VWidth) {
            s = (t.VThresh + t.VWidth - t.AShift - t.AMin) / t.VWidth;
         }
      }
   }
   else {
      int const last = t.numVertices - 1;
      float vLast = t.vV[0], sLast = t.slopes[1];
#pragma unroll
      for (int i = 0; i < PVHIP_MAX_VERTICES; i++) {
         if (i == last) {
            vLast = t.vV[i];
            sLast = t.slopes[i + 1];
         }
      }
      if (V < t.vV[0]) {
         s = t.slopes[0];
      }
      else if (V >= vLast) {
         s = sLast;
      }
      else {
#pragma unroll
         for (int v = 0; v < PVHIP_MAX_VERTICES - 1; v++) {
            if (v < last && V >= t.vV[v] && V < t.vV[v + 1]) {
               s = t.slopes[v + 1];
            }
         }
      }
   }
   if (t.mode == PVHIP_TRANSFER_VTHRESH && t.AMax < FLT_MAX && apply_transfer(t, V) >= t.AMax) {
      s = 0.0f; // clamped
   }
   return s;
}

struct AdjParams {
   Geo loc;
   Transfer transfer;
   const float *expTau;
   float selfInteract;
   const float *V;
   const float *gA;
   float l1;
   float *gV;
   float *gG;
   float *gAprev;
};

__global__ void __launch_bounds__(256) lca_adjoint_kernel(AdjParams p) {
   size_t const n = (size_t)p.loc.nbatch * p.loc.numRes();
   size_t const i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (i >= n) {
      return;
   }
   int const b   = (int)(i / p.loc.numRes());
   int const k   = (int)(i % p.loc.numRes());
   float const e = p.expTau[b];
   float const v = p.V[i];
   float ga      = p.gA ? p.gA[i] : 0.0f;
   if (p.l1 != 0.0f) {
      float const a = apply_transfer(p.transfer, v);
      ga += a > 0.0f ? p.l1 : (a < 0.0f ? -p.l1 : 0.0f);
   }
   float const gvt = p.gV[i] + transfer_slope(p.transfer, v) * ga;
   float const gin = (1.0f - e) * gvt;
   p.gG[(size_t)b * p.loc.numExt() + p.loc.kex(k)] = gin;
   if (p.gAprev) {
      p.gAprev[i] = p.selfInteract * gin;
   }
   p.gV[i] = e * gvt;
}

__global__ void __launch_bounds__(256)
      scale_to_extended_kernel(Geo loc, float scale, const float *src, float *dst, int accumulate) {
   size_t const n = (size_t)loc.nbatch * loc.numRes();
   size_t const i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (i >= n) {
      return;
   }
   int const b    = (int)(i / loc.numRes());
   int const k    = (int)(i % loc.numRes());
   size_t const d = (size_t)b * loc.numExt() + loc.kex(k);
   float const v  = scale * src[i];
   dst[d]         = accumulate ? dst[d] + v : v;
}

__global__ void __launch_bounds__(256) axpby_kernel(size_t n, float a, const float *x, float b, float *y) {
   size_t const i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (i < n) {
      y[i] = a * x[i] + b * y[i];
   }
}

inline unsigned nblocks(size_t n) { return (unsigned)((n + 255) / 256); }

} // namespace
} // namespace pvhip

using namespace pvhip;

extern "C" {

int pvhip_lca_adjoint_step(const pvhip_lca_adjoint_args *a, pvhip_stream stream) {
   PVHIP_REQUIRE(a != nullptr, "lca_adjoint_step: NULL args");
   PVHIP_REQUIRE(loc_valid(a->loc), "lca_adjoint_step: invalid loc");
   PVHIP_REQUIRE(a->V && a->exp_tau && a->gV && a->gG, "lca_adjoint_step: NULL buffer");
   AdjParams p;
   p.loc = geo(a->loc);
   int st = make_transfer(a->transfer, &p.transfer);
   if (st) {
      return st;
   }
   p.expTau       = a->exp_tau;
   p.selfInteract = a->self_interact ? 1.0f : 0.0f;
   p.V            = a->V;
   p.gA           = a->gA;
   p.l1           = a->l1;
   p.gV           = a->gV;
   p.gG           = a->gG;
   p.gAprev       = a->gA_prev;
   size_t n       = (size_t)p.loc.nbatch * p.loc.numRes();
   hipLaunchKernelGGL(lca_adjoint_kernel, dim3(nblocks(n)), dim3(256), 0, (hipStream_t)stream, p);
   PVHIP_LAUNCHED("lca_adjoint");
   return PVHIP_OK;
}

int pvhip_scale_to_extended(
      const pvhip_loc *loc, float scale, const float *src, float *dst, int accumulate,
      pvhip_stream stream) {
   PVHIP_REQUIRE(loc && src && dst, "scale_to_extended: NULL argument");
   PVHIP_REQUIRE(loc_valid(*loc), "scale_to_extended: invalid loc");
   Geo g    = geo(*loc);
   size_t n = (size_t)g.nbatch * g.numRes();
   hipLaunchKernelGGL(scale_to_extended_kernel, dim3(nblocks(n)), dim3(256), 0, (hipStream_t)stream,
         g, scale, src, dst, accumulate);
   PVHIP_LAUNCHED("scale_to_extended");
   return PVHIP_OK;
}

int pvhip_axpby(size_t n, float a, const float *x, float b, float *y, pvhip_stream stream) {
   PVHIP_REQUIRE(x && y, "axpby: NULL argument");
   if (n == 0) {
      return PVHIP_OK;
   }
   hipLaunchKernelGGL(axpby_kernel, dim3(nblocks(n)), dim3(256), 0, (hipStream_t)stream, n, a, x, b, y);
   PVHIP_LAUNCHED("axpby");
   return PVHIP_OK;
}

} // extern "C"
