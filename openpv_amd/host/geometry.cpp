// Host-side patch geometry for the device deliveries.
//
// Restates OpenPV's connection geometry (src/components/PatchGeometry.cpp,
// src/components/Weights.cpp:205-231, src/components/PatchSize.cpp:98-130,
// src/utils/requiredConvolveMargin.hpp:14-40, src/utils/TransposeWeights.cpp)
// and turns it into the per-axis and coverage tables that libpvhip kernels
// consume.  Every layout PatchGeometry produces is separable in x and y, so
// the device never needs the reference's per-patch arrays (numPatches
// entries, e.g. 1.1 M for a 256x256x16 pre layer): a few hundred ints per axis
// describe the same thing.
#include "pvhost.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <vector>

namespace {

// One axis of a connection: n neurons pre (restricted), n neurons post.
struct Axis {
   int nPre, nPost;
   int preLo;   // pre border used to turn extended into restricted index
   int postLo;  // post border used for extended post coordinates
   int patch;   // pre-perspective patch size along the axis
};

// Start of the (unshrunken) patch of restricted pre index i in restricted
// post coordinates (PatchGeometry::calcPatchStartInPost, :334-389).
int patchStartInPost(int i, int patch, int nPre, int nPost) {
   if (nPre == nPost) {
      return i - (patch - 1) / 2;
   }
   if (nPre < nPost) {
      int t = nPost / nPre;
      return i * t - (patch - t) / 2;
   }
   int s    = nPre / nPost;
   int half = (s / 2) * (patch - 1);
   // floor division of a possibly negative numerator, in float as the reference
   return (int)std::floor((float)(i - half) / (float)s);
}

struct AxisEntry {
   int dim;           // kept patch extent (Patch::nx / ny)
   int patchStart;    // first kept patch index (offset component)
   int postStartRes;  // restricted post index of the first kept element
   int postStartExt;  // extended post index of the first kept element
   int unshrunk;      // extended post index of the unshrunken patch start
};

// PatchGeometry::calcPatchData (:391-453) for extended pre index e.
AxisEntry axisEntry(const Axis &a, int e) {
   int start = patchStartInPost(e - a.preLo, a.patch, a.nPre, a.nPost);
   AxisEntry r;
   r.unshrunk = start + a.postLo;
   int lo     = start;
   int hi     = start + a.patch;
   int dim    = a.patch;
   int first  = 0;
   if (hi < 0) {
      lo += -hi;
      hi = 0;
   }
   if (lo > a.nPost) {
      hi -= lo - a.nPost;
      lo = a.nPost;
   }
   if (lo < 0) {
      dim -= -lo;
      first = -lo;
      lo    = 0;
   }
   if (hi > a.nPost) {
      dim -= hi - a.nPost;
      hi = a.nPost;
   }
   r.dim          = std::max(dim, 0);
   r.patchStart   = first;
   r.postStartRes = lo;
   r.postStartExt = lo + a.postLo;
   return r;
}

// Axis descriptors exactly as PatchGeometry::setPatchGeometry builds them
// (:186-224): x uses the pre lt border and -- as the reference does -- the
// post dn border; y uses the pre dn and post dn borders.
Axis axisOf(const pvhip_loc &pre, const pvhip_loc &post, int nxp, int nyp, int axis) {
   if (axis == 0) {
      return Axis{pre.nx, post.nx, pre.halo.lt, post.halo.dn, nxp};
   }
   return Axis{pre.ny, post.ny, pre.halo.dn, post.halo.dn, nyp};
}

bool locOk(const pvhip_loc *l) {
   return l && l->nbatch > 0 && l->nx > 0 && l->ny > 0 && l->nf > 0 && l->halo.lt >= 0
          && l->halo.rt >= 0 && l->halo.dn >= 0 && l->halo.up >= 0;
}

bool scaleOk(int a, int b) {
   int hi = std::max(a, b), lo = std::min(a, b);
   if (hi % lo) {
      return false;
   }
   int q = hi / lo;
   return (q & (q - 1)) == 0;
}

int kernelsAlong(int nPre, int nPost) { return nPre > nPost ? nPre / nPost : 1; }

int posMod(int a, int m) {
   int r = a % m;
   return r < 0 ? r + m : r;
}

} // namespace

// TransposeWeights::transposeShared (src/utils/TransposeWeights.cpp:56-113)
// as a visit of (post-weight index, pre-weight index) pairs: the post
// (post-perspective) weights of a shared-weight connection pre -> post with
// pre-perspective patch nxp x nyp x nfp are a permutation of its pre weights.
template <typename Visit>
int transposeItems(const pvhip_loc *pre, const pvhip_loc *post, int nxp, int nyp, int nfp, Visit visit) {
   if (!locOk(pre) || !locOk(post) || nfp != post->nf) {
      return -1;
   }
   // Strides between the two layers (PatchGeometry::setTransposeItemIndices :281-332).
   int sx  = pre->nx > post->nx ? pre->nx / post->nx : 1;  // many-to-one stride
   int sy  = pre->ny > post->ny ? pre->ny / post->ny : 1;
   int tx  = post->nx > pre->nx ? post->nx / pre->nx : 1;  // one-to-many stride
   int ty  = post->ny > pre->ny ? post->ny / pre->ny : 1;
   int nKx = sx, nKy = sy, nKf = pre->nf;
   int pxPost = nxp * sx / tx;
   int pyPost = nyp * sy / ty;
   int pKx    = tx; // data patches of the post weights along x (kernelsAlong(post, pre))
   int pKy    = ty;
   int patchPre  = nxp * nyp * nfp;
   int patchPost = pxPost * pyPost * pre->nf;
   for (int kPre = 0; kPre < nKx * nKy * nKf; kPre++) {
      int kf  = kPre % nKf;
      int kxp = (kPre / nKf) % nKx;
      int kyp = kPre / (nKf * nKx);
      for (int item = 0; item < patchPre; item++) {
         int f  = item % nfp;
         int ix = (item / nfp) % nxp;
         int iy = item / (nfp * nxp);
         // mirrored item position and the sub-cell phase of even many-to-one patches
         int phaseX = posMod(kxp - (nxp - 1) * sx / 2, sx);
         int phaseY = posMod(kyp - (nyp - 1) * sy / 2, sy);
         int ixPost = (sx * (nxp - 1 - ix) + phaseX) / tx;
         int iyPost = (sy * (nyp - 1 - iy) + phaseY) / ty;
         int itemPost = kf + (ixPost + iyPost * pxPost) * pre->nf;
         // data patch of the post weights (TransposeWeights.cpp:73-99)
         int dpx = pKx > 1 ? posMod(-(nxp - pKx) / 2 + ix, pKx) : 0;
         int dpy = pKy > 1 ? posMod(-(nyp - pKy) / 2 + iy, pKy) : 0;
         int dataPost = f + (dpx + dpy * pKx) * nfp;
         visit((size_t)dataPost * patchPost + itemPost, (size_t)kPre * patchPre + item);
      }
   }
   return 0;
}

extern "C" {

int pvhost_required_margin(int nPre, int nPost, int patch) {
   if (nPre <= 0 || nPost <= 0 || patch <= 0 || !scaleOk(nPre, nPost)) {
      return -1;
   }
   if (nPre == nPost) {
      return patch % 2 == 1 ? (patch - 1) / 2 : -1;
   }
   if (nPre > nPost) {
      return (patch - 1) * (nPre / nPost) / 2;
   }
   int ratio = nPost / nPre;
   if (patch % ratio) {
      return -1;
   }
   return (patch / ratio) / 2;
}

int pvhost_post_patch_size(int prePatch, int nPre, int nPost) {
   if (nPre <= 0 || nPost <= 0 || prePatch <= 0) {
      return -1;
   }
   if (nPre == nPost) {
      return prePatch;
   }
   if (nPre > nPost) {
      return nPre % nPost ? -1 : prePatch * (nPre / nPost);
   }
   if (nPost % nPre) {
      return -1;
   }
   int s = nPost / nPre;
   return prePatch % s ? -1 : prePatch / s;
}

int pvhost_num_kernels(const pvhip_loc *pre, const pvhip_loc *post, int *nkx, int *nky) {
   if (!locOk(pre) || !locOk(post)) {
      return -1;
   }
   if (nkx) *nkx = kernelsAlong(pre->nx, post->nx);
   if (nky) *nky = kernelsAlong(pre->ny, post->ny);
   return 0;
}

int pvhost_patch_axis(const pvhip_loc *pre, const pvhip_loc *post, int nxp, int nyp, int axis,
      int32_t *dim, int32_t *pstart, int32_t *postRes, int32_t *postExt, int32_t *unshrunk) {
   if (!locOk(pre) || !locOk(post) || nxp <= 0 || nyp <= 0 || (axis != 0 && axis != 1)) {
      return -1;
   }
   Axis a = axisOf(*pre, *post, nxp, nyp, axis);
   int n  = axis == 0 ? pre->nx + pre->halo.lt + pre->halo.rt : pre->ny + pre->halo.dn + pre->halo.up;
   for (int e = 0; e < n; e++) {
      AxisEntry r = axisEntry(a, e);
      if (dim) dim[e] = r.dim;
      if (pstart) pstart[e] = r.patchStart;
      if (postRes) postRes[e] = r.postStartRes;
      if (postExt) postExt[e] = r.postStartExt;
      if (unshrunk) unshrunk[e] = r.unshrunk;
   }
   return 0;
}

int pvhost_hebbian_windows(const pvhip_loc *pre, const pvhip_loc *post, int nxp, int nyp,
      int *stride_x, int *start_x, int *stride_y, int *start_y) {
   if (!locOk(pre) || !locOk(post) || nxp <= 0 || nyp <= 0) {
      return -1;
   }
   for (int axis = 0; axis < 2; axis++) {
      Axis a      = axisOf(*pre, *post, nxp, nyp, axis);
      int n       = axis == 0 ? pre->nx + pre->halo.lt + pre->halo.rt : pre->ny + pre->halo.dn + pre->halo.up;
      int p       = axis == 0 ? nxp : nyp;
      int lo_res  = axis == 0 ? post->halo.lt : post->halo.up;
      int n_res   = axis == 0 ? post->nx : post->ny;
      int u0      = axisEntry(a, 0).unshrunk;
      int stride  = n > 1 ? axisEntry(a, 1).unshrunk - u0 : 1;
      for (int e = 0; e < n; e++) {
         AxisEntry r = axisEntry(a, e);
         if (r.unshrunk != u0 + stride * e) {
            return -2; // window starts are not affine in the extended index
         }
         int lo = std::max(r.unshrunk, lo_res);
         int hi = std::min(r.unshrunk + p, lo_res + n_res);
         int d  = std::max(0, hi - lo);
         if (d != r.dim || (d > 0 && (r.postStartExt != lo || r.patchStart != lo - r.unshrunk))) {
            return -3; // shrunken patch is not the window clipped to the restricted layer
         }
      }
      if (axis == 0) {
         if (stride_x) *stride_x = stride;
         if (start_x) *start_x = u0;
      }
      else {
         if (stride_y) *stride_y = stride;
         if (start_y) *start_y = u0;
      }
   }
   return 0;
}

int pvhost_recv_post_tables(const pvhip_loc *source, const pvhip_loc *target, int nxp, int nyp,
      int32_t *start_x, int32_t *start_y, int32_t *kernel_x, int32_t *kernel_y, int *nkx, int *nky) {
   if (!locOk(source) || !locOk(target) || nxp <= 0 || nyp <= 0) {
      return -1;
   }
   // Post weights: geometry with pre = target, post = source (WeightsPair.cpp:188-205).
   Axis ax = axisOf(*target, *source, nxp, nyp, 0);
   Axis ay = axisOf(*target, *source, nxp, nyp, 1);
   int kx  = kernelsAlong(target->nx, source->nx);
   int ky  = kernelsAlong(target->ny, source->ny);
   int nxExtS = source->nx + source->halo.lt + source->halo.rt;
   int nyExtS = source->ny + source->halo.dn + source->halo.up;
   for (int x = 0; x < target->nx; x++) {
      int e = x + target->halo.lt; // kIndexExtended x offset
      int s = axisEntry(ax, e).unshrunk;
      if (s < 0 || s + nxp > nxExtS) {
         return -2;
      }
      if (start_x) start_x[x] = s;
      if (kernel_x) kernel_x[x] = posMod(e - target->halo.lt, kx);
   }
   for (int y = 0; y < target->ny; y++) {
      int e = y + target->halo.up; // kIndexExtended y offset
      int s = axisEntry(ay, e).unshrunk;
      if (s < 0 || s + nyp > nyExtS) {
         return -2;
      }
      if (start_y) start_y[y] = s;
      if (kernel_y) kernel_y[y] = posMod(e - target->halo.up, ky);
   }
   if (nkx) *nkx = kx;
   if (nky) *nky = ky;
   return 0;
}

int pvhost_recv_post_inverse(const pvhip_loc *source, const pvhip_loc *target, int nxp, int nyp,
      int32_t *inverse_x, int32_t *inverse_y, int *window_x0, int *window_nx, int *window_y0,
      int *window_ny) {
   if (!locOk(source) || !locOk(target) || !window_x0 || !window_nx || !window_y0 || !window_ny) {
      return -1;
   }
   std::vector<int32_t> sx(target->nx), sy(target->ny), kx(target->nx), ky(target->ny);
   int nkx, nky;
   int st = pvhost_recv_post_tables(source, target, nxp, nyp, sx.data(), sy.data(), kx.data(),
         ky.data(), &nkx, &nky);
   if (st != 0) {
      return st;
   }
   auto fill = [](const std::vector<int32_t> &start, const std::vector<int32_t> &ker, int nk,
                     int32_t *inv, int *w0, int *wn) {
      int lo = *std::min_element(start.begin(), start.end());
      int hi = *std::max_element(start.begin(), start.end());
      *w0    = lo;
      *wn    = hi - lo + 1;
      if (inv) {
         std::fill(inv, inv + (size_t)nk * (*wn), -1);
         for (size_t t = 0; t < start.size(); t++) {
            int32_t &slot = inv[(size_t)ker[t] * (*wn) + (start[t] - lo)];
            if (slot >= 0) {
               return -3; // two targets share (kernel, window): not a bijection
            }
            slot = (int32_t)t;
         }
      }
      return 0;
   };
   st = fill(sx, kx, nkx, inverse_x, window_x0, window_nx);
   if (st != 0) {
      return st;
   }
   return fill(sy, ky, nky, inverse_y, window_y0, window_ny);
}

int pvhost_recv_pre_tables(const pvhip_loc *pre, const pvhip_loc *post, int nxp, int nyp,
      int32_t *col_ptr, int32_t *col_xe, int32_t *col_wx, int32_t *col_kx, int *ncol,
      int32_t *row_ptr, int32_t *row_ye, int32_t *row_wy, int32_t *row_ky, int *nrow) {
   if (!locOk(pre) || !locOk(post) || nxp <= 0 || nyp <= 0 || !ncol || !nrow) {
      return -1;
   }
   Axis ax    = axisOf(*pre, *post, nxp, nyp, 0);
   Axis ay    = axisOf(*pre, *post, nxp, nyp, 1);
   int kx     = kernelsAlong(pre->nx, post->nx);
   int ky     = kernelsAlong(pre->ny, post->ny);
   int nxExt  = pre->nx + pre->halo.lt + pre->halo.rt;
   int nyExt  = pre->ny + pre->halo.dn + pre->halo.up;
   bool fill  = col_ptr != nullptr;
   // Columns: pre extended columns whose kept patch covers post column X.
   std::vector<AxisEntry> ex(nxExt), ey(nyExt);
   for (int e = 0; e < nxExt; e++) ex[e] = axisEntry(ax, e);
   for (int e = 0; e < nyExt; e++) ey[e] = axisEntry(ay, e);
   int n = 0;
   for (int X = 0; X < post->nx; X++) {
      if (fill) col_ptr[X] = n;
      for (int xe = 0; xe < nxExt; xe++) {
         int rel = X - ex[xe].postStartRes;
         if (rel >= 0 && rel < ex[xe].dim) {
            if (fill) {
               col_xe[n] = xe;
               col_wx[n] = ex[xe].patchStart + rel;
               col_kx[n] = posMod(xe - pre->halo.lt, kx);
            }
            n++;
         }
      }
   }
   if (fill) col_ptr[post->nx] = n;
   *ncol = n;
   // Rows: for post row Y and patch-row loop index r (PresynapticPerspective-
   // ConvolveDelivery.cpp:152-165), the pre extended rows whose r-th kept row
   // lands on Y.
   n = 0;
   for (int Y = 0; Y < post->ny; Y++) {
      for (int r = 0; r < nyp; r++) {
         if (fill) row_ptr[Y * nyp + r] = n;
         for (int ye = 0; ye < nyExt; ye++) {
            if (r < ey[ye].dim && ey[ye].postStartRes + r == Y) {
               if (fill) {
                  row_ye[n] = ye;
                  row_wy[n] = ey[ye].patchStart + r;
                  row_ky[n] = posMod(ye - pre->halo.up, ky);
               }
               n++;
            }
         }
      }
   }
   if (fill) row_ptr[post->ny * nyp] = n;
   *nrow = n;
   return 0;
}

int pvhost_recv_pre_dense_start(const pvhip_loc *pre, const pvhip_loc *post, int nxp, int nyp,
      int *stride, int *start_x0, int *start_y0) {
   if (!locOk(pre) || !locOk(post) || nxp <= 0 || nyp <= 0 || !start_x0 || !start_y0) {
      return -1;
   }
   if (post->nx % pre->nx || post->ny % pre->ny || post->nx / pre->nx != post->ny / pre->ny) {
      return -1; // a coarser target, or different strides along x and y
   }
   int const s = post->nx / pre->nx;
   int starts[2];
   for (int axis = 0; axis < 2; axis++) {
      Axis a    = axisOf(*pre, *post, nxp, nyp, axis);
      int nExt  = axis == 0 ? pre->nx + pre->halo.lt + pre->halo.rt : pre->ny + pre->halo.dn + pre->halo.up;
      int patch = axis == 0 ? nxp : nyp;
      int nPost = axis == 0 ? post->nx : post->ny;
      bool have = false;
      int s0    = 0;
      for (int e = 0; e < nExt; e++) {
         AxisEntry en = axisEntry(a, e);
         if (en.dim <= 0) {
            continue;
         }
         int start = en.postStartRes - en.patchStart; // restricted target index of patch element 0
         if (!have) {
            s0   = start - s * e;
            have = true;
         }
         else if (start != s * e + s0) {
            return -1;
         }
      }
      if (!have) {
         return -1;
      }
      // columns whose kept patch is empty must not reach the target either
      for (int e = 0; e < nExt; e++) {
         if (axisEntry(a, e).dim <= 0 && s * e + s0 + patch > 0 && s * e + s0 < nPost) {
            return -1;
         }
      }
      starts[axis] = s0;
   }
   if (stride) {
      *stride = s;
   }
   *start_x0 = starts[0];
   *start_y0 = starts[1];
   return 0;
}

int pvhost_make_transfer(float VThresh, float AMin, float AMax, float AShift, float VWidth,
      int mode, pvhip_transfer *out) {
   if (!out || (mode != PVHIP_TRANSFER_VTHRESH && mode != PVHIP_TRANSFER_VERTICES)) {
      return -1;
   }
   *out         = pvhip_transfer{};
   out->mode    = mode;
   out->VThresh = VThresh;
   out->AMin    = AMin;
   out->AMax    = AMax;
   out->AShift  = AShift;
   out->VWidth  = VWidth;
   // Vertices of the piecewise-linear transfer function
   // (ANNActivityBuffer::setVertices :227-370, setSlopes :372-397).
   float vt = VThresh, vw = VWidth;
   if (vw < 0) {
      vt += vw;
      vw = -vw;
   }
   std::vector<float> V, A;
   float slopeNeg, slopePos = 1.0f;
   if (vt <= -0.999f * FLT_MAX) {
      V.push_back(0.0f);
      A.push_back(-AShift);
      slopeNeg = 1.0f;
   }
   else {
      if (vw == 0.0f && vt - AShift == AMin) {
         V.push_back(vt);
         A.push_back(AMin);
      }
      else {
         V.push_back(vt);
         V.push_back(vt + vw);
         A.push_back(AMin);
         A.push_back(vt + vw - AShift);
      }
      slopeNeg = 0.0f;
   }
   if (AMax < 0.999f * FLT_MAX) {
      if (A.back() < AMax) {
         float gap = AMax - A.back();
         V.push_back(V.back() + gap);
         A.push_back(AMax);
      }
      else {
         int keep = -1;
         for (int v = (int)A.size() - 1; v >= 0; v--) {
            if (A[v] < AMax) {
               keep = v;
               break;
            }
         }
         if (keep >= 0) {
            float gap = AMax - A[keep];
            V.resize(keep + 1);
            A.resize(keep + 1);
            V.push_back(V[keep] + gap);
            A.push_back(AMax);
         }
         else {
            V.resize(1);
            A.resize(1);
            if (slopeNeg > 0) {
               float dA = A[0] - AMax;
               V[0]     = V[0] - (float)(dA / slopeNeg);
               A[0]     = AMax;
            }
            else {
               V[0]     = 0.0f;
               A[0]     = AMax;
               slopeNeg = 0.0f;
            }
         }
      }
      slopePos = 0.0f;
   }
   int nv = (int)V.size();
   if (nv > PVHIP_MAX_VERTICES) {
      return -2;
   }
   out->num_vertices = nv;
   for (int i = 0; i < nv; i++) {
      out->vertices_v[i] = V[i];
      out->vertices_a[i] = A[i];
   }
   out->slopes[0] = slopeNeg;
   for (int k = 1; k < nv; k++) {
      float dv = V[k] - V[k - 1];
      if (V[k - 1] != V[k]) {
         out->slopes[k] = (A[k] - A[k - 1]) / dv;
      }
      else {
         out->slopes[k] = A[k] > A[k - 1] ? INFINITY : (A[k] < A[k - 1] ? -INFINITY : NAN);
      }
   }
   out->slopes[nv] = slopePos;
   return 0;
}

int pvhost_transpose_weights(const pvhip_loc *pre, const pvhip_loc *post, int nxp, int nyp,
      int nfp, const float *preW, float *postW) {
   if (!preW || !postW) {
      return -1;
   }
   return transposeItems(pre, post, nxp, nyp, nfp, [&](size_t dst, size_t src) { postW[dst] = preW[src]; });
}

int pvhost_transpose_weights_index(const pvhip_loc *pre, const pvhip_loc *post, int nxp, int nyp,
      int nfp, int32_t *index, size_t n) {
   if (!index) {
      return -1;
   }
   for (size_t i = 0; i < n; i++) {
      index[i] = -1;
   }
   bool ok = true;
   int st  = transposeItems(pre, post, nxp, nyp, nfp, [&](size_t dst, size_t src) {
      if (dst < n && src <= (size_t)INT32_MAX) {
         index[dst] = (int32_t)src;
      }
      else {
         ok = false;
      }
   });
   return st != 0 || !ok ? -1 : 0;
}

} // extern "C"
