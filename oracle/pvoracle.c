/*
 * pvoracle.c -- CPU restatement of OpenPV's delivery + LCA path.
 *
 * TEST INFRASTRUCTURE ONLY (the parity oracle and the timed host baseline);
 * never linked into the product.  Compile with -O3 -ffp-contract=off, no
 * fast-math: the reference Release build is -O3 without -march on x86-64
 * (cmake/PVConfigProject.cmake:59), i.e. SSE scalar float, one rounding per
 * operation, no FMA contraction.
 */
#include "pvoracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------------------------------------------------------- index math */

int or_kIndex(int kx, int ky, int kf, int nx, int ny, int nf) {
   (void)ny;
   return kf + (kx + ky * nx) * nf; /* conversions.hpp:251-253 */
}

static int kxPos(int k, int nx, int ny, int nf) { (void)ny; return (k / nf) % nx; }
static int kyPos(int k, int nx, int ny, int nf) { return k / (nx * nf) % ny; }
static int featureIndex(int k, int nf) { return k % nf; }

int or_kIndexExtended(int k, int nx, int ny, int nf, int lt, int rt, int dn, int up) {
   /* conversions.hpp:401-407: x offset lt, y offset up */
   int const kx_ex = lt + kxPos(k, nx, ny, nf);
   int const ky_ex = up + kyPos(k, nx, ny, nf);
   int const kf    = featureIndex(k, nf);
   return or_kIndex(kx_ex, ky_ex, kf, nx + lt + rt, ny + dn + up, nf);
}

static int locNumExt(const or_loc *l) { return (l->nx + l->lt + l->rt) * (l->ny + l->dn + l->up) * l->nf; }
static int locNumRes(const or_loc *l) { return l->nx * l->ny * l->nf; }

int or_requiredConvolveMargin(int nPre, int nPost, int patchSize) {
   /* utils/requiredConvolveMargin.hpp:14-40 */
   int margin = 0;
   if (nPre == nPost) {
      margin = (patchSize - 1) / 2;
   }
   else if (nPre > nPost) {
      int densityRatio = nPre / nPost;
      margin           = (patchSize - 1) * densityRatio / 2;
   }
   else {
      int densityRatio = nPost / nPre;
      int numCells     = patchSize / densityRatio;
      margin           = numCells / 2;
   }
   return margin;
}

int or_calcPostPatchSize(int prePatchSize, int nPre, int nPost) {
   /* components/PatchSize.cpp:98-130 */
   if (nPre == nPost) {
      return prePatchSize;
   }
   else if (nPre > nPost) {
      return prePatchSize * (nPre / nPost);
   }
   return prePatchSize / (nPost / nPre);
}

/* ---------------------------------------------------------------- geometry */

int or_calcPatchStartInPost(int indexRestrictedPre, int patchSize, int nPre, int nPost) {
   /* PatchGeometry.cpp:334-389 */
   int patchStartInPost;
   if (nPre == nPost) {
      int extentOneSide = (patchSize - 1) / 2;
      patchStartInPost  = indexRestrictedPre - extentOneSide;
   }
   else if (nPre < nPost) {
      int tstride       = nPost / nPre;
      int extentOneSide = (patchSize - tstride) / 2;
      patchStartInPost  = indexRestrictedPre * tstride - extentOneSide;
   }
   else {
      int stride           = nPre / nPost;
      int extentOneSide    = (stride / 2) * (patchSize - 1);
      float fStride        = (float)stride;
      float fPatchStartPre = (float)(indexRestrictedPre - extentOneSide);
      patchStartInPost     = (int)floorf(fPatchStartPre / fStride);
   }
   return patchStartInPost;
}

void or_calcPatchData(
      int index, int numPreRestricted, int preStartBorder, int preEndBorder,
      int numPostRestricted, int postStartBorder, int postEndBorder, int patchSize,
      int *patchDim, int *patchStart, int *postPatchStartRestricted,
      int *postPatchStartExtended, int *postPatchUnshrunkenStart) {
   /* PatchGeometry.cpp:391-453 */
   (void)preEndBorder;
   (void)postEndBorder;
   int lPatchDim       = patchSize;
   int lPatchStart     = 0;
   int restrictedIndex = index - preStartBorder;
   int lPostPatchStartRes =
         or_calcPatchStartInPost(restrictedIndex, patchSize, numPreRestricted, numPostRestricted);
   *postPatchUnshrunkenStart = lPostPatchStartRes + postStartBorder;
   int lPostPatchEndRes      = lPostPatchStartRes + patchSize;

   if (lPostPatchEndRes < 0) {
      int excess = -lPostPatchEndRes;
      lPostPatchStartRes += excess;
      lPostPatchEndRes = 0;
   }
   if (lPostPatchStartRes > numPostRestricted) {
      int excess         = lPostPatchStartRes - numPostRestricted;
      lPostPatchStartRes = numPostRestricted;
      lPostPatchEndRes -= excess;
   }
   if (lPostPatchStartRes < 0) {
      int excess         = -lPostPatchStartRes;
      lPostPatchStartRes = 0;
      lPatchDim -= excess;
      lPatchStart = excess;
   }
   if (lPostPatchEndRes > numPostRestricted) {
      int excess       = lPostPatchEndRes - numPostRestricted;
      lPostPatchEndRes = numPostRestricted;
      lPatchDim -= excess;
   }
   if (lPatchDim < 0) {
      lPatchDim = 0;
   }
   *patchDim                 = lPatchDim;
   *patchStart               = lPatchStart;
   *postPatchStartRestricted = lPostPatchStartRes;
   *postPatchStartExtended   = lPostPatchStartRes + postStartBorder;
}

void or_patch_axis(
      const or_loc *pre, const or_loc *post, int nxp, int nyp, int axis,
      int *dim, int *pstart, int *postStartRes, int *postStartExt, int *unshrunkenStart) {
   /* setPatchGeometry's per-axis loops, PatchGeometry.cpp:186-224.  For x the
    * reference passes the post dn/up halo as borders (:198-199). */
   int n = axis == 0 ? pre->nx + pre->lt + pre->rt : pre->ny + pre->dn + pre->up;
   for (int i = 0; i < n; i++) {
      int d, s, pr, pe, us;
      if (axis == 0) {
         or_calcPatchData(i, pre->nx, pre->lt, pre->rt, post->nx, post->dn, post->up, nxp,
               &d, &s, &pr, &pe, &us);
      }
      else {
         or_calcPatchData(i, pre->ny, pre->dn, pre->up, post->ny, post->dn, post->up, nyp,
               &d, &s, &pr, &pe, &us);
      }
      if (dim) dim[i] = d;
      if (pstart) pstart[i] = s;
      if (postStartRes) postStartRes[i] = pr;
      if (postStartExt) postStartExt[i] = pe;
      if (unshrunkenStart) unshrunkenStart[i] = us;
   }
}

void or_patch_geometry(
      const or_loc *pre, const or_loc *post, int nxp, int nyp, int nfp,
      or_patch *patches, uint64_t *gSynPatchStart, uint64_t *aPostOffset,
      uint64_t *unshrunkenStart) {
   /* PatchGeometry::setPatchGeometry, PatchGeometry.cpp:179-279 */
   int nxExt = pre->nx + pre->lt + pre->rt;
   int nyExt = pre->ny + pre->dn + pre->up;
   int nf    = pre->nf;
   int *ax   = (int *)malloc(sizeof(int) * 5 * (size_t)nxExt);
   int *ay   = (int *)malloc(sizeof(int) * 5 * (size_t)nyExt);
   or_patch_axis(pre, post, nxp, nyp, 0, ax, ax + nxExt, ax + 2 * nxExt, ax + 3 * nxExt, ax + 4 * nxExt);
   or_patch_axis(pre, post, nxp, nyp, 1, ay, ay + nyExt, ay + 2 * nyExt, ay + 3 * nyExt, ay + 4 * nyExt);
   int nxExtPost = post->nx + post->lt + post->rt;
   int nyExtPost = post->ny + post->dn + post->up;
   int numPatches = nxExt * nyExt * nf;
   for (int p = 0; p < numPatches; p++) {
      int xi = kxPos(p, nxExt, nyExt, nf);
      int yi = kyPos(p, nxExt, nyExt, nf);
      if (patches) {
         patches[p].nx     = (uint16_t)ax[xi];
         patches[p].ny     = (uint16_t)ay[yi];
         patches[p].offset = (uint32_t)or_kIndex(ax[nxExt + xi], ay[nyExt + yi], 0, nxp, nyp, nfp);
      }
      if (gSynPatchStart) {
         gSynPatchStart[p] = (uint64_t)or_kIndex(
               ax[2 * nxExt + xi], ay[2 * nyExt + yi], 0, post->nx, post->ny, post->nf);
      }
      if (aPostOffset) {
         aPostOffset[p] = (uint64_t)or_kIndex(
               ax[3 * nxExt + xi], ay[3 * nyExt + yi], 0, nxExtPost, nyExtPost, post->nf);
      }
      if (unshrunkenStart) {
         unshrunkenStart[p] = (uint64_t)or_kIndex(
               ax[4 * nxExt + xi], ay[4 * nyExt + yi], 0, nxExtPost, nyExtPost, post->nf);
      }
   }
   free(ax);
   free(ay);
}

void or_num_kernels(const or_loc *pre, const or_loc *post, int *kx, int *ky, int *kf) {
   /* PatchGeometry.cpp:53-55 */
   *kx = pre->nx > post->nx ? pre->nx / post->nx : 1;
   *ky = pre->ny > post->ny ? pre->ny / post->ny : 1;
   *kf = pre->nf;
}

int or_data_index(const or_loc *pre, int numKernelsX, int numKernelsY, int patchIndex) {
   /* Weights::calcDataIndexFromPatchIndex, Weights.cpp:205-231 (shared) */
   int nxExt = pre->nx + pre->lt + pre->rt;
   int nyExt = pre->ny + pre->dn + pre->up;
   int nf    = pre->nf;
   int xIndex = kxPos(patchIndex, nxExt, nyExt, nf);
   xIndex     = (xIndex - pre->lt) % numKernelsX;
   if (xIndex < 0) xIndex += numKernelsX;
   int yIndex = kyPos(patchIndex, nxExt, nyExt, nf);
   yIndex     = (yIndex - pre->up) % numKernelsY;
   if (yIndex < 0) yIndex += numKernelsY;
   int fIndex = featureIndex(patchIndex, nf);
   return or_kIndex(xIndex, yIndex, fIndex, numKernelsX, numKernelsY, nf);
}

int or_transpose_item_index(const or_loc *pre, const or_loc *post, int nxp, int nyp, int nfp,
      int kernelIndexPre, int itemInPatchPre) {
   /* PatchGeometry::setTransposeItemIndices, PatchGeometry.cpp:281-332 */
   int nKx, nKy, nKf;
   or_num_kernels(pre, post, &nKx, &nKy, &nKf);
   int const xStride  = pre->nx > post->nx ? pre->nx / post->nx : 1;
   int const yStride  = pre->ny > post->ny ? pre->ny / post->ny : 1;
   int const xTStride = post->nx > pre->nx ? post->nx / pre->nx : 1;
   int const yTStride = post->ny > pre->ny ? post->ny / pre->ny : 1;
   int const patchSizeXPost = nxp * xStride / xTStride;
   int const patchSizeYPost = nyp * yStride / yTStride;

   int const kernelIndexXPre  = kxPos(kernelIndexPre, nKx, nKy, nKf);
   int const itemInPatchXPre  = kxPos(itemInPatchPre, nxp, nyp, nfp);
   int const itemInPatchXConj = nxp - 1 - itemInPatchXPre;
   int const extentOneSideX   = (nxp - 1) * xStride / 2;
   int kernelStartX           = (kernelIndexXPre - extentOneSideX) % xStride;
   if (kernelStartX < 0) kernelStartX += xStride;
   int const itemInPatchXPost = (xStride * itemInPatchXConj + kernelStartX) / xTStride;

   int const kernelIndexYPre  = kyPos(kernelIndexPre, nKx, nKy, nKf);
   int const itemInPatchYPre  = kyPos(itemInPatchPre, nxp, nyp, nfp);
   int const itemInPatchYConj = nyp - 1 - itemInPatchYPre;
   int const extentOneSideY   = (nyp - 1) * yStride / 2;
   int kernelStartY           = (kernelIndexYPre - extentOneSideY) % yStride;
   if (kernelStartY < 0) kernelStartY += yStride;
   int const itemInPatchYPost = (yStride * itemInPatchYConj + kernelStartY) / yTStride;

   int const kernelIndexFPre = featureIndex(kernelIndexPre, nKf);
   return or_kIndex(itemInPatchXPost, itemInPatchYPost, kernelIndexFPre,
         patchSizeXPost, patchSizeYPost, nKf);
}

void or_transpose_weights_shared(const or_loc *pre, const or_loc *post, int nxp, int nyp, int nfp,
      const float *preWeights, float *postWeights) {
   /* TransposeWeights::transposeShared, src/utils/TransposeWeights.cpp:56-113.
    * Post weights geometry: pre = post layer, post = pre layer. */
   int nKx, nKy, nKf;
   or_num_kernels(pre, post, &nKx, &nKy, &nKf);
   int const numPatchesPre = nKx * nKy * nKf;
   int const patchSizePre  = nxp * nyp * nfp;
   int pKx, pKy, pKf;
   or_num_kernels(post, pre, &pKx, &pKy, &pKf);
   int const nxpPost = or_calcPostPatchSize(nxp, pre->nx, post->nx);
   int const nypPost = or_calcPostPatchSize(nyp, pre->ny, post->ny);
   int const patchSizePost = nxpPost * nypPost * pre->nf;
   for (int patchIndexPre = 0; patchIndexPre < numPatchesPre; patchIndexPre++) {
      for (int itemInPatchPre = 0; itemInPatchPre < patchSizePre; itemInPatchPre++) {
         int itemInPatchPost =
               or_transpose_item_index(pre, post, nxp, nyp, nfp, patchIndexPre, itemInPatchPre);
         int patchIndexXPost = 0;
         if (pKx > 1) {
            int const itemInPatchXPre = kxPos(itemInPatchPre, nxp, nyp, nfp);
            patchIndexXPost           = -(nxp - pKx) / 2 + itemInPatchXPre;
            patchIndexXPost %= pKx;
            if (patchIndexXPost < 0) patchIndexXPost += pKx;
         }
         int patchIndexYPost = 0;
         if (pKy > 1) {
            int const itemInPatchYPre = kyPos(itemInPatchPre, nxp, nyp, nfp);
            patchIndexYPost           = -(nyp - pKy) / 2 + itemInPatchYPre;
            patchIndexYPost %= pKy;
            if (patchIndexYPost < 0) patchIndexYPost += pKy;
         }
         int patchIndexFPost = featureIndex(itemInPatchPre, nfp);
         int patchIndexPost  = or_kIndex(patchIndexXPost, patchIndexYPost, patchIndexFPost, pKx, pKy, pKf);
         postWeights[(size_t)patchIndexPost * patchSizePost + itemInPatchPost] =
               preWeights[(size_t)patchIndexPre * patchSizePre + itemInPatchPre];
      }
   }
}

/* ---------------------------------------------------------------- delivery */

void or_recv_post(const or_loc *S, const or_loc *T, int nxp, int nyp,
      const float *preActivity, const float *W, float *gsyn, float dtFactor, int num_threads) {
   /* PostsynapticPerspectiveConvolveDelivery::deliver, :56-151.  The post
    * weights' geometry has pre = target, post = source (WeightsPair.cpp:188-205). */
   int const nfp               = S->nf;
   int const numPostRestricted = locNumRes(T);
   int const sy                = (S->nx + S->lt + S->rt) * S->nf;
   int const syp               = nxp * nfp;
   int const numPerStride      = nxp * nfp;
   int const neuronIndexStride = T->nf < 4 ? 1 : T->nf / 4;
   int const patchSize         = nxp * nyp * nfp;
   int const nxExtT            = T->nx + T->lt + T->rt;
   int const nyExtT            = T->ny + T->dn + T->up;
   int const nxExtS            = S->nx + S->lt + S->rt;
   int const nyExtS            = S->ny + S->dn + S->up;
   int const sourceNumExtended = nxExtS * nyExtS * S->nf;
   int nKx, nKy, nKf;
   or_num_kernels(T, S, &nKx, &nKy, &nKf);
   int *ux = (int *)malloc(sizeof(int) * (size_t)nxExtT);
   int *uy = (int *)malloc(sizeof(int) * (size_t)nyExtT);
   or_patch_axis(T, S, nxp, nyp, 0, NULL, NULL, NULL, NULL, ux);
   or_patch_axis(T, S, nxp, nyp, 1, NULL, NULL, NULL, NULL, uy);
#ifdef _OPENMP
   if (num_threads <= 0) num_threads = omp_get_max_threads();
#else
   (void)num_threads;
#endif
   for (int b = 0; b < T->nbatch; b++) {
      float const *activityBatch = preActivity + (size_t)b * sourceNumExtended;
      float *gSynPatchHeadBatch  = gsyn + (size_t)b * numPostRestricted;
      for (int ky = 0; ky < nyp; ky++) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(num_threads)
#endif
         for (int feature = 0; feature < neuronIndexStride; feature++) {
            for (int idx = feature; idx < numPostRestricted; idx += neuronIndexStride) {
               float *gSyn = gSynPatchHeadBatch + idx;
               int kTargetExt = or_kIndexExtended(idx, T->nx, T->ny, T->nf, T->lt, T->rt, T->dn, T->up);
               int xe = kxPos(kTargetExt, nxExtT, nyExtT, T->nf);
               int ye = kyPos(kTargetExt, nxExtT, nyExtT, T->nf);
               int startSourceExt = or_kIndex(ux[xe], uy[ye], 0, nxExtS, nyExtS, S->nf);
               float const *a     = activityBatch + startSourceExt + ky * sy;
               int dataIndex      = or_data_index(T, nKx, nKy, kTargetExt);
               float const *weightValues = W + (size_t)dataIndex * patchSize + ky * syp;
               float dv = 0.0f;
               for (int k = 0; k < numPerStride; ++k) {
                  dv += a[k] * weightValues[k];
               }
               *gSyn += dtFactor * dv;
            }
         }
      }
   }
   free(ux);
   free(uy);
}

void or_recv_pre(const or_loc *pre, const or_loc *post, int nxp, int nyp,
      const float *preActivity, const or_entry *entries, const int64_t *numActive,
      const float *W, float *gsyn, float dtFactor, int num_threads) {
   /* PresynapticPerspectiveConvolveDelivery::deliver, :61-192 */
   int const nfp                = post->nf;
   int const nxPreExtended      = pre->nx + pre->rt + pre->rt; /* as written at :73 */
   int const nyPreExtended      = pre->ny + pre->dn + pre->up;
   int const numPreExtended     = nxPreExtended * nyPreExtended * pre->nf;
   int const numPreExtTrue      = locNumExt(pre);
   int const numPostRestricted  = locNumRes(post);
   int const sy                 = post->nx * post->nf;
   int const syw                = nxp * nfp;
   int const patchSize          = nxp * nyp * nfp;
   int nKx, nKy, nKf;
   or_num_kernels(pre, post, &nKx, &nKy, &nKf);
   or_patch *patches = (or_patch *)malloc(sizeof(or_patch) * (size_t)numPreExtTrue);
   uint64_t *gSynPatchStart = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)numPreExtTrue);
   or_patch_geometry(pre, post, nxp, nyp, nfp, patches, gSynPatchStart, NULL, NULL);
   int *dataIndex = (int *)malloc(sizeof(int) * (size_t)numPreExtTrue);
   for (int p = 0; p < numPreExtTrue; p++) dataIndex[p] = or_data_index(pre, nKx, nKy, p);

#ifdef _OPENMP
   if (num_threads <= 0) num_threads = omp_get_max_threads();
#else
   num_threads = 1;
#endif
   /* BaseDelivery::allocateThreadGSyn: per-thread buffers only when > 1 thread */
   float *threadGSyn = NULL;
   if (num_threads > 1) {
      threadGSyn = (float *)malloc(sizeof(float) * (size_t)num_threads * numPostRestricted);
   }
   for (int b = 0; b < pre->nbatch; b++) {
      size_t batchOffset         = (size_t)b * numPreExtended;
      float const *activityBatch = preActivity ? preActivity + batchOffset : NULL;
      float *gSynPatchHeadBatch  = gsyn + (size_t)b * numPostRestricted;
      or_entry const *activeIndicesBatch = NULL;
      int numNeurons;
      if (entries) {
         activeIndicesBatch = entries + batchOffset;
         numNeurons         = (int)numActive[b];
      }
      else {
         numNeurons = numPreExtTrue;
      }
      if (threadGSyn) {
         memset(threadGSyn, 0, sizeof(float) * (size_t)num_threads * numPostRestricted);
      }
      for (int y = 0; y < nyp; y++) {
#ifdef _OPENMP
#pragma omp parallel for schedule(guided) num_threads(num_threads)
#endif
         for (int idx = 0; idx < numNeurons; idx++) {
            int kPreExt;
            float a;
            if (activeIndicesBatch) {
               kPreExt = (int)activeIndicesBatch[idx].index;
            }
            else {
               kPreExt = idx;
            }
            or_patch const *patch = &patches[kPreExt];
            if (y >= patch->ny) continue;
            a = activeIndicesBatch ? activeIndicesBatch[idx].value : activityBatch[kPreExt];
            if (a == 0.0f) continue;
            a *= dtFactor;
            float *gSynPatchHead = gSynPatchHeadBatch;
#ifdef _OPENMP
            if (threadGSyn) {
               gSynPatchHead = threadGSyn + (size_t)omp_get_thread_num() * numPostRestricted;
            }
#endif
            float *postPatchStart = &gSynPatchHead[gSynPatchStart[kPreExt]];
            int const nk          = patch->nx * nfp;
            float const *weightDataStart = W + (size_t)dataIndex[kPreExt] * patchSize + patch->offset;
            float *v                     = postPatchStart + y * sy;
            float const *weightValues    = weightDataStart + y * syw;
            for (int k = 0; k < nk; k++) {
               v[k] += a * weightValues[k];
            }
         }
      }
      if (threadGSyn) { /* BaseDelivery::accumulateThreadGSyn, :117-134 */
         for (int ti = 0; ti < num_threads; ti++) {
            float const *threadData = threadGSyn + (size_t)ti * numPostRestricted;
#ifdef _OPENMP
#pragma omp parallel for num_threads(num_threads)
#endif
            for (int ni = 0; ni < numPostRestricted; ni++) {
               gSynPatchHeadBatch[ni] += threadData[ni];
            }
         }
      }
   }
   free(threadGSyn);
   free(patches);
   free(gSynPatchStart);
   free(dataIndex);
}

void or_ident_deliver(const or_loc *loc, const float *preActivity, float *gsyn) {
   /* IdentDelivery::deliver dense branch, IdentDelivery.cpp:146-162 */
   int const nx = loc->nx, ny = loc->ny, nf = loc->nf;
   int const nxPreExtended  = nx + loc->lt + loc->rt;
   int const nyPreExtended  = ny + loc->dn + loc->up;
   int const numPreExtended = nxPreExtended * nyPreExtended * nf;
   int const numPostRestricted = nx * ny * nf;
   for (int b = 0; b < loc->nbatch; b++) {
      float const *preActivityBuffer = preActivity + (size_t)b * numPreExtended;
      float *postGSynBuffer          = gsyn + (size_t)b * numPostRestricted;
      int const nk                   = nx * nf;
      for (int y = 0; y < ny; y++) {
         int preLineIndex = or_kIndex(loc->lt, y + loc->up, 0, nxPreExtended, nyPreExtended, nf);
         float const *preActivityLine = &preActivityBuffer[preLineIndex];
         int postLineIndex            = or_kIndex(0, y, 0, nx, ny, nf);
         float *postGSynLine          = &postGSynBuffer[postLineIndex];
         for (int k = 0; k < nk; k++) {
            postGSynLine[k] += preActivityLine[k];
         }
      }
   }
}

void or_active_indices(const or_loc *loc, const float *A, or_entry *entries, int64_t *numActive) {
   /* DataStore::updateActiveIndices, DataStore.cpp:46-66 */
   int const numExt = locNumExt(loc);
   for (int b = 0; b < loc->nbatch; b++) {
      float const *activity = A + (size_t)b * numExt;
      or_entry *activeIndices = entries + (size_t)b * numExt;
      int n = 0;
      for (int kex = 0; kex < numExt; kex++) {
         float a = activity[kex];
         if (a != 0.0f) {
            activeIndices[n].index = (uint32_t)kex;
            activeIndices[n].value = a;
            n++;
         }
      }
      numActive[b] = n;
   }
}

/* ---------------------------------------------------------------- layer updates */

void or_gsyn_accumulate(int numNeuronsAcrossBatch, int numChannels, const float *coeffs,
      const float *layerInput, float *out) {
   /* GSynAccumulator.kpp:21-42 */
   for (int index = 0; index < numNeuronsAcrossBatch; index++) {
      float value = 0.0f;
      for (int ch = 0; ch < numChannels; ch++) {
         value += coeffs[ch] * layerInput[(size_t)ch * numNeuronsAcrossBatch + index];
      }
      out[index] = value;
   }
}

void or_update_lca(const or_loc *loc, int selfInteract, const double *dtAdapt, float tau,
      const float *accumulatedGSyn, const float *A, float *V) {
   /* HyPerLCAInternalStateBuffer.kpp:45-60 */
   int const numNeurons = locNumRes(loc);
   int const numExt     = locNumExt(loc);
   for (int index = 0; index < numNeurons * loc->nbatch; index++) {
      int b = index / numNeurons;
      int k = index % numNeurons;
      float exp_tau       = (float)exp(-dtAdapt[b] / (double)tau);
      float *VBatch       = V + (size_t)b * numNeurons;
      float const gSyn    = accumulatedGSyn[(size_t)b * numNeurons + k];
      float const *ABatch = A + (size_t)b * numExt;
      int kex   = or_kIndexExtended(k, loc->nx, loc->ny, loc->nf, loc->lt, loc->rt, loc->dn, loc->up);
      VBatch[k] = exp_tau * VBatch[k] + (1.0f - exp_tau) * (gSyn + (float)selfInteract * ABatch[kex]);
   }
}

void or_activity_copy(const or_loc *loc, const float *V, float *A) {
   /* HyPerActivityBuffer::updateBufferCPU, HyPerActivityBuffer.cpp:36-54 */
   int const numNeurons = locNumRes(loc);
   int const numExt     = locNumExt(loc);
   for (int b = 0; b < loc->nbatch; b++) {
      for (int k = 0; k < numNeurons; k++) {
         int kex = or_kIndexExtended(k, loc->nx, loc->ny, loc->nf, loc->lt, loc->rt, loc->dn, loc->up);
         A[(size_t)b * numExt + kex] = V[(size_t)b * numNeurons + k];
      }
   }
}

void or_apply_vthresh(const or_loc *loc, const float *V, float VThresh, float AMin, float AShift,
      float VWidth, float *activity) {
   /* ANNActivityBuffer::applyVThresh, ANNActivityBuffer.cpp:490-528 */
   if (!(VThresh > -FLT_MAX)) return;
   int const numNeurons = locNumRes(loc);
   int const numExt     = locNumExt(loc);
   for (int kbatch = 0; kbatch < numNeurons * loc->nbatch; kbatch++) {
      int b = kbatch / numNeurons;
      int k = kbatch % numNeurons;
      float const *VBatch  = V + (size_t)b * numNeurons;
      float *activityBatch = activity + (size_t)b * numExt;
      int kex = or_kIndexExtended(k, loc->nx, loc->ny, loc->nf, loc->lt, loc->rt, loc->dn, loc->up);
      if (VBatch[k] < VThresh) {
         activityBatch[kex] = AMin;
      }
      else if (VBatch[k] < VThresh + VWidth) {
         activityBatch[kex] = AMin + (VThresh + VWidth - AShift - AMin) * (VBatch[k] - VThresh) / VWidth;
      }
      else {
         activityBatch[kex] -= AShift;
      }
   }
}

void or_apply_amax(const or_loc *loc, float AMax, float *activity) {
   /* ANNActivityBuffer::applyAMax, ANNActivityBuffer.cpp:530-556 */
   if (!(AMax < FLT_MAX)) return;
   int const numNeurons = locNumRes(loc);
   int const numExt     = locNumExt(loc);
   for (int kbatch = 0; kbatch < numNeurons * loc->nbatch; kbatch++) {
      int b = kbatch / numNeurons;
      int k = kbatch % numNeurons;
      float *activityBatch = activity + (size_t)b * numExt;
      int kex = or_kIndexExtended(k, loc->nx, loc->ny, loc->nf, loc->lt, loc->rt, loc->dn, loc->up);
      if (activityBatch[kex] > AMax) {
         activityBatch[kex] = AMax;
      }
   }
}

int or_set_vertices(float VThresh, float AMin, float AMax, float AShift, float VWidth,
      float *verticesV, float *verticesA, float *slopes) {
   /* ANNActivityBuffer::setVertices (:227-370) and setSlopes (:372-397) */
   if (VWidth < 0) {
      VThresh += VWidth;
      VWidth = -VWidth;
   }
   float vV[9], vA[9];
   int numVertices;
   float slopeNegInf, slopePosInf = 1.0f;
   if (VThresh <= -(float)0.999 * FLT_MAX) {
      numVertices = 1;
      vV[0]       = 0.0f;
      vA[0]       = -AShift;
      slopeNegInf = 1.0f;
   }
   else {
      if (VWidth == 0.0f && VThresh - AShift == AMin) {
         numVertices = 1;
         vV[0]       = VThresh;
         vA[0]       = AMin;
      }
      else {
         numVertices = 2;
         vV[0]       = VThresh;
         vV[1]       = VThresh + VWidth;
         vA[0]       = AMin;
         vA[1]       = VThresh + VWidth - AShift;
      }
      slopeNegInf = 0.0f;
   }
   if (AMax < (float)0.999 * FLT_MAX) {
      if (vA[numVertices - 1] < AMax) {
         float interval   = AMax - vA[numVertices - 1];
         vV[numVertices] = vV[numVertices - 1] + interval;
         vA[numVertices] = AMax;
         numVertices++;
      }
      else {
         int found = 0;
         int v;
         for (v = numVertices - 1; v >= 0; v--) {
            if (vA[v] < AMax) {
               found = 1;
               break;
            }
         }
         if (found) {
            float interval = AMax - vA[v];
            numVertices    = v + 1;
            vV[numVertices] = vV[v] + interval;
            vA[numVertices] = AMax;
            numVertices++;
         }
         else {
            numVertices = 1;
            if (slopeNegInf > 0) {
               float intervalA = vA[0] - AMax;
               float intervalV = (float)(intervalA / slopeNegInf);
               vV[0]           = vV[0] - intervalV;
               vA[0]           = AMax;
            }
            else {
               vV[0]       = 0.0f;
               vA[0]       = AMax;
               slopeNegInf = 0;
            }
         }
      }
      slopePosInf = 0.0f;
   }
   for (int i = 0; i < numVertices; i++) {
      verticesV[i] = vV[i];
      verticesA[i] = vA[i];
   }
   slopes[0] = slopeNegInf;
   for (int k = 1; k < numVertices; k++) {
      float V1 = verticesV[k - 1];
      float V2 = verticesV[k];
      if (V1 != V2) {
         slopes[k] = (verticesA[k] - verticesA[k - 1]) / (V2 - V1);
      }
      else {
         slopes[k] = verticesA[k] > verticesA[k - 1] ? INFINITY
                     : verticesA[k] < verticesA[k - 1] ? -INFINITY
                                                       : NAN;
      }
   }
   slopes[numVertices] = slopePosInf;
   return numVertices;
}

void or_apply_vertices(const or_loc *loc, int numVertices, const float *verticesV,
      const float *verticesA, const float *slopes, const float *V, float *A) {
   /* applyVerticesANNActivityBufferOnCPU, ANNActivityBuffer.kpp:36-77 */
   int const numNeurons = locNumRes(loc);
   int const numExt     = locNumExt(loc);
   for (int index = 0; index < numNeurons * loc->nbatch; index++) {
      int last            = numVertices - 1;
      int b               = index / numNeurons;
      int k               = index % numNeurons;
      float const *VBatch = V + (size_t)b * numNeurons;
      float *ABatch       = A + (size_t)b * numExt;
      int kex = or_kIndexExtended(k, loc->nx, loc->ny, loc->nf, loc->lt, loc->rt, loc->dn, loc->up);
      float potential = VBatch[k];
      float activity  = 0.0f;
      if (potential < verticesV[0]) {
         activity = verticesA[0] + slopes[0] * (potential - verticesV[0]);
      }
      else if (potential >= verticesV[last]) {
         activity = verticesA[last] + slopes[numVertices] * (potential - verticesV[last]);
      }
      else {
         for (int v = 0; v < last; v++) {
            if (potential < verticesV[v]) break;
            if (potential == verticesV[v]) {
               activity = verticesA[v];
            }
            else if (potential > verticesV[v] && potential < verticesV[v + 1]) {
               activity = verticesA[v] + slopes[v + 1] * (potential - verticesV[v]);
            }
         }
      }
      ABatch[kex] = activity;
   }
}

void or_errscale(int n, float errScale, float *V) {
   /* ErrScaleInternalStateBuffer::updateBufferCPU, :41-53 */
   for (int k = 0; k < n; k++) V[k] *= errScale;
}

/* ---------------------------------------------------------------- plasticity */

void or_hebbian_update_dw(const or_loc *pre, const or_loc *post, int nxp, int nyp, int nfp,
      float dWMax, const float *preActivity, const float *postActivity, float *dw, int64_t *counts) {
   /* HebbianUpdater::update_dW (:476-538), shared branch, and updateInd_dW (:564-627) */
   int const numPreExt  = locNumExt(pre);
   int const numPostExt = locNumExt(post);
   int xCellSize = pre->nx > post->nx ? pre->nx / post->nx : 1; /* zUnitCellSize */
   int yCellSize = pre->ny > post->ny ? pre->ny / post->ny : 1;
   int nxExt     = pre->nx + pre->lt + pre->rt;
   int nyExt     = pre->ny + pre->up + pre->dn;
   int nf        = pre->nf;
   int nKx, nKy, nKf;
   or_num_kernels(pre, post, &nKx, &nKy, &nKf);
   int numKernels = nKx * nKy * nKf;
   int const patchSizeOverall = nxp * nyp * nfp;
   or_patch *patches = (or_patch *)malloc(sizeof(or_patch) * (size_t)numPreExt);
   uint64_t *aPostOffset = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)numPreExt);
   or_patch_geometry(pre, post, nxp, nyp, nfp, patches, NULL, aPostOffset, NULL);
   int const syp = nxp * nfp;
   int const sya = post->nf * (post->nx + post->lt + post->rt);
   for (int b = 0; b < pre->nbatch; b++) {
#ifdef _OPENMP
#pragma omp parallel for
#endif
      for (int kernelIdx = 0; kernelIdx < numKernels; kernelIdx++) {
         int kxCellIdx = kxPos(kernelIdx, xCellSize, yCellSize, nf);
         int kyCellIdx = kyPos(kernelIdx, xCellSize, yCellSize, nf);
         int kfIdx     = featureIndex(kernelIdx, nf);
         for (int kyIdx = kyCellIdx; kyIdx < nyExt; kyIdx += yCellSize) {
            for (int kxIdx = kxCellIdx; kxIdx < nxExt; kxIdx += xCellSize) {
               int kExt = or_kIndex(kxIdx, kyIdx, kfIdx, nxExt, nyExt, nf);
               /* updateInd_dW */
               float preact = preActivity[(size_t)b * numPreExt + kExt];
               if (preact == 0.0f) continue;
               or_patch const *patch = &patches[kExt];
               int ny = patch->ny;
               int nk = patch->nx * nfp;
               if (ny == 0 || nk == 0) continue;
               float const *postactRef = postActivity + (size_t)b * numPostExt + aPostOffset[kExt];
               int dataIndex = or_data_index(pre, nKx, nKy, kExt);
               float *dwdata = dw + (size_t)dataIndex * patchSizeOverall + patch->offset;
               int64_t *activations = counts ? counts + (size_t)dataIndex * patchSizeOverall + patch->offset : NULL;
               int lineoffsetw = 0, lineoffseta = 0;
               for (int y = 0; y < ny; y++) {
                  for (int k = 0; k < nk; k++) {
                     float aPost = postactRef[lineoffseta + k];
                     if (activations) activations[lineoffsetw + k]++;
                     dwdata[lineoffsetw + k] += dWMax * preact * aPost;
                  }
                  lineoffsetw += syp;
                  lineoffseta += sya;
               }
            }
         }
      }
   }
   free(patches);
   free(aPostOffset);
}

void or_hebbian_normalize_apply(float *w, float *dw, const int64_t *counts, size_t n, int normalizeDw) {
   /* normalize_dW (:773-807) then updateWeights (:822-833) */
   if (normalizeDw) {
      for (size_t i = 0; i < n; i++) {
         long divisor = (long)counts[i];
         if (divisor != 0) {
            dw[i] /= divisor;
         }
         else {
            dw[i] = 0;
         }
      }
   }
   for (size_t i = 0; i < n; i++) w[i] += dw[i];
}

int or_momentum_factors(int method, float tau, float decay, float dwMax, float *dwFactor,
      float *wFactor) {
   /* MomentumUpdater::applyMomentum (MomentumUpdater.cpp:296-307) after the
    * parameter checks of checkTimeConstantTau (:82-112) and
    * ioParam_momentumDecay (:138-145).  method 0 viscosity, 1 simple, 2 alex.
    * Returns nonzero for a parameter the reference rejects. */
   if (decay < 0.0f || decay > 1.0f) return 1;
   switch (method) {
      case 0:
         if (tau < 0) return 1;
         *dwFactor = tau ? expf(-1.0f / tau) : 0.0f;
         *wFactor  = decay;
         return 0;
      case 1:
         if (tau < 0 || tau >= 1) return 1;
         *dwFactor = tau;
         *wFactor  = decay;
         return 0;
      case 2:
         if (tau < 0 || tau >= 1) return 1;
         *dwFactor = tau;
         *wFactor  = decay * dwMax;
         return 0;
      default: return 1;
   }
}

int or_momentum_factors_deprecated(int method, float momentumTau, float decay, float dwMax,
      float *dwFactor, float *wFactor) {
   /* MomentumUpdater::applyMomentumDeprecated (MomentumUpdater.cpp:333-343):
    * momentumTau as given, no range check; momentumDecay checked by
    * ioParam_momentumDecay (:138-145). */
   if (decay < 0.0f || decay > 1.0f) return 1;
   switch (method) {
      case 0: *dwFactor = expf(-1.0f / momentumTau); *wFactor = decay; return 0;
      case 1: *dwFactor = momentumTau; *wFactor = decay; return 0;
      case 2: *dwFactor = momentumTau; *wFactor = decay * dwMax; return 0;
      default: return 1;
   }
}

void or_momentum_apply(float *w, float *dw, float *prevDw, const int64_t *counts, size_t n,
      int normalizeDw, float dwFactor, float wFactor, int deprecatedTau) {
   /* HebbianUpdater::normalize_dW (:773-807), then MomentumUpdater::updateWeights
    * (MomentumUpdater.cpp:276-294): applyMomentum (:309-329, or the deprecated
    * momentumTau form :346-360), prev_dW = dW, then HebbianUpdater::updateWeights
    * (w += dW, HebbianUpdater.cpp:822-833). */
   if (normalizeDw) {
      for (size_t i = 0; i < n; i++) {
         long divisor = (long)counts[i];
         if (divisor != 0) {
            dw[i] /= divisor;
         }
         else {
            dw[i] = 0;
         }
      }
   }
   for (size_t k = 0; k < n; k++) {
      if (deprecatedTau) {
         dw[k] += dwFactor * prevDw[k] - wFactor * w[k];
      }
      else {
         dw[k] *= 1 - dwFactor;
         dw[k] += dwFactor * prevDw[k];
         dw[k] -= wFactor * w[k];
      }
   }
   memcpy(prevDw, dw, sizeof(float) * n);
   for (size_t i = 0; i < n; i++) w[i] += dw[i];
}

void or_update_momentum_lca(const or_loc *loc, int selfInteract, float momentumRate,
      const double *dtAdapt, float tau, const float *accumulatedGSyn, const float *A,
      float *prevDrive, float *V) {
   /* MomentumLCAInternalStateBuffer.kpp:44-63 */
   int const numNeurons = locNumRes(loc);
   int const numExt     = locNumExt(loc);
   for (int index = 0; index < numNeurons * loc->nbatch; index++) {
      int b = index / numNeurons;
      int k = index % numNeurons;
      float exp_tau         = (float)exp(-dtAdapt[b] / (double)tau);
      float *VBatch         = V + (size_t)b * numNeurons;
      float const gSyn      = accumulatedGSyn[(size_t)b * numNeurons + k];
      float *prevDriveBatch = prevDrive + (size_t)b * numNeurons;
      float const *ABatch   = A + (size_t)b * numExt;
      int kex = or_kIndexExtended(k, loc->nx, loc->ny, loc->nf, loc->lt, loc->rt, loc->dn, loc->up);
      float currentDrive = (1.0f - exp_tau) * (gSyn + (float)selfInteract * ABatch[kex]);
      VBatch[k]          = exp_tau * VBatch[k] + currentDrive + momentumRate * prevDriveBatch[k];
      prevDriveBatch[k]  = currentDrive;
   }
}

void or_update_ista(const or_loc *loc, float VThresh, const double *dtAdapt, float tau,
      const float *accumulatedGSyn, const float *A, float *V) {
   /* ISTAInternalStateBuffer.kpp:34-57: V += (dt/tau)*(G - VThresh*sign(A)) */
   int const numNeurons = locNumRes(loc);
   int const numExt     = locNumExt(loc);
   for (int index = 0; index < numNeurons * loc->nbatch; index++) {
      int b = index / numNeurons;
      int k = index % numNeurons;
      float *VBatch       = V + (size_t)b * numNeurons;
      float const gSyn    = accumulatedGSyn[(size_t)b * numNeurons + k];
      float const *ABatch = A + (size_t)b * numExt;
      int kex = or_kIndexExtended(k, loc->nx, loc->ny, loc->nf, loc->lt, loc->rt, loc->dn, loc->up);
      float sign = 0.0f;
      if (ABatch[kex] != 0.0f) {
         sign = ABatch[kex] / fabsf(ABatch[kex]);
      }
      float const rate = (float)dtAdapt[b] / tau;
      float const drive = gSyn - VThresh * sign;
      VBatch[k] += rate * drive;
   }
}

void or_normalize_l2(float *w, int numPatches, int patchSize, float strength, float minL2) {
   /* NormalizeL2::normalizeWeights non-individual-arbor branch, NormalizeL2.cpp:93-118 */
   float scaleFactor = 1.0f * strength;
   for (int patchindex = 0; patchindex < numPatches; patchindex++) {
      float *data = w + (size_t)patchindex * patchSize;
      float sumsq = 0.0f;
      for (int k = 0; k < patchSize; k++) { /* NormalizeBase::accumulateSumSquared :184-193 */
         float x = data[k];
         sumsq += x * x;
      }
      float l2norm = sqrtf(sumsq);
      if (fabsf(sumsq) <= minL2) {
         break;
      }
      float multiplier = scaleFactor / l2norm;
      for (int k = 0; k < patchSize; k++) data[k] *= multiplier; /* normalizePatch */
   }
}

void or_normalize_l2_f64(float *w, int numPatches, int patchSize, double strength, double minL2) {
   for (int p = 0; p < numPatches; p++) {
      float *data = w + (size_t)p * patchSize;
      double s = 0.0;
      for (int k = 0; k < patchSize; k++) s += (double)data[k] * (double)data[k];
      if (s <= minL2) continue;
      double m = strength / sqrt(s);
      for (int k = 0; k < patchSize; k++) data[k] = (float)((double)data[k] * m);
   }
}

double or_cost_sum(const or_loc *loc, const float *buf, int extended, int batch, int kind,
      double p0, double p1) {
   /* CostFunctionSum::calculateSumNoMask, CostFunctionSum.hpp:87-104 */
   int const numNeurons = locNumRes(loc);
   int const stride     = extended ? locNumExt(loc) : numNeurons;
   float const *buffer  = buf + (size_t)batch * stride;
   double threshPlusWidth = p0 + p1;
   double maxCost = 0.5 * threshPlusWidth, quadCoeff = 0.5 / threshPlusWidth;
   double sum = 0.0;
   for (int k = 0; k < numNeurons; ++k) {
      int kex = extended ? or_kIndexExtended(k, loc->nx, loc->ny, loc->nf, loc->lt, loc->rt, loc->dn, loc->up) : k;
      double v = (double)buffer[kex];
      double c;
      switch (kind) {
         case 0: c = v * v; break;
         case 1: c = fabs(v); break;
         case 2: {
            double absx = fabs(v);
            c = absx < threshPlusWidth ? absx * (1.0 - quadCoeff * absx) : maxCost;
            break;
         }
         default: c = fabs(v) > p0 ? 1.0 : 0.0; break;
      }
      sum += c;
   }
   return sum;
}

/* ------------------------------------------------- host-side loop controllers */

void or_timescale_init(or_timescale *ts, int nbatch, double baseMax, double baseMin) {
   /* AdaptiveTimeScaleController.cpp:17-47 */
   for (int b = 0; b < nbatch; b++) {
      ts[b].timeScale     = baseMin;
      ts[b].timeScaleMax  = baseMax;
      ts[b].timeScaleTrue = -1.0;
   }
}

void or_timescale_calc(or_timescale *ts, int nbatch, const double *values, double baseMax,
      double baseMin, double tauFactor, double growthFactor, double *dt_out) {
   /* AdaptiveTimeScaleController::calcTimesteps, AdaptiveTimeScaleController.cpp:60-88 */
   for (int b = 0; b < nbatch; b++) {
      double const E_0  = ts[b].timeScaleTrue; /* the copy taken at :67 */
      ts[b].timeScaleTrue = values[b];
      double const E_dt = values[b];
      if (E_dt != E_0) {
         double const dE_dt_scaled = (E_0 - E_dt) / ts[b].timeScale;
         if (dE_dt_scaled < 0.0 || E_0 <= 0.0 || E_dt <= 0.0) {
            ts[b].timeScale    = baseMin;
            ts[b].timeScaleMax = baseMax;
         }
         else {
            double const tau_eff_scaled = E_0 / dE_dt_scaled;
            ts[b].timeScale = tauFactor * tau_eff_scaled;
            if (ts[b].timeScale >= ts[b].timeScaleMax) {
               ts[b].timeScale = ts[b].timeScaleMax;
               ts[b].timeScaleMax *= (1 + growthFactor);
            }
         }
      }
      dt_out[b] = ts[b].timeScale;
   }
}

int or_need_update(double simTime, double deltaTime, double lastUpdateTime, double period) {
   /* LayerUpdateController::needUpdate, LayerUpdateController.cpp:318-335 (no trigger) */
   if (period <= 0) return 0;
   double const numUpdates = (simTime - lastUpdateTime) / period;
   double const closest    = fabs(numUpdates - nearbyint(numUpdates)) * period;
   return closest < 0.5 * deltaTime;
}

/* ---------------------------------------------------------------- LCA loop */

void or_lca_run(const or_lca_config *cfg, const float *weights, const float *input,
      const float *V0, int nsteps, float *V1_V, float *V1_A, float *recon,
      double *energy, int64_t *numActive) {
   or_loc const in = cfg->input;
   or_loc const R  = cfg->res;
   or_loc const L  = cfg->v1;
   or_loc recL     = R; /* Recon: same size as Residual, no halo */
   recL.lt = recL.rt = recL.dn = recL.up = 0;
   int const nb    = in.nbatch;
   int const nR    = locNumRes(&R), nRx = locNumExt(&R);
   int const nL    = locNumRes(&L), nLx = locNumExt(&L);
   int const nfIn  = in.nf;
   int const nxp = cfg->nxp, nyp = cfg->nyp;

   float *inA   = (float *)calloc((size_t)nb * nR, sizeof(float)); /* Input A, no halo */
   float *resG  = (float *)calloc((size_t)2 * nb * nR, sizeof(float));
   float *resV  = (float *)calloc((size_t)nb * nR, sizeof(float));
   float *resA  = (float *)calloc((size_t)nb * nRx, sizeof(float));
   float *v1G   = (float *)calloc((size_t)nb * nL, sizeof(float));
   float *v1Acc = (float *)calloc((size_t)nb * nL, sizeof(float));
   float *v1V   = (float *)calloc((size_t)nb * nL, sizeof(float));
   float *v1A   = (float *)calloc((size_t)nb * nLx, sizeof(float));
   or_entry *ent = (or_entry *)calloc((size_t)nb * nLx, sizeof(or_entry));
   int64_t *nAct = (int64_t *)calloc((size_t)nb, sizeof(int64_t));
   float *recG  = (float *)calloc((size_t)nb * nR, sizeof(float));
   float *recV  = (float *)calloc((size_t)nb * nR, sizeof(float));
   float *recA  = (float *)calloc((size_t)nb * nR, sizeof(float));
   float *resAcc = (float *)calloc((size_t)nb * nR, sizeof(float));
   double *dt   = (double *)malloc(sizeof(double) * (size_t)nb);
   (void)nfIn;

   memcpy(inA, input, sizeof(float) * (size_t)nb * nR);
   if (V0) memcpy(v1V, V0, sizeof(float) * (size_t)nb * nL);
   float const coeffs[2] = {1.0f, -1.0f};
   float const one[1]    = {1.0f};
   float tauScaled = (float)(cfg->tau / cfg->dt); /* initializeState :104-112 */
   float vv[8], va[8], sl[9];
   int nv = or_set_vertices(cfg->VThresh, cfg->AMin, cfg->AMax, cfg->AShift, cfg->VWidth, vv, va, sl);

#define V1_TRANSFER()                                                                       \
   do {                                                                                     \
      if (cfg->transfer_vertices) {                                                         \
         or_apply_vertices(&L, nv, vv, va, sl, v1V, v1A);                                   \
      }                                                                                     \
      else {                                                                                \
         or_activity_copy(&L, v1V, v1A);                                                    \
         or_apply_vthresh(&L, v1V, cfg->VThresh, cfg->AMin, cfg->AShift, cfg->VWidth, v1A); \
         or_apply_amax(&L, cfg->AMax, v1A);                                                 \
      }                                                                                     \
   } while (0)

   /* InitializeState: V1 A from V0; Residual/Recon start at zero. */
   V1_TRANSFER();
   for (int b = 0; b < nb; b++) dt[b] = cfg->dt;

   for (int step = 0; step < nsteps; step++) {
      /* phase 1: Residual (ANNErrorLayer) */
      memset(resG, 0, sizeof(float) * (size_t)2 * nb * nR);
      or_loc inNoHalo = in;
      inNoHalo.lt = inNoHalo.rt = inNoHalo.dn = inNoHalo.up = 0;
      or_ident_deliver(&inNoHalo, inA, resG);
      or_ident_deliver(&recL, recA, resG + (size_t)nb * nR);
      or_gsyn_accumulate(nb * nR, 2, coeffs, resG, resAcc);
      memcpy(resV, resAcc, sizeof(float) * (size_t)nb * nR);
      or_errscale(nb * nR, cfg->errScale, resV);
      or_activity_copy(&R, resV, resA);

      /* phase 2: V1 (HyPerLCALayer) */
      memset(v1G, 0, sizeof(float) * (size_t)nb * nL);
      {
         int nxpPost = or_calcPostPatchSize(nxp, L.nx, R.nx);
         int nypPost = or_calcPostPatchSize(nyp, L.ny, R.ny);
         (void)nxpPost;
         (void)nypPost;
         /* TransposeConn: post weights = original's pre weights, patch nxp x nyp x nf_in */
         or_recv_post(&R, &L, nxp, nyp, resA, weights, v1G, 1.0f, cfg->num_threads);
      }
      or_gsyn_accumulate(nb * nL, 1, one, v1G, v1Acc);
      or_update_lca(&L, cfg->selfInteract, dt, tauScaled, v1Acc, v1A, v1V);
      V1_TRANSFER();
      or_active_indices(&L, v1A, ent, nAct);

      /* phase 3: Recon (ANNLayer) */
      memset(recG, 0, sizeof(float) * (size_t)nb * nR);
      or_recv_pre(&L, &recL, nxp, nyp, v1A, ent, nAct, weights, recG, 1.0f, cfg->num_threads);
      or_gsyn_accumulate(nb * nR, 1, one, recG, recV);
      memcpy(recA, recV, sizeof(float) * (size_t)nb * nR);

      for (int b = 0; b < nb; b++) {
         if (energy) {
            energy[(size_t)step * nb + b] = 0.5 * or_cost_sum(&R, resA, 1, b, 0, 0, 0)
                  + (double)cfg->VThresh * or_cost_sum(&L, v1A, 1, b, 1, 0, 0);
         }
         if (numActive) numActive[(size_t)step * nb + b] = nAct[b];
      }
   }
#undef V1_TRANSFER
   if (V1_V) memcpy(V1_V, v1V, sizeof(float) * (size_t)nb * nL);
   if (V1_A) memcpy(V1_A, v1A, sizeof(float) * (size_t)nb * nLx);
   if (recon) memcpy(recon, recA, sizeof(float) * (size_t)nb * nR);
   free(inA); free(resG); free(resV); free(resA); free(v1G); free(v1Acc); free(v1V); free(v1A);
   free(ent); free(nAct); free(recG); free(recV); free(recA); free(resAcc); free(dt);
}
