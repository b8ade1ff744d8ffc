/*
 * pvoracle.h -- CPU restatement of OpenPV's synaptic-delivery + LCA path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the parity oracle: plain C (C11 +
 * optional OpenMP), compiled with -O3 -ffp-contract=off and no fast-math,
 * which is the reference's own Release arithmetic on x86-64 (strict IEEE
 * float, mul and add rounded separately, sequential reductions).  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * only as the checker or as the timed host baseline; the product path never
 * links it.
 *
 * Each function cites the reference file:line it restates (paths relative to
 * the reference checkout).  Pinning of each function (golden vectors, the
 * compiled reference pieces under oracle/_ref, closed forms) is listed in the
 * tests that exercise it.
 */
#ifndef PVORACLE_H_
#define PVORACLE_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
   int nbatch, nx, ny, nf;
   int lt, rt, dn, up;
} or_loc;

typedef struct {
   uint16_t nx, ny;
   uint32_t offset;
} or_patch;

typedef struct {
   uint32_t index;
   float value;
} or_entry;

/* --- index math (src/utils/conversions.hpp) --- */
int or_kIndex(int kx, int ky, int kf, int nx, int ny, int nf);                 /* :251-253 */
int or_kIndexExtended(int k, int nx, int ny, int nf, int lt, int rt, int dn, int up); /* :401-407 */
int or_requiredConvolveMargin(int nPre, int nPost, int patchSize);         /* utils/requiredConvolveMargin.hpp:14-40 */
int or_calcPostPatchSize(int prePatchSize, int nPre, int nPost);          /* components/PatchSize.cpp:98-130 */

/* --- PatchGeometry (src/components/PatchGeometry.cpp) --- */
int or_calcPatchStartInPost(int indexRestrictedPre, int patchSize, int nPre, int nPost); /* :334-389 */
void or_calcPatchData(                                                     /* :391-453 */
      int index, int numPreRestricted, int preStartBorder, int preEndBorder,
      int numPostRestricted, int postStartBorder, int postEndBorder, int patchSize,
      int *patchDim, int *patchStart, int *postPatchStartRestricted,
      int *postPatchStartExtended, int *postPatchUnshrunkenStart);

/* Full per-patch tables over the pre extended space (setPatchGeometry :179-279).
 * Arrays sized numPatches = nxExtPre*nyExtPre*nfPre. Any output may be NULL. */
void or_patch_geometry(
      const or_loc *pre, const or_loc *post, int nxp, int nyp, int nfp,
      or_patch *patches, uint64_t *gSynPatchStart, uint64_t *aPostOffset,
      uint64_t *unshrunkenStart);

/* Per-axis form of the same tables (x uses the pre lt/rt and, as the reference
 * does at :198-211, the post dn/up borders; y uses dn/up for both). */
void or_patch_axis(
      const or_loc *pre, const or_loc *post, int nxp, int nyp, int axis,
      int *dim, int *pstart, int *postStartRes, int *postStartExt, int *unshrunkenStart);

/* Weights::calcDataIndexFromPatchIndex for shared weights (Weights.cpp:205-231). */
int or_data_index(const or_loc *pre, int numKernelsX, int numKernelsY, int patchIndex);
void or_num_kernels(const or_loc *pre, const or_loc *post, int *kx, int *ky, int *kf); /* PatchGeometry.cpp:53-55 */

/* PatchGeometry::setTransposeItemIndices (:281-332) for one kernel/item. */
int or_transpose_item_index(const or_loc *pre, const or_loc *post, int nxp, int nyp, int nfp,
      int kernelIndexPre, int itemInPatchPre);
/* TransposeWeights::transposeShared (src/utils/TransposeWeights.cpp:56-113):
 * post-perspective weights from pre-perspective shared weights. */
void or_transpose_weights_shared(const or_loc *pre, const or_loc *post, int nxp, int nyp, int nfp,
      const float *preWeights, float *postWeights);

/* --- delivery --- */
/* PostsynapticPerspectiveConvolveDelivery::deliver
 * (src/delivery/PostsynapticPerspectiveConvolveDelivery.cpp:56-151).
 * `source` is the pre layer, `target` the post layer, patch = post-perspective
 * patch (nxpPost, nypPost, nfp = source nf); postWeights laid out
 * [numKernelsPost][nyp][nxp][nfp].  GSyn accumulated into gsyn (caller zeros).
 * num_threads <= 0: all OpenMP threads; the result does not depend on it. */
void or_recv_post(const or_loc *source, const or_loc *target, int nxpPost, int nypPost,
      const float *preActivity, const float *postWeights, float *gsyn, float dtFactor,
      int num_threads);

/* PresynapticPerspectiveConvolveDelivery::deliver
 * (src/delivery/PresynapticPerspectiveConvolveDelivery.cpp:61-192) with
 * per-thread GSyn copies of BaseDelivery (src/delivery/BaseDelivery.cpp:82-144).
 * If entries != NULL the sparse branch is used (active list per batch at offset
 * b*numPreExt, counts numActive[b]); otherwise the dense branch reads preActivity.
 * num_threads == 1 is the deterministic reference order. */
void or_recv_pre(const or_loc *pre, const or_loc *post, int nxp, int nyp,
      const float *preActivity, const or_entry *entries, const int64_t *numActive,
      const float *preWeights, float *gsyn, float dtFactor, int num_threads);

/* IdentDelivery::deliver dense branch (src/delivery/IdentDelivery.cpp:99-165). */
void or_ident_deliver(const or_loc *loc, const float *preActivity, float *gsyn);

/* DataStore::updateActiveIndices (src/columns/DataStore.cpp:46-66), per batch
 * element; entries at offset b*numExt. */
void or_active_indices(const or_loc *loc, const float *A, or_entry *entries, int64_t *numActive);

/* --- layer updates --- */
/* updateGSynAccumulatorOnCPU (src/components/GSynAccumulator.kpp:21-42). */
void or_gsyn_accumulate(int numNeuronsAcrossBatch, int numChannels, const float *coeffs,
      const float *layerInput, float *out);
/* updateHyPerLCAOnCPU (src/components/HyPerLCAInternalStateBuffer.kpp:18-61). */
void or_update_lca(const or_loc *loc, int selfInteract, const double *dtAdapt, float tau,
      const float *accumulatedGSyn, const float *A, float *V);
/* HyPerActivityBuffer::updateBufferCPU (src/components/HyPerActivityBuffer.cpp:36-54). */
void or_activity_copy(const or_loc *loc, const float *V, float *A);
/* ANNActivityBuffer::applyVThresh / applyAMax (src/components/ANNActivityBuffer.cpp:490-556). */
void or_apply_vthresh(const or_loc *loc, const float *V, float VThresh, float AMin, float AShift,
      float VWidth, float *A);
void or_apply_amax(const or_loc *loc, float AMax, float *A);
/* ANNActivityBuffer::setVertices + setSlopes (ANNActivityBuffer.cpp:227-397).
 * Returns numVertices; arrays must hold 8 (vertices) and 9 (slopes). */
int or_set_vertices(float VThresh, float AMin, float AMax, float AShift, float VWidth,
      float *verticesV, float *verticesA, float *slopes);
/* applyVerticesANNActivityBufferOnCPU (src/components/ANNActivityBuffer.kpp:16-77). */
void or_apply_vertices(const or_loc *loc, int numVertices, const float *verticesV,
      const float *verticesA, const float *slopes, const float *V, float *A);
/* ErrScaleInternalStateBuffer::updateBufferCPU (ErrScaleInternalStateBuffer.cpp:41-53). */
void or_errscale(int n, float errScale, float *V);

/* --- plasticity / normalisation / probes --- */
/* HebbianUpdater::update_dW shared-weights branch + updateInd_dW
 * (src/weightupdaters/HebbianUpdater.cpp:476-627).  dw and counts accumulate
 * (caller clears, as clear_dW / clearNumActivations do); counts may be NULL. */
void or_hebbian_update_dw(const or_loc *pre, const or_loc *post, int nxp, int nyp, int nfp,
      float dWMax, const float *preActivity, const float *postActivity, float *dw, int64_t *counts);
/* The factors of MomentumUpdater::applyMomentumDeprecated (MomentumUpdater.cpp:333-343)
 * for the deprecated momentumTau: no range check on it. */
int or_momentum_factors_deprecated(int method, float momentumTau, float decay, float dwMax,
      float *dwFactor, float *wFactor);
/* normalize_dW (:773-807) and updateWeights (:822-833). */
void or_hebbian_normalize_apply(float *w, float *dw, const int64_t *counts, size_t n, int normalizeDw);
int or_momentum_factors(int method, float tau, float decay, float dwMax, float *dwFactor,
      float *wFactor);
void or_momentum_apply(float *w, float *dw, float *prevDw, const int64_t *counts, size_t n,
      int normalizeDw, float dwFactor, float wFactor, int deprecatedTau);
void or_update_momentum_lca(const or_loc *loc, int selfInteract, float momentumRate,
      const double *dtAdapt, float tau, const float *accumulatedGSyn, const float *A,
      float *prevDrive, float *V);
/* ISTAInternalStateBuffer.kpp:34-57 (the reference's K9 updateISTAInternalStateBufferOnGPU):
 * V += ((float)dt[b]/tau)*(G - VThresh*sign(A[kex])), tau = timeConstantTau/dt. */
void or_update_ista(const or_loc *loc, float VThresh, const double *dtAdapt, float tau,
      const float *accumulatedGSyn, const float *A, float *V);
/* NormalizeL2::normalizeWeights, one arbor, one connection, not from post perspective
 * (src/normalizers/NormalizeL2.cpp:36-120; NormalizeBase.cpp:184-193). FP32 like the reference. */
void or_normalize_l2(float *w, int numPatches, int patchSize, float strength, float minL2);
/* Same in FP64 (the build's device form). */
void or_normalize_l2_f64(float *w, int numPatches, int patchSize, double strength, double minL2);
/* CostFunctionSum::calculateSumNoMask (src/probes/CostFunctionSum.hpp:87-104) with
 * the cost functions of src/probes/CostFunctions.hpp:18-50. kind as pvhip_cost_sum_f64. */
double or_cost_sum(const or_loc *loc, const float *buf, int extended, int batch, int kind,
      double p0, double p1);

/* --- host-side controllers of the LCA loop --- */
/* AdaptiveTimeScaleController state of one batch element
 * (src/components/AdaptiveTimeScaleController.hpp TimeScaleData). */
typedef struct {
   double timeScale, timeScaleMax, timeScaleTrue;
} or_timescale;
/* AdaptiveTimeScaleController ctor (AdaptiveTimeScaleController.cpp:17-47). */
void or_timescale_init(or_timescale *ts, int nbatch, double baseMax, double baseMin);
/* AdaptiveTimeScaleController::calcTimesteps (AdaptiveTimeScaleController.cpp:60-88):
 * values[b] is the target energy probe's value (or -1 on a trigger step);
 * writes the new time scale of each batch element to dt_out[b]. */
void or_timescale_calc(or_timescale *ts, int nbatch, const double *values, double baseMax,
      double baseMin, double tauFactor, double growthFactor, double *dt_out);
/* LayerUpdateController::needUpdate without a trigger layer
 * (src/components/LayerUpdateController.cpp:318-335). */
int or_need_update(double simTime, double deltaTime, double lastUpdateTime, double period);

/* --- whole LCA loop --- */
/* The reference's 4-phase LCA column (tests/GPULCATest/input/GPULCATest.params,
 * tutorials/LCA-CIFAR): Input -IdentConn ch0-> Residual (ANNErrorLayer, errScale)
 * <-IdentConn ch1- Recon; V1 (HyPerLCALayer) <-TransposeConn (post perspective)-
 * Residual; Recon (ANNLayer) <-CloneConn (pre perspective, sparse V1)- V1.
 * Runs `nsteps` timesteps of HyPerCol::advanceTime (src/columns/HyPerCol.cpp:659-791)
 * with a fixed dt and no adaptive time scale. */
typedef struct {
   or_loc input;  /* nx, ny, nf, nbatch; halo of the Residual is taken from res */
   or_loc res;    /* Residual loc (same nx,ny,nf as input), with its halo */
   or_loc v1;     /* V1 loc with its halo */
   int nxp, nyp;  /* pre-perspective patch of V1 -> Residual (nfp = input nf) */
   float VThresh, AMin, AMax, AShift, VWidth; /* V1 ANNActivityBuffer */
   int transfer_vertices; /* 1: GPU vertices formula, 0: CPU applyVThresh */
   double tau, dt; /* timeConstantTau and the column dt, read as double from params */
   int selfInteract;
   float errScale;
   int num_threads; /* OpenMP threads of the deliveries; 1 = the bit-reproducible order */
} or_lca_config;

/* Runs the loop. weights: V1->Residual pre-perspective [nf_v1][nyp][nxp][nf_in];
 * input restricted [nb][ny][nx][nf]; V0 initial V1 potential restricted.
 * Outputs (any may be NULL): final V1 V (restricted) and A (extended), final
 * Recon A (extended, Recon halo 0), per-step energy [nsteps][nb]
 * (0.5*sum r^2 + VThresh*sum|a|) and numActive [nsteps][nb]. */
void or_lca_run(const or_lca_config *cfg, const float *weights, const float *input,
      const float *V0, int nsteps, float *V1_V, float *V1_A, float *recon,
      double *energy, int64_t *numActive);

#ifdef __cplusplus
}
#endif

#endif /* PVORACLE_H_ */
