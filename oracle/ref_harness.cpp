/*
 * ref_harness.cpp -- TEST INFRASTRUCTURE ONLY.
 *
 * A thin extern "C" wrapper, written for this project, around the pieces of
 * the reference that compile from their own source files (built by
 * `make ref` from the reference checkout where it lies, never copied):
 *   - PV::PatchGeometry        (src/components/PatchGeometry.cpp)
 *   - updateHyPerLCAOnCPU      (src/components/HyPerLCAInternalStateBuffer.kpp)
 *   - updateMomentumLCAOnCPU   (src/components/MomentumLCAInternalStateBuffer.kpp)
 *   - updateISTAInternalStateBufferOnCPU (src/components/ISTAInternalStateBuffer.kpp)
 *   - applyVerticesANNActivityBufferOnCPU (src/components/ANNActivityBuffer.kpp)
 *   - updateGSynAccumulatorOnCPU (src/components/GSynAccumulator.kpp)
 *   - requiredConvolveMargin   (src/utils/requiredConvolveMargin.hpp)
 *   - index math               (src/utils/conversions.hpp)
 *   - cost functions           (src/probes/CostFunctions.hpp)
 * The oracle's C restatement is checked against these in tests/.
 */
#include "components/PatchGeometry.hpp"
#include "include/PVLayerLoc.h"
#include "probes/CostFunctions.hpp"
#include "utils/conversions.hpp"
#include "utils/requiredConvolveMargin.hpp"

#undef PV_RUN_ON_GPU
#include "components/ANNActivityBuffer.kpp"
#include "components/GSynAccumulator.kpp"
#include "components/HyPerLCAInternalStateBuffer.kpp"
#include "components/MomentumLCAInternalStateBuffer.kpp"
#include "components/ISTAInternalStateBuffer.kpp"

#include <cstdint>
#include <cstring>

namespace {
PVLayerLoc makeLoc(const int *l) {
   /* l = {nbatch, nx, ny, nf, lt, rt, dn, up} */
   PVLayerLoc loc;
   std::memset(&loc, 0, sizeof(loc));
   loc.nbatch       = l[0];
   loc.nx           = l[1];
   loc.ny           = l[2];
   loc.nf           = l[3];
   loc.nbatchGlobal = l[0];
   loc.nxGlobal     = l[1];
   loc.nyGlobal     = l[2];
   loc.halo.lt      = l[4];
   loc.halo.rt      = l[5];
   loc.halo.dn      = l[6];
   loc.halo.up      = l[7];
   return loc;
}
} // namespace

extern "C" {

int ref_patch_geometry(const int *pre, const int *post, int nxp, int nyp, int nfp,
      uint16_t *pnx, uint16_t *pny, uint32_t *poffset, uint64_t *gSynPatchStart,
      uint64_t *aPostOffset, uint64_t *unshrunkenStart, int *transposeItemIndex) {
   PVLayerLoc preLoc  = makeLoc(pre);
   PVLayerLoc postLoc = makeLoc(post);
   try {
      PV::PatchGeometry g("ref", nxp, nyp, nfp, &preLoc, &postLoc);
      g.allocateDataStructures();
      int n = g.getNumPatches();
      for (int p = 0; p < n; p++) {
         PV::Patch const &patch = g.getPatch(p);
         pnx[p]                 = patch.nx;
         pny[p]                 = patch.ny;
         poffset[p]             = patch.offset;
         gSynPatchStart[p]      = g.getGSynPatchStart(p);
         aPostOffset[p]         = g.getAPostOffset(p);
         unshrunkenStart[p]     = g.getUnshrunkenStart(p);
      }
      if (transposeItemIndex) {
         int nk = g.getNumKernels();
         int ps = g.getPatchSizeOverall();
         for (int k = 0; k < nk; k++) {
            for (int i = 0; i < ps; i++) {
               transposeItemIndex[k * ps + i] = g.getTransposeItemIndex(k, i);
            }
         }
      }
      return n;
   } catch (std::exception const &) {
      return -1;
   }
}

int ref_required_margin(int nPre, int nPost, int patchSize) {
   return requiredConvolveMargin(nPre, nPost, patchSize);
}

int ref_kindex_extended(int k, int nx, int ny, int nf, int lt, int rt, int dn, int up) {
   return PV::kIndexExtended(k, nx, ny, nf, lt, rt, dn, up);
}

void ref_update_lca(const int *loc, int selfInteract, const double *dtAdapt, float tau,
      const float *gsyn, const float *A, float *V) {
   PV::updateHyPerLCAOnCPU(loc[0], loc[1] * loc[2] * loc[3], loc[1], loc[2], loc[3], loc[4],
         loc[5], loc[6], loc[7], selfInteract != 0, dtAdapt, tau, gsyn, A, V);
}

void ref_update_momentum_lca(const int *loc, int selfInteract, float momentumRate,
      const double *dtAdapt, float tau, const float *gsyn, const float *A, float *prevDrive,
      float *V) {
   PV::updateMomentumLCAOnCPU(loc[0], loc[1] * loc[2] * loc[3], loc[1], loc[2], loc[3], loc[4],
         loc[5], loc[6], loc[7], selfInteract != 0, momentumRate, dtAdapt, tau, gsyn, A,
         prevDrive, V);
}

void ref_update_ista(const int *loc, float VThresh, const double *dtAdapt, float tau,
      const float *gsyn, const float *A, float *V) {
   PV::updateISTAInternalStateBufferOnCPU(loc[0], loc[1] * loc[2] * loc[3], loc[1], loc[2], loc[3],
         loc[4], loc[5], loc[6], loc[7], VThresh, dtAdapt, tau, gsyn, A, V);
}

void ref_apply_vertices(const int *loc, int numVertices, const float *vV, const float *vA,
      const float *slopes, const float *V, float *A) {
   PV::applyVerticesANNActivityBufferOnCPU(loc[0], loc[1] * loc[2] * loc[3], loc[1], loc[2],
         loc[3], loc[4], loc[5], loc[6], loc[7], numVertices, vV, vA, slopes, V, A);
}

void ref_gsyn_accumulate(int n, int numChannels, const float *coeffs, const float *in, float *out) {
   PV::updateGSynAccumulatorOnCPU(n, numChannels, coeffs, in, out);
}

double ref_cost(int kind, double x, double p0, double p1) {
   switch (kind) {
      case 0: return PV::L2CostFunction()(x);
      case 1: return PV::L1CostFunction()(x);
      case 2: return PV::FirmThresholdCostFunction(p0, p1)(x);
      default: return PV::L0CostFunction(p0)(x);
   }
}

} // extern "C"
