/*
 * pvhost.h -- C-ABI of libpvhost, the C++ host side above libpvhip: patch
 * geometry and the device tables of the deliveries, weight transposition,
 * transfer-function vertices.  The .params reader (pvparams.h) is part of the
 * same library; the pvrun executable builds and runs a column from a .params
 * file over both libraries.
 *
 * The reference computes these tables on the host as well
 * (PatchGeometry::setPatchGeometry, src/components/PatchGeometry.cpp:179-279;
 * PostsynapticPerspectiveGPUDelivery::initializeRecvKernelArgs,
 * src/delivery/PostsynapticPerspectiveGPUDelivery.cpp:111-237;
 * PresynapticPerspectiveGPUDelivery::initializeRecvKernelArgs,
 * src/delivery/PresynapticPerspectiveGPUDelivery.cpp:120-196).
 * All functions return 0 on success and a negative value on invalid input.
 */
#ifndef PVHOST_H_
#define PVHOST_H_

#include "pvhip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* requiredConvolveMargin (src/utils/requiredConvolveMargin.hpp:14-40). */
int pvhost_required_margin(int n_pre, int n_post, int patch_size);
/* PatchSize::calcPostPatchSize (src/components/PatchSize.cpp:98-130). */
int pvhost_post_patch_size(int pre_patch_size, int n_pre, int n_post);
/* Shared data patches along x, y (PatchGeometry.cpp:53-55). */
int pvhost_num_kernels(const pvhip_loc *pre, const pvhip_loc *post, int *nkx, int *nky);

/* PatchGeometry::calcPatchData over one axis of the pre extended space
 * (axis 0 = x, 1 = y); arrays sized nxExt / nyExt of pre; any may be NULL. */
int pvhost_patch_axis(const pvhip_loc *pre, const pvhip_loc *post, int nxp, int nyp, int axis,
      int32_t *dim, int32_t *patch_start, int32_t *post_start_res, int32_t *post_start_ext,
      int32_t *unshrunken_start);

/* Window geometry of the MFMA forms of pvhip_hebbian_dw: the unshrunken
 * patch of pre extended position e starts at post extended coordinate
 * start + stride * e along each axis, and its shrunken patch (PatchGeometry)
 * is that window clipped to the post restricted region.  Returns 0 when both
 * hold for every pre extended position, a negative value otherwise (the EXACT
 * form then applies). */
int pvhost_hebbian_windows(const pvhip_loc *pre, const pvhip_loc *post, int nxp, int nyp,
      int *stride_x, int *start_x, int *stride_y, int *start_y);

/* Tables of pvhip_recv_post_args for a post-perspective delivery from
 * `source` (pre layer) into `target` (post layer) with post-perspective patch
 * nxp x nyp: start_x/kernel_x [target.nx], start_y/kernel_y [target.ny].
 * Returns a negative value if a patch would reach outside the source's
 * extended buffer (halo smaller than the required margin). */
int pvhost_recv_post_tables(const pvhip_loc *source, const pvhip_loc *target, int nxp, int nyp,
      int32_t *start_x, int32_t *start_y, int32_t *kernel_x, int32_t *kernel_y, int *nkx, int *nky);

/* Inverse tables of pvhip_recv_post_args for the MFMA path (any number of
 * data patches per feature): *window_x0 and *window_nx (and y) receive the range
 * of window starts; inverse_x (nkx*window_nx ints) and inverse_y may be NULL
 * on a first call that only queries the ranges. */
int pvhost_recv_post_inverse(const pvhip_loc *source, const pvhip_loc *target, int nxp, int nyp,
      int32_t *inverse_x, int32_t *inverse_y, int *window_x0, int *window_nx, int *window_y0,
      int *window_ny);

/* Coverage tables of pvhip_recv_pre_args (CSR).  Call once with NULL arrays
 * to get *ncol and *nrow, then with arrays of those sizes (col_ptr has
 * post.nx + 1 entries, row_ptr post.ny*nyp + 1). */
int pvhost_recv_pre_tables(const pvhip_loc *pre, const pvhip_loc *post, int nxp, int nyp,
      int32_t *col_ptr, int32_t *col_xe, int32_t *col_wx, int32_t *col_kx, int *ncol,
      int32_t *row_ptr, int32_t *row_ye, int32_t *row_wy, int32_t *row_ky, int *nrow);

/* Patch starts of pvhip_recv_pre_dense_args from PatchGeometry::calcPatchData
 * (src/components/PatchGeometry.cpp:391-453): 0 and the restricted target
 * x (y) of the unshrunken patch start of source extended column (row) 0 when
 * every source extended column xe (row ye) starts at stride*xe + *start_x0
 * with stride = post.nx / pre.nx (resp. y) and one data patch along the axis
 * (target no coarser than the source); -1 otherwise.  stride may be NULL. */
int pvhost_recv_pre_dense_start(const pvhip_loc *pre, const pvhip_loc *post, int nxp, int nyp,
      int *stride, int *start_x0, int *start_y0);

/* ANNActivityBuffer::setVertices + setSlopes (src/components/ANNActivityBuffer.cpp:227-397)
 * filled into a transfer descriptor; mode selects VTHRESH or VERTICES. */
int pvhost_make_transfer(float VThresh, float AMin, float AMax, float AShift, float VWidth,
      int mode, pvhip_transfer *out);

/* TransposeWeights::transposeShared (src/utils/TransposeWeights.cpp:56-113):
 * post-perspective weights of a shared-weight connection pre -> post with
 * pre-perspective patch nxp x nyp x nfp. */
int pvhost_transpose_weights(const pvhip_loc *pre, const pvhip_loc *post, int nxp, int nyp,
      int nfp, const float *pre_weights, float *post_weights);
/* The same transposition as a table for the device (pvhip_permute_weights):
 * post_weights[i] = pre_weights[index[i]], index[i] = -1 where no pre weight
 * lands (the post weight stays 0).  index holds n entries, the size of the
 * post weights. */
int pvhost_transpose_weights_index(const pvhip_loc *pre, const pvhip_loc *post, int nxp, int nyp,
      int nfp, int32_t *index, size_t n);

#ifdef __cplusplus
}
#endif

#endif /* PVHOST_H_ */
