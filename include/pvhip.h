/*
 * pvhip.h -- C-ABI of libpvhip, the MI355X (gfx950) device layer and hot-path
 * kernels of OpenPV's synaptic-delivery + LCA path.
 *
 * Every entry point is `extern "C"`, takes plain pointers, sizes and POD
 * structs, and returns an int status (PVHIP_OK == 0).  A non-zero status
 * leaves a message retrievable with pvhip_last_error().  The reference turns
 * every CUDA error into Fatal() (src/arch/cuda/cuda_util.hpp:14-27); the host
 * wrapper of this library does the same with the status codes.
 *
 * Device pointers are plain `void*`/`float*` into HBM.  Streams are HIP
 * streams passed as `pvhip_stream` (NULL = the null stream).  Nothing here
 * allocates on a kernel call: buffers are owned by the caller, as in the
 * reference where layers, publishers and weights own their CudaBuffers
 * (src/components/ComponentBuffer.cpp:106, src/components/BasePublisherComponent.cpp:109-127,
 * src/components/Weights.cpp:114-159).
 *
 * Data layouts (all as the reference's PV layout, feature index fastest,
 * src/utils/conversions.hpp:251-253 and :401-407):
 *   restricted buffer  [nbatch][ny][nx][nf]
 *   extended buffer    [nbatch][ny+dn+up][nx+lt+rt][nf]
 *   GSyn channel       [nbatch][ny][nx][nf]   (channels outermost: [nch][nbatch]...)
 *   weights            [numKernels][nyp][nxp][nfp]   (one arbor)
 *   active list        pvhip_sparse_entry[nbatch][numExtended], count int64[nbatch]
 */
#ifndef PVHIP_H_
#define PVHIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PVHIP_ABI_VERSION 8

enum {
   PVHIP_OK              = 0,
   PVHIP_ERR_INVALID     = 1, /* bad argument or shape */
   PVHIP_ERR_RUNTIME     = 2, /* HIP runtime call failed */
   PVHIP_ERR_UNSUPPORTED = 3, /* valid request the library does not implement */
   PVHIP_ERR_NOMEM       = 4
};

/* Arithmetic form of a kernel. */
typedef enum {
   PVHIP_ARITH_EXACT = 0, /* FP32, the reference CPU summation order, mul and add rounded separately */
   PVHIP_ARITH_FP32  = 1, /* FP32 inputs, FP32 FMA / MFMA accumulation, any order */
   PVHIP_ARITH_BF16  = 2, /* BF16 inputs on MFMA, FP32 accumulate and FP32 state */
   PVHIP_ARITH_FP8   = 3  /* FP8 e4m3 inputs with per-feature weight scales, FP32 accumulate */
} pvhip_arith;

/* ------------------------------------------------------------------------ */
/* Types mirroring the reference's structs                                   */
/* ------------------------------------------------------------------------ */

/* PVHalo (src/include/PVLayerLoc.h:14-16). */
typedef struct {
   int lt, rt, dn, up;
} pvhip_halo;

/* PVLayerLoc (src/include/PVLayerLoc.h:21-26) for one rank: global == local. */
typedef struct {
   int nbatch, nx, ny, nf;
   pvhip_halo halo;
} pvhip_loc;

/* Patch (src/structures/Patch.hpp:15-18): same field order and sizes. */
typedef struct {
   uint16_t nx, ny;
   uint32_t offset;
} pvhip_patch;

/* SparseList<float>::Entry (src/structures/SparseList.hpp:19-22). */
typedef struct {
   uint32_t index;
   float value;
} pvhip_sparse_entry;

/* ------------------------------------------------------------------------ */
/* Device layer: replaces src/arch/cuda (CudaDevice, CudaBuffer, CudaTimer)  */
/* ------------------------------------------------------------------------ */

typedef struct pvhip_device pvhip_device;
typedef struct pvhip_buffer pvhip_buffer;
typedef struct pvhip_event pvhip_event;
typedef struct pvhip_graph pvhip_graph;
typedef void *pvhip_stream; /* hipStream_t */

typedef struct {
   char name[256];
   char arch[64]; /* gcnArchName, e.g. "gfx950:sramecc+:xnack-" */
   int ordinal;
   int compute_units;
   int wavefront_size;
   int max_threads_per_block;
   int max_grid_x;
   int clock_khz;
   int memory_clock_khz;
   int memory_bus_width;
   int l2_cache_bytes;
   size_t lds_per_block;
   size_t total_global_mem;
} pvhip_props;

/* Error text of the last failing call on this thread (never NULL). */
const char *pvhip_last_error(void);
int pvhip_abi_version(void);

/* CudaDevice(int) ctor: set device, create the one stream, query props
 * (src/arch/cuda/CudaDevice.cpp:17-94). */
int pvhip_device_create(int ordinal, pvhip_device **out);
/* ~CudaDevice (src/arch/cuda/CudaDevice.cpp:96-109). */
int pvhip_device_destroy(pvhip_device *dev);
/* CudaDevice::syncDevice (src/arch/cuda/CudaDevice.cpp:96), then the check of
 * pvhip_check_faults. */
int pvhip_device_sync(pvhip_device *dev);
/* CudaDevice getters (src/arch/cuda/CudaDevice.cpp:167-183). */
int pvhip_device_props(pvhip_device *dev, pvhip_props *out);
/* CudaDevice::getStream (src/arch/cuda/CudaDevice.hpp). */
pvhip_stream pvhip_device_stream(pvhip_device *dev);
/* Bytes allocated through this device (CudaDevice VRAM bookkeeping, :111-123). */
size_t pvhip_device_allocated_bytes(pvhip_device *dev);

/* CudaDevice::createBuffer / CudaBuffer ctor (src/arch/cuda/CudaBuffer.cpp:16-30). */
int pvhip_buffer_alloc(pvhip_device *dev, size_t bytes, pvhip_buffer **out);
/* ~CudaBuffer (src/arch/cuda/CudaBuffer.cpp:32-35). */
int pvhip_buffer_free(pvhip_buffer *buf);
/* CudaBuffer::copyToDevice(in, size, offset) (src/arch/cuda/CudaBuffer.cpp:37-49). */
int pvhip_buffer_h2d(
      pvhip_buffer *buf, const void *host, size_t bytes, size_t offset, pvhip_stream stream);
/* CudaBuffer::copyFromDevice (src/arch/cuda/CudaBuffer.cpp:51-62). */
int pvhip_buffer_d2h(
      pvhip_buffer *buf, void *host, size_t bytes, size_t offset, pvhip_stream stream);
void *pvhip_buffer_ptr(pvhip_buffer *buf);
size_t pvhip_buffer_size(pvhip_buffer *buf);

/* Zero-fill with a kernel (graph-capturable; no hipMemsetAsync node). */
int pvhip_fill_zero(void *dptr, size_t bytes, pvhip_stream stream);
/* Device-to-device copy on a stream: the Publisher's copy of a layer's
 * activity into a DataStore level (src/columns/Publisher.cpp:104-135). */
int pvhip_copy_device(void *dst, const void *src, size_t bytes, pvhip_stream stream);
/* hipStreamSynchronize, then the check of pvhip_check_faults. */
int pvhip_stream_sync(pvhip_stream stream);
/* Waits for the device, then reports (and clears) tiles a delivery kernel
 * refused since the last check: geometry tables that do not belong to the
 * locs of their call leave those outputs unwritten, and the next
 * pvhip_stream_sync / pvhip_device_sync / pvhip_check_faults returns
 * PVHIP_ERR_INVALID with a message.  The reference has no such case: its
 * tables are built inside the delivery (CudaRecvPost.cpp:51-389). */
int pvhip_check_faults(void);

/* CudaTimer (src/arch/cuda/CudaTimer.cpp:41-72): events on a stream. */
int pvhip_event_create(pvhip_event **out);
int pvhip_event_destroy(pvhip_event *ev);
int pvhip_event_record(pvhip_event *ev, pvhip_stream stream);
int pvhip_event_elapsed_ms(pvhip_event *start, pvhip_event *stop, float *ms);

/* hipGraph capture of one timestep (no reference counterpart: the reference
 * launches every kernel each step).  begin/end bracket stream capture. */
int pvhip_graph_begin(pvhip_stream stream);
int pvhip_graph_end(pvhip_stream stream, pvhip_graph **out);
int pvhip_graph_launch(pvhip_graph *graph, pvhip_stream stream);
int pvhip_graph_destroy(pvhip_graph *graph);

/* ------------------------------------------------------------------------ */
/* Post-perspective gather: replaces CudaRecvPost / cudnnConvolutionForward  */
/* (src/cudakernels/CudaRecvPost.cpp:357-389) and is bit-identical, in EXACT */
/* form, to PostsynapticPerspectiveConvolveDelivery::deliver                 */
/* (src/delivery/PostsynapticPerspectiveConvolveDelivery.cpp:56-151).        */
/* ------------------------------------------------------------------------ */

/* Threshold / transfer function of ANNActivityBuffer. */
typedef enum {
   /* CPU semantics: A = V, applyVThresh, applyAMax (src/components/ANNActivityBuffer.cpp:432-556). */
   PVHIP_TRANSFER_VTHRESH = 0,
   /* GPU semantics: piecewise-linear vertices (src/components/ANNActivityBuffer.kpp:16-77). */
   PVHIP_TRANSFER_VERTICES = 1
} pvhip_transfer_mode;

#define PVHIP_MAX_VERTICES 8
#define PVHIP_MAX_CHANNELS 4

typedef struct {
   int mode;        /* pvhip_transfer_mode */
   float VThresh;   /* -FLT_MAX disables thresholding (ANNActivityBuffer.cpp:507) */
   float AMin;
   float AMax;      /* FLT_MAX disables the clamp (ANNActivityBuffer.cpp:543) */
   float AShift;
   float VWidth;
   int num_vertices; /* PVHIP_TRANSFER_VERTICES: setVertices/setSlopes (ANNActivityBuffer.cpp:227-397) */
   float vertices_v[PVHIP_MAX_VERTICES];
   float vertices_a[PVHIP_MAX_VERTICES];
   float slopes[PVHIP_MAX_VERTICES + 1];
} pvhip_transfer;

/* HyPerLCA V/A update, the reference's K8+K6+K7 kernels in one pass:
 *   G  = sum_ch coeff[ch] * GSyn[ch]              (GSynAccumulator.kpp:21-42)
 *   V  = e*V + (1-e)*(G + selfInteract*A[kex])    (HyPerLCAInternalStateBuffer.kpp:48-59)
 *   A[kex] = transfer(V)                          (ANNActivityBuffer)
 * with e = exp_tau[b] = (float)exp(-dt[b]/(double)tau) computed on the host
 * by pvhip_lca_exp_tau. */
typedef struct {
   pvhip_loc loc;
   int num_channels;
   float channel_coeffs[PVHIP_MAX_CHANNELS];
   const float *gsyn;    /* [num_channels][nbatch][ny][nx][nf] */
   float *V;             /* restricted */
   float *A;             /* extended; halo untouched.  May be NULL when A_bf16c is
                          * set and activity_from_v holds: only the copy is written,
                          * and A = transfer(V) is derived when a reader needs it. */
   const float *exp_tau; /* device [nbatch] */
   int self_interact;
   pvhip_transfer transfer;
   /* Nonzero: the caller guarantees A[kex] == transfer(V) bit for bit for every
    * restricted neuron on entry -- the state this update leaves behind and the
    * state the initial activity update (ANNActivityBuffer) gives.  A(t-1) in the
    * self-interaction term is then recomputed from V instead of read, which
    * saves one read of A per neuron with identical results. */
   int activity_from_v;
   /* Optional (NULL: none): also write each updated activity, rounded to BF16
    * (round to nearest even), into this copy of A in the chunked BF16 layout
    * of pvhip_activity_to_bf16c (loc.nf a multiple of 32).  A BF16 channel GEMM
    * reads it through pvhip_recv_post_args.pre_data_bf16c.  Only restricted
    * neurons are written; the halo keeps what the caller put there. */
   void *A_bf16c;
} pvhip_lca_update_args;

/* Chunked BF16 copy of a layer's extended activity, the source format of the
 * BF16 channel GEMM: [nbatch][nf/32][nyExt][nxExt][32] values, feature f of
 * extended position (xe, ye) of batch b at
 *   ((b*(nf/32) + f/32)*nyExt + ye)*nxExt + xe)*32 + f%32,
 * each the activity rounded to BF16 (round to nearest even).  A chunk of 32
 * features of a row of positions is one contiguous run, which is what the GEMM
 * stages.  Writes the whole buffer (halo included): nbatch*numExt*2 bytes.
 * No counterpart in the reference (its GPU delivery reads the FP32 activity
 * through cuDNN, CudaRecvPost.cpp:51-389). */
int pvhip_activity_to_bf16c(const pvhip_loc *loc, const float *A, void *out, pvhip_stream stream);

/* Host helper: exp_tau[b] = (float)exp(-dt[b]/(double)tau), the reference's
 * double-precision exp then cast (HyPerLCAInternalStateBuffer.kpp:51). */
void pvhip_lca_exp_tau(const double *dt, int nbatch, float tau, float *exp_tau_out);

typedef struct {
   pvhip_loc pre;  /* source layer; its extended activity is read */
   pvhip_loc post; /* target layer; a restricted GSyn channel is written */
   int nxp, nyp, nfp; /* post-perspective patch (PatchSize::calcPostPatchSize), nfp == pre.nf */
   int num_kernels_x, num_kernels_y; /* shared data patches along x/y; kernels = kx*ky*post.nf */
   /* Per-axis tables of PatchGeometry::getUnshrunkenStart (extended source
    * coordinate of the patch start) and of Weights::calcDataIndexFromPatchIndex
    * (data-patch x/y index), indexed by restricted target x / y.
    * (src/components/PatchGeometry.cpp:179-279, src/components/Weights.cpp:205-231) */
   const int32_t *start_x; /* [post.nx] */
   const int32_t *start_y; /* [post.ny] */
   const int32_t *kernel_x; /* [post.nx] */
   const int32_t *kernel_y; /* [post.ny] */
   float dt_factor;         /* HyPerDelivery::mDeltaTimeFactor; 1 for convolve */
   const float *pre_data;   /* extended source activity */
   const float *weights;    /* [kernels][nyp][nxp][nfp] post-perspective */
   float *post_gsyn;        /* GSyn channel slice [nbatch][ny][nx][nf] */
   int accumulate;          /* 1: GSyn += ..., 0: GSyn was reset (reference resetGSynBuffers) */
   int arith;               /* pvhip_arith */
   /* BF16 form: weights converted by pvhip_recv_post_bf16_weights, layout
    * [kernels][Kpad] with Kpad = patch size rounded up to 32 (the shadow copy
    * plays the role of the reference's cuDNN weight copy, Weights.cpp:333-341). */
   const void *weights_lp;
   /* Targets larger than the source (data patches per feature > 1) on the MFMA
    * path: the window starts of start_x / start_y span [window_x0,
    * window_x0 + window_nx) (resp. y); inverse_x[kx*window_nx + (sx - window_x0)]
    * is the target x with kernel x-index kx and window start sx, or -1
    * (pvhost_recv_post_inverse).  May be NULL otherwise. */
   const int32_t *inverse_x;
   const int32_t *inverse_y;
   int window_x0, window_nx, window_y0, window_ny;
   /* Optional fused LCA epilogue (requires accumulate == 0, one channel,
    * coeff 1): the gather result is the accumulated G of the update.  post_gsyn
    * may then be NULL (GSyn is not materialised). */
   const pvhip_lca_update_args *fused_update;
   /* Optional, with fused_update: also publish the active list of the updated
    * activity, as pvhip_active_compact would after the update (HyPerLayer
    * publish of a sparseLayer: Publisher::updateActiveIndices,
    * src/columns/Publisher.cpp:85-134).  NULL: no active list. */
   const struct pvhip_active_out *active_out;
   /* Optional device scratch for the channel GEMM (FP32/BF16 forms with
    * several data patches per feature): the channels are split over
    * workgroups and their FP32 partial sums added in split order.  Size from
    * pvhip_recv_post_scratch_bytes(args); NULL or smaller runs unsplit.  The
    * library allocates nothing itself. */
   void *scratch;
   size_t scratch_bytes;
   /* Channel GEMM launch shape, 0 = the library's default: window starts per
    * workgroup (64, 128 or 256) and channel ranges (1 .. nfp / 32 for BF16,
    * nfp / 16 for FP32; more than one needs the scratch).  Picked by an
    * autotuner; results differ only in the FP32 summation order. */
   int cgemm_tile;
   int cgemm_splits;
   /* FP8 form (one data patch per feature): the source activity is
    * multiplied by this power of two before its e4m3 conversion, so that it
    * uses the format's range (0 = 1).  weights_lp then holds the FP8 shadow of
    * pvhip_recv_post_fp8_weights, which carries its own per-kernel scales. */
   float lp_source_scale;
   /* Optional, BF16 channel GEMM only (NULL: convert pre_data while staging):
    * pre_data in the chunked BF16 layout of pvhip_activity_to_bf16c (pre.nf a
    * multiple of 32), e.g. written by the producing layer's update
    * (pvhip_lca_update_args.A_bf16c).  The GEMM then reads half the bytes, in
    * contiguous runs, with the same results: it rounds pre_data to BF16 the
    * same way.  The other forms ignore it. */
   const void *pre_data_bf16c;
   /* Optional, channel GEMM only (targets with several data patches per
    * feature), with accumulate == 0 and no fused_update: instead of writing
    * post_gsyn, apply the target layer's update to the delivered values, as
    * pvhip_layer_update does with this delivery as its only channel
    * (num_channels 1; its gsyn is ignored).  The same bits as the delivery
    * followed by that update, without the GSyn round trip. */
   const struct pvhip_layer_update_args *fused_layer;
} pvhip_recv_post_args;

/* Active-list outputs of a fused delivery (same layout and contents as
 * pvhip_active_compact over fused_update->loc and fused_update->A). */
typedef struct pvhip_active_out {
   pvhip_sparse_entry *entries; /* [nbatch][numExt] */
   int64_t *num_active;         /* [nbatch] */
   uint32_t *pos_start;         /* [nbatch][nxExt*nyExt + 1], or NULL */
   void *scratch;               /* device memory, pvhip_recv_post_active_scratch_bytes() bytes */
   size_t scratch_bytes;
   /* 1: the caller guarantees the activity halo holds only zeros (no border
    * exchange or mirror boundary writes it; one rank with valueBC 0).  The
    * list is then emitted from the delivery's epilogue; otherwise it is built
    * by a separate ordered compaction of the whole extended buffer. */
   int halo_is_zero;
} pvhip_active_out;

int pvhip_recv_post(const pvhip_recv_post_args *args, pvhip_stream stream);
/* Scratch bytes pvhip_active_out needs for these arguments (fused_update set). */
size_t pvhip_recv_post_active_scratch_bytes(const pvhip_recv_post_args *args);
/* Scratch bytes of pvhip_recv_post_args.scratch for these arguments (0: the
 * launch is not split).  Depends on the geometry and the form only; the data
 * pointers may still be NULL.  An extension of the reference interface: the
 * reference's cuDNN call takes its workspace the same way
 * (CudaRecvPost.cpp:227-249). */
size_t pvhip_recv_post_scratch_bytes(const pvhip_recv_post_args *args);

/* Low-precision weight shadow for the BF16 form; call after every weight change. */
int pvhip_recv_post_bf16_weights(
      const float *weights, void *weights_bf16, int num_kernels, int patch_size, pvhip_stream stream);
size_t pvhip_recv_post_bf16_weights_bytes(int num_kernels, int patch_size);
/* FP8 (e4m3) weight shadow: per kernel (data patch), the weights times the
 * power of two that brings max |w| into [120, 240), as [kernels][Kpad] bytes
 * followed by the float scales.  An extension of the reference interface (its
 * GPU path has no FP8 form); call after every weight change. */
int pvhip_recv_post_fp8_weights(
      const float *weights, void *weights_fp8, int num_kernels, int patch_size, pvhip_stream stream);
size_t pvhip_recv_post_fp8_weights_bytes(int num_kernels, int patch_size);

/* ------------------------------------------------------------------------ */
/* Pre-perspective delivery: replaces CudaRecvPre / HyPerLayer_recv_pre      */
/* (src/cudakernels/CudaRecvPre.cu:10-101); EXACT form is bit-identical to   */
/* PresynapticPerspectiveConvolveDelivery::deliver with one thread           */
/* (src/delivery/PresynapticPerspectiveConvolveDelivery.cpp:61-192).         */
/* ------------------------------------------------------------------------ */

typedef struct {
   pvhip_loc pre;
   pvhip_loc post;
   int nxp, nyp, nfp; /* pre-perspective patch; nfp == post.nf */
   int num_kernels_x, num_kernels_y;
   /* Coverage tables built on the host from PatchGeometry (see
    * pvhost_recv_pre_tables): for restricted post x, the pre extended columns
    * whose (shrunken) patch covers x, ascending, with the weight column index;
    * for (post y, patch row r), the pre extended rows whose patch row r lands
    * on y, ascending, with the weight row index.  CSR layout. */
   const int32_t *col_ptr; /* [post.nx + 1] */
   const int32_t *col_xe;  /* pre extended x */
   const int32_t *col_wx;  /* weight x inside the patch */
   const int32_t *col_kx;  /* data-patch x index of that pre column */
   const int32_t *row_ptr; /* [post.ny * nyp + 1], row (y, r) at y*nyp + r */
   const int32_t *row_ye;
   const int32_t *row_wy;
   const int32_t *row_ky;
   float dt_factor;
   /* Sparse pre (sparseLayer = true): the active list of pvhip_active_compact
    * and its position segments.  Dense pre: pos_start == NULL, pre_data used. */
   const pvhip_sparse_entry *active; /* [nbatch][numPreExt] */
   const uint32_t *pos_start;        /* [nbatch][nyExt*nxExt + 1] */
   const float *pre_data;            /* [nbatch][numPreExt] (dense mode) */
   const float *weights;             /* [kernels][nyp][nxp][nfp] pre-perspective */
   float *post_gsyn;
   int accumulate;
   int arith; /* EXACT (ordered) or FP32 (same kernel, FMA) */
} pvhip_recv_pre_args;

int pvhip_recv_pre(const pvhip_recv_pre_args *args, pvhip_stream stream);

/* The BF16 (or FP32) form of the same delivery for a dense source whose neurons each
 * send one shared patch (one data patch per source feature) into a target 4
 * times finer along x and y (the reconstruction of an LCA column): for every
 * source position p and patch offset o, P[p][o] = sum_c a[p][c]*w[c][o] on
 * the matrix cores (BF16 or FP32 operands, FP32 sums), added to the target neuron at
 * the patch start of p plus o, in an order fixed by the launch (the result
 * does not depend on scheduling).  Replaces PresynapticPerspectiveGPUDelivery::
 * deliver (PresynapticPerspectiveGPUDelivery.cpp:198-270) for this geometry;
 * the additions are those of PresynapticPerspectiveConvolveDelivery.cpp:61-192
 * in another order.  Requires: target nx, ny = 4 x source; (nxp/4)*(nyp/4)*nfp
 * == 16; pre.nf 128, 256 or 512. */
typedef struct {
   pvhip_loc pre;
   pvhip_loc post;
   int nxp, nyp, nfp; /* pre-perspective patch; nfp == post.nf */
   /* Restricted target x (y) of the unshrunken patch start of source extended
    * column (row) 0; column xe starts at 4*xe + start_x0
    * (pvhost_recv_pre_dense_start). */
   int start_x0, start_y0;
   /* 1: the caller guarantees the source halo holds only zeros; only the
    * restricted positions are multiplied.  0: every extended position. */
   int halo_is_zero;
   const void *pre_data_bf16c; /* the source activity in the pvhip_activity_to_bf16c layout */
   const void *weights_bf16;   /* pvhip_recv_pre_dense_weights of the weights */
   float dt_factor;
   float *post_gsyn;
   int accumulate;
   /* Optional, with accumulate == 0: the target layer's update applied to the
    * delivered values (as pvhip_recv_post_args.fused_layer); post_gsyn is
    * then not written. */
   const struct pvhip_layer_update_args *fused_layer;
   /* 1: the FP32 form: FP32 products and sums on the matrix cores
    * (v_mfma_f32_16x16x4f32), reading pre_data (the FP32 activity,
    * [nbatch][numPreExt]) and weights_f32 (pvhip_recv_pre_dense_weights_f32)
    * instead of the two BF16 buffers. */
   int fp32;
   const float *pre_data;
   const float *weights_f32;
} pvhip_recv_pre_dense_args;

int pvhip_recv_pre_dense(const pvhip_recv_pre_dense_args *args, pvhip_stream stream);
/* 1 when pvhip_recv_pre_dense can run these arguments' geometry and form
 * (pre, post, nxp, nyp, nfp, start_x0/y0, halo_is_zero, fp32 are read; the
 * workgroup's weights and target tile fit its LDS), else 0.  Host only, no
 * device call; the choice of the dense form before any buffer exists. */
int pvhip_recv_pre_dense_fits(const pvhip_recv_pre_dense_args *args);
/* BF16 weights of pvhip_recv_pre_dense: pre-perspective weights
 * [num_features][patch_size] (patch_size = nyp*nxp*nfp) transposed to
 * [patch_size][num_features], rounded to nearest even; call after every
 * weight change (the role of the reference's GPU weight copy,
 * Weights.cpp:333-341). */
int pvhip_recv_pre_dense_weights(
      const float *weights, void *weights_bf16, int num_features, int patch_size, pvhip_stream stream);
size_t pvhip_recv_pre_dense_weights_bytes(int num_features, int patch_size);
/* FP32 weights of the FP32 form: the same transposition, [patch_size][num_features]
 * floats (num_features * patch_size * 4 bytes). */
int pvhip_recv_pre_dense_weights_f32(
      const float *weights, float *weights_f32, int num_features, int patch_size, pvhip_stream stream);

/* ------------------------------------------------------------------------ */
/* Layer updates                                                             */
/* ------------------------------------------------------------------------ */

/* Fused K8+K6+K7 (see pvhip_lca_update_args). */
int pvhip_lca_update(const pvhip_lca_update_args *args, pvhip_stream stream);

/* MomentumLCALayer V update (MomentumLCAInternalStateBuffer.kpp:44-63, the
 * reference's K10 updateMomentumLCAOnGPU) with the GSyn accumulation and the
 * transfer fused as in pvhip_lca_update:
 *   drive = (1-e)*(G + selfInteract*A[kex])
 *   V     = e*V + drive + momentum_rate*prev_drive
 *   prev_drive = drive;  A[kex] = transfer(V)
 * prev_drive: restricted [nbatch][ny][nx][nf], read and rewritten
 * (MomentumLCALayer's "prevDrive" buffer, zero at start).  momentum_rate is
 * the layer's LCAMomentumRate. */
int pvhip_momentum_lca_update(
      const pvhip_lca_update_args *args, float momentum_rate, float *prev_drive,
      pvhip_stream stream);

/* ISTALayer V update (ISTAInternalStateBuffer.kpp:34-57, the reference's K9
 * updateISTAInternalStateBufferOnGPU) with the GSyn accumulation and the
 * transfer fused as in pvhip_lca_update:
 *   V += rate[b]*(G - vthresh*sign(A[kex])),  A[kex] = transfer(V)
 * where args->exp_tau holds the per-batch rates (float)dt[b]/tau of
 * pvhip_ista_rate (tau = timeConstantTau/dt) and vthresh is the layer's
 * VThresh (ANNActivityBuffer::getVThresh).  self_interact is not used. */
int pvhip_ista_update(const pvhip_lca_update_args *args, float vthresh, pvhip_stream stream);
/* Host helper: rate_out[b] = (float)dt[b] / tau (ISTAInternalStateBuffer.kpp:56). */
void pvhip_ista_rate(const double *dt, int nbatch, float tau, float *rate_out);

/* GSynAccumulator (src/components/GSynAccumulator.kpp:21-42):
 * out[i] = sum_ch coeff[ch]*gsyn[ch*n + i], starting from 0.0f. */
int pvhip_gsyn_accumulate(
      int num_neurons_across_batch,
      int num_channels,
      const float *channel_coeffs_host,
      const float *gsyn,
      float *out,
      pvhip_stream stream);

/* Non-LCA layer update: V = scale * accumulate(GSyn) (HyPerInternalStateBuffer +
 * ErrScaleInternalStateBuffer.cpp:41-53; scale = errScale, 1 for ANNLayer),
 * then A[kex] = transfer(V) (ANNActivityBuffer; HyPerActivityBuffer when the
 * transfer is disabled). */
typedef struct pvhip_layer_update_args {
   pvhip_loc loc;
   int num_channels;
   float channel_coeffs[PVHIP_MAX_CHANNELS];
   const float *gsyn;
   float *V;
   float *A;
   float scale;
   int apply_scale; /* 0: V = G exactly (no multiply) */
   pvhip_transfer transfer;
} pvhip_layer_update_args;

int pvhip_layer_update(const pvhip_layer_update_args *args, pvhip_stream stream);

/* Transfer only: A[kex] = transfer(V[k]) (K7 / applyVThresh+applyAMax). */
int pvhip_ann_activity(
      const pvhip_loc *loc, const pvhip_transfer *transfer, const float *V, float *A,
      pvhip_stream stream);

/* A residual layer (ANNErrorLayer, ErrScaleInternalStateBuffer.cpp:41-53)
 * fed by two IdentConn: channel 0 from `in`, channel 1 from `rec`
 * (IdentDelivery.cpp:146-162 onto reset GSyn, GSynAccumulator.kpp:21-42,
 * errScale, then the transfer).  One pass with the operations and order of
 * pvhip_ident_deliver x2 + pvhip_layer_update (bit-identical), without
 * materialising GSyn.  A fusion of the reference's components, no single
 * reference counterpart. */
typedef struct {
   pvhip_loc loc;     /* the error layer (halo of its activity) */
   pvhip_loc in_loc;  /* channel-0 source, same nbatch/nx/ny/nf */
   const float *in;   /* extended activity of in_loc */
   pvhip_loc rec_loc; /* channel-1 source */
   const float *rec;
   float channel_coeffs[2]; /* GSynAccumulator coefficients, {1, -1} */
   float scale;
   int apply_scale;   /* 0: V = G exactly (no multiply) */
   pvhip_transfer transfer;
   float *V;          /* restricted */
   float *A;          /* extended */
} pvhip_error_update_args;

int pvhip_error_update(const pvhip_error_update_args *args, pvhip_stream stream);

/* IdentDelivery::deliver (src/delivery/IdentDelivery.cpp:99-165), dense form:
 * GSyn[b][k] += A_ext[b][kex(k)]. */
int pvhip_ident_deliver(
      const pvhip_loc *loc, const float *pre_data, float *post_gsyn, int accumulate,
      pvhip_stream stream);

/* DataStore::updateActiveIndices (src/columns/DataStore.cpp:46-66): ascending
 * scan of the extended buffer, entries {kex, a} for a != 0.0f, count per batch
 * element written as int64.  pos_start[b][p] receives the first entry of each
 * extended (x,y) position p and pos_start[b][nPos] the count, so that
 * [pos_start[p], pos_start[p+1]) are the active features at p.  pos_start may
 * be NULL.  chunk_scratch is caller-owned device memory of
 * pvhip_active_compact_scratch_bytes(loc) bytes (no allocation inside). */
int pvhip_active_compact(
      const pvhip_loc *loc,
      const float *A,
      pvhip_sparse_entry *entries,
      int64_t *num_active,
      uint32_t *pos_start,
      uint32_t *chunk_scratch,
      pvhip_stream stream);
size_t pvhip_active_compact_scratch_bytes(const pvhip_loc *loc);

/* pvhip_active_compact without a caller workspace: the per-chunk counts live
 * in the library's own device memory (a fixed 256 KiB array per GPU, held
 * until the process ends), so the call allocates nothing.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under hipGraph capture, serving loops, harnesses that forbid hipMalloc).
 * Limit: pvhip_active_compact_scratch_bytes(loc) <= 262144, i.e. nbatch times
 * ceil(numExtended / 16384) at most 65536; larger requests return
 * PVHIP_ERR_INVALID.  Every call on one GPU shares the workspace: calls that
 * overlap on one GPU (different streams, concurrent host threads) race. */
int pvhip_active_compact_noalloc(
      const pvhip_loc *loc,
      const float *A,
      pvhip_sparse_entry *entries,
      int64_t *num_active,
      uint32_t *pos_start,
      pvhip_stream stream);

/* ------------------------------------------------------------------------ */
/* Plasticity, normalisation, energy                                         */
/* ------------------------------------------------------------------------ */

/* HebbianUpdater::update_dW / updateInd_dW for shared weights
 * (src/weightupdaters/HebbianUpdater.cpp:476-627): for every pre extended
 * neuron with a_pre != 0, dW[kernel][patch item] += dWMax*a_pre*a_post and
 * counts[item]++ over the shrunken patch.  Accumulation order per dW element:
 * batch ascending, then pre position ascending (the reference's).  Per-axis
 * geometry tables over the pre extended axes come from PatchGeometry. */
typedef struct {
   pvhip_loc pre, post;
   int nxp, nyp, nfp; /* pre-perspective patch */
   int num_kernels_x, num_kernels_y;
   const int32_t *dim_x;       /* [nxExtPre] patch.nx */
   const int32_t *dim_y;       /* [nyExtPre] patch.ny */
   const int32_t *pstart_x;    /* [nxExtPre] first patch column kept (patch.offset x) */
   const int32_t *pstart_y;    /* [nyExtPre] */
   const int32_t *apost_x;     /* [nxExtPre] extended post x of the patch (aPostOffset) */
   const int32_t *apost_y;     /* [nyExtPre] */
   float dw_max;
   const float *pre_data;  /* extended [nbatch][numPreExt] */
   const float *post_data; /* extended [nbatch][numPostExt] */
   float *dw;              /* [kernels][nyp][nxp][nfp], accumulated into (caller clears) */
   int64_t *counts;        /* same shape, or NULL when normalizeDw is false */
   int arith;              /* EXACT (reference order) or FP32/BF16 (MFMA correlation) */
   /* MFMA forms (FP32, BF16): dW = dWMax * sum over batch and pre extended
    * positions of a_pre (x) the post window, a dense correlation on the matrix
    * cores, for connections with one data patch per feature (pre no larger
    * than post).  The window of pre extended (xe, ye) starts at post extended
    * (window_start_x + window_stride_x*xe, window_start_y + window_stride_y*ye)
    * and is clipped to the post restricted region (pvhost_hebbian_windows).
    * Partial sums over split-K chunks go to scratch
    * (pvhip_hebbian_dw_scratch_bytes) and are added in chunk order. */
   int window_stride_x, window_start_x, window_stride_y, window_start_y;
   void *scratch;
   size_t scratch_bytes;
   /* MFMA forms: pre rows (batch x extended y) per split-K chunk, 0 = the
    * library's default (about 512 workgroups, at least 4 rows).  Picked by an
    * autotuner; results differ only in the FP32 summation order. */
   int rows_per_chunk;
} pvhip_hebbian_args;

int pvhip_hebbian_dw(const pvhip_hebbian_args *args, pvhip_stream stream);
size_t pvhip_hebbian_dw_scratch_bytes(const pvhip_hebbian_args *args);

/* HebbianUpdater::normalize_dW (HebbianUpdater.cpp:773-807): dw /= count, 0
 * where count == 0; then updateWeights (:822-833): w += dw. */
int pvhip_hebbian_apply(
      float *weights, float *dw, const int64_t *counts, size_t n, int normalize_dw,
      pvhip_stream stream);

/* MomentumUpdater (src/weightupdaters/MomentumUpdater.cpp).  Method codes of
 * momentumMethod: */
enum { PVHIP_MOMENTUM_VISCOSITY = 0, PVHIP_MOMENTUM_SIMPLE = 1, PVHIP_MOMENTUM_ALEX = 2 };

/* Host only: the factors of applyMomentum (MomentumUpdater.cpp:296-307) from
 * timeConstantTau, momentumDecay and dWMax -- viscosity: exp(-1/tau) (0 for
 * tau == 0) and decay; simple: tau and decay; alex: tau and decay*dWMax --
 * after the reference's parameter checks (:82-112, :138-145): a rejected
 * value returns PVHIP_ERR_INVALID. */
int pvhip_momentum_factors(
      int method, float tau, float decay, float dw_max, float *dw_factor, float *w_factor);

/* The factors of applyMomentumDeprecated (MomentumUpdater.cpp:333-343) for the
 * deprecated momentumTau parameter (used when timeConstantTau is absent):
 * viscosity exp(-1/momentumTau), simple momentumTau, alex momentumTau with
 * decay*dWMax.  The reference applies no range check to momentumTau (only
 * momentumDecay's, :138-145), and neither does this. */
int pvhip_momentum_factors_deprecated(
      int method, float momentum_tau, float decay, float dw_max, float *dw_factor, float *w_factor);

/* MomentumUpdater::updateWeights in one pass (MomentumUpdater.cpp:276-294):
 * normalize_dW as pvhip_hebbian_apply (normalize_dw, counts), then
 *   dw = dw*(1-dw_factor) + dw_factor*prev_dw - w_factor*w   (:309-329), or,
 *   with deprecated_tau (the momentumTau parameter, :346-360),
 *   dw = dw + (dw_factor*prev_dw - w_factor*w);
 * then prev_dw = dw and w += dw (HebbianUpdater::updateWeights).  Every
 * array holds n values; prev_dw starts at zero (or initPrev_dWFile). */
int pvhip_momentum_apply(
      float *weights, float *dw, float *prev_dw, const int64_t *counts, size_t n,
      int normalize_dw, float dw_factor, float w_factor, int deprecated_tau,
      pvhip_stream stream);

/* TransposeWeights::transposeShared on the device (src/utils/TransposeWeights.cpp:56-113,
 * called by WeightsPair::finalizeUpdate after each weight update,
 * src/components/WeightsPair.cpp:261-280): dst[i] = src[index[i]], 0 where
 * index[i] < 0, with index from pvhost_transpose_weights_index.  n = the size
 * of dst. */
int pvhip_permute_weights(
      const float *src, const int32_t *index, float *dst, size_t n, pvhip_stream stream);

/* NormalizeL2 per data patch in the reference's float arithmetic and order
 * (src/normalizers/NormalizeL2.cpp:93-125): sumsq += w*w, w *= strength /
 * sqrtf(sumsq); a patch with sumsq <= min_l2 ends the loop, leaving it and all
 * later patches unchanged, as the reference does. */
int pvhip_normalize_l2(
      float *weights, int num_patches, int patch_size, float strength, float min_l2,
      pvhip_stream stream);

/* NormalizeL2 per data patch (src/normalizers/NormalizeL2.cpp:36-120) with the
 * sum of squares and scale in FP64: w *= strength / sqrt(sum w^2). Patches
 * whose sum of squares is <= min_l2 are left unchanged. */
int pvhip_normalize_l2_f64(
      float *weights, int num_patches, int patch_size, double strength, double min_l2,
      pvhip_stream stream);

/* CostFunctionSum::calculateSumNoMask (src/probes/CostFunctionSum.hpp:87-104)
 * in FP64 over the restricted part of a (possibly extended) buffer:
 *   kind 0: L2  sum x^2, kind 1: L1 sum |x|, kind 2: firm threshold cost
 *   (src/probes/CostFunctions.hpp:18-50, params p0 = VThresh, p1 = VWidth),
 *   kind 3: L0 count of |x| > p0.
 * Writes out[b] (device double[nbatch]). */
int pvhip_cost_sum_f64(
      const pvhip_loc *loc, const float *buffer, int extended, int kind, double p0, double p1,
      double *out, pvhip_stream stream);

/* ------------------------------------------------------------------------ */
/* Adjoint of the unrolled LCA loop (an extension: the reference has no      */
/* backward pass).  One backward step on the LCA layer's grid:               */
/*   gVt = gV + T'(V_t) * (gA + l1*sign(T(V_t)))                             */
/*   gG[kex] = (1-e)*gVt   (extended, the gather's output gradient)          */
/*   gA_prev = selfInteract*(1-e)*gVt   (restricted; may be NULL)            */
/*   gV = e*gVt                                                              */
/* the derivatives of HyPerLCAInternalStateBuffer.kpp:48-59 and of the       */
/* transfer (ANNActivityBuffer.cpp:227-556).                                 */
/* ------------------------------------------------------------------------ */
typedef struct {
   pvhip_loc loc;             /* the LCA layer (halo of its activity) */
   pvhip_transfer transfer;
   const float *exp_tau;      /* device [nbatch] */
   int self_interact;
   const float *V;            /* V_t, restricted */
   const float *gA;           /* dE/dA_t from the later steps, restricted; may be NULL */
   float l1;                  /* coefficient of sum |A_t| in E (0: none) */
   float *gV;                 /* in: dE/dV_t carried from step t+1 (e*gVt); out: e*gVt */
   float *gG;                 /* out: dE/dG_t, extended (halo left as it is) */
   float *gA_prev;            /* out: self-interaction part of dE/dA_{t-1}, restricted */
} pvhip_lca_adjoint_args;

int pvhip_lca_adjoint_step(const pvhip_lca_adjoint_args *args, pvhip_stream stream);

/* dst[b][kex(k)] (+)= scale * src[b][k]: restricted buffer into an extended one. */
int pvhip_scale_to_extended(
      const pvhip_loc *loc, float scale, const float *src, float *dst, int accumulate,
      pvhip_stream stream);

/* y = a*x + b*y over n floats. */
int pvhip_axpby(size_t n, float a, const float *x, float b, float *y, pvhip_stream stream);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* PVHIP_H_ */
