"""Python face of the device path: the reference's delivery / layer-update /
plasticity objects over libpvhip.

Names follow the reference (PostsynapticPerspectiveGPUDelivery,
PresynapticPerspectiveGPUDelivery, HyPerLCAInternalStateBuffer, ...).  Device
memory and the stream come from torch (plumbing only); every computation is a
libpvhip kernel.  Nothing here falls back to the CPU: a missing library or a
non-GPU tensor raises.
"""
import ctypes as C

import numpy as np
import torch

from . import abi

FLT_MAX = float(np.finfo(np.float32).max)


def _stream(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def _ptr(t):
    if t is None:
        return None
    if not t.is_cuda:
        raise abi.PVHIPError("device tensor required, got a %s tensor" % t.device)
    if not t.is_contiguous():
        raise abi.PVHIPError("contiguous tensor required")
    return C.c_void_p(t.data_ptr())


def _dev_i32(a, device):
    return torch.as_tensor(np.ascontiguousarray(a, np.int32), device=device)


def loc(nbatch, nx, ny, nf, halo=0):
    return abi.Loc.make(nbatch, nx, ny, nf, halo)


def make_transfer(VThresh=-FLT_MAX, AMin=-FLT_MAX, AMax=FLT_MAX, AShift=0.0, VWidth=0.0,
                  vertices=False):
    """ANNActivityBuffer transfer (VTHRESH = CPU semantics, vertices = GPU)."""
    t = abi.Transfer()
    mode = abi.TRANSFER_VERTICES if vertices else abi.TRANSFER_VTHRESH
    r = abi.host().pvhost_make_transfer(VThresh, AMin, AMax, AShift, VWidth, mode, C.byref(t))
    if r != 0:
        raise abi.PVHIPError("pvhost_make_transfer failed (%d)" % r)
    return t


def exp_tau(dt, tau, nbatch):
    """(float)exp(-dt/(double)tau) per batch element, on the host."""
    dts = (C.c_double * nbatch)(*([float(dt)] * nbatch) if np.isscalar(dt) else list(map(float, dt)))
    out = (C.c_float * nbatch)()
    abi.hip().pvhip_lca_exp_tau(dts, nbatch, C.c_float(tau), out)
    return np.array(out[:], np.float32)


def fp8_scale_for(bound):
    """The power of two s with bound*s in [120, 240): the scale that puts values
    of magnitude up to `bound` into the e4m3 range with headroom."""
    bound = float(bound)
    if not bound > 0.0 or not np.isfinite(bound):
        return 1.0
    m, e = np.frexp(240.0 / bound)   # 240/bound = m * 2^e, m in [0.5, 1)
    return float(np.ldexp(1.0, int(e) - 1))


def scaled_tau(timeConstantTau, deltaTime):
    """mScaledTimeConstantTau = (float)(tau / dt) (HyPerLCAInternalStateBuffer.cpp:104-112)."""
    return float(np.float32(float(timeConstantTau) / float(deltaTime)))


class PostsynapticPerspectiveGPUDelivery:
    """Gather delivery (PostsynapticPerspectiveGPUDelivery + CudaRecvPost).

    source: pre layer loc; target: post layer loc; nxp, nyp: post-perspective
    patch; weights: device tensor [kernels][nyp][nxp][source.nf].
    """

    def __init__(self, source, target, nxp, nyp, weights, arith=abi.ARITH_EXACT, dt_factor=1.0,
                 split=True):
        """split: give the channel GEMM (several data patches per feature) its
        scratch, so the channels are split over workgroups."""
        self.source, self.target = source, target
        self.nxp, self.nyp, self.nfp = nxp, nyp, source.nf
        self.arith = arith
        self.dt_factor = dt_factor
        dev = weights.device
        sx = np.zeros(target.nx, np.int32)
        sy = np.zeros(target.ny, np.int32)
        kx = np.zeros(target.nx, np.int32)
        ky = np.zeros(target.ny, np.int32)
        nkx, nky = C.c_int(), C.c_int()
        r = abi.host().pvhost_recv_post_tables(
            C.byref(source), C.byref(target), nxp, nyp, sx.ctypes.data_as(C.c_void_p),
            sy.ctypes.data_as(C.c_void_p), kx.ctypes.data_as(C.c_void_p),
            ky.ctypes.data_as(C.c_void_p), C.byref(nkx), C.byref(nky))
        if r != 0:
            raise abi.PVHIPError("recv_post tables: halo of %r too small for patch %dx%d (%d)"
                                 % (source, nxp, nyp, r))
        self.nkx, self.nky = nkx.value, nky.value
        self.tables = [_dev_i32(a, dev) for a in (sx, sy, kx, ky)]
        self.num_kernels = self.nkx * self.nky * target.nf
        self.patch_size = nxp * nyp * self.nfp
        if weights.numel() != self.num_kernels * self.patch_size:
            raise abi.PVHIPError("weights hold %d values, expected %d x %d"
                                 % (weights.numel(), self.num_kernels, self.patch_size))
        self.weights = weights
        self.weights_lp = None
        self.inverse = None
        self.windows = (0, 0, 0, 0)
        if self.nkx * self.nky > 1 and arith in (abi.ARITH_BF16, abi.ARITH_FP32):
            # several data patches per feature on the MFMA path: inverse tables
            x0, nx_, y0, ny_ = C.c_int(), C.c_int(), C.c_int(), C.c_int()
            H = abi.host()
            r = H.pvhost_recv_post_inverse(C.byref(source), C.byref(target), nxp, nyp, None, None,
                                           C.byref(x0), C.byref(nx_), C.byref(y0), C.byref(ny_))
            if r != 0:
                raise abi.PVHIPError("recv_post inverse tables failed (%d)" % r)
            ix = np.zeros(self.nkx * nx_.value, np.int32)
            iy = np.zeros(self.nky * ny_.value, np.int32)
            r = H.pvhost_recv_post_inverse(C.byref(source), C.byref(target), nxp, nyp,
                                           ix.ctypes.data_as(C.c_void_p), iy.ctypes.data_as(C.c_void_p),
                                           C.byref(x0), C.byref(nx_), C.byref(y0), C.byref(ny_))
            if r != 0:
                raise abi.PVHIPError("recv_post inverse tables failed (%d)" % r)
            self.inverse = (_dev_i32(ix, dev), _dev_i32(iy, dev))
            self.windows = (x0.value, nx_.value, y0.value, ny_.value)
        self.lp_source_scale = 0.0   # FP8 form: power-of-two source scale (0 = 1)
        if arith in (abi.ARITH_BF16, abi.ARITH_FP8):
            self.refresh_weights()
        self.scratch = None
        self.split = bool(split)
        self.cgemm_tile, self.cgemm_splits = 0, 0   # channel GEMM launch shape, 0 = default
        self._alloc_scratch()

    def _alloc_scratch(self):
        """(Re)allocate the channel GEMM's split scratch for the current launch
        shape; call outside graph capture."""
        if not self.split or self.inverse is None:
            return
        n = abi.hip().pvhip_recv_post_scratch_bytes(C.byref(self.args(None, None)))
        if n and (self.scratch is None or self.scratch.numel() * 4 < n):
            self.scratch = torch.empty((n + 3) // 4, dtype=torch.float32, device=self.weights.device)

    def set_launch(self, cgemm_tile=0, cgemm_splits=0):
        """Channel GEMM launch shape (window starts per workgroup, channel
        ranges); 0 keeps the library's default.  Used by the autotuner."""
        self.cgemm_tile, self.cgemm_splits = int(cgemm_tile), int(cgemm_splits)
        self._alloc_scratch()

    def refresh_weights(self, stream=None):
        """Re-derive the low-precision shadow (BF16, or FP8 with its per-kernel
        scales) after a weight change."""
        if self.arith not in (abi.ARITH_BF16, abi.ARITH_FP8):
            return
        H = abi.hip()
        kind = "bf16" if self.arith == abi.ARITH_BF16 else "fp8"
        nbytes = getattr(H, "pvhip_recv_post_%s_weights_bytes" % kind)(self.num_kernels, self.patch_size)
        if self.weights_lp is None:
            self.weights_lp = torch.empty((nbytes + 15) // 16 * 4, dtype=torch.int32,
                                          device=self.weights.device)
        abi.check(getattr(H, "pvhip_recv_post_%s_weights" % kind)(
            _ptr(self.weights), _ptr(self.weights_lp), self.num_kernels, self.patch_size,
            _stream(stream)), "pvhip_recv_post_%s_weights" % kind)

    def set_source_scale(self, scale):
        """FP8 form: the power of two the source activity is multiplied by before
        its e4m3 conversion (fp8_scale_for(bound on |source|))."""
        self.lp_source_scale = float(scale)

    def args(self, pre_activity, gsyn, accumulate=False, fused=None, active=None,
             halo_is_zero=False, fused_layer=None):
        a = abi.RecvPostArgs()
        a.pre, a.post = self.source, self.target
        a.nxp, a.nyp, a.nfp = self.nxp, self.nyp, self.nfp
        a.num_kernels_x, a.num_kernels_y = self.nkx, self.nky
        a.start_x, a.start_y, a.kernel_x, a.kernel_y = [_ptr(t) for t in self.tables]
        a.dt_factor = self.dt_factor
        a.pre_data = _ptr(pre_activity)
        a.weights = _ptr(self.weights)
        a.post_gsyn = _ptr(gsyn)
        a.accumulate = int(accumulate)
        a.arith = self.arith
        a.weights_lp = _ptr(self.weights_lp)
        if self.inverse is not None:
            a.inverse_x, a.inverse_y = _ptr(self.inverse[0]), _ptr(self.inverse[1])
            a.window_x0, a.window_nx, a.window_y0, a.window_ny = self.windows
        if getattr(self, "scratch", None) is not None:
            a.scratch, a.scratch_bytes = _ptr(self.scratch), self.scratch.numel() * 4
        a.cgemm_tile = getattr(self, "cgemm_tile", 0)
        a.cgemm_splits = getattr(self, "cgemm_splits", 0)
        a.lp_source_scale = getattr(self, "lp_source_scale", 0.0)
        a.pre_data_bf16c = _ptr(getattr(self, "pre_bf16c", None))
        if fused is not None:
            a.fused_update = C.pointer(fused)
        if fused_layer is not None:
            a.fused_layer = C.pointer(fused_layer)
        if active is not None:
            if fused is None:
                raise abi.PVHIPError("an active list needs the fused update")
            nbytes = abi.hip().pvhip_recv_post_active_scratch_bytes(C.byref(a))
            scratch = active.delivery_scratch(nbytes)
            o = abi.ActiveOut()
            o.entries, o.num_active = _ptr(active.entries), _ptr(active.num_active)
            o.pos_start = _ptr(active.pos_start)
            o.scratch, o.scratch_bytes = _ptr(scratch), scratch.numel() * 4
            o.halo_is_zero = int(halo_is_zero)
            a.active_out = C.pointer(o)
            a._keep = o  # the struct must outlive the call
        return a

    def deliver(self, pre_activity, gsyn, accumulate=False, fused=None, active=None,
                halo_is_zero=False, stream=None, fused_layer=None):
        """GSyn (+)= gather; with `fused` the LCA update is applied instead of
        writing GSyn, and with `active` (an ActiveList of the updated layer) the
        active list is published in the same call (halo_is_zero: the layer's
        activity halo holds only zeros, which lets the epilogue emit it).
        fused_layer (channel GEMM only): the target layer's update
        (layer_update_args) applied instead of writing GSyn."""
        self.launch(self.args(pre_activity, gsyn, accumulate, fused, active, halo_is_zero,
                              fused_layer), stream)

    @staticmethod
    def launch(args, stream=None):
        """Runs a prepared argument block (see args()); used where the block is
        built once, e.g. outside a hipGraph capture."""
        abi.check(abi.hip().pvhip_recv_post(C.byref(args), _stream(stream)), "pvhip_recv_post")


class ActiveList:
    """The publisher's sparse activity (DataStore active indices) on the device."""

    ENTRY = np.dtype([("index", np.uint32), ("value", np.float32)])

    def __init__(self, loc_, device):
        self.loc = loc_
        nb = loc_.nbatch
        self.entries = torch.zeros(nb * loc_.num_ext * 2, dtype=torch.int32, device=device)
        self.num_active = torch.zeros(nb, dtype=torch.int64, device=device)
        self.pos_start = torch.zeros(nb * (loc_.nx_ext * loc_.ny_ext + 1), dtype=torch.int32,
                                     device=device)
        nscr = abi.hip().pvhip_active_compact_scratch_bytes(C.byref(loc_))
        self.scratch = torch.zeros(max(1, nscr // 4), dtype=torch.int32, device=device)

    def delivery_scratch(self, nbytes):
        """Device scratch of a delivery that publishes this list (grown on demand)."""
        n = max(1, (int(nbytes) + 3) // 4)
        if getattr(self, "_dscratch", None) is None or self._dscratch.numel() < n:
            self._dscratch = torch.empty(n, dtype=torch.int32, device=self.entries.device)
        return self._dscratch

    def compact(self, A, stream=None):
        abi.check(abi.hip().pvhip_active_compact(
            C.byref(self.loc), _ptr(A), _ptr(self.entries), _ptr(self.num_active),
            _ptr(self.pos_start), _ptr(self.scratch), _stream(stream)), "pvhip_active_compact")

    def host_entries(self):
        """(entries structured array per batch [nb][numExt], counts) copied to the host."""
        e = self.entries.cpu().numpy().view(self.ENTRY).reshape(self.loc.nbatch, self.loc.num_ext)
        return e, self.num_active.cpu().numpy()


class PresynapticPerspectiveGPUDelivery:
    """Scatter delivery (PresynapticPerspectiveGPUDelivery + CudaRecvPre).  Three
    ways to run it, chosen by `mode`:
      "ordered" -- an ordered gather over the active list (EXACT, or FP32 with FMA);
      "route"   -- FP32/BF16: a post-perspective gather with the transposed
                   weights (a channel GEMM over the dense activity);
      "dense"   -- BF16/FP32, one data patch per source feature into a target
                   4x finer (pvhip_recv_pre_dense): a GEMM of the source
                   positions by the patch offsets over the channels, then the
                   overlap-add; BF16 reads the source's chunked BF16 copy
                   (bind_pre_bf16c), FP32 the dense activity.
    The default is the first of dense, route, ordered that the geometry allows.
    pre_halo_is_zero: the caller guarantees the source activity's halo holds
    only zeros (the dense mode then multiplies restricted positions only)."""

    def __init__(self, pre, post, nxp, nyp, weights, arith=abi.ARITH_EXACT, dt_factor=1.0):
        self.pre, self.post = pre, post
        self.nxp, self.nyp, self.nfp = nxp, nyp, post.nf
        # the scatter's MFMA route (channel GEMM) has BF16 and FP32 forms: an
        # FP8 request runs it in BF16
        arith = abi.ARITH_BF16 if arith == abi.ARITH_FP8 else arith
        self.arith = arith
        self.dt_factor = dt_factor
        dev = weights.device
        H = abi.host()
        ncol, nrow = C.c_int(), C.c_int()
        r = H.pvhost_recv_pre_tables(C.byref(pre), C.byref(post), nxp, nyp, None, None, None, None,
                                     C.byref(ncol), None, None, None, None, C.byref(nrow))
        if r != 0:
            raise abi.PVHIPError("recv_pre tables failed (%d)" % r)
        cp = np.zeros(post.nx + 1, np.int32)
        cx, cw, ck = (np.zeros(max(1, ncol.value), np.int32) for _ in range(3))
        rp = np.zeros(post.ny * nyp + 1, np.int32)
        ry, rw, rk = (np.zeros(max(1, nrow.value), np.int32) for _ in range(3))
        P = lambda a: a.ctypes.data_as(C.c_void_p)
        r = H.pvhost_recv_pre_tables(C.byref(pre), C.byref(post), nxp, nyp, P(cp), P(cx), P(cw),
                                     P(ck), C.byref(ncol), P(rp), P(ry), P(rw), P(rk),
                                     C.byref(nrow))
        if r != 0:
            raise abi.PVHIPError("recv_pre tables failed (%d)" % r)
        self.tables = [_dev_i32(a, dev) for a in (cp, cx, cw, ck, rp, ry, rw, rk)]
        nkx, nky = C.c_int(), C.c_int()
        H.pvhost_num_kernels(C.byref(pre), C.byref(post), C.byref(nkx), C.byref(nky))
        self.nkx, self.nky = nkx.value, nky.value
        self.weights = weights
        self.transposed = None
        self.dense = None
        self._pre_bf16c = None
        self.pre_halo_is_zero = False
        self._mode = None
        if arith in (abi.ARITH_BF16, abi.ARITH_FP32):
            self.transposed = self._transposed_route()
            self.dense = self._dense_route()

    def _dense_route(self):
        """The geometry of pvhip_recv_pre_dense (see the class docstring), with
        its BF16 (FP32 form: FP32) weights; None when the geometry does not fit."""
        pre, post = self.pre, self.post
        if (post.nx != 4 * pre.nx or post.ny != 4 * pre.ny or self.nxp % 4 or self.nyp % 4
                or (self.nxp // 4) * (self.nyp // 4) * self.nfp != 16 or pre.nf not in (128, 256, 512)
                or self.nkx != 1 or self.nky != 1):
            return None
        x0, y0 = C.c_int(), C.c_int()
        if abi.host().pvhost_recv_pre_dense_start(C.byref(pre), C.byref(post), self.nxp, self.nyp, None,
                                                  C.byref(x0), C.byref(y0)) != 0:
            return None
        patch = self.nxp * self.nyp * self.nfp
        f32 = self.arith == abi.ARITH_FP32
        # the workgroup's weights and target tile must fit its LDS, checked for
        # every extended position (the restricted-only mode needs less)
        q = abi.RecvPreDenseArgs()
        q.pre, q.post = pre, post
        q.nxp, q.nyp, q.nfp = self.nxp, self.nyp, self.nfp
        q.start_x0, q.start_y0 = x0.value, y0.value
        q.halo_is_zero = 0
        q.fp32 = int(f32)
        if not abi.hip().pvhip_recv_pre_dense_fits(C.byref(q)):
            return None
        wb = torch.zeros(patch * pre.nf, dtype=torch.float32 if f32 else torch.bfloat16,
                         device=self.weights.device)
        self.dense = {"start": (x0.value, y0.value), "weights": wb, "patch": patch, "fp32": f32}
        self._refresh_dense(None)
        return self.dense

    def _refresh_dense(self, stream):
        d = self.dense
        if d["fp32"]:
            abi.check(abi.hip().pvhip_recv_pre_dense_weights_f32(
                _ptr(self.weights), _ptr(d["weights"]), self.pre.nf, d["patch"], _stream(stream)),
                "pvhip_recv_pre_dense_weights_f32")
            return
        abi.check(abi.hip().pvhip_recv_pre_dense_weights(_ptr(self.weights), _ptr(d["weights"]), self.pre.nf,
                                                         d["patch"], _stream(stream)),
                  "pvhip_recv_pre_dense_weights")

    def _dense_ready(self):
        """The dense mode's source is there: FP32 reads the activity passed to
        deliver, BF16 the bound chunked copy."""
        return self.dense is not None and (self.dense["fp32"] or self._pre_bf16c is not None)

    @property
    def mode(self):
        if self._mode is not None:
            return self._mode
        if self._dense_ready():
            return "dense"
        return "route" if self.transposed is not None else "ordered"

    @mode.setter
    def mode(self, m):
        if m not in (None, "dense", "route", "ordered"):
            raise abi.PVHIPError("unknown scatter mode %r" % (m,))
        if m == "dense" and not self._dense_ready():
            raise abi.PVHIPError("the dense scatter needs its geometry (and, in BF16, the source's chunked "
                                 "BF16 copy)")
        if m == "route" and self.transposed is None:
            raise abi.PVHIPError("this scatter has no transposed route")
        self._mode = m

    @property
    def use_route(self):
        """True when the delivery reads the dense activity (dense or route mode)."""
        return self.mode != "ordered"

    @use_route.setter
    def use_route(self, v):
        if not v:
            self._mode = "ordered"
        else:
            self._mode = None
            if self.mode == "ordered":
                raise abi.PVHIPError("this scatter has no dense route")

    def _transposed_route(self):
        """BF16 / FP32 forms: the scatter as a post-perspective gather with the
        transposed weights (TransposeWeights::transposeShared), which for a
        target larger than the source is a stride-1 convolution over the pre
        grid with (phase, feature) output channels -- an MFMA implicit GEMM.
        Returns None when the geometry does not fit that route."""
        pre, post = self.pre, self.post
        H = abi.host()
        ck = 32 if self.arith == abi.ARITH_BF16 else 16
        nxq = H.pvhost_post_patch_size(self.nxp, pre.nx, post.nx)
        nyq = H.pvhost_post_patch_size(self.nyp, pre.ny, post.ny)
        pkx, pky = C.c_int(), C.c_int()
        H.pvhost_num_kernels(C.byref(post), C.byref(pre), C.byref(pkx), C.byref(pky))
        m = pkx.value * pky.value * post.nf
        single = pkx.value * pky.value == 1  # target no larger than the source: a plain gather
        if nxq <= 0 or nyq <= 0 or (not single and (pre.nf % ck or m > 64)):
            return None
        # The reference's post-perspective patch geometry offsets the source's x
        # axis by its dn halo and the target's y axis by its dn halo
        # (PatchGeometry.cpp:186-224), while the buffers are offset by lt and up:
        # the transposed gather equals the scatter only where those halos agree,
        # which holds when each layer's halo is the same on all four sides.
        # Other halos keep the ordered scatter.
        for l in (pre, post):
            if len({l.halo.lt, l.halo.rt, l.halo.dn, l.halo.up}) > 1:
                return None
        sx = np.zeros(post.nx, np.int32)
        sy = np.zeros(post.ny, np.int32)
        if H.pvhost_recv_post_tables(C.byref(pre), C.byref(post), nxq, nyq,
                                     sx.ctypes.data_as(C.c_void_p), sy.ctypes.data_as(C.c_void_p),
                                     None, None, None, None) != 0:
            return None
        self.transposed_shape = (m, nyq, nxq, pre.nf)
        wt = torch.zeros(m * nyq * nxq * pre.nf, dtype=torch.float32, device=self.weights.device)
        route = PostsynapticPerspectiveGPUDelivery(pre, post, nxq, nyq, wt, arith=self.arith,
                                                   dt_factor=self.dt_factor)
        self._route = route
        self.refresh_weights()
        return route

    def refresh_weights(self, stream=None):
        """Re-derive the transposed weights (and their low-precision shadow)
        after a weight change, on the device: WeightsPair::finalizeUpdate's
        transposition as a permutation table (graph-capturable, no host copy),
        and the dense mode's BF16 weights."""
        if getattr(self, "dense", None) is not None:
            self._refresh_dense(stream)
        if self.transposed is None and getattr(self, "_route", None) is None:
            return
        route = self._route
        if getattr(self, "_tindex", None) is None:
            n = route.weights.numel()
            idx = np.zeros(n, np.int32)
            r = abi.host().pvhost_transpose_weights_index(
                C.byref(self.pre), C.byref(self.post), self.nxp, self.nyp, self.nfp,
                idx.ctypes.data_as(C.c_void_p), n)
            if r != 0:
                raise abi.PVHIPError("pvhost_transpose_weights_index failed (%d)" % r)
            self._tindex = _dev_i32(idx, route.weights.device)
        abi.check(abi.hip().pvhip_permute_weights(_ptr(self.weights), _ptr(self._tindex),
                                                  _ptr(route.weights), route.weights.numel(),
                                                  _stream(stream)), "pvhip_permute_weights")
        route.refresh_weights(stream)

    def bind_pre_bf16c(self, buf, halo_is_zero=False):
        """Chunked BF16 copy of the pre layer's activity (pvhip_activity_to_bf16c
        layout), kept current by the pre layer's update: the BF16 channel GEMM of
        the transposed route then reads it instead of converting the FP32
        activity (same results, half the bytes), and the dense mode becomes
        available.  halo_is_zero: the caller guarantees the activity's halo
        holds only zeros (the dense mode then multiplies restricted positions
        only).  None unbinds.  Returns True when a route can use it."""
        ok = False
        t = self.transposed
        if t is not None and t.inverse is not None and t.arith == abi.ARITH_BF16 and self.pre.nf % 32 == 0:
            t.pre_bf16c = buf
            ok = True
        if self.dense is not None and not self.dense["fp32"]:
            self._pre_bf16c = buf
            self.pre_halo_is_zero = bool(halo_is_zero)
            ok = True
        if buf is None and self._mode == "dense":
            self._mode = None
        return ok

    def _deliver_dense(self, gsyn, accumulate, stream, fused_layer, pre_activity=None):
        d = self.dense
        a = abi.RecvPreDenseArgs()
        a.pre, a.post = self.pre, self.post
        a.nxp, a.nyp, a.nfp = self.nxp, self.nyp, self.nfp
        a.start_x0, a.start_y0 = d["start"]
        a.halo_is_zero = int(self.pre_halo_is_zero)
        if d["fp32"]:
            if pre_activity is None:
                raise abi.PVHIPError("the FP32 dense scatter reads the dense pre activity")
            a.fp32 = 1
            a.pre_data = _ptr(pre_activity)
            a.weights_f32 = _ptr(d["weights"])
        else:
            a.pre_data_bf16c = _ptr(self._pre_bf16c)
            a.weights_bf16 = _ptr(d["weights"])
        a.dt_factor = self.dt_factor
        a.post_gsyn = _ptr(gsyn)
        a.accumulate = int(accumulate)
        if fused_layer is not None:
            a.fused_layer = C.pointer(fused_layer)
        abi.check(abi.hip().pvhip_recv_pre_dense(C.byref(a), _stream(stream)), "pvhip_recv_pre_dense")

    def deliver(self, gsyn, active=None, pre_activity=None, accumulate=False, stream=None,
                fused_layer=None):
        """GSyn (+)= scatter.  fused_layer: the target layer's update
        (layer_update_args, reading `gsyn` as its only channel) follows the
        delivery; the dense route's channel GEMM applies it in its own epilogue
        (no GSyn round trip), the other paths write GSyn and then update."""
        mode = self.mode
        if mode == "dense":
            fl = fused_layer if not accumulate else None
            self._deliver_dense(gsyn, accumulate, stream, fl, pre_activity)
            if fused_layer is not None and fl is None:
                launch_layer_update(fused_layer, stream)
            return
        if self.transposed is not None and (mode == "route" or active is None):
            if pre_activity is None:
                raise abi.PVHIPError("the %s scatter reads the dense pre activity"
                                     % ("BF16" if self.arith == abi.ARITH_BF16 else "FP32"))
            if fused_layer is not None and self.transposed.inverse is not None and not accumulate:
                self.transposed.deliver(pre_activity, gsyn, stream=stream, fused_layer=fused_layer)
                return
            self.transposed.deliver(pre_activity, gsyn, accumulate=accumulate, stream=stream)
            if fused_layer is not None:
                launch_layer_update(fused_layer, stream)
            return
        a = abi.RecvPreArgs()
        a.pre, a.post = self.pre, self.post
        a.nxp, a.nyp, a.nfp = self.nxp, self.nyp, self.nfp
        a.num_kernels_x, a.num_kernels_y = self.nkx, self.nky
        (a.col_ptr, a.col_xe, a.col_wx, a.col_kx,
         a.row_ptr, a.row_ye, a.row_wy, a.row_ky) = [_ptr(t) for t in self.tables]
        a.dt_factor = self.dt_factor
        if active is not None:
            a.active = _ptr(active.entries)
            a.pos_start = _ptr(active.pos_start)
        else:
            a.pre_data = _ptr(pre_activity)
        a.weights = _ptr(self.weights)
        a.post_gsyn = _ptr(gsyn)
        a.accumulate = int(accumulate)
        # the ordered scatter has EXACT and FP32 forms; a BF16 delivery whose
        # geometry has no transposed MFMA route runs it in FP32
        a.arith = abi.ARITH_EXACT if self.arith == abi.ARITH_EXACT else abi.ARITH_FP32
        abi.check(abi.hip().pvhip_recv_pre(C.byref(a), _stream(stream)), "pvhip_recv_pre")
        if fused_layer is not None:
            launch_layer_update(fused_layer, stream)


class HyPerLCAUpdate:
    """HyPerLCALayer update: GSynAccumulator + HyPerLCAInternalStateBuffer +
    ANNActivityBuffer in one kernel."""

    def __init__(self, loc_, transfer, timeConstantTau=1.0, dt=1.0, self_interact=True,
                 channel_coeffs=(1.0,), activity_from_v=False, device="cuda"):
        """activity_from_v: A == transfer(V) holds on entry to every update (the layer's
        activity is written only by this update and its initial A = T(V)), so the
        kernels recompute A(t-1) from V instead of reading it."""
        self.loc = loc_
        self.activity_from_v = bool(activity_from_v)
        self.transfer = transfer
        self.tau = scaled_tau(timeConstantTau, dt)
        self.exp_tau = torch.as_tensor(exp_tau(dt, self.tau, loc_.nbatch), device=device)
        self.self_interact = bool(self_interact)
        self.coeffs = list(channel_coeffs)

    def set_dt(self, dt):
        """Per-batch dt (adaptive time scales): recompute exp(-dt/tau) on the host."""
        self.exp_tau.copy_(torch.as_tensor(exp_tau(dt, self.tau, self.loc.nbatch)))

    def args(self, gsyn, V, A, num_channels=None, A_bf16c=None):
        """A_bf16c: optional chunked BF16 copy of A that the update also writes."""
        a = abi.LCAUpdateArgs()
        a.loc = self.loc
        nch = len(self.coeffs) if num_channels is None else num_channels
        a.num_channels = nch
        for i, c in enumerate(self.coeffs[:abi.MAX_CHANNELS]):
            a.channel_coeffs[i] = c
        a.gsyn = _ptr(gsyn)
        a.V, a.A = _ptr(V), _ptr(A)
        a.exp_tau = _ptr(self.exp_tau)
        a.self_interact = int(self.self_interact)
        a.transfer = self.transfer
        a.activity_from_v = int(self.activity_from_v)
        a.A_bf16c = _ptr(A_bf16c)
        return a

    def update(self, gsyn, V, A, stream=None, A_bf16c=None):
        a = self.args(gsyn, V, A, A_bf16c=A_bf16c)
        abi.check(abi.hip().pvhip_lca_update(C.byref(a), _stream(stream)), "pvhip_lca_update")


class MomentumLCAUpdate(HyPerLCAUpdate):
    """MomentumLCALayer update (MomentumLCAInternalStateBuffer.kpp:44-63): the
    HyPerLCA update plus LCAMomentumRate times the previous step's drive, which
    the layer keeps in its "prevDrive" buffer (restricted, zero at start)."""

    def __init__(self, loc_, transfer, timeConstantTau=1.0, dt=1.0, self_interact=True,
                 LCAMomentumRate=0.0, channel_coeffs=(1.0,), activity_from_v=False, device="cuda"):
        super().__init__(loc_, transfer, timeConstantTau, dt, self_interact, channel_coeffs,
                         activity_from_v, device)
        self.momentum_rate = float(LCAMomentumRate)
        self.prev_drive = torch.zeros(loc_.nbatch * loc_.num_res, dtype=torch.float32, device=device)

    def update(self, gsyn, V, A, stream=None, A_bf16c=None):
        a = self.args(gsyn, V, A, A_bf16c=A_bf16c)
        abi.check(abi.hip().pvhip_momentum_lca_update(C.byref(a), self.momentum_rate,
                                                      _ptr(self.prev_drive), _stream(stream)),
                  "pvhip_momentum_lca_update")


class ISTAUpdate(HyPerLCAUpdate):
    """ISTALayer update (ISTAInternalStateBuffer.kpp:34-57):
    V += (dt/tau)*(G - VThresh*sign(A)), then the transfer; VThresh is the
    layer's (its ANNActivityBuffer's)."""

    def __init__(self, loc_, transfer, VThresh, timeConstantTau=1.0, dt=1.0, channel_coeffs=(1.0,),
                 activity_from_v=False, device="cuda"):
        super().__init__(loc_, transfer, timeConstantTau, dt, True, channel_coeffs, activity_from_v,
                         device)
        self.vthresh = float(VThresh)
        self.set_dt(np.full(loc_.nbatch, dt))

    def set_dt(self, dt):
        """Per-batch rates (float)dt/tau (pvhip_ista_rate)."""
        dt = np.ascontiguousarray(np.broadcast_to(np.asarray(dt, np.float64), (self.loc.nbatch,)))
        out = (C.c_float * self.loc.nbatch)()
        abi.hip().pvhip_ista_rate(dt.ctypes.data_as(C.POINTER(C.c_double)), self.loc.nbatch, self.tau, out)
        self.exp_tau.copy_(torch.as_tensor(np.ctypeslib.as_array(out).copy()))

    def update(self, gsyn, V, A, stream=None, A_bf16c=None):
        a = self.args(gsyn, V, A, A_bf16c=A_bf16c)
        abi.check(abi.hip().pvhip_ista_update(C.byref(a), self.vthresh, _stream(stream)),
                  "pvhip_ista_update")


def bf16c_buffer(loc_, device="cuda"):
    """A zeroed chunked BF16 activity buffer for layer loc_ (nf % 32 == 0)."""
    if loc_.nf % 32:
        raise abi.PVHIPError("the chunked BF16 layout needs nf % 32 == 0 (nf %d)" % loc_.nf)
    return torch.zeros(loc_.nbatch * loc_.num_ext, dtype=torch.bfloat16, device=device)


def activity_to_bf16c(loc_, A, out, stream=None):
    """out = the chunked BF16 copy of the extended activity A (halo included)."""
    abi.check(abi.hip().pvhip_activity_to_bf16c(C.byref(loc_), _ptr(A), _ptr(out), _stream(stream)),
              "pvhip_activity_to_bf16c")


def layer_update_args(loc_, gsyn, V, A, transfer, channel_coeffs=(1.0,), err_scale=None):
    """Argument block of pvhip_layer_update (also the fused_layer of a delivery)."""
    a = abi.LayerUpdateArgs()
    a.loc = loc_
    a.num_channels = len(channel_coeffs)
    for i, c in enumerate(channel_coeffs):
        a.channel_coeffs[i] = c
    a.gsyn, a.V, a.A = _ptr(gsyn), _ptr(V), _ptr(A)
    a.scale = 1.0 if err_scale is None else err_scale
    a.apply_scale = 0 if err_scale is None else 1
    a.transfer = transfer
    return a


def launch_layer_update(args, stream=None):
    abi.check(abi.hip().pvhip_layer_update(C.byref(args), _stream(stream)), "pvhip_layer_update")


def layer_update(loc_, gsyn, V, A, transfer, channel_coeffs=(1.0,), err_scale=None, stream=None):
    """ANNLayer / ANNErrorLayer / HyPerLayer update: V = [errScale*] sum coeff*GSyn; A = T(V)."""
    launch_layer_update(layer_update_args(loc_, gsyn, V, A, transfer, channel_coeffs, err_scale),
                        stream)


def error_update(loc_, in_loc, in_activity, rec_loc, rec_activity, V, A, transfer,
                 channel_coeffs=(1.0, -1.0), err_scale=None, stream=None):
    """Residual layer fed by two identity connections (channel 0 from in, channel
    1 from rec) in one pass: the same result as two ident_deliver calls onto a
    reset GSyn followed by layer_update, bit for bit."""
    a = abi.ErrorUpdateArgs()
    a.loc, a.in_loc, a.rec_loc = loc_, in_loc, rec_loc
    a.in_, a.rec = _ptr(in_activity), _ptr(rec_activity)
    a.channel_coeffs[0], a.channel_coeffs[1] = channel_coeffs
    a.scale = 1.0 if err_scale is None else err_scale
    a.apply_scale = 0 if err_scale is None else 1
    a.transfer = transfer
    a.V, a.A = _ptr(V), _ptr(A)
    abi.check(abi.hip().pvhip_error_update(C.byref(a), _stream(stream)), "pvhip_error_update")


def ident_deliver(loc_, pre_activity, gsyn, accumulate=False, stream=None):
    abi.check(abi.hip().pvhip_ident_deliver(C.byref(loc_), _ptr(pre_activity), _ptr(gsyn),
                                            int(accumulate), _stream(stream)), "pvhip_ident_deliver")


def ann_activity(loc_, transfer, V, A, stream=None):
    abi.check(abi.hip().pvhip_ann_activity(C.byref(loc_), C.byref(transfer), _ptr(V), _ptr(A),
                                           _stream(stream)), "pvhip_ann_activity")


def fill_zero(t, stream=None):
    abi.check(abi.hip().pvhip_fill_zero(_ptr(t), t.numel() * t.element_size(), _stream(stream)),
              "pvhip_fill_zero")


def cost_sum_f64(loc_, buf, extended, kind, p0=0.0, p1=0.0, out=None, stream=None):
    """CostFunctionSum in FP64 per batch element (kind 0 L2, 1 L1, 2 firm, 3 L0)."""
    if out is None:
        out = torch.zeros(loc_.nbatch, dtype=torch.float64, device=buf.device)
    abi.check(abi.hip().pvhip_cost_sum_f64(C.byref(loc_), _ptr(buf), int(extended), kind, p0, p1,
                                           _ptr(out), _stream(stream)), "pvhip_cost_sum_f64")
    return out


def normalize_l2_f64(weights, num_patches, patch_size, strength=1.0, min_l2=0.0, stream=None):
    abi.check(abi.hip().pvhip_normalize_l2_f64(_ptr(weights), num_patches, patch_size, strength,
                                               min_l2, _stream(stream)), "pvhip_normalize_l2_f64")


class HebbianUpdater:
    """HebbianUpdater dW for a shared-weight connection pre -> post with
    pre-perspective patch nxp x nyp x post.nf."""

    def __init__(self, pre, post, nxp, nyp, dw_max, normalize_dw=True, device="cuda"):
        self.pre, self.post = pre, post
        self.nxp, self.nyp, self.nfp = nxp, nyp, post.nf
        self.dw_max = dw_max
        self.normalize_dw = normalize_dw
        H = abi.host()
        tabs = []
        for axis, n in ((0, pre.nx_ext), (1, pre.ny_ext)):
            dim, ps, pr, pe, us = (np.zeros(n, np.int32) for _ in range(5))
            P = lambda a: a.ctypes.data_as(C.c_void_p)
            H.pvhost_patch_axis(C.byref(pre), C.byref(post), nxp, nyp, axis, P(dim), P(ps), P(pr),
                                P(pe), P(us))
            tabs.append((dim, ps, pe))
        (dx, psx, pex), (dy, psy, pey) = tabs
        self.tables = [_dev_i32(a, device) for a in (dx, dy, psx, psy, pex, pey)]
        nkx, nky = C.c_int(), C.c_int()
        H.pvhost_num_kernels(C.byref(pre), C.byref(post), C.byref(nkx), C.byref(nky))
        self.nkx, self.nky = nkx.value, nky.value
        self.size = self.nkx * self.nky * pre.nf * nxp * nyp * self.nfp
        self.dw = torch.zeros(self.size, dtype=torch.float32, device=device)
        self.counts = torch.zeros(self.size, dtype=torch.int64, device=device) if normalize_dw else None
        # MFMA forms: window geometry (None when only the EXACT form applies)
        w = [C.c_int() for _ in range(4)]
        r = H.pvhost_hebbian_windows(C.byref(pre), C.byref(post), nxp, nyp, *[C.byref(v) for v in w])
        self.windows = tuple(v.value for v in w) if r == 0 and self.nkx * self.nky == 1 else None
        self._scratch = None
        self.rows_per_chunk = 0   # MFMA split-K chunk rows, 0 = the library's default (autotuner)

    def update_dw(self, pre_activity, post_activity, clear=True, arith=abi.ARITH_EXACT, stream=None):
        if clear:
            fill_zero(self.dw, stream)
            if self.counts is not None:
                fill_zero(self.counts, stream)
        a = abi.HebbianArgs()
        a.pre, a.post = self.pre, self.post
        a.nxp, a.nyp, a.nfp = self.nxp, self.nyp, self.nfp
        a.num_kernels_x, a.num_kernels_y = self.nkx, self.nky
        a.dim_x, a.dim_y, a.pstart_x, a.pstart_y, a.apost_x, a.apost_y = [_ptr(t) for t in self.tables]
        a.dw_max = self.dw_max
        a.pre_data, a.post_data = _ptr(pre_activity), _ptr(post_activity)
        a.dw = _ptr(self.dw)
        a.counts = _ptr(self.counts)
        a.arith = arith
        a.rows_per_chunk = self.rows_per_chunk
        if arith in (abi.ARITH_FP32, abi.ARITH_BF16):
            if self.windows is None:
                raise abi.PVHIPError("the MFMA forms of the Hebbian update need one data patch per "
                                     "feature and windows clipped to the post layer")
            a.window_stride_x, a.window_start_x, a.window_stride_y, a.window_start_y = self.windows
            n = abi.hip().pvhip_hebbian_dw_scratch_bytes(C.byref(a))
            if self._scratch is None or self._scratch.numel() * 4 < n:
                self._scratch = torch.empty(max(1, (n + 3) // 4), dtype=torch.float32,
                                            device=self.dw.device)
            a.scratch, a.scratch_bytes = _ptr(self._scratch), self._scratch.numel() * 4
        abi.check(abi.hip().pvhip_hebbian_dw(C.byref(a), _stream(stream)), "pvhip_hebbian_dw")

    def apply(self, weights, stream=None):
        abi.check(abi.hip().pvhip_hebbian_apply(_ptr(weights), _ptr(self.dw), _ptr(self.counts),
                                                self.size, int(self.normalize_dw), _stream(stream)),
                  "pvhip_hebbian_apply")


_MOMENTUM_METHODS = {"viscosity": abi.MOMENTUM_VISCOSITY, "simple": abi.MOMENTUM_SIMPLE,
                     "alex": abi.MOMENTUM_ALEX}


def momentum_factors(method, tau, decay=0.0, dw_max=1.0, deprecated_tau=False):
    """(dwFactor, wFactor) of MomentumUpdater::applyMomentum for momentumMethod
    `method` ("viscosity", "simple", "alex" or its code); PVHIPError where the
    reference's parameter checks reject the values.  deprecated_tau: `tau` is
    the deprecated momentumTau (applyMomentumDeprecated's factors, which the
    reference does not range-check)."""
    code = _MOMENTUM_METHODS[method] if isinstance(method, str) else int(method)
    f, wf = C.c_float(), C.c_float()
    fn = "pvhip_momentum_factors_deprecated" if deprecated_tau else "pvhip_momentum_factors"
    abi.check(getattr(abi.hip(), fn)(code, tau, decay, dw_max, C.byref(f), C.byref(wf)), fn)
    return f.value, wf.value


class MomentumUpdater(HebbianUpdater):
    """MomentumUpdater (src/weightupdaters/MomentumUpdater.cpp): the Hebbian dW,
    then, at each weight update, dW blended with the previous update's dW and
    decayed by the weights, prev_dW = dW, w += dW.  Parameters as the reference's
    .params keys: momentumMethod, timeConstantTau (its default per method:
    viscosity 100, simple 0.25, alex 0.9), momentumDecay, and momentumTau, the
    deprecated spelling that selects the older blend (MomentumUpdater.cpp:346-360),
    whose factors the reference computes without a range check
    (applyMomentumDeprecated, :333-343); only momentumDecay is checked then."""

    _TAU_DEFAULT = {"viscosity": 100.0, "simple": 0.25, "alex": 0.9}

    def __init__(self, pre, post, nxp, nyp, dw_max, momentumMethod="viscosity",
                 timeConstantTau=None, momentumDecay=0.0, momentumTau=None, normalize_dw=True,
                 device="cuda"):
        super().__init__(pre, post, nxp, nyp, dw_max, normalize_dw=normalize_dw, device=device)
        self.method = momentumMethod
        self.deprecated_tau = timeConstantTau is None and momentumTau is not None
        if self.deprecated_tau:
            tau = float(momentumTau)
        else:
            tau = self._TAU_DEFAULT[momentumMethod] if timeConstantTau is None else float(timeConstantTau)
        self.tau, self.decay = tau, float(momentumDecay)
        self.dw_factor, self.w_factor = momentum_factors(momentumMethod, tau, self.decay, dw_max,
                                                         deprecated_tau=self.deprecated_tau)
        self.prev_dw = torch.zeros(self.size, dtype=torch.float32, device=self.dw.device)

    def apply(self, weights, stream=None):
        abi.check(abi.hip().pvhip_momentum_apply(
            _ptr(weights), _ptr(self.dw), _ptr(self.prev_dw), _ptr(self.counts), self.size,
            int(self.normalize_dw), self.dw_factor, self.w_factor, int(self.deprecated_tau),
            _stream(stream)), "pvhip_momentum_apply")
