"""ctypes mirror of include/pvhip.h and include/pvhost.h.

The product's compute runs in libpvhip.so (HIP kernels for gfx950); this
module only loads it and declares its C ABI.  Loading fails loudly when the
library is missing: there is no CPU fallback.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(HERE, "lib")

PVHIP_OK = 0
ARITH_EXACT, ARITH_FP32, ARITH_BF16, ARITH_FP8 = 0, 1, 2, 3
MOMENTUM_VISCOSITY, MOMENTUM_SIMPLE, MOMENTUM_ALEX = 0, 1, 2  # momentumMethod codes
TRANSFER_VTHRESH, TRANSFER_VERTICES = 0, 1
MAX_VERTICES = 8
MAX_CHANNELS = 4

vp = C.c_void_p
i32 = C.c_int
f32 = C.c_float


class Halo(C.Structure):
    _fields_ = [("lt", i32), ("rt", i32), ("dn", i32), ("up", i32)]


class Loc(C.Structure):
    """pvhip_loc -- PVLayerLoc of one rank (src/include/PVLayerLoc.h:21-26)."""
    _fields_ = [("nbatch", i32), ("nx", i32), ("ny", i32), ("nf", i32), ("halo", Halo)]

    @classmethod
    def make(cls, nbatch, nx, ny, nf, halo=0):
        if isinstance(halo, int):
            halo = (halo,) * 4
        return cls(nbatch, nx, ny, nf, Halo(*halo))

    @property
    def nx_ext(self):
        return self.nx + self.halo.lt + self.halo.rt

    @property
    def ny_ext(self):
        return self.ny + self.halo.dn + self.halo.up

    @property
    def num_ext(self):
        return self.nx_ext * self.ny_ext * self.nf

    @property
    def num_res(self):
        return self.nx * self.ny * self.nf

    def halo_tuple(self):
        return (self.halo.lt, self.halo.rt, self.halo.dn, self.halo.up)

    def __repr__(self):
        return "Loc(nb=%d, %dx%dx%d, halo=%s)" % (self.nbatch, self.nx, self.ny, self.nf,
                                                  self.halo_tuple())


class Props(C.Structure):
    _fields_ = [("name", C.c_char * 256), ("arch", C.c_char * 64), ("ordinal", i32),
                ("compute_units", i32), ("wavefront_size", i32), ("max_threads_per_block", i32),
                ("max_grid_x", i32), ("clock_khz", i32), ("memory_clock_khz", i32),
                ("memory_bus_width", i32), ("l2_cache_bytes", i32), ("lds_per_block", C.c_size_t),
                ("total_global_mem", C.c_size_t)]


class Transfer(C.Structure):
    _fields_ = [("mode", i32), ("VThresh", f32), ("AMin", f32), ("AMax", f32), ("AShift", f32),
                ("VWidth", f32), ("num_vertices", i32), ("vertices_v", f32 * MAX_VERTICES),
                ("vertices_a", f32 * MAX_VERTICES), ("slopes", f32 * (MAX_VERTICES + 1))]


class LCAUpdateArgs(C.Structure):
    _fields_ = [("loc", Loc), ("num_channels", i32), ("channel_coeffs", f32 * MAX_CHANNELS),
                ("gsyn", vp), ("V", vp), ("A", vp), ("exp_tau", vp), ("self_interact", i32),
                ("transfer", Transfer), ("activity_from_v", i32), ("A_bf16c", vp)]


class ActiveOut(C.Structure):
    _fields_ = [("entries", vp), ("num_active", vp), ("pos_start", vp), ("scratch", vp),
                ("scratch_bytes", C.c_size_t), ("halo_is_zero", i32)]


class LayerUpdateArgs(C.Structure):
    _fields_ = [("loc", Loc), ("num_channels", i32), ("channel_coeffs", f32 * MAX_CHANNELS),
                ("gsyn", vp), ("V", vp), ("A", vp), ("scale", f32), ("apply_scale", i32),
                ("transfer", Transfer)]


class RecvPostArgs(C.Structure):
    _fields_ = [("pre", Loc), ("post", Loc), ("nxp", i32), ("nyp", i32), ("nfp", i32),
                ("num_kernels_x", i32), ("num_kernels_y", i32),
                ("start_x", vp), ("start_y", vp), ("kernel_x", vp), ("kernel_y", vp),
                ("dt_factor", f32), ("pre_data", vp), ("weights", vp), ("post_gsyn", vp),
                ("accumulate", i32), ("arith", i32), ("weights_lp", vp),
                ("inverse_x", vp), ("inverse_y", vp), ("window_x0", i32), ("window_nx", i32),
                ("window_y0", i32), ("window_ny", i32),
                ("fused_update", C.POINTER(LCAUpdateArgs)),
                ("active_out", C.POINTER(ActiveOut)), ("scratch", vp), ("scratch_bytes", C.c_size_t),
                ("cgemm_tile", i32), ("cgemm_splits", i32), ("lp_source_scale", f32),
                ("pre_data_bf16c", vp), ("fused_layer", C.POINTER(LayerUpdateArgs))]


class RecvPreArgs(C.Structure):
    _fields_ = [("pre", Loc), ("post", Loc), ("nxp", i32), ("nyp", i32), ("nfp", i32),
                ("num_kernels_x", i32), ("num_kernels_y", i32),
                ("col_ptr", vp), ("col_xe", vp), ("col_wx", vp), ("col_kx", vp),
                ("row_ptr", vp), ("row_ye", vp), ("row_wy", vp), ("row_ky", vp),
                ("dt_factor", f32), ("active", vp), ("pos_start", vp), ("pre_data", vp),
                ("weights", vp), ("post_gsyn", vp), ("accumulate", i32), ("arith", i32)]


class RecvPreDenseArgs(C.Structure):
    _fields_ = [("pre", Loc), ("post", Loc), ("nxp", i32), ("nyp", i32), ("nfp", i32),
                ("start_x0", i32), ("start_y0", i32), ("halo_is_zero", i32),
                ("pre_data_bf16c", vp), ("weights_bf16", vp), ("dt_factor", f32),
                ("post_gsyn", vp), ("accumulate", i32), ("fused_layer", C.POINTER(LayerUpdateArgs)),
                ("fp32", i32), ("pre_data", vp), ("weights_f32", vp)]


class ErrorUpdateArgs(C.Structure):
    _fields_ = [("loc", Loc), ("in_loc", Loc), ("in_", vp), ("rec_loc", Loc), ("rec", vp),
                ("channel_coeffs", f32 * 2), ("scale", f32), ("apply_scale", i32),
                ("transfer", Transfer), ("V", vp), ("A", vp)]


class LCAAdjointArgs(C.Structure):
    _fields_ = [("loc", Loc), ("transfer", Transfer), ("exp_tau", vp), ("self_interact", i32),
                ("V", vp), ("gA", vp), ("l1", f32), ("gV", vp), ("gG", vp), ("gA_prev", vp)]


class HebbianArgs(C.Structure):
    _fields_ = [("pre", Loc), ("post", Loc), ("nxp", i32), ("nyp", i32), ("nfp", i32),
                ("num_kernels_x", i32), ("num_kernels_y", i32),
                ("dim_x", vp), ("dim_y", vp), ("pstart_x", vp), ("pstart_y", vp),
                ("apost_x", vp), ("apost_y", vp), ("dw_max", f32), ("pre_data", vp),
                ("post_data", vp), ("dw", vp), ("counts", vp), ("arith", i32),
                ("window_stride_x", i32), ("window_start_x", i32), ("window_stride_y", i32),
                ("window_start_y", i32), ("scratch", vp), ("scratch_bytes", C.c_size_t),
                ("rows_per_chunk", i32)]


# Every symbol include/pvhip.h declares, with its ctypes signature.
PVHIP_SYMBOLS = {
    "pvhip_last_error": (C.c_char_p, []),
    "pvhip_abi_version": (i32, []),
    "pvhip_device_create": (i32, [i32, C.POINTER(vp)]),
    "pvhip_device_destroy": (i32, [vp]),
    "pvhip_device_sync": (i32, [vp]),
    "pvhip_device_props": (i32, [vp, C.POINTER(Props)]),
    "pvhip_device_stream": (vp, [vp]),
    "pvhip_device_allocated_bytes": (C.c_size_t, [vp]),
    "pvhip_buffer_alloc": (i32, [vp, C.c_size_t, C.POINTER(vp)]),
    "pvhip_buffer_free": (i32, [vp]),
    "pvhip_buffer_h2d": (i32, [vp, vp, C.c_size_t, C.c_size_t, vp]),
    "pvhip_buffer_d2h": (i32, [vp, vp, C.c_size_t, C.c_size_t, vp]),
    "pvhip_buffer_ptr": (vp, [vp]),
    "pvhip_buffer_size": (C.c_size_t, [vp]),
    "pvhip_fill_zero": (i32, [vp, C.c_size_t, vp]),
    "pvhip_stream_sync": (i32, [vp]),
    "pvhip_check_faults": (i32, []),
    "pvhip_copy_device": (i32, [vp, vp, C.c_size_t, vp]),
    "pvhip_event_create": (i32, [C.POINTER(vp)]),
    "pvhip_event_destroy": (i32, [vp]),
    "pvhip_event_record": (i32, [vp, vp]),
    "pvhip_event_elapsed_ms": (i32, [vp, vp, C.POINTER(f32)]),
    "pvhip_graph_begin": (i32, [vp]),
    "pvhip_graph_end": (i32, [vp, C.POINTER(vp)]),
    "pvhip_graph_launch": (i32, [vp, vp]),
    "pvhip_graph_destroy": (i32, [vp]),
    "pvhip_lca_exp_tau": (None, [C.POINTER(C.c_double), i32, f32, C.POINTER(f32)]),
    "pvhip_recv_post": (i32, [C.POINTER(RecvPostArgs), vp]),
    "pvhip_recv_post_active_scratch_bytes": (C.c_size_t, [C.POINTER(RecvPostArgs)]),
    "pvhip_recv_post_scratch_bytes": (C.c_size_t, [C.POINTER(RecvPostArgs)]),
    "pvhip_recv_post_fp8_weights": (i32, [vp, vp, i32, i32, vp]),
    "pvhip_recv_post_fp8_weights_bytes": (C.c_size_t, [i32, i32]),
    "pvhip_recv_post_bf16_weights": (i32, [vp, vp, i32, i32, vp]),
    "pvhip_recv_post_bf16_weights_bytes": (C.c_size_t, [i32, i32]),
    "pvhip_recv_pre": (i32, [C.POINTER(RecvPreArgs), vp]),
    "pvhip_recv_pre_dense": (i32, [C.POINTER(RecvPreDenseArgs), vp]),
    "pvhip_recv_pre_dense_weights": (i32, [vp, vp, i32, i32, vp]),
    "pvhip_recv_pre_dense_weights_bytes": (C.c_size_t, [i32, i32]),
    "pvhip_recv_pre_dense_weights_f32": (i32, [vp, vp, i32, i32, vp]),
    "pvhip_recv_pre_dense_fits": (i32, [C.POINTER(RecvPreDenseArgs)]),
    "pvhip_lca_update": (i32, [C.POINTER(LCAUpdateArgs), vp]),
    "pvhip_momentum_lca_update": (i32, [C.POINTER(LCAUpdateArgs), C.c_float, vp, vp]),
    "pvhip_ista_update": (i32, [C.POINTER(LCAUpdateArgs), C.c_float, vp]),
    "pvhip_ista_rate": (None, [C.POINTER(C.c_double), i32, f32, C.POINTER(f32)]),
    "pvhip_gsyn_accumulate": (i32, [i32, i32, C.POINTER(f32), vp, vp, vp]),
    "pvhip_layer_update": (i32, [C.POINTER(LayerUpdateArgs), vp]),
    "pvhip_ann_activity": (i32, [C.POINTER(Loc), C.POINTER(Transfer), vp, vp, vp]),
    "pvhip_error_update": (i32, [C.POINTER(ErrorUpdateArgs), vp]),
    "pvhip_activity_to_bf16c": (i32, [C.POINTER(Loc), vp, vp, vp]),
    "pvhip_ident_deliver": (i32, [C.POINTER(Loc), vp, vp, i32, vp]),
    "pvhip_active_compact": (i32, [C.POINTER(Loc), vp, vp, vp, vp, vp, vp]),
    "pvhip_active_compact_noalloc": (i32, [C.POINTER(Loc), vp, vp, vp, vp, vp]),
    "pvhip_active_compact_scratch_bytes": (C.c_size_t, [C.POINTER(Loc)]),
    "pvhip_hebbian_dw": (i32, [C.POINTER(HebbianArgs), vp]),
    "pvhip_hebbian_dw_scratch_bytes": (C.c_size_t, [C.POINTER(HebbianArgs)]),
    "pvhip_hebbian_apply": (i32, [vp, vp, vp, C.c_size_t, i32, vp]),
    "pvhip_momentum_factors": (i32, [i32, C.c_float, C.c_float, C.c_float, C.POINTER(C.c_float),
                                     C.POINTER(C.c_float)]),
    "pvhip_momentum_factors_deprecated": (i32, [i32, C.c_float, C.c_float, C.c_float,
                                                C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "pvhip_momentum_apply": (i32, [vp, vp, vp, vp, C.c_size_t, i32, C.c_float, C.c_float, i32, vp]),
    "pvhip_permute_weights": (i32, [vp, vp, vp, C.c_size_t, vp]),
    "pvhip_normalize_l2_f64": (i32, [vp, i32, i32, C.c_double, C.c_double, vp]),
    "pvhip_normalize_l2": (i32, [vp, i32, i32, f32, f32, vp]),
    "pvhip_cost_sum_f64": (i32, [C.POINTER(Loc), vp, i32, i32, C.c_double, C.c_double, vp, vp]),
    "pvhip_lca_adjoint_step": (i32, [C.POINTER(LCAAdjointArgs), vp]),
    "pvhip_scale_to_extended": (i32, [C.POINTER(Loc), f32, vp, vp, i32, vp]),
    "pvhip_axpby": (i32, [C.c_size_t, f32, vp, f32, vp, vp]),
}

i32p = C.POINTER(C.c_int32)
PVHOST_SYMBOLS = {
    "pvhost_required_margin": (i32, [i32, i32, i32]),
    "pvhost_post_patch_size": (i32, [i32, i32, i32]),
    "pvhost_num_kernels": (i32, [C.POINTER(Loc), C.POINTER(Loc), C.POINTER(i32), C.POINTER(i32)]),
    "pvhost_patch_axis": (i32, [C.POINTER(Loc), C.POINTER(Loc), i32, i32, i32, vp, vp, vp, vp, vp]),
    "pvhost_recv_post_tables": (i32, [C.POINTER(Loc), C.POINTER(Loc), i32, i32, vp, vp, vp, vp,
                                      C.POINTER(i32), C.POINTER(i32)]),
    "pvhost_recv_post_inverse": (i32, [C.POINTER(Loc), C.POINTER(Loc), i32, i32, vp, vp,
                                       C.POINTER(i32), C.POINTER(i32), C.POINTER(i32),
                                       C.POINTER(i32)]),
    "pvhost_recv_pre_tables": (i32, [C.POINTER(Loc), C.POINTER(Loc), i32, i32, vp, vp, vp, vp,
                                     C.POINTER(i32), vp, vp, vp, vp, C.POINTER(i32)]),
    "pvhost_recv_pre_dense_start": (i32, [C.POINTER(Loc), C.POINTER(Loc), i32, i32, C.POINTER(i32),
                                          C.POINTER(i32), C.POINTER(i32)]),
    "pvhost_make_transfer": (i32, [f32, f32, f32, f32, f32, i32, C.POINTER(Transfer)]),
    "pvhost_transpose_weights": (i32, [C.POINTER(Loc), C.POINTER(Loc), i32, i32, i32, vp, vp]),
    "pvhost_transpose_weights_index": (i32, [C.POINTER(Loc), C.POINTER(Loc), i32, i32, i32, vp,
                                             C.c_size_t]),
    "pvhost_hebbian_windows": (i32, [C.POINTER(Loc), C.POINTER(Loc), i32, i32, C.POINTER(i32),
                                     C.POINTER(i32), C.POINTER(i32), C.POINTER(i32)]),
}

# the .params reader (pvparams.h), exported by libpvhost.so
PVPARAMS_SYMBOLS = {
    "pvparams_load": (i32, [C.c_char_p, C.POINTER(vp)]),
    "pvparams_parse": (i32, [C.c_char_p, C.POINTER(vp)]),
    "pvparams_error": (C.c_char_p, []),
    "pvparams_free": (None, [vp]),
    "pvparams_num_groups": (i32, [vp]),
    "pvparams_group_keyword": (C.c_char_p, [vp, i32]),
    "pvparams_group_name": (C.c_char_p, [vp, i32]),
    "pvparams_find_group": (i32, [vp, C.c_char_p]),
    "pvparams_number": (i32, [vp, i32, C.c_char_p, C.POINTER(C.c_double)]),
    "pvparams_string": (i32, [vp, i32, C.c_char_p, C.POINTER(C.c_char_p)]),
    "pvparams_array": (i32, [vp, i32, C.c_char_p, C.POINTER(C.c_double), i32]),
    "pvparams_directive": (i32, [vp, C.c_char_p, C.POINTER(C.c_double)]),
}

_libs = {}


class PVHIPError(RuntimeError):
    pass


def _load(name, symbols):
    if name in _libs:
        return _libs[name]
    path = os.path.join(LIBDIR, name)
    if not os.path.exists(path):
        raise PVHIPError("%s is not built (run __graft_entry__.build() or python -m openpv_amd.build)"
                         % path)
    lib = C.CDLL(path)
    for sym, (res, args) in symbols.items():
        fn = getattr(lib, sym)
        fn.restype = res
        fn.argtypes = args
    _libs[name] = lib
    return lib


def hip():
    """libpvhip.so (device layer + kernels)."""
    return _load("libpvhip.so", PVHIP_SYMBOLS)


def host():
    """libpvhost.so (host geometry / tables, and the .params reader)."""
    return _load("libpvhost.so", dict(PVHOST_SYMBOLS, **PVPARAMS_SYMBOLS))


def pvrun_path():
    """The pvrun executable (runs a .params file on the device)."""
    path = os.path.join(LIBDIR, "pvrun")
    if not os.path.exists(path):
        raise PVHIPError("%s is not built (run __graft_entry__.build() or python -m openpv_amd.build)"
                         % path)
    return path


def check(status, what=""):
    if status != PVHIP_OK:
        msg = hip().pvhip_last_error().decode(errors="replace")
        raise PVHIPError("%s failed (status %d): %s" % (what or "pvhip call", status, msg))


def library_paths():
    return [os.path.join(LIBDIR, n) for n in ("libpvhip.so", "libpvhost.so")]
