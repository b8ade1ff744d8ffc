"""CPU oracle for OpenPV's delivery + LCA path -- TEST INFRASTRUCTURE ONLY.

Thin numpy/ctypes wrappers over ``lib/liboracle.so`` (the plain-C restatement
in ``pvoracle.c``) and, where it has been built, ``_ref/libpvref.so`` (the
reference's own PatchGeometry / .kpp kernels / cost functions compiled from the
reference checkout by ``make ref``).  Only tests/, ``__graft_entry__.smoke()``
and bench.py's cpu_baseline leg import this package; the product never does.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "lib", "liboracle.so")
_REF = os.path.join(_HERE, "_ref", "libpvref.so")

f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
u16p = np.ctypeslib.ndpointer(np.uint16, flags="C_CONTIGUOUS")

ENTRY_DTYPE = np.dtype([("index", np.uint32), ("value", np.float32)])
PATCH_DTYPE = np.dtype([("nx", np.uint16), ("ny", np.uint16), ("offset", np.uint32)])
entp = np.ctypeslib.ndpointer(ENTRY_DTYPE, flags="C_CONTIGUOUS")
patp = np.ctypeslib.ndpointer(PATCH_DTYPE, flags="C_CONTIGUOUS")


class Loc(C.Structure):
    """or_loc: nbatch, nx, ny, nf, lt, rt, dn, up (PVLayerLoc for one rank)."""
    _fields_ = [(n, C.c_int) for n in ("nbatch", "nx", "ny", "nf", "lt", "rt", "dn", "up")]

    @classmethod
    def make(cls, nbatch, nx, ny, nf, halo=0):
        if isinstance(halo, int):
            halo = (halo, halo, halo, halo)
        return cls(nbatch, nx, ny, nf, *halo)

    @property
    def nx_ext(self):
        return self.nx + self.lt + self.rt

    @property
    def ny_ext(self):
        return self.ny + self.dn + self.up

    @property
    def num_ext(self):
        return self.nx_ext * self.ny_ext * self.nf

    @property
    def num_res(self):
        return self.nx * self.ny * self.nf

    def as_list(self):
        return [self.nbatch, self.nx, self.ny, self.nf, self.lt, self.rt, self.dn, self.up]


class LCAConfig(C.Structure):
    _fields_ = [
        ("input", Loc), ("res", Loc), ("v1", Loc),
        ("nxp", C.c_int), ("nyp", C.c_int),
        ("VThresh", C.c_float), ("AMin", C.c_float), ("AMax", C.c_float),
        ("AShift", C.c_float), ("VWidth", C.c_float),
        ("transfer_vertices", C.c_int),
        ("tau", C.c_double), ("dt", C.c_double),
        ("selfInteract", C.c_int), ("errScale", C.c_float), ("num_threads", C.c_int),
    ]


def build(force=False):
    """Compile liboracle.so (and _ref when the reference checkout is present)."""
    if force or not os.path.exists(_LIB):
        subprocess.check_call(["make", "-s", "-C", _HERE, "lib/liboracle.so"])
    if os.path.isdir("/root/reference/src") and (force or not os.path.exists(_REF)):
        subprocess.check_call(["make", "-s", "-C", _HERE, "ref"])


_lib = None
_ref = None

LP = C.POINTER(Loc)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        L.or_kIndexExtended.restype = C.c_int
        L.or_requiredConvolveMargin.restype = C.c_int
        L.or_calcPostPatchSize.restype = C.c_int
        L.or_calcPatchStartInPost.restype = C.c_int
        L.or_data_index.restype = C.c_int
        L.or_transpose_item_index.restype = C.c_int
        L.or_patch_geometry.argtypes = [LP, LP, C.c_int, C.c_int, C.c_int, patp, u64p, u64p, u64p]
        L.or_patch_axis.argtypes = [LP, LP, C.c_int, C.c_int, C.c_int, i32p, i32p, i32p, i32p, i32p]
        L.or_transpose_weights_shared.argtypes = [LP, LP, C.c_int, C.c_int, C.c_int, f32p, f32p]
        L.or_recv_post.argtypes = [LP, LP, C.c_int, C.c_int, f32p, f32p, f32p, C.c_float, C.c_int]
        L.or_recv_pre.argtypes = [LP, LP, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                  f32p, f32p, C.c_float, C.c_int]
        L.or_ident_deliver.argtypes = [LP, f32p, f32p]
        L.or_active_indices.argtypes = [LP, f32p, entp, i64p]
        L.or_gsyn_accumulate.argtypes = [C.c_int, C.c_int, f32p, f32p, f32p]
        L.or_update_lca.argtypes = [LP, C.c_int, f64p, C.c_float, f32p, f32p, f32p]
        L.or_activity_copy.argtypes = [LP, f32p, f32p]
        L.or_apply_vthresh.argtypes = [LP, f32p, C.c_float, C.c_float, C.c_float, C.c_float, f32p]
        L.or_apply_amax.argtypes = [LP, C.c_float, f32p]
        L.or_set_vertices.argtypes = [C.c_float] * 5 + [f32p, f32p, f32p]
        L.or_set_vertices.restype = C.c_int
        L.or_apply_vertices.argtypes = [LP, C.c_int, f32p, f32p, f32p, f32p, f32p]
        L.or_errscale.argtypes = [C.c_int, C.c_float, f32p]
        L.or_hebbian_update_dw.argtypes = [LP, LP, C.c_int, C.c_int, C.c_int, C.c_float,
                                           f32p, f32p, f32p, C.c_void_p]
        L.or_hebbian_normalize_apply.argtypes = [f32p, f32p, C.c_void_p, C.c_size_t, C.c_int]
        L.or_momentum_factors.argtypes = [C.c_int, C.c_float, C.c_float, C.c_float,
                                          C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.or_momentum_factors.restype = C.c_int
        L.or_momentum_factors_deprecated.argtypes = L.or_momentum_factors.argtypes
        L.or_momentum_factors_deprecated.restype = C.c_int
        L.or_momentum_apply.argtypes = [f32p, f32p, f32p, C.c_void_p, C.c_size_t, C.c_int,
                                        C.c_float, C.c_float, C.c_int]
        L.or_update_momentum_lca.argtypes = [LP, C.c_int, C.c_float, f64p, C.c_float, f32p, f32p,
                                             f32p, f32p]
        L.or_normalize_l2.argtypes = [f32p, C.c_int, C.c_int, C.c_float, C.c_float]
        L.or_update_ista.argtypes = [LP, C.c_float, f64p, C.c_float, f32p, f32p, f32p]
        L.or_normalize_l2_f64.argtypes = [f32p, C.c_int, C.c_int, C.c_double, C.c_double]
        L.or_cost_sum.argtypes = [LP, f32p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double]
        L.or_cost_sum.restype = C.c_double
        L.or_lca_run.argtypes = [C.POINTER(LCAConfig), f32p, f32p, C.c_void_p, C.c_int,
                                 C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.or_timescale_init.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double]
        L.or_timescale_calc.argtypes = [C.c_void_p, C.c_int, f64p, C.c_double, C.c_double,
                                        C.c_double, C.c_double, f64p]
        L.or_need_update.argtypes = [C.c_double] * 4
        L.or_need_update.restype = C.c_int
        _lib = L
    return _lib


def ref_available():
    return os.path.exists(_REF)


def ref():
    """The compiled reference pieces (None when not built)."""
    global _ref
    if _ref is None and os.path.exists(_REF):
        R = C.CDLL(_REF)
        R.ref_patch_geometry.argtypes = [i32p, i32p, C.c_int, C.c_int, C.c_int, u16p, u16p, u32p,
                                         u64p, u64p, u64p, C.c_void_p]
        R.ref_patch_geometry.restype = C.c_int
        R.ref_update_lca.argtypes = [i32p, C.c_int, f64p, C.c_float, f32p, f32p, f32p]
        R.ref_update_momentum_lca.argtypes = [i32p, C.c_int, C.c_float, f64p, C.c_float, f32p,
                                              f32p, f32p, f32p]
        R.ref_apply_vertices.argtypes = [i32p, C.c_int, f32p, f32p, f32p, f32p, f32p]
        R.ref_update_ista.argtypes = [i32p, C.c_float, f64p, C.c_float, f32p, f32p, f32p]
        R.ref_gsyn_accumulate.argtypes = [C.c_int, C.c_int, f32p, f32p, f32p]
        R.ref_cost.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double]
        R.ref_cost.restype = C.c_double
        R.ref_required_margin.restype = C.c_int
        R.ref_kindex_extended.restype = C.c_int
        _ref = R
    return _ref


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


# ----------------------------------------------------------------- wrappers

def patch_geometry(pre, post, nxp, nyp, nfp):
    n = pre.num_ext
    patches = np.zeros(n, PATCH_DTYPE)
    gs = np.zeros(n, np.uint64)
    ap = np.zeros(n, np.uint64)
    us = np.zeros(n, np.uint64)
    lib().or_patch_geometry(C.byref(pre), C.byref(post), nxp, nyp, nfp, patches, gs, ap, us)
    return patches, gs, ap, us


def patch_axis(pre, post, nxp, nyp, axis):
    n = pre.nx_ext if axis == 0 else pre.ny_ext
    out = [np.zeros(n, np.int32) for _ in range(5)]
    lib().or_patch_axis(C.byref(pre), C.byref(post), nxp, nyp, axis, *out)
    return dict(zip(("dim", "pstart", "post_start_res", "post_start_ext", "unshrunken"), out))


def num_kernels(pre, post):
    kx = pre.nx // post.nx if pre.nx > post.nx else 1
    ky = pre.ny // post.ny if pre.ny > post.ny else 1
    return kx, ky, pre.nf


def recv_post(source, target, nxp, nyp, pre_activity, post_weights, gsyn=None, dt_factor=1.0,
              num_threads=0):
    if gsyn is None:
        gsyn = np.zeros(target.nbatch * target.num_res, np.float32)
    lib().or_recv_post(C.byref(source), C.byref(target), nxp, nyp,
                       np.ascontiguousarray(pre_activity, np.float32).ravel(),
                       np.ascontiguousarray(post_weights, np.float32).ravel(), gsyn,
                       dt_factor, num_threads)
    return gsyn


def recv_pre(pre, post, nxp, nyp, weights, pre_activity=None, entries=None, num_active=None,
             gsyn=None, dt_factor=1.0, num_threads=1):
    if gsyn is None:
        gsyn = np.zeros(post.nbatch * post.num_res, np.float32)
    pa = None if pre_activity is None else np.ascontiguousarray(pre_activity, np.float32).ravel()
    lib().or_recv_pre(C.byref(pre), C.byref(post), nxp, nyp, _ptr(pa), _ptr(entries),
                      _ptr(num_active), np.ascontiguousarray(weights, np.float32).ravel(), gsyn,
                      dt_factor, num_threads)
    return gsyn


def active_indices(loc, A):
    ent = np.zeros(loc.nbatch * loc.num_ext, ENTRY_DTYPE)
    n = np.zeros(loc.nbatch, np.int64)
    lib().or_active_indices(C.byref(loc), np.ascontiguousarray(A, np.float32).ravel(), ent, n)
    return ent, n


def set_vertices(VThresh, AMin, AMax, AShift, VWidth):
    vv = np.zeros(8, np.float32)
    va = np.zeros(8, np.float32)
    sl = np.zeros(9, np.float32)
    nv = lib().or_set_vertices(VThresh, AMin, AMax, AShift, VWidth, vv, va, sl)
    return nv, vv, va, sl


def update_lca(loc, self_interact, dt, tau, gsyn, A, V):
    V = V.copy()
    lib().or_update_lca(C.byref(loc), int(self_interact), np.asarray(dt, np.float64), tau,
                        np.ascontiguousarray(gsyn, np.float32), np.ascontiguousarray(A, np.float32), V)
    return V


def update_ista(loc, VThresh, dt, tau, gsyn, A, V):
    """ISTA V update (ISTAInternalStateBuffer.kpp); returns the new V."""
    V = np.array(V, np.float32)
    lib().or_update_ista(C.byref(loc), VThresh, np.asarray(dt, np.float64), tau,
                         np.ascontiguousarray(gsyn, np.float32), np.ascontiguousarray(A, np.float32), V)
    return V


def update_momentum_lca(loc, self_interact, momentum_rate, dt, tau, gsyn, A, prev_drive, V):
    """MomentumLCA V update; returns (V, prev_drive), both new arrays."""
    V, prev = V.copy(), np.array(prev_drive, np.float32)
    lib().or_update_momentum_lca(C.byref(loc), int(self_interact), momentum_rate,
                                 np.asarray(dt, np.float64), tau,
                                 np.ascontiguousarray(gsyn, np.float32),
                                 np.ascontiguousarray(A, np.float32), prev, V)
    return V, prev


def momentum_factors(method, tau, decay=0.0, dw_max=1.0, deprecated_tau=False):
    """(dwFactor, wFactor) of MomentumUpdater::applyMomentum (or, with
    deprecated_tau, applyMomentumDeprecated); method 0 viscosity, 1 simple,
    2 alex.  Raises ValueError where the reference rejects the parameters."""
    f, wf = C.c_float(), C.c_float()
    fn = lib().or_momentum_factors_deprecated if deprecated_tau else lib().or_momentum_factors
    if fn(method, tau, decay, dw_max, C.byref(f), C.byref(wf)):
        raise ValueError("momentum parameters rejected")
    return f.value, wf.value


def momentum_apply(w, dw, prev_dw, counts, dw_factor, w_factor, deprecated_tau=False):
    """normalize_dW (counts given) + applyMomentum + prev_dW = dW + w += dW;
    returns new (w, dw, prev_dw)."""
    w, dw, prev = (np.array(x, np.float32) for x in (w, dw, prev_dw))
    cnt = None if counts is None else np.ascontiguousarray(counts, np.int64)
    lib().or_momentum_apply(w, dw, prev, None if cnt is None else cnt.ctypes.data, w.size,
                            int(cnt is not None), dw_factor, w_factor, int(deprecated_tau))
    return w, dw, prev


def transfer(loc, V, VThresh, AMin, AMax, AShift, VWidth, vertices=False, A=None):
    """A (extended) = transfer(V); CPU applyVThresh path or GPU vertices path."""
    if A is None:
        A = np.zeros(loc.nbatch * loc.num_ext, np.float32)
    else:
        A = A.copy()
    V = np.ascontiguousarray(V, np.float32)
    if vertices:
        nv, vv, va, sl = set_vertices(VThresh, AMin, AMax, AShift, VWidth)
        lib().or_apply_vertices(C.byref(loc), nv, vv, va, sl, V, A)
    else:
        lib().or_activity_copy(C.byref(loc), V, A)
        lib().or_apply_vthresh(C.byref(loc), V, VThresh, AMin, AShift, VWidth, A)
        lib().or_apply_amax(C.byref(loc), AMax, A)
    return A


def hebbian_dw(pre, post, nxp, nyp, nfp, dw_max, pre_act, post_act, counts=True):
    nkx, nky, nkf = num_kernels(pre, post)
    n = nkx * nky * nkf * nxp * nyp * nfp
    dw = np.zeros(n, np.float32)
    cnt = np.zeros(n, np.int64) if counts else None
    lib().or_hebbian_update_dw(C.byref(pre), C.byref(post), nxp, nyp, nfp, dw_max,
                               np.ascontiguousarray(pre_act, np.float32).ravel(),
                               np.ascontiguousarray(post_act, np.float32).ravel(), dw, _ptr(cnt))
    return dw, cnt


def cost_sum(loc, buf, extended, batch, kind, p0=0.0, p1=0.0):
    return lib().or_cost_sum(C.byref(loc), np.ascontiguousarray(buf, np.float32).ravel(),
                             int(extended), batch, kind, p0, p1)


def lca_run(cfg, weights, inp, V0, nsteps):
    nb = cfg.input.nbatch
    V = np.zeros(nb * cfg.v1.num_res, np.float32)
    A = np.zeros(nb * cfg.v1.num_ext, np.float32)
    rec = np.zeros(nb * cfg.res.num_res, np.float32)
    en = np.zeros((nsteps, nb), np.float64)
    na = np.zeros((nsteps, nb), np.int64)
    v0 = None if V0 is None else np.ascontiguousarray(V0, np.float32).ravel()
    lib().or_lca_run(C.byref(cfg), np.ascontiguousarray(weights, np.float32).ravel(),
                     np.ascontiguousarray(inp, np.float32).ravel(), _ptr(v0), nsteps,
                     _ptr(V), _ptr(A), _ptr(rec), _ptr(en), _ptr(na))
    return dict(V=V, A=A, recon=rec, energy=en, num_active=na)


class AdaptiveTimeScales:
    """The oracle's AdaptiveTimeScaleController (or_timescale_*): calc(values)
    returns the per-batch time scales for the coming step."""

    def __init__(self, nbatch, base_max, base_min, tau_factor, growth_factor):
        self.nb = nbatch
        self.p = (float(base_max), float(base_min), float(tau_factor), float(growth_factor))
        self.state = np.zeros(3 * nbatch, np.float64)
        lib().or_timescale_init(self.state.ctypes.data, nbatch, self.p[0], self.p[1])

    def calc(self, values):
        out = np.zeros(self.nb, np.float64)
        lib().or_timescale_calc(self.state.ctypes.data, self.nb,
                                np.ascontiguousarray(values, np.float64), *self.p, out)
        return out


def need_update(sim_time, dt, last_update_time, period):
    return bool(lib().or_need_update(sim_time, dt, last_update_time, period))
