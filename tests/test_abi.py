"""The C-ABI libraries load and export every declared symbol; host-side tables
agree with the oracle's geometry.  CPU only (no compute calls on a device)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle as O
from openpv_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FMAX = float(np.finfo(np.float32).max)


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pv(?:hip|host|params)_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("header,loader,table", [
    ("pvhip.h", abi.hip, abi.PVHIP_SYMBOLS), ("pvhost.h", abi.host, abi.PVHOST_SYMBOLS),
    ("pvparams.h", abi.host, abi.PVPARAMS_SYMBOLS)])
def test_library_exports_every_declared_symbol(header, loader, table):
    lib = loader()
    names = declared(header)
    assert names, header
    for n in names:
        assert hasattr(lib, n), "%s declares %s but the library does not export it" % (header, n)
        assert n in table, "%s not bound in openpv_amd.abi" % n


def test_abi_version_and_error_string():
    assert abi.hip().pvhip_abi_version() == 8
    assert isinstance(abi.hip().pvhip_last_error(), bytes)


def test_invalid_arguments_are_rejected_without_a_device():
    H = abi.hip()
    assert H.pvhip_recv_post(None, None) == 1
    assert b"NULL" in H.pvhip_last_error()
    args = abi.RecvPreArgs()
    assert H.pvhip_recv_pre(C.byref(args), None) == 1  # invalid loc
    bad = abi.Loc.make(0, 4, 4, 1)
    assert H.pvhip_active_compact_scratch_bytes(C.byref(bad)) == 0


@pytest.mark.parametrize("arith", [abi.ARITH_EXACT, abi.ARITH_FP32, abi.ARITH_BF16, abi.ARITH_FP8])
@pytest.mark.parametrize("case", [
    # source (nb, nx, ny, nf), target, post-perspective patch, required margin
    ((1, 128, 128, 1), (1, 32, 32, 512), 16, 6),   # the LCA gather: many to one, ratio 4
    ((1, 256, 256, 16), (1, 128, 128, 256), 12, 5),
    ((1, 32, 32, 3), (1, 32, 32, 8), 5, 2),         # one to one
    ((1, 16, 16, 64), (1, 64, 64, 2), 4, 2),        # one to many (the transposed scatter)
])
def test_recv_post_refuses_a_source_halo_below_the_margin(case, arith):
    """pvhip_recv_post checks on the host that every side of the source halo
    covers the patch's reach (the margin the connection requires of its pre
    layer) and returns PVHIP_ERR_INVALID before any launch when it does not;
    the margin itself is accepted up to the launch (which needs a device).
    The margins are pvhost_required_margin's for the connection's own
    pre-perspective patch, and pvhost_recv_post_tables refuses the same
    halos."""
    src, tgt, P, m = case
    H, Ho = abi.hip(), abi.host()
    fake = C.c_void_p(4096)   # never dereferenced: the call must stop before a launch

    def call(halo):
        a = abi.RecvPostArgs()
        a.pre = abi.Loc.make(*src, halo=halo)
        a.post = abi.Loc.make(*tgt, halo=0)
        a.nxp = a.nyp = P
        a.nfp = src[3]
        a.num_kernels_x = a.num_kernels_y = 1
        for f in ("start_x", "start_y", "kernel_x", "kernel_y", "pre_data", "weights", "post_gsyn"):
            setattr(a, f, fake)
        a.arith = arith
        return H.pvhip_recv_post(C.byref(a), None), a

    for halo in ((m - 1,) * 4, (m - 1, m, m, m), (m, m - 1, m, m), (m, m, m - 1, m), (m, m, m, m - 1)):
        st, a = call(halo)
        assert st == 1, halo
        assert b"halo" in H.pvhip_last_error()
        n = np.zeros(max(tgt[1], tgt[2]), np.int32)
        ptr = lambda x: x.ctypes.data_as(C.c_void_p)
        nk = C.c_int()
        assert Ho.pvhost_recv_post_tables(C.byref(a.pre), C.byref(a.post), P, P, ptr(n), ptr(n), ptr(n),
                                          ptr(n), C.byref(nk), C.byref(nk)) < 0, halo
    # the required margin passes the halo check: a BF16 / FP8 call without its
    # weight shadow (or, for several data patches per feature, without inverse
    # tables) then stops at that check, still before any launch
    if arith in (abi.ARITH_BF16, abi.ARITH_FP8):
        st, _ = call((m,) * 4)
        msg = H.pvhip_last_error()
        assert st == 1 and b"halo" not in msg and (b"weights_lp" in msg or b"inverse" in msg
                                                   or b"FP8" in msg), msg


def test_check_faults_is_exported_and_needs_a_device():
    """pvhip_check_faults reports refused tiles after a device synchronisation;
    without a device the synchronisation itself fails (a runtime error, not a
    silent success)."""
    assert abi.hip().pvhip_check_faults() in (0, 2)


def test_channel_gemm_scratch_query_follows_the_launch_shape():
    """pvhip_recv_post_scratch_bytes (host logic only): the stride-4 transposed
    gather (32x32x512 -> 128x128x1) needs splits x GSyn floats of scratch:
    FP32 by default splits into ranges of >= 8 of its 32 chunks (4), BF16 does
    not split by default, an explicit cgemm_splits is honoured (clamped to the
    chunk count), EXACT has no channel GEMM, and a bad tile is refused."""
    H = abi.hip()
    stub = np.zeros(64, np.int32)          # table pointers: never read by the query
    P = stub.ctypes.data_as(C.c_void_p)
    a = abi.RecvPostArgs()
    a.pre, a.post = abi.Loc.make(1, 32, 32, 512, 2), abi.Loc.make(1, 128, 128, 1, 0)
    a.nxp, a.nyp, a.nfp = 4, 4, 512
    a.num_kernels_x, a.num_kernels_y = 4, 4
    a.start_x = a.start_y = a.kernel_x = a.kernel_y = P
    a.inverse_x = a.inverse_y = P
    a.window_x0, a.window_nx, a.window_y0, a.window_ny = 0, 33, 0, 33
    a.dt_factor = 1.0
    gsyn = 128 * 128 * 4
    q = lambda: H.pvhip_recv_post_scratch_bytes(C.byref(a))
    a.arith = abi.ARITH_FP32
    assert q() == 4 * gsyn
    a.cgemm_splits = 2
    assert q() == 2 * gsyn
    a.cgemm_splits = 1000                  # clamped to 512 / 16 chunks
    assert q() == 32 * gsyn
    a.cgemm_splits, a.cgemm_tile = 0, 256
    assert q() == 4 * gsyn
    a.arith, a.cgemm_tile = abi.ARITH_BF16, 0
    assert q() == 0
    a.cgemm_splits = 8
    assert q() == 8 * gsyn
    a.arith = abi.ARITH_EXACT
    assert q() == 0
    a.arith, a.cgemm_tile = abi.ARITH_FP32, 100
    assert q() == 0


CASES = [
    ((1, 16, 16, 128, 2), (1, 32, 32, 3, 3), 8, 8),
    ((1, 32, 32, 512, 2), (1, 128, 128, 1, 6), 16, 16),
    ((1, 256, 256, 16, 6), (1, 128, 128, 256, 0), 6, 6),
    ((1, 16, 16, 2, (3, 2, 1, 4)), (1, 8, 8, 5, (1, 2, 0, 3)), 6, 4),
    ((1, 8, 8, 2, 3), (1, 32, 32, 4, 1), 8, 12),
    ((1, 32, 32, 2, 9), (1, 8, 8, 4, 0), 5, 5),
]


def both(t):
    return abi.Loc.make(*t[:4], halo=t[4]), O.Loc.make(*t[:4], halo=t[4])


@pytest.mark.parametrize("case", CASES)
def test_host_patch_axis_matches_oracle(case):
    (hp, op), (hq, oq), nxp, nyp = both(case[0]), both(case[1]), case[2], case[3]
    for axis in (0, 1):
        n = hp.nx_ext if axis == 0 else hp.ny_ext
        outs = [np.zeros(n, np.int32) for _ in range(5)]
        assert abi.host().pvhost_patch_axis(C.byref(hp), C.byref(hq), nxp, nyp, axis,
                                            *[o.ctypes.data_as(C.c_void_p) for o in outs]) == 0
        ref = O.patch_axis(op, oq, nxp, nyp, axis)
        for o, k in zip(outs, ("dim", "pstart", "post_start_res", "post_start_ext", "unshrunken")):
            assert (o == ref[k]).all(), (axis, k)


@pytest.mark.parametrize("case", CASES)
def test_recv_pre_coverage_tables_reproduce_the_scatter_order(case):
    """Walking the CSR tables visits exactly the (pre neuron, patch row) pairs of the
    reference loop, in its order, for every post neuron."""
    (hp, op), (hq, oq), nxp, nyp = both(case[0]), both(case[1]), case[2], case[3]
    H = abi.host()
    nc, nr = C.c_int(), C.c_int()
    assert H.pvhost_recv_pre_tables(C.byref(hp), C.byref(hq), nxp, nyp, None, None, None, None,
                                    C.byref(nc), None, None, None, None, C.byref(nr)) == 0
    cp = np.zeros(hq.nx + 1, np.int32)
    cx, cw, ck = (np.zeros(max(1, nc.value), np.int32) for _ in range(3))
    rp = np.zeros(hq.ny * nyp + 1, np.int32)
    ry, rw, rk = (np.zeros(max(1, nr.value), np.int32) for _ in range(3))
    P = lambda a: a.ctypes.data_as(C.c_void_p)
    assert H.pvhost_recv_pre_tables(C.byref(hp), C.byref(hq), nxp, nyp, P(cp), P(cx), P(cw), P(ck),
                                    C.byref(nc), P(rp), P(ry), P(rw), P(rk), C.byref(nr)) == 0
    pa, gs, _, _ = O.patch_geometry(op, oq, nxp, nyp, oq.nf)
    nxe, nye = op.nx_ext, op.ny_ext
    # reference visit list per post (X, Y) with nf = 1 view: (y, ye, xe, weight index)
    ref = {}
    for y in range(nyp):
        for ye in range(nye):
            for xe in range(nxe):
                kex = (ye * nxe + xe) * op.nf
                pt = pa[kex]
                if y >= pt["ny"]:
                    continue
                start = int(gs[kex]) // oq.nf
                Y0, X0 = divmod(start, oq.nx)
                oy, ox = divmod(int(pt["offset"]) // oq.nf, nxp)
                for kx in range(pt["nx"]):
                    ref.setdefault((Y0 + y, X0 + kx), []).append((y, ye, xe, oy + y, ox + kx))
    got = {}
    for Y in range(hq.ny):
        for X in range(hq.nx):
            seq = []
            for r in range(nyp):
                for ri in range(rp[Y * nyp + r], rp[Y * nyp + r + 1]):
                    for ci in range(cp[X], cp[X + 1]):
                        seq.append((r, int(ry[ri]), int(cx[ci]), int(rw[ri]), int(cw[ci])))
            if seq:
                got[(Y, X)] = seq
    assert got == ref


@pytest.mark.parametrize("case", CASES[:3])
def test_recv_post_tables_match_unshrunken_start(case):
    (hs, os_), (ht, ot), _, _ = both(case[1]), both(case[0]), case[2], case[3]
    # gather from source=case[1] into target=case[0] with the original pre patch
    nxp, nyp = case[2], case[3]
    sx, sy, kx, ky = (np.zeros(n, np.int32) for n in (ht.nx, ht.ny, ht.nx, ht.ny))
    nkx, nky = C.c_int(), C.c_int()
    P = lambda a: a.ctypes.data_as(C.c_void_p)
    r = abi.host().pvhost_recv_post_tables(C.byref(hs), C.byref(ht), nxp, nyp, P(sx), P(sy), P(kx),
                                           P(ky), C.byref(nkx), C.byref(nky))
    if r != 0:
        pytest.skip("halo of source too small for this patch")
    us_x = O.patch_axis(ot, os_, nxp, nyp, 0)["unshrunken"][ot.lt:ot.lt + ot.nx]
    us_y = O.patch_axis(ot, os_, nxp, nyp, 1)["unshrunken"][ot.up:ot.up + ot.ny]
    assert (sx == us_x).all() and (sy == us_y).all()


@pytest.mark.parametrize("params", [
    (0.1, 0.0, FMAX, 0.1, 0.0), (0.025, 0.0, FMAX, 0.0, 0.025), (0.1, 0.0, 0.5, 0.1, 0.0),
    (-FMAX, -FMAX, FMAX, 0.0, 0.0), (0.2, -0.1, 1.0, 0.05, 0.3), (0.3, 0.0, 0.1, 0.0, 0.0)])
def test_make_transfer_matches_oracle_vertices(params):
    t = abi.Transfer()
    assert abi.host().pvhost_make_transfer(*params, abi.TRANSFER_VERTICES, C.byref(t)) == 0
    nv, vv, va, sl = O.set_vertices(*params)
    assert t.num_vertices == nv
    assert np.array_equal(np.array(t.vertices_v[:nv], np.float32), vv[:nv])
    assert np.array_equal(np.array(t.vertices_a[:nv], np.float32), va[:nv])
    assert np.array_equal(np.array(t.slopes[:nv + 1], np.float32), sl[:nv + 1], equal_nan=True)


@pytest.mark.parametrize("case", [
    ((1, 32, 32, 6, 5), (1, 16, 16, 10, 2), 6, 6),   # many-to-one
    ((1, 16, 16, 12, 2), (1, 32, 32, 3, 3), 8, 8),   # one-to-many
    ((1, 16, 16, 4, 2), (1, 16, 16, 5, 2), 5, 5),    # one-to-one
    ((1, 16, 16, 4, 4), (1, 64, 64, 2, 0), 8, 8),    # one-to-many stride 4
])
def test_transpose_weights_matches_oracle(case):
    (hp, op), (hq, oq), nxp, nyp = both(case[0]), both(case[1]), case[2], case[3]
    kx, ky, kf = O.num_kernels(op, oq)
    rng = np.random.default_rng(3)
    W = rng.standard_normal(kx * ky * kf * nxp * nyp * oq.nf).astype(np.float32)
    nxq = O.lib().or_calcPostPatchSize(nxp, op.nx, oq.nx)
    nyq = O.lib().or_calcPostPatchSize(nyp, op.ny, oq.ny)
    pkx, pky, pkf = O.num_kernels(oq, op)
    n = pkx * pky * pkf * nxq * nyq * op.nf
    a = np.full(n, np.nan, np.float32)
    b = np.full(n, np.nan, np.float32)
    O.lib().or_transpose_weights_shared(C.byref(op), C.byref(oq), nxp, nyp, oq.nf, W, a)
    assert abi.host().pvhost_transpose_weights(C.byref(hp), C.byref(hq), nxp, nyp, oq.nf,
                                               W.ctypes.data_as(C.c_void_p),
                                               b.ctypes.data_as(C.c_void_p)) == 0
    assert np.array_equal(a, b, equal_nan=True)
    # the permutation table of the device refresh gives the same weights
    # (entries no pre weight reaches are 0 there, as in a zero-initialised copy)
    idx = np.zeros(n, np.int32)
    assert abi.host().pvhost_transpose_weights_index(C.byref(hp), C.byref(hq), nxp, nyp, oq.nf,
                                                     idx.ctypes.data_as(C.c_void_p), n) == 0
    perm = np.where(idx >= 0, W[np.maximum(idx, 0)], np.float32(0.0))
    assert np.array_equal(perm, np.where(np.isnan(a), np.float32(0.0), a))
    hit = idx[idx >= 0]
    assert len(np.unique(hit)) == len(hit)      # a permutation: no pre weight used twice


@pytest.mark.parametrize("case", [
    ((1, 16, 16, 128, 2), (1, 32, 32, 3, 3), 8),
    ((1, 32, 32, 512, 2), (1, 128, 128, 1, 6), 16),
    ((1, 128, 128, 16, 3), (1, 512, 512, 1, 6), 16),
    ((1, 8, 8, 5, 1), (1, 8, 8, 2, 1), 3),
])
def test_hebbian_windows_match_patch_geometry(case):
    """pvhost_hebbian_windows: affine unshrunken windows, and the oracle's
    shrunken patches are those windows clipped to the post restricted region."""
    pre, post, p = abi.Loc.make(*case[0][:4], halo=case[0][4]), abi.Loc.make(*case[1][:4], halo=case[1][4]), case[2]
    w = [C.c_int() for _ in range(4)]
    assert abi.host().pvhost_hebbian_windows(C.byref(pre), C.byref(post), p, p, *[C.byref(v) for v in w]) == 0
    sx, cx, sy, cy = (v.value for v in w)
    opre = O.Loc.make(*case[0][:4], halo=case[0][4])
    opost = O.Loc.make(*case[1][:4], halo=case[1][4])
    for axis, (s, c, lo_res, n_res) in enumerate([(sx, cx, post.halo.lt, post.nx), (sy, cy, post.halo.up, post.ny)]):
        t = O.patch_axis(opre, opost, p, p, axis)
        for e in range(len(t["dim"])):
            u = c + s * e
            lo, hi = max(u, lo_res), min(u + p, lo_res + n_res)
            assert t["dim"][e] == max(0, hi - lo)
            if hi > lo:
                assert t["post_start_ext"][e] == lo and t["pstart"][e] == lo - u


def test_hebbian_windows_refuse_unclipped_geometry():
    """Asymmetric halos: the reference's shrunken patches are not the clipped
    windows (its per-axis border quirks), so only the EXACT Hebbian form applies."""
    pre = abi.Loc.make(1, 12, 10, 16, (2, 1, 3, 2))
    post = abi.Loc.make(1, 24, 20, 5, (3, 4, 4, 3))
    w = [C.c_int() for _ in range(4)]
    assert abi.host().pvhost_hebbian_windows(C.byref(pre), C.byref(post), 6, 6, *[C.byref(v) for v in w]) < 0


def test_momentum_factors_host_only():
    """pvhip_momentum_factors runs on the host: the oracle's factors (the
    reference's applyMomentum) and the reference's parameter checks."""
    import oracle as O
    from openpv_amd import runtime as R
    for method, code, tau, decay, dwmax in [("viscosity", 0, 100.0, 0.0, 1.0), ("viscosity", 0, 0.0, 0.1, 1.0),
                                            ("simple", 1, 0.75, 0.01, 1.0), ("alex", 2, 0.9, 0.5, 0.25)]:
        assert R.momentum_factors(method, tau, decay, dwmax) == O.momentum_factors(code, tau, decay, dwmax)
    for bad in [("simple", 1.0, 0.0), ("alex", -0.5, 0.0), ("viscosity", -1.0, 0.0),
                ("viscosity", 100.0, 2.0), (7, 0.5, 0.0)]:
        with pytest.raises(abi.PVHIPError):
            R.momentum_factors(*bad)


def test_momentum_factors_of_the_deprecated_momentum_tau():
    """momentumTau (deprecated, used when timeConstantTau is absent) takes the
    factors of applyMomentumDeprecated without timeConstantTau's range check:
    viscosity with a negative momentumTau, simple and alex with momentumTau >= 1
    are accepted, as the reference accepts them; the closed forms
    exp(-1/momentumTau), momentumTau and decay*dWMax hold; momentumDecay is still
    checked.  The runtime's MomentumUpdater builds from such values."""
    import oracle as O
    from openpv_amd import runtime as R
    for method, code, tau, decay, dwmax in [("viscosity", 0, -2.0, 0.0, 1.0), ("viscosity", 0, 4.0, 0.25, 1.0),
                                            ("simple", 1, 1.5, 0.0, 1.0), ("alex", 2, 2.0, 0.5, 0.25)]:
        got = R.momentum_factors(method, tau, decay, dwmax, deprecated_tau=True)
        assert got == O.momentum_factors(code, tau, decay, dwmax, deprecated_tau=True)
        want = np.float32(np.exp(np.float32(-1.0) / np.float32(tau))) if code == 0 else np.float32(tau)
        assert got[0] == pytest.approx(float(want), rel=1e-6)
        assert got[1] == pytest.approx(decay * (dwmax if code == 2 else 1.0))
    with pytest.raises(abi.PVHIPError):
        R.momentum_factors("simple", 1.5, 1.5, 1.0, deprecated_tau=True)
    with pytest.raises(abi.PVHIPError):     # the same value as timeConstantTau is rejected
        R.momentum_factors("simple", 1.5, 0.0, 1.0)


@pytest.mark.parametrize("case", [((2, 32, 32, 512, 2), (2, 128, 128, 1, 0), 16),
                                  ((9, 6, 5, 128, 3), (9, 24, 20, 1, 0), 16),
                                  ((2, 16, 16, 256, (2, 1, 3, 2)), (2, 64, 64, 4, 0), 8),
                                  ((1, 10, 10, 32, 1), (1, 20, 20, 1, 0), 6)])
def test_recv_pre_dense_start_matches_patch_geometry(case):
    """pvhost_recv_pre_dense_start: every source neuron with a kept patch has
    its unshrunken patch start (the oracle's PatchGeometry: kept start minus
    kept offset) at stride*extended index + start, along x and y."""
    (hp, op), (hq, oq), p = both(case[0]), both(case[1]), case[2]
    s, x0, y0 = C.c_int(), C.c_int(), C.c_int()
    assert abi.host().pvhost_recv_pre_dense_start(C.byref(hp), C.byref(hq), p, p, C.byref(s), C.byref(x0),
                                                  C.byref(y0)) == 0
    assert s.value == hq.nx // hp.nx
    pa, gs, _, _ = O.patch_geometry(op, oq, p, p, oq.nf)
    seen = 0
    for ye in range(op.ny_ext):
        for xe in range(op.nx_ext):
            kex = (ye * op.nx_ext + xe) * op.nf
            pt = pa[kex]
            if pt["nx"] == 0 or pt["ny"] == 0:
                continue
            Y0, X0 = divmod(int(gs[kex]) // oq.nf, oq.nx)
            oy, ox = divmod(int(pt["offset"]) // oq.nf, p)
            assert X0 - ox == s.value * xe + x0.value and Y0 - oy == s.value * ye + y0.value
            seen += 1
    assert seen > 0


def test_recv_pre_dense_start_refuses_coarser_targets():
    s, x0, y0 = C.c_int(), C.c_int(), C.c_int()
    for pre, post in (((1, 64, 64, 32, 2), (1, 32, 32, 1, 0)), ((1, 16, 16, 32, 2), (1, 64, 32, 1, 0))):
        (hp, _), (hq, _) = both(pre), both(post)
        assert abi.host().pvhost_recv_pre_dense_start(C.byref(hp), C.byref(hq), 8, 8, C.byref(s), C.byref(x0),
                                                      C.byref(y0)) == -1


def _dense_fits(pre, post, nxp, halo_is_zero, fp32):
    x0, y0 = C.c_int(), C.c_int()
    assert abi.host().pvhost_recv_pre_dense_start(C.byref(pre), C.byref(post), nxp, nxp, None, C.byref(x0),
                                                  C.byref(y0)) == 0
    q = abi.RecvPreDenseArgs()
    q.pre, q.post = pre, post
    q.nxp = q.nyp = nxp
    q.nfp = post.nf
    q.start_x0, q.start_y0 = x0.value, y0.value
    q.halo_is_zero, q.fp32 = halo_is_zero, fp32
    return abi.hip().pvhip_recv_pre_dense_fits(C.byref(q))


def test_recv_pre_dense_fits():
    """pvhip_recv_pre_dense_fits (host only): the batched LCA's reconstruction
    fits both forms with or without its halo; a 512x512 target does not (its
    tile rows of 528 values exceed the LDS beside the weights); nor does a
    source whose feature count is not 128, 256 or 512."""
    pre, post = abi.Loc.make(64, 32, 32, 512, halo=2), abi.Loc.make(64, 128, 128, 1, halo=0)
    for halo_is_zero in (0, 1):
        for fp32 in (0, 1):
            assert _dense_fits(pre, post, 16, halo_is_zero, fp32) == 1
    pre, post = abi.Loc.make(8, 128, 128, 512, halo=2), abi.Loc.make(8, 512, 512, 1, halo=0)
    assert _dense_fits(pre, post, 16, 1, 0) == 0
    assert _dense_fits(pre, post, 16, 1, 1) == 0
    pre, post = abi.Loc.make(2, 16, 16, 64, halo=2), abi.Loc.make(2, 64, 64, 1, halo=0)
    assert _dense_fits(pre, post, 16, 1, 0) == 0
    assert abi.hip().pvhip_recv_pre_dense_fits(None) == 0
