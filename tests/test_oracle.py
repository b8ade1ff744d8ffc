"""Pins the CPU oracle (oracle/pvoracle.c) -- CPU only.

Each oracle function is checked against at least one of:
  * the reference's own code compiled from its sources (oracle/_ref: PatchGeometry,
    the .kpp kernels, the cost functions), bit for bit;
  * golden data the reference's tests hold (tests/PatchGeometryTest expectations);
  * a library routine / closed form chosen so that a dropped term, wrong sign,
    wrong index or transposed operand fails (float64 torch conv2d and
    conv_transpose2d, autograd weight gradients, exponential relaxation).
"""
import json
import os

import numpy as np
import pytest
import torch

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
FMAX = float(np.finfo(np.float32).max)

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built")

GEOM_CASES = [
    # (pre nb,nx,ny,nf,halo), (post ...), nxp, nyp, nfp
    ((1, 16, 16, 3, 2), (1, 16, 16, 10, 2), 5, 5, 10),
    ((1, 16, 16, 128, 2), (1, 32, 32, 3, 3), 8, 8, 3),       # C0 V1 -> Residual
    ((1, 32, 32, 3, 3), (1, 16, 16, 128, 2), 4, 4, 128),     # C0 transpose
    ((1, 32, 32, 512, 2), (1, 128, 128, 1, 6), 16, 16, 1),   # C2 V1 -> Residual
    ((1, 128, 128, 1, 6), (1, 32, 32, 512, 2), 4, 4, 512),   # C2 transpose
    ((1, 256, 256, 16, 6), (1, 128, 128, 256, 0), 6, 6, 256),  # C1 pre perspective
    ((1, 16, 16, 2, (3, 2, 1, 4)), (1, 8, 8, 5, (1, 2, 0, 3)), 6, 4, 5),  # asymmetric halos, even patch
    ((1, 8, 8, 2, 3), (1, 32, 32, 4, 1), 8, 12, 4),          # one-to-many stride 4
    ((1, 32, 32, 2, 9), (1, 8, 8, 4, 0), 5, 5, 4),           # many-to-one stride 4
]


def L(t):
    return O.Loc.make(*t[:4], halo=t[4])


@needs_ref
@pytest.mark.parametrize("case", GEOM_CASES)
def test_patch_geometry_matches_reference(case):
    pre, post, nxp, nyp, nfp = L(case[0]), L(case[1]), *case[2:]
    pa, gs, ap, us = O.patch_geometry(pre, post, nxp, nyp, nfp)
    n = pre.num_ext
    nx_ = np.zeros(n, np.uint16)
    ny_ = np.zeros(n, np.uint16)
    of = np.zeros(n, np.uint32)
    g, a, u = (np.zeros(n, np.uint64) for _ in range(3))
    kx, ky, kf = O.num_kernels(pre, post)
    tii = np.zeros(kx * ky * kf * nxp * nyp * nfp, np.int32)
    m = O.ref().ref_patch_geometry(np.array(pre.as_list(), np.int32), np.array(post.as_list(), np.int32),
                                   nxp, nyp, nfp, nx_, ny_, of, g, a, u, tii.ctypes.data)
    assert m == n
    assert (pa["nx"] == nx_).all() and (pa["ny"] == ny_).all() and (pa["offset"] == of).all()
    assert (gs == g).all() and (ap == a).all() and (us == u).all()
    mine = np.array([O.lib().or_transpose_item_index(O.C.byref(pre), O.C.byref(post), nxp, nyp, nfp,
                                                     k, i)
                     for k in range(kx * ky * kf) for i in range(nxp * nyp * nfp)], np.int32)
    assert (mine == tii).all()


def test_patch_geometry_golden_cases():
    cases = json.load(open(os.path.join(HERE, "golden", "patch_geometry_cases.json")))["cases"]
    for c in cases:
        pre, post = L(c["pre"]), L(c["post"])
        for axis in (0, 1):
            t = O.patch_axis(pre, post, c["nxp"], c["nxp"], axis)
            assert t["dim"].tolist() == c["sizes"], c["name"]
            assert t["pstart"].tolist() == c["starts"], c["name"]
            assert t["post_start_res"].tolist() == c["post_res"], c["name"]
            assert t["post_start_ext"].tolist() == c["post_ext"], c["name"]


@needs_ref
def test_margins_and_index_math_match_reference():
    R = O.ref()
    for npre, npost, p in [(32, 16, 4), (16, 32, 8), (128, 32, 4), (32, 128, 16), (16, 16, 5),
                           (256, 128, 6), (64, 16, 3)]:
        assert O.lib().or_requiredConvolveMargin(npre, npost, p) == R.ref_required_margin(npre, npost, p)
    for k in range(0, 3 * 7 * 5, 7):
        assert (O.lib().or_kIndexExtended(k, 5, 7, 3, 1, 2, 3, 4)
                == R.ref_kindex_extended(k, 5, 7, 3, 1, 2, 3, 4))


def _near_threshold_values(rng, n, vt):
    v = rng.standard_normal(n).astype(np.float32) * 0.2
    v[::7] = np.float32(vt)
    v[1::11] = np.nextafter(np.float32(vt), np.float32(1))
    v[2::13] = np.nextafter(np.float32(vt), np.float32(-1))
    return v


@needs_ref
@pytest.mark.parametrize("self_interact", [True, False])
def test_update_lca_matches_reference(self_interact):
    rng = np.random.default_rng(1)
    loc = O.Loc.make(3, 7, 5, 4, (1, 2, 2, 1))
    n = loc.nbatch * loc.num_res
    G = rng.standard_normal(n).astype(np.float32)
    A = rng.standard_normal(loc.nbatch * loc.num_ext).astype(np.float32)
    V = rng.standard_normal(n).astype(np.float32)
    dt = np.array([1.0, 0.5, 0.25])
    mine = O.update_lca(loc, self_interact, dt, 100.0, G, A, V)
    ref = V.copy()
    O.ref().ref_update_lca(np.array(loc.as_list(), np.int32), int(self_interact), dt, 100.0, G, A, ref)
    assert np.array_equal(mine.view(np.uint32), ref.view(np.uint32))


@needs_ref
@pytest.mark.parametrize("params", [
    (0.1, 0.0, FMAX, 0.1, 0.0),       # soft threshold (LCA)
    (0.025, 0.0, FMAX, 0.0, 0.025),   # firm threshold (GPULCATest)
    (0.1, 0.0, 0.5, 0.1, 0.0),        # with AMax
    (-FMAX, -FMAX, FMAX, 0.0, 0.0),   # identity
    (0.2, -0.1, 1.0, 0.05, 0.3),
])
def test_vertices_match_reference(params):
    rng = np.random.default_rng(2)
    loc = O.Loc.make(2, 9, 4, 3, 1)
    V = _near_threshold_values(rng, loc.nbatch * loc.num_res, params[0] if params[0] > -1e30 else 0.0)
    nv, vv, va, sl = O.set_vertices(*params)
    mine = O.transfer(loc, V, *params, vertices=True)
    ref = np.zeros_like(mine)
    O.ref().ref_apply_vertices(np.array(loc.as_list(), np.int32), nv, vv, va, sl, V, ref)
    assert np.array_equal(mine.view(np.uint32), ref.view(np.uint32))


def test_vthresh_closed_form():
    rng = np.random.default_rng(3)
    loc = O.Loc.make(1, 10, 10, 4, 2)
    V = _near_threshold_values(rng, loc.num_res, 0.1)
    A = O.transfer(loc, V, 0.1, 0.0, FMAX, 0.1, 0.0).reshape(14, 14, 4)[2:12, 2:12].ravel()
    expect = np.where(V >= np.float32(0.1), V - np.float32(0.1), np.float32(0)).astype(np.float32)
    assert np.array_equal(A, expect)
    # firm threshold: linear ramp from (VT, AMin) to (VT+VW, VT+VW-AShift), then V - AShift
    A = O.transfer(loc, V, 0.1, 0.0, FMAX, 0.0, 0.2).reshape(14, 14, 4)[2:12, 2:12].ravel()
    v = V.astype(np.float64)
    ramp = (0.3 * (v - 0.1) / 0.2)
    expect = np.where(v < 0.1, 0.0, np.where(v < 0.3, ramp, v))
    assert np.allclose(A, expect, rtol=1e-6, atol=1e-7)
    # soft threshold: CPU formula and GPU vertices are bit-identical (VWidth = 0)
    Acpu = O.transfer(loc, V, 0.1, 0.0, FMAX, 0.1, 0.0)
    Agpu = O.transfer(loc, V, 0.1, 0.0, FMAX, 0.1, 0.0, vertices=True)
    assert np.array_equal(Acpu.view(np.uint32), Agpu.view(np.uint32))


@needs_ref
def test_accumulate_and_costs_match_reference():
    rng = np.random.default_rng(4)
    g = rng.standard_normal(3 * 50).astype(np.float32)
    co = np.array([1.0, -1.0, 0.5], np.float32)
    out = np.zeros(50, np.float32)
    ref = np.zeros(50, np.float32)
    O.lib().or_gsyn_accumulate(50, 3, co, g, out)
    O.ref().ref_gsyn_accumulate(50, 3, co, g, ref)
    assert np.array_equal(out, ref)
    loc = O.Loc.make(2, 6, 5, 3, 1)
    buf = rng.standard_normal(loc.nbatch * loc.num_ext).astype(np.float32)
    for kind, p0, p1 in [(0, 0, 0), (1, 0, 0), (2, 0.025, 0.025), (3, 0.1, 0)]:
        for b in range(2):
            mine = O.cost_sum(loc, buf, True, b, kind, p0, p1)
            x = buf.reshape(2, 7, 8, 3)[b, 1:6, 1:7].ravel().astype(np.float64)
            ref = sum(O.ref().ref_cost(kind, float(v), p0, p1) for v in x)
            assert mine == pytest.approx(ref, rel=1e-12, abs=1e-12)


@pytest.mark.parametrize("case", [
    # source (nb,nx,ny,nf,halo), target (...), post-perspective patch
    ((2, 32, 32, 3, 3), (2, 16, 16, 128, 2), 8),      # C0 shape (Residual -> V1)
    ((1, 64, 64, 1, 6), (1, 16, 16, 32, 2), 16),      # C2-like, stride 4
    ((1, 16, 16, 4, 2), (1, 16, 16, 8, 0), 5),        # one-to-one
    ((1, 32, 32, 16, 6), (1, 16, 16, 24, 0), 12),     # C1-like many-to-one
])
def test_recv_post_equals_conv2d(case):
    S, T, p = L(case[0]), L(case[1]), case[2]
    rng = np.random.default_rng(5)
    A = np.zeros((S.nbatch, S.ny_ext, S.nx_ext, S.nf), np.float32)
    A[:, S.up:S.up + S.ny, S.lt:S.lt + S.nx] = rng.standard_normal((S.nbatch, S.ny, S.nx, S.nf))
    W = rng.standard_normal((T.nf, p, p, S.nf)).astype(np.float32)
    g = O.recv_post(S, T, p, p, A, W).reshape(T.nbatch, T.ny, T.nx, T.nf)
    # unshrunken start of target x in source extended coordinates
    ux = O.patch_axis(T, S, p, p, 0)["unshrunken"][T.lt:T.lt + T.nx]
    uy = O.patch_axis(T, S, p, p, 1)["unshrunken"][T.up:T.up + T.ny]
    stride = ux[1] - ux[0]
    x = torch.tensor(A, dtype=torch.float64).permute(0, 3, 1, 2)
    x = x[:, :, uy[0]:, ux[0]:]
    w = torch.tensor(W, dtype=torch.float64).permute(0, 3, 1, 2)
    y = torch.nn.functional.conv2d(x, w, stride=int(stride))[:, :, :T.ny, :T.nx].permute(0, 2, 3, 1)
    y = y.numpy()
    assert np.abs(g - y).max() <= 1e-5 * np.abs(y).max()
    # mistake detector: a transposed weight patch must not agree
    g2 = O.recv_post(S, T, p, p, A, np.ascontiguousarray(W.transpose(0, 2, 1, 3)))
    assert np.abs(g2.reshape(y.shape) - y).max() > 1e-3 * np.abs(y).max()


@pytest.mark.parametrize("case", [
    ((2, 16, 16, 128, 2), (2, 32, 32, 3, 0), 8),      # C0 V1 -> Recon
    ((1, 16, 16, 32, 2), (1, 64, 64, 1, 0), 16),      # C2-like
    ((1, 32, 32, 16, 6), (1, 16, 16, 24, 0), 6),      # C1-like many-to-one
    ((1, 12, 12, 3, 2), (1, 12, 12, 4, 0), 5),        # one-to-one
])
def test_recv_pre_equals_conv_transpose_and_post(case):
    pre, post, p = L(case[0]), L(case[1]), case[2]
    rng = np.random.default_rng(6)
    A = np.zeros((pre.nbatch, pre.ny_ext, pre.nx_ext, pre.nf), np.float32)
    act = rng.standard_normal((pre.nbatch, pre.ny, pre.nx, pre.nf)).astype(np.float32)
    act[rng.random(act.shape) < 0.9] = 0
    A[:, pre.up:pre.up + pre.ny, pre.lt:pre.lt + pre.nx] = act
    kx, ky, kf = O.num_kernels(pre, post)
    W = rng.standard_normal((kx * ky * pre.nf, p, p, post.nf)).astype(np.float32)
    ent, cnt = O.active_indices(pre, A)
    g_sparse = O.recv_pre(pre, post, p, p, W, entries=ent, num_active=cnt)
    g_dense = O.recv_pre(pre, post, p, p, W, pre_activity=A)
    assert np.array_equal(g_sparse, g_dense)  # same order, same bits
    # multithreaded form (per-thread GSyn, guided schedule: the split varies from
    # run to run) agrees to rounding relative to the largest sum
    g_mt = O.recv_pre(pre, post, p, p, W, entries=ent, num_active=cnt, num_threads=4)
    assert np.abs(g_mt - g_sparse).max() <= 1e-5 * np.abs(g_sparse).max()
    # equals float64 brute force over the patch tables
    pa, gs, _, _ = O.patch_geometry(pre, post, p, p, post.nf)
    ref = np.zeros((pre.nbatch, post.num_res))
    W2 = W.reshape(-1, p * p * post.nf).astype(np.float64)
    for b in range(pre.nbatch):
        for i in range(cnt[b]):
            kex, a = int(ent[b * pre.num_ext + i]["index"]), float(ent[b * pre.num_ext + i]["value"])
            d = O.lib().or_data_index(O.C.byref(pre), kx, ky, kex)
            pt = pa[kex]
            nk = int(pt["nx"]) * post.nf
            for yy in range(pt["ny"]):
                o = int(gs[kex]) + yy * post.nx * post.nf
                w0 = int(pt["offset"]) + yy * p * post.nf
                ref[b, o:o + nk] += a * W2[d, w0:w0 + nk]
    assert np.abs(g_sparse - ref.ravel()).max() <= 1e-5 * max(1.0, np.abs(ref).max())


def test_post_perspective_equals_pre_perspective():
    """ReceiveFromPostTest invariant: gather with transposed weights == scatter."""
    rng = np.random.default_rng(8)
    nxp = 6                                 # many-to-one, pre-perspective patch
    pre = O.Loc.make(1, 32, 32, 6, O.lib().or_requiredConvolveMargin(32, 16, nxp))  # 5
    post = O.Loc.make(1, 16, 16, 10, 2)
    pre_act = np.zeros((1, pre.ny_ext, pre.nx_ext, pre.nf), np.float32)
    pre_act[:, 5:37, 5:37] = rng.standard_normal((1, 32, 32, 6))
    W = rng.standard_normal((2 * 2 * 6, nxp, nxp, post.nf)).astype(np.float32)
    g_pre = O.recv_pre(pre, post, nxp, nxp, W, pre_activity=pre_act)
    nxpPost = O.lib().or_calcPostPatchSize(nxp, pre.nx, post.nx)
    Wpost = np.zeros((post.nf, nxpPost, nxpPost, pre.nf), np.float32)
    O.lib().or_transpose_weights_shared(O.C.byref(pre), O.C.byref(post), nxp, nxp, post.nf,
                                        W.ravel(), Wpost.ravel())
    g_post = O.recv_post(pre, post, nxpPost, nxpPost, pre_act, Wpost)
    assert np.abs(g_post - g_pre).max() <= 3e-3  # the reference test's tolerance
    assert np.abs(g_post - g_pre).max() <= 1e-5 * np.abs(g_pre).max()


def test_single_impulse_scatter_is_patch_copy():
    pre = O.Loc.make(1, 8, 8, 4, 2)
    post = O.Loc.make(1, 16, 16, 3, 0)
    W = np.random.default_rng(9).standard_normal((4, 8, 8, 3)).astype(np.float32)
    A = np.zeros((1, 12, 12, 4), np.float32)
    A[0, 2 + 0, 2 + 5, 2] = 1.5   # border neuron: shrunken patch
    g = O.recv_pre(pre, post, 8, 8, W, pre_activity=A).reshape(16, 16, 3)
    start = 2 * 5 - 3
    exp = np.zeros((16 + 16, 16 + 16, 3), np.float32)
    exp[8 - 3:8 + 5, 8 + start:8 + start + 8] = np.float32(1.5) * W[2]
    assert np.array_equal(g, exp[8:24, 8:24])


def test_active_indices_definition():
    rng = np.random.default_rng(10)
    loc = O.Loc.make(3, 9, 7, 5, 2)
    A = rng.standard_normal(loc.nbatch * loc.num_ext).astype(np.float32)
    A[rng.random(A.shape) < 0.8] = 0
    A[5] = -0.0  # negative zero is not active
    ent, cnt = O.active_indices(loc, A)
    for b in range(3):
        Ab = A[b * loc.num_ext:(b + 1) * loc.num_ext]
        nz = np.nonzero(Ab != 0)[0]
        assert cnt[b] == len(nz)
        e = ent[b * loc.num_ext:b * loc.num_ext + cnt[b]]
        assert (e["index"] == nz).all() and np.array_equal(e["value"], Ab[nz])


def test_lca_relaxation_closed_form():
    """Constant input G, no self interaction: V(t) -> G with V(t) = G + (V0-G) e^{-t dt/tau}."""
    loc = O.Loc.make(1, 4, 4, 2, 0)
    G = np.full(32, 0.7, np.float32)
    V = np.full(32, -0.2, np.float32)
    A = np.zeros(32, np.float32)
    for t in range(1, 201):
        V = O.update_lca(loc, False, np.array([1.0]), 10.0, G, A, V)
    expect = 0.7 + (-0.2 - 0.7) * np.exp(-200 / 10.0)
    assert np.allclose(V, expect, rtol=1e-5)
    # fixed point with self interaction: V = G + A
    V2 = O.update_lca(loc, True, np.array([1.0]), 10.0, G, np.full(32, 0.3, np.float32),
                      np.full(32, 1.0, np.float32))
    assert np.allclose(V2, 1.0, rtol=1e-6)


@pytest.mark.parametrize("case", [
    ((2, 16, 16, 12, 2), (2, 32, 32, 3, 3), 8),      # one-to-many (V1 -> Residual of C0)
    ((1, 8, 8, 5, 1), (1, 8, 8, 2, 1), 3),           # one-to-one
])
def test_hebbian_equals_weight_gradient(case):
    """dW (normalizeDw off) = dWMax * sum_b a_pre (x) a_post = conv weight gradient."""
    pre, post, p = L(case[0]), L(case[1]), case[2]
    rng = np.random.default_rng(11)
    a = np.zeros((pre.nbatch, pre.ny_ext, pre.nx_ext, pre.nf), np.float32)
    av = rng.standard_normal((pre.nbatch, pre.ny, pre.nx, pre.nf)).astype(np.float32)
    av[rng.random(av.shape) < 0.7] = 0
    a[:, pre.up:pre.up + pre.ny, pre.lt:pre.lt + pre.nx] = av
    r = np.zeros((post.nbatch, post.ny_ext, post.nx_ext, post.nf), np.float32)
    r[:, post.up:post.up + post.ny, post.lt:post.lt + post.nx] = rng.standard_normal(
        (post.nbatch, post.ny, post.nx, post.nf))
    dw, cnt = O.hebbian_dw(pre, post, p, p, post.nf, 0.5, a, r)
    # float64 brute force through the geometry: d/dW of sum_post g * r with g = scatter(a, W)
    Wt = torch.zeros((pre.nf, p, p, post.nf), dtype=torch.float64, requires_grad=True)
    pa, _, apo, _ = O.patch_geometry(pre, post, p, p, post.nf)
    kx, ky, kf = O.num_kernels(pre, post)
    assert kx == ky == 1
    tot = torch.zeros((), dtype=torch.float64)
    rflat = torch.tensor(r.reshape(post.nbatch, -1), dtype=torch.float64)
    sya = post.nx_ext * post.nf
    for b in range(pre.nbatch):
        for kex in np.nonzero(a[b].ravel())[0]:
            pt = pa[kex]
            f = kex % pre.nf
            oy, ox = divmod(int(pt["offset"]) // post.nf, p)
            for yy in range(pt["ny"]):
                seg = rflat[b, int(apo[kex]) + yy * sya: int(apo[kex]) + yy * sya + pt["nx"] * post.nf]
                tot = tot + float(a[b].ravel()[kex]) * (
                    Wt[f, oy + yy, ox:ox + pt["nx"], :].reshape(-1) * seg).sum()
    tot.backward()
    ref = 0.5 * Wt.grad.numpy().ravel()
    assert np.abs(dw - ref).max() <= 1e-5 * max(1.0, np.abs(ref).max())
    assert cnt.sum() == sum(int(pa[k]["nx"]) * int(pa[k]["ny"]) * post.nf
                            for b in range(pre.nbatch) for k in np.nonzero(a[b].ravel())[0])


def test_normalize_l2():
    rng = np.random.default_rng(12)
    w = rng.standard_normal(24 * 50).astype(np.float32)
    w32 = w.copy()
    O.lib().or_normalize_l2(w32, 24, 50, 1.0, 0.0)
    w64 = w.copy()
    O.lib().or_normalize_l2_f64(w64, 24, 50, 2.0, 0.0)
    n32 = np.sqrt((w32.astype(np.float64).reshape(24, 50) ** 2).sum(1))
    n64 = np.sqrt((w64.astype(np.float64).reshape(24, 50) ** 2).sum(1))
    assert np.allclose(n32, 1.0, rtol=1e-6) and np.allclose(n64, 2.0, rtol=1e-7)
    # direction unchanged
    assert np.allclose(w64.reshape(24, 50) / n64[:, None], w.reshape(24, 50) /
                       np.sqrt((w.astype(np.float64).reshape(24, 50) ** 2).sum(1))[:, None], atol=1e-6)


def test_normalize_l2_degenerate_patch_ends_the_loop():
    """NormalizeL2.cpp:107-116 (shared weights, arbors together): a patch whose
    sum of squares is within minL2NormTolerated of zero breaks the patch loop,
    so that patch and every later one keep their weights; earlier patches are
    normalised."""
    rng = np.random.default_rng(21)
    w = rng.standard_normal((6, 40)).astype(np.float32)
    w[3] = 0.0
    got = w.copy()
    O.lib().or_normalize_l2(got.ravel(), 6, 40, 1.5, 0.0)
    n = np.sqrt((got[:3].astype(np.float64) ** 2).sum(1))
    assert np.allclose(n, 1.5, rtol=1e-6)
    assert np.array_equal(got[3:], w[3:])
    # a tolerance above a patch's sum of squares acts the same way
    w2 = rng.standard_normal((4, 40)).astype(np.float32)
    w2[1] *= 1e-3
    got2 = w2.copy()
    O.lib().or_normalize_l2(got2.ravel(), 4, 40, 1.0, 1e-3)
    assert np.allclose(np.sqrt((got2[0].astype(np.float64) ** 2).sum()), 1.0, rtol=1e-6)
    assert np.array_equal(got2[1:], w2[1:])


def test_lca_loop_descends_energy_and_is_sparse():
    rng = np.random.default_rng(13)
    nb, n, nf, nv, nfv, p = 1, 16, 3, 8, 24, 8
    img = rng.standard_normal((nb, n, n, nf)).astype(np.float32)
    img -= img.mean()
    W = rng.standard_normal((nfv, p, p, nf)).astype(np.float32)
    W /= np.sqrt((W.reshape(nfv, -1) ** 2).sum(1)).reshape(-1, 1, 1, 1)
    cfg = lca_cfg(nb, n, nf, nv, nfv, p)
    out = O.lca_run(cfg, W, img, None, 60)
    e = out["energy"][:, 0]
    assert e[-1] < e[5] < e[0]
    assert 0 < out["num_active"][-1, 0] < nv * nv * nfv


def lca_cfg(nb, n, nf, nv, nfv, p, vt=0.1, tau=100.0):
    cfg = O.LCAConfig()
    cfg.input = O.Loc.make(nb, n, n, nf, 0)
    mres = O.lib().or_requiredConvolveMargin(n, nv, O.lib().or_calcPostPatchSize(p, nv, n))
    mv1 = O.lib().or_requiredConvolveMargin(nv, n, p)
    cfg.res = O.Loc.make(nb, n, n, nf, mres)
    cfg.v1 = O.Loc.make(nb, nv, nv, nfv, mv1)
    cfg.nxp = cfg.nyp = p
    cfg.VThresh, cfg.AMin, cfg.AMax, cfg.AShift, cfg.VWidth = vt, 0.0, FMAX, vt, 0.0
    cfg.transfer_vertices = 0
    cfg.tau, cfg.dt = tau, 1.0
    cfg.selfInteract = 1
    cfg.errScale = 1.0
    cfg.num_threads = 1
    return cfg


def test_hebbian_normalize_apply_closed_form():
    """normalize_dW divides by the activation count (0 where the count is 0), then W += dW;
    without normalizeDw, W += dW unchanged."""
    rng = np.random.default_rng(14)
    n = 64
    w = rng.standard_normal(n).astype(np.float32)
    dw = rng.standard_normal(n).astype(np.float32)
    cnt = rng.integers(0, 4, n).astype(np.int64)
    w1, dw1 = w.copy(), dw.copy()
    O.lib().or_hebbian_normalize_apply(w1, dw1, cnt.ctypes.data, n, 1)
    expect_dw = np.where(cnt > 0, dw / np.maximum(cnt, 1).astype(np.float32), np.float32(0))
    assert np.array_equal(dw1, expect_dw.astype(np.float32))
    assert np.array_equal(w1, (w + expect_dw).astype(np.float32))
    w2, dw2 = w.copy(), dw.copy()
    O.lib().or_hebbian_normalize_apply(w2, dw2, cnt.ctypes.data, n, 0)
    assert np.array_equal(dw2, dw) and np.array_equal(w2, w + dw)


def test_ident_deliver_and_errscale_closed_form():
    loc = O.Loc.make(2, 5, 3, 4, (1, 2, 2, 1))
    rng = np.random.default_rng(15)
    A = rng.standard_normal(loc.nbatch * loc.num_ext).astype(np.float32)
    g = np.zeros(loc.nbatch * loc.num_res, np.float32)
    O.lib().or_ident_deliver(O.C.byref(loc), A, g)
    # restricted window of the extended buffer: rows up.., columns lt..
    Ar = A.reshape(2, loc.ny_ext, loc.nx_ext, 4)[:, loc.up:loc.up + 3,
                                                 loc.lt:loc.lt + 5]
    assert np.array_equal(g, np.ascontiguousarray(Ar).ravel())
    V = rng.standard_normal(30).astype(np.float32)
    V1 = V.copy()
    O.lib().or_errscale(30, np.float32(0.5), V1)
    assert np.array_equal(V1, (V * np.float32(0.5)).astype(np.float32))


def test_lca_loop_fixed_point():
    """The converged loop satisfies the LCA fixed point V = G + A with
    G = gather(x - scatter(A)), the scatter being the gather's adjoint; both maps
    here come from float64 torch conv2d (and its autograd), not from the oracle."""
    rng = np.random.default_rng(16)
    nb, n, nf, nv, nfv, p = 1, 16, 3, 8, 24, 8
    img = rng.standard_normal((nb, n, n, nf)).astype(np.float32)
    img -= img.mean()
    W = rng.standard_normal((nfv, p, p, nf)).astype(np.float32)
    W /= np.sqrt((W.reshape(nfv, -1) ** 2).sum(1)).reshape(-1, 1, 1, 1)
    cfg = lca_cfg(nb, n, nf, nv, nfv, p, vt=0.1, tau=4.0)
    out = O.lca_run(cfg, W, img, None, 400)
    S, T = cfg.res, cfg.v1
    ux = O.patch_axis(T, S, p, p, 0)["unshrunken"][T.lt:T.lt + T.nx]
    uy = O.patch_axis(T, S, p, p, 1)["unshrunken"][T.up:T.up + T.ny]
    stride = int(ux[1] - ux[0])
    w = torch.tensor(W, dtype=torch.float64).permute(0, 3, 1, 2)

    def gather(r):  # r: restricted residual [nb, nf, n, n] -> V1 input [nb, nfv, nv, nv]
        ext = torch.zeros((nb, nf, S.ny_ext, S.nx_ext), dtype=torch.float64)
        ext[:, :, S.up:S.up + n, S.lt:S.lt + n] = r
        return torch.nn.functional.conv2d(ext[:, :, uy[0]:, ux[0]:], w, stride=stride)[:, :, :nv, :nv]

    A = out["A"].reshape(nb, T.ny_ext, T.nx_ext, nfv)[:, T.up:T.up + nv, T.lt:T.lt + nv]
    a = torch.tensor(np.ascontiguousarray(A), dtype=torch.float64).permute(0, 3, 1, 2)
    r0 = torch.zeros((nb, nf, n, n), dtype=torch.float64, requires_grad=True)
    (gather(r0) * a).sum().backward()
    recon = r0.grad                                  # scatter(A) = gather^T A
    rec_or = torch.tensor(out["recon"].reshape(nb, n, n, nf)).permute(0, 3, 1, 2).double()
    assert (recon - rec_or).abs().max() <= 1e-5 * recon.abs().max()
    x = torch.tensor(img, dtype=torch.float64).permute(0, 3, 1, 2)
    G = gather(x - recon)
    V = torch.tensor(out["V"].reshape(nb, nv, nv, nfv), dtype=torch.float64).permute(0, 3, 1, 2)
    assert (V - (G + a)).abs().max() <= 2e-3 * V.abs().max()
    assert 0 < out["num_active"][-1, 0] < nv * nv * nfv


def test_adaptive_timescale_controller_cases():
    """AdaptiveTimeScaleController::calcTimesteps by cases: the first value and any
    rise of the energy reset to baseMin; a fall sets tauFactor*E0/((E0-E)/ts),
    capped at timeScaleMax which then grows by growthFactor; equal values keep
    the state."""
    from openpv_amd.timescale import AdaptiveTimeScaleController, PeriodicUpdate
    c = AdaptiveTimeScaleController(1, 0.55, 0.5, 0.1, 0.01)
    assert c.calc_timesteps([-1.0])[0] == 0.5          # equal to the initial -1: unchanged
    assert c.calc_timesteps([10.0])[0] == 0.5          # E_0 = -1 <= 0: reset
    ts = c.calc_timesteps([9.0])[0]                     # fall: 0.1 * 10 / ((10-9)/0.5) = 0.5
    assert ts == 0.1 * 10.0 / ((10.0 - 9.0) / 0.5)
    ts = c.calc_timesteps([8.99])[0]                    # steep: capped at 0.55, max grows
    assert ts == 0.55 and c.time_scale_max[0] == 0.55 * 1.01
    assert c.calc_timesteps([9.5])[0] == 0.5 and c.time_scale_max[0] == 0.55   # rise: reset
    p = PeriodicUpdate(20)
    hits = [t for t in range(1, 60) if p.need_update(t) and (p.updated(t) or True)]
    assert hits == [1, 21, 41]


def test_pvp_roundtrip_and_reference_fixtures(tmp_path):
    """PVP reader on the reference's own fixture files (headers and sizes) and a
    write/read round trip."""
    from openpv_amd import pvp
    g = os.path.join(HERE, "golden", "gpulca")
    h, t, d = pvp.read_activity(os.path.join(g, "inputImages.pvp"))
    assert (h["fileType"], h["nx"], h["ny"], h["nf"]) == (4, 128, 64, 1) and d.shape == (2, 64, 128, 1)
    h, geo, w = pvp.read_weights(os.path.join(g, "weights.pvp"))
    assert (h["nxp"], h["nyp"], h["nfp"], h["numPatches"]) == (18, 18, 1, 24)
    assert (geo[0, :, :2] == 18).all() and w.shape == (1, 24, 18, 18, 1)
    assert np.allclose(np.sqrt((w[0].reshape(24, -1).astype(np.float64) ** 2).sum(1)), 1.0, atol=1e-5)
    assert w.min() >= h["wMin"] - 1e-7 and w.max() <= h["wMax"] + 1e-7
    rng = np.random.default_rng(17)
    fr = rng.standard_normal((3, 5, 7, 2)).astype(np.float32)
    pvp.write_activity(str(tmp_path / "a.pvp"), fr, [1.0, 2.0, 3.5])
    h, t, d = pvp.read_activity(str(tmp_path / "a.pvp"))
    assert list(t) == [1.0, 2.0, 3.5] and np.array_equal(d, fr)


def _firm_fixture():
    """tests/FirmThresholdCostTest: 16x16x3 inputs (batch 2, displayPeriod 0, so
    batch element b shows frame b throughout) and FirmThresholdCostLayer's
    per-neuron cost with VThresh 0.5, VWidth 0.3 (in float), 10 steps."""
    from openpv_amd import pvp
    g = os.path.join(HERE, "golden", "firmthresh")
    _, _, x = pvp.read_activity(os.path.join(g, "firmthreshinput.pvp"))
    _, t, y = pvp.read_activity(os.path.join(g, "correctoutput.pvp"))
    return x, t, y


def test_firm_threshold_cost_matches_reference_fixture():
    """The oracle's firm-threshold cost (CostFunctions.hpp:37-50, FP64), element
    by element and summed, against the reference's own expected output: equal
    up to the float rounding of the reference layer's evaluation."""
    x, t, y = _firm_fixture()
    assert y.shape == (20, 16, 16, 3) and list(t) == [s for s in range(1, 11) for _ in range(2)]
    one = O.Loc.make(1, 1, 1, 1, 0)
    for b in range(2):
        xb = x[b].ravel()
        cost = np.array([O.cost_sum(one, xb[i:i + 1], False, 0, 2, 0.5, 0.3) for i in range(xb.size)])
        for step in range(10):
            np.testing.assert_allclose(cost, y[2 * step + b].ravel(), rtol=0, atol=1e-7)
    loc = O.Loc.make(2, 16, 16, 3, 0)
    for b in range(2):
        total = O.cost_sum(loc, x[:2].ravel(), False, b, 2, 0.5, 0.3)
        assert abs(total - y[b].astype(np.float64).sum()) <= 1e-6 * total


@needs_ref
@pytest.mark.parametrize("self_interact", [True, False])
def test_momentum_lca_matches_reference(self_interact):
    """The oracle's MomentumLCA V update against the reference's own
    updateMomentumLCAOnCPU (MomentumLCAInternalStateBuffer.kpp), bit for bit over
    three steps that carry prevDrive."""
    rng = np.random.default_rng(41)
    loc = O.Loc.make(3, 7, 5, 4, (1, 2, 2, 1))
    n = loc.nbatch * loc.num_res
    V = rng.standard_normal(n).astype(np.float32)
    prev = np.zeros(n, np.float32)
    Vr, prevr = V.copy(), prev.copy()
    dt = np.array([1.0, 0.5, 0.25])
    for _ in range(3):
        G = rng.standard_normal(n).astype(np.float32)
        A = rng.standard_normal(loc.nbatch * loc.num_ext).astype(np.float32)
        V, prev = O.update_momentum_lca(loc, self_interact, 0.4, dt, 100.0, G, A, prev, V)
        O.ref().ref_update_momentum_lca(np.array(loc.as_list(), np.int32), int(self_interact), 0.4,
                                        dt, 100.0, G, A, prevr, Vr)
        assert np.array_equal(V.view(np.uint32), Vr.view(np.uint32))
        assert np.array_equal(prev.view(np.uint32), prevr.view(np.uint32))


@needs_ref
def test_ista_update_matches_reference():
    """The oracle's ISTA V update against the reference's own
    updateISTAInternalStateBufferOnCPU (ISTAInternalStateBuffer.kpp), bit for
    bit over three steps, with activity of both signs and zeros, and per-batch
    time scales."""
    rng = np.random.default_rng(43)
    loc = O.Loc.make(3, 7, 5, 4, (1, 2, 2, 1))
    n = loc.nbatch * loc.num_res
    V = rng.standard_normal(n).astype(np.float32)
    Vr = V.copy()
    dt = np.array([1.0, 0.5, 0.3])
    for _ in range(3):
        G = rng.standard_normal(n).astype(np.float32)
        A = rng.standard_normal(loc.nbatch * loc.num_ext).astype(np.float32)
        A[rng.random(A.size) < 0.4] = 0.0
        V = O.update_ista(loc, 0.25, dt, 100.0, G, A, V)
        O.ref().ref_update_ista(np.array(loc.as_list(), np.int32), 0.25, dt, 100.0, G, A, Vr)
        assert np.array_equal(V.view(np.uint32), Vr.view(np.uint32))


def test_ista_update_closed_forms():
    """Where A = 0 the step is V + (dt/tau)*G; where G = VThresh*sign(A) it is a
    fixed point; a sign error (+VThresh*sign) would move those points."""
    loc = O.Loc.make(1, 4, 3, 2, 1)
    n = loc.num_res
    rng = np.random.default_rng(44)
    V = rng.standard_normal(n).astype(np.float32)
    A = np.zeros(loc.num_ext, np.float32)
    G = rng.standard_normal(n).astype(np.float32)
    dt = np.array([2.0])
    got = O.update_ista(loc, 0.5, dt, 10.0, G, A, V)
    rate = np.float32(np.float32(2.0) / np.float32(10.0))
    assert np.array_equal(got.view(np.uint32), (V + rate * G).astype(np.float32).view(np.uint32))
    # activity of both signs at every restricted neuron, G = VThresh*sign(A)
    Aext = np.zeros((loc.ny_ext, loc.nx_ext, loc.nf), np.float32)
    signs = np.where(rng.random((loc.ny, loc.nx, loc.nf)) < 0.5, -1.0, 1.0).astype(np.float32)
    Aext[1:1 + loc.ny, 1:1 + loc.nx] = signs * rng.uniform(0.1, 2.0, signs.shape).astype(np.float32)
    G = (np.float32(0.5) * signs).ravel()
    got = O.update_ista(loc, 0.5, dt, 10.0, G, Aext.ravel(), V)
    assert np.array_equal(got.view(np.uint32), V.view(np.uint32))


def test_momentum_lca_reduces_to_lca():
    """Rate 0 is the HyPerLCA update bit for bit; the drive it stores is that
    update from V = 0; a rate r adds r*prevDrive as one rounded term."""
    rng = np.random.default_rng(42)
    loc = O.Loc.make(2, 6, 5, 3, 1)
    n = loc.nbatch * loc.num_res
    G = rng.standard_normal(n).astype(np.float32)
    A = rng.standard_normal(loc.nbatch * loc.num_ext).astype(np.float32)
    V = rng.standard_normal(n).astype(np.float32)
    dt = np.array([1.0, 0.5])
    prev0 = rng.standard_normal(n).astype(np.float32)
    plain = O.update_lca(loc, True, dt, 100.0, G, A, V)
    V0, drive = O.update_momentum_lca(loc, True, 0.0, dt, 100.0, G, A, prev0, V)
    assert np.array_equal(V0.view(np.uint32), plain.view(np.uint32))
    from_zero = O.update_lca(loc, True, dt, 100.0, G, A, np.zeros(n, np.float32))
    assert np.array_equal(drive.view(np.uint32), from_zero.view(np.uint32))
    V1, _ = O.update_momentum_lca(loc, True, 0.4, dt, 100.0, G, A, prev0, V)
    expect = plain + np.float32(0.4) * prev0
    assert np.array_equal(V1.view(np.uint32), expect.astype(np.float32).view(np.uint32))


def test_momentum_simple_matches_reference_correct_state():
    """MomentumConnSimpleCheckpointerTest's expected state (its CorrectState.cpp):
    a 1x1 identity weight, input ceil(t/4), output input*w, momentum "simple" with
    timeConstantTau 0.75 and no decay, an update every fourth step with the
    Hebbian dW input*output (one count each, normalizeDw).  The oracle's
    momentum step gives the test's weights and dW bit for bit."""
    tau = np.float32(0.75)
    f, wf = O.momentum_factors(1, 0.75, 0.0, 1.0)
    assert (f, wf) == (0.75, 0.0)
    cw, cdw = np.float32(1.0), np.float32(0.0)     # CorrectState
    w, prev = np.array([1.0], np.float32), np.zeros(1, np.float32)
    updates = 0
    for t in range(1, 41):
        inp = np.float32(np.ceil(np.float32(t) / np.float32(4.0)))
        out = inp * cw
        if abs(np.fmod(t + 1.0, 4.0) - 1.0) < 0.5:
            base = inp * out
            cdw = (np.float32(1) - tau) * base + tau * cdw
            cw = cw + cdw
            w, dw, prev = O.momentum_apply(w, [base], prev, [1], f, wf)
            updates += 1
            assert dw[0].view(np.uint32) == np.float32(cdw).view(np.uint32)
            assert w[0].view(np.uint32) == np.float32(cw).view(np.uint32)
    assert updates == 10


def test_momentum_factors_and_forms():
    """applyMomentum's factors per method and the reference's parameter checks;
    both blends on values whose every product and sum is exact."""
    f, wf = O.momentum_factors(0, 100.0, 0.25)
    assert f == pytest.approx(np.exp(-0.01), rel=1e-7) and wf == 0.25
    assert O.momentum_factors(0, 0.0) == (0.0, 0.0)
    assert O.momentum_factors(2, 0.5, 0.25, 2.0) == (0.5, 0.5)
    for bad in [(0, -1.0, 0.0), (1, 1.0, 0.0), (2, -0.1, 0.0), (1, 0.5, 1.5), (3, 0.5, 0.0)]:
        with pytest.raises(ValueError):
            O.momentum_factors(*bad)
    # dw 1, prev 2, w 4, f 0.5, wf 0.25: 1*0.5 + 0.5*2 - 0.25*4 = 0.5
    w, dw, prev = O.momentum_apply([4.0], [1.0], [2.0], None, 0.5, 0.25)
    assert (w[0], dw[0], prev[0]) == (4.5, 0.5, 0.5)
    # deprecated momentumTau blend: 1 + (0.5*2 - 0.25*4) = 1
    w, dw, prev = O.momentum_apply([4.0], [1.0], [2.0], None, 0.5, 0.25, deprecated_tau=True)
    assert (w[0], dw[0], prev[0]) == (5.0, 1.0, 1.0)
    # normalize_dW first: counts 4 and 0
    w, dw, prev = O.momentum_apply([0.0, 0.0], [8.0, 3.0], [0.0, 0.0], [4, 0], 0.5, 0.0)
    assert dw.tolist() == [1.0, 0.0] and w.tolist() == [1.0, 0.0]
