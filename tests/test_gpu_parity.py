"""GPU parity of the HIP path against the CPU oracle, through the C ABI.

Bar: EXACT forms are bit-identical to the oracle (V, A, GSyn, dW, active
lists, counts); FP32 (FMA) forms within 1e-5 relative to the largest
magnitude; BF16 forms within the tolerance written at each test.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest
import torch

import oracle as O
from openpv_amd import abi, runtime as R
from openpv_amd.lca import LCAColumn

pytestmark = pytest.mark.gpu
FMAX = float(np.finfo(np.float32).max)
DEV = "cuda"


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def same(a, b):
    return np.array_equal(bits(a), bits(b))


def both(t):
    return abi.Loc.make(*t[:4], halo=t[4]), O.Loc.make(*t[:4], halo=t[4])


def dev(a, dtype=torch.float32):
    return torch.as_tensor(np.ascontiguousarray(a)).to(DEV, dtype=dtype).contiguous()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def ext_random(loc, rng, density=1.0, scale=1.0):
    """Extended buffer with zero halo and random restricted values."""
    A = np.zeros((loc.nbatch, loc.ny_ext, loc.nx_ext, loc.nf), np.float32)
    v = (rng.standard_normal((loc.nbatch, loc.ny, loc.nx, loc.nf)) * scale).astype(np.float32)
    if density < 1.0:
        v[rng.random(v.shape) >= density] = 0
    A[:, loc.halo.up:loc.halo.up + loc.ny, loc.halo.lt:loc.halo.lt + loc.nx] = v
    return A.ravel()


def test_device_layer():
    H = abi.hip()
    d = C.c_void_p()
    abi.check(H.pvhip_device_create(0, C.byref(d)), "device_create")
    p = abi.Props()
    abi.check(H.pvhip_device_props(d, C.byref(p)), "props")
    print("device:", p.name.decode(), p.arch.decode(), "CUs", p.compute_units, "LDS/block",
          p.lds_per_block, flush=True)
    assert p.arch.decode().startswith("gfx950") and p.wavefront_size == 64
    b = C.c_void_p()
    abi.check(H.pvhip_buffer_alloc(d, 4099, C.byref(b)), "alloc")
    src = np.arange(4099, dtype=np.uint8)
    out = np.zeros(4099, np.uint8)
    s = H.pvhip_device_stream(d)
    abi.check(H.pvhip_buffer_h2d(b, src.ctypes.data_as(C.c_void_p), 4099, 0, s), "h2d")
    abi.check(H.pvhip_buffer_d2h(b, out.ctypes.data_as(C.c_void_p), 4099, 0, s), "d2h")
    abi.check(H.pvhip_device_sync(d), "sync")
    assert (out == src).all()
    # zero-fill an unaligned sub-range
    base = H.pvhip_buffer_ptr(b)
    abi.check(H.pvhip_fill_zero(C.c_void_p(base + 3), 1001, s), "fill_zero")
    abi.check(H.pvhip_buffer_d2h(b, out.ctypes.data_as(C.c_void_p), 4099, 0, s), "d2h")
    abi.check(H.pvhip_device_sync(d), "sync")
    exp = src.copy()
    exp[3:1004] = 0
    assert (out == exp).all()
    assert H.pvhip_buffer_h2d(b, src.ctypes.data_as(C.c_void_p), 4099, 8, s) == abi.PVHIP_OK + 1
    abi.check(H.pvhip_buffer_free(b), "free")
    abi.check(H.pvhip_device_destroy(d), "destroy")


@pytest.mark.parametrize("vertices,params", [
    (False, (0.1, 0.0, FMAX, 0.1, 0.0)),
    (True, (0.1, 0.0, FMAX, 0.1, 0.0)),
    (False, (0.025, 0.0, FMAX, 0.0, 0.025)),
    (True, (0.025, 0.0, FMAX, 0.0, 0.025)),
    (False, (0.1, 0.0, 0.4, 0.05, 0.1)),
])
@pytest.mark.parametrize("from_v", [False, True])
def test_lca_update_bit_exact(vertices, params, from_v):
    rng = np.random.default_rng(21)
    hl, ol = both((3, 11, 7, 5, (2, 1, 1, 2)))
    n = hl.nbatch * hl.num_res
    G = (rng.standard_normal(2 * n) * 0.3).astype(np.float32)
    V = (rng.standard_normal(n) * 0.2).astype(np.float32)
    V[::5] = np.float32(params[0])
    A = O.transfer(ol, V, *params, vertices=vertices)
    # oracle: accumulate {1, -0.5}, update, transfer
    acc = np.zeros(n, np.float32)
    O.lib().or_gsyn_accumulate(n, 2, np.array([1.0, -0.5], np.float32), G, acc)
    Vr = O.update_lca(ol, True, np.array([1.0, 0.5, 2.0]), 100.0, acc, A, V)
    Ar = O.transfer(ol, Vr, *params, vertices=vertices, A=A)
    # A == T(V) on entry, so recomputing A(t-1) from V must give the same bits
    upd = R.HyPerLCAUpdate(hl, R.make_transfer(*params, vertices=vertices), 100.0, 1.0, True,
                           channel_coeffs=(1.0, -0.5), activity_from_v=from_v, device=DEV)
    upd.set_dt([1.0, 0.5, 2.0])
    Vd, Ad = dev(V), dev(A)
    upd.update(dev(G), Vd, Ad)
    assert same(host(Vd), Vr) and same(host(Ad), Ar)


@pytest.mark.parametrize("nx,ny,nf,halo", [
    (9, 7, 3, 3),     # rows of 27 values: the one-value-per-thread form
    (16, 6, 1, 6),    # rows of 16, A's rows start on 2 mod 4: 16-byte reads, 8-byte A stores
    (16, 5, 4, 2),    # A's rows on multiples of 4 values: 16-byte accesses throughout
])
def test_error_update_equals_ident_ident_layer_update(nx, ny, nf, halo):
    """pvhip_error_update (the residual layer in one pass) against the separate
    pvhip_ident_deliver x2 + pvhip_layer_update, bit for bit, signed zeros
    included, with and without errScale and with a non-identity transfer, in
    each of its access forms."""
    rng = np.random.default_rng(24)
    hl = abi.Loc.make(2, nx, ny, nf, halo)
    hi = abi.Loc.make(2, nx, ny, nf, 0)
    hr = abi.Loc.make(2, nx, ny, nf, (1, 2, 0, 1) if nf != 1 else (4, 0, 0, 2))
    n = hl.nbatch * hl.num_res
    xin = (rng.standard_normal(hi.nbatch * hi.num_ext) * 0.5).astype(np.float32)
    xrec = (rng.standard_normal(hr.nbatch * hr.num_ext) * 0.5).astype(np.float32)
    xin[::7] = -0.0
    xrec[::5] = -0.0
    xrec[::11] = 0.0
    for scale, tr in ((None, R.make_transfer()), (0.75, R.make_transfer()),
                      (1.25, R.make_transfer(0.1, 0.0, FMAX, 0.05, 0.1))):
        g = torch.zeros(2 * n, device=DEV)
        V1, A1 = torch.zeros(n, device=DEV), torch.full((hl.nbatch * hl.num_ext,), 3.0, device=DEV)
        R.ident_deliver(hi, dev(xin), g[:n], accumulate=False)
        R.ident_deliver(hr, dev(xrec), g[n:], accumulate=False)
        R.layer_update(hl, g, V1, A1, tr, channel_coeffs=(1.0, -1.0), err_scale=scale)
        V2, A2 = torch.zeros(n, device=DEV), torch.full((hl.nbatch * hl.num_ext,), 3.0, device=DEV)
        R.error_update(hl, hi, dev(xin), hr, dev(xrec), V2, A2, tr, err_scale=scale)
        assert same(host(V1), host(V2)) and same(host(A1), host(A2))


def test_layer_updates_and_ident_bit_exact():
    rng = np.random.default_rng(22)
    hl, ol = both((2, 9, 6, 3, 3))
    n = hl.nbatch * hl.num_res
    G = rng.standard_normal(2 * n).astype(np.float32)
    acc = np.zeros(n, np.float32)
    O.lib().or_gsyn_accumulate(n, 2, np.array([1.0, -1.0], np.float32), G, acc)
    Vr = acc.copy()
    O.lib().or_errscale(n, np.float32(0.75), Vr)
    Ar = O.transfer(ol, Vr, -FMAX, -FMAX, FMAX, 0.0, 0.0)
    Vd = torch.zeros(n, device=DEV)
    Ad = torch.zeros(hl.nbatch * hl.num_ext, device=DEV)
    R.layer_update(hl, dev(G), Vd, Ad, R.make_transfer(), channel_coeffs=(1.0, -1.0), err_scale=0.75)
    assert same(host(Vd), Vr) and same(host(Ad), Ar)
    # IdentDelivery from an extended buffer, fresh and accumulating
    pre = ext_random(hl, rng)
    g = np.zeros(n, np.float32)
    O.lib().or_ident_deliver(O.C.byref(ol), pre, g)
    gd = torch.full((n,), 7.0, device=DEV)
    R.ident_deliver(hl, dev(pre), gd, accumulate=False)
    assert same(host(gd), g)
    O.lib().or_ident_deliver(O.C.byref(ol), pre, g)
    R.ident_deliver(hl, dev(pre), gd, accumulate=True)
    assert same(host(gd), g)


@pytest.mark.parametrize("shape,density", [
    ((2, 16, 16, 128, 2), 0.05), ((3, 13, 7, 3, (1, 2, 3, 0)), 0.3), ((1, 36, 36, 512, 2), 0.02),
    ((2, 8, 8, 5, 1), 0.0), ((1, 9, 9, 64, 0), 1.0)])
def test_active_compact_bit_exact(shape, density):
    rng = np.random.default_rng(23)
    hl, ol = both(shape)
    A = ext_random(hl, rng, density)
    if A.size > 11:
        A[11] = -0.0  # negative zero is inactive
    ent, cnt = O.active_indices(ol, A)
    al = R.ActiveList(hl, DEV)
    al.compact(dev(A))
    e, n = al.host_entries()
    assert (n == cnt).all()
    ps = host(al.pos_start).astype(np.int64).reshape(hl.nbatch, -1)
    npos = hl.nx_ext * hl.ny_ext
    for b in range(hl.nbatch):
        eo = ent[b * ol.num_ext:b * ol.num_ext + cnt[b]]
        assert (e[b, :cnt[b]]["index"] == eo["index"]).all()
        assert same(e[b, :cnt[b]]["value"], eo["value"])
        # pos_start[p] = number of active entries before position p
        Ab = A[b * ol.num_ext:(b + 1) * ol.num_ext].reshape(npos, hl.nf)
        counts = (Ab != 0).sum(1)
        assert (ps[b, :npos] == np.concatenate([[0], np.cumsum(counts)[:-1]])).all()
        assert ps[b, npos] == cnt[b]


GATHER_CASES = [
    # source, target, post-perspective patch
    ((2, 32, 32, 3, 3), (2, 16, 16, 128, 2), 8),        # config[0] Residual -> V1
    ((2, 64, 64, 1, 6), (2, 16, 16, 96, 2), 16),        # config[2] geometry, smaller
    ((1, 64, 64, 16, 6), (1, 32, 32, 40, 0), 12),       # config[1] geometry, smaller
    ((1, 16, 16, 4, 2), (1, 16, 16, 8, 1), 5),          # one-to-one
    ((1, 16, 16, 4, 2), (1, 32, 32, 3, 0), 4),          # target larger: direct kernel
    ((1, 20, 12, 2, (2, 3, 3, 2)), (1, 10, 6, 7, 1), 6),  # asymmetric halo, ragged tiles
]


def gather_setup(case, rng, density=1.0):
    (hs, os_), (ht, ot), p = both(case[0]), both(case[1]), case[2]
    nkx = ot.nx // os_.nx if ot.nx > os_.nx else 1
    nky = ot.ny // os_.ny if ot.ny > os_.ny else 1
    W = (rng.standard_normal((nkx * nky * ot.nf, p, p, os_.nf)) * 0.05).astype(np.float32)
    A = ext_random(hs, rng, density)
    return hs, os_, ht, ot, p, W, A


@pytest.mark.parametrize("case", GATHER_CASES)
def test_recv_post_exact_bit_exact(case):
    rng = np.random.default_rng(24)
    hs, os_, ht, ot, p, W, A = gather_setup(case, rng)
    ref = O.recv_post(os_, ot, p, p, A, W)
    d = R.PostsynapticPerspectiveGPUDelivery(hs, ht, p, p, dev(W), arith=abi.ARITH_EXACT)
    g = torch.full((ht.nbatch * ht.num_res,), 3.0, device=DEV)
    d.deliver(dev(A), g, accumulate=False)
    assert same(host(g), ref)
    # accumulate into existing GSyn: second delivery adds in the same order
    ref2 = O.recv_post(os_, ot, p, p, A, W, gsyn=ref.copy())
    d.deliver(dev(A), g, accumulate=True)
    assert same(host(g), ref2)


@pytest.mark.parametrize("case", GATHER_CASES[:4])
def test_recv_post_fp32_and_bf16(case):
    rng = np.random.default_rng(25)
    hs, os_, ht, ot, p, W, A = gather_setup(case, rng)
    ref = O.recv_post(os_, ot, p, p, A, W).astype(np.float64)
    scale = np.abs(ref).max()
    d32 = R.PostsynapticPerspectiveGPUDelivery(hs, ht, p, p, dev(W), arith=abi.ARITH_FP32)
    g = torch.zeros(ht.nbatch * ht.num_res, device=DEV)
    d32.deliver(dev(A), g)
    assert np.abs(host(g) - ref).max() <= 1e-5 * scale
    d16 = R.PostsynapticPerspectiveGPUDelivery(hs, ht, p, p, dev(W), arith=abi.ARITH_BF16)
    g.zero_()
    d16.deliver(dev(A), g)
    err = np.abs(host(g) - ref)
    # BF16 inputs (8-bit mantissa), FP32 accumulation over K terms: relative
    # error of a K-term dot product stays within ~2^-8 * sqrt(K) of the scale.
    K = p * p * os_.nf
    assert err.max() <= 2.0 ** -8 * np.sqrt(K) * np.abs(W).max() * np.abs(A).max() * 4
    assert np.sqrt((err ** 2).mean()) <= 1e-2 * np.sqrt((ref ** 2).mean())


@pytest.mark.parametrize("case", GATHER_CASES[:3])
@pytest.mark.parametrize("arith", [abi.ARITH_EXACT, abi.ARITH_BF16])
@pytest.mark.parametrize("from_v,vertices", [(False, False), (True, False), (True, True)])
def test_recv_post_fused_update(case, arith, from_v, vertices):
    """Gather with the LCA update in its epilogue == gather then update; with
    activity_from_v the epilogue recomputes A(t-1) = T(V(t-1)) instead of reading it."""
    rng = np.random.default_rng(26)
    hs, os_, ht, ot, p, W, A = gather_setup(case, rng)
    n = ht.nbatch * ht.num_res
    V0 = (rng.standard_normal(n) * 0.1).astype(np.float32)
    tr = R.make_transfer(0.05, 0.0, FMAX, 0.05, 0.0, vertices=vertices)
    A0 = O.transfer(ot, V0, 0.05, 0.0, FMAX, 0.05, 0.0, vertices=vertices)
    d = R.PostsynapticPerspectiveGPUDelivery(hs, ht, p, p, dev(W), arith=arith)
    upd = R.HyPerLCAUpdate(ht, tr, 100.0, 1.0, True, device=DEV)
    fus = R.HyPerLCAUpdate(ht, tr, 100.0, 1.0, True, activity_from_v=from_v, device=DEV)
    g = torch.zeros(n, device=DEV)
    Va, Aa = dev(V0), dev(A0)
    d.deliver(dev(A), g)
    upd.update(g, Va, Aa)
    Vb, Ab = dev(V0), dev(A0)
    d.deliver(dev(A), None, fused=fus.args(None, Vb, Ab, num_channels=1))
    assert same(host(Va), host(Vb)) and same(host(Aa), host(Ab))


SCATTER_CASES = [
    # pre, post, pre-perspective patch, density
    ((2, 16, 16, 128, 2), (2, 32, 32, 3, 0), 8, 0.05),     # config[0] V1 -> Recon
    ((2, 16, 16, 96, 2), (2, 64, 64, 1, 0), 16, 0.03),     # config[2] geometry
    ((1, 64, 64, 16, 5), (1, 32, 32, 64, 0), 6, 0.02),     # config[1] geometry
    ((1, 12, 12, 3, 2), (1, 12, 12, 4, 1), 5, 0.3),        # one-to-one
    ((1, 10, 14, 5, (2, 1, 3, 2)), (1, 20, 28, 3, 0), 6, 0.2),  # asymmetric halo
]


@pytest.mark.parametrize("case", SCATTER_CASES)
def test_recv_pre_exact_bit_exact(case):
    rng = np.random.default_rng(27)
    (hp, op), (hq, oq), p, dens = both(case[0]), both(case[1]), case[2], case[3]
    kx, ky, kf = O.num_kernels(op, oq)
    W = (rng.standard_normal((kx * ky * kf, p, p, oq.nf)) * 0.1).astype(np.float32)
    A = ext_random(hp, rng, dens)
    ent, cnt = O.active_indices(op, A)
    ref = O.recv_pre(op, oq, p, p, W, entries=ent, num_active=cnt, num_threads=1)
    d = R.PresynapticPerspectiveGPUDelivery(hp, hq, p, p, dev(W), arith=abi.ARITH_EXACT)
    al = R.ActiveList(hp, DEV)
    al.compact(dev(A))
    g = torch.full((hq.nbatch * hq.num_res,), -1.0, device=DEV)
    d.deliver(g, active=al)
    assert same(host(g), ref)
    # dense pre layer (sparseLayer = false) reads A directly: same order, same bits
    g2 = torch.zeros_like(g)
    d.deliver(g2, pre_activity=dev(A))
    assert same(host(g2), ref)
    # FP32 (FMA) form
    d32 = R.PresynapticPerspectiveGPUDelivery(hp, hq, p, p, dev(W), arith=abi.ARITH_FP32)
    d32.deliver(g2, active=al, pre_activity=dev(A))
    assert np.abs(host(g2) - ref).max() <= 1e-5 * max(1e-30, np.abs(ref).max())


def test_recv_pre_exact_batches_with_unequal_side_halos():
    """lt != rt with several batch elements: each batch element of the EXACT
    scatter equals, bit for bit, the oracle run on that element alone.  The
    library steps batches by the extended buffer's true size; the reference's
    CPU scatter (and the oracle, which keeps its arithmetic) by nx + rt + rt
    (PresynapticPerspectiveConvolveDelivery.cpp:73), which differs only for
    lt != rt, a halo the reference never builds (its margins are equal on both
    sides of an axis)."""
    rng = np.random.default_rng(29)
    pre, post, p = (2, 10, 14, 5, (2, 1, 3, 2)), (2, 20, 28, 3, 0), 6
    (hp, op), (hq, oq) = both(pre), both(post)
    kx, ky, kf = O.num_kernels(op, oq)
    W = (rng.standard_normal((kx * ky * kf, p, p, oq.nf)) * 0.1).astype(np.float32)
    A = ext_random(hp, rng, 0.2)
    d = R.PresynapticPerspectiveGPUDelivery(hp, hq, p, p, dev(W), arith=abi.ARITH_EXACT)
    al = R.ActiveList(hp, DEV)
    al.compact(dev(A))
    g = torch.zeros(hq.nbatch * hq.num_res, device=DEV)
    d.deliver(g, active=al)
    got = host(g).reshape(2, -1)
    op1 = O.Loc.make(1, *pre[1:4], halo=pre[4])
    oq1 = O.Loc.make(1, *post[1:4], halo=post[4])
    for b in range(2):
        Ab = A.reshape(2, -1)[b].copy()
        ent, cnt = O.active_indices(op1, Ab)
        ref = O.recv_pre(op1, oq1, p, p, W, entries=ent, num_active=cnt, num_threads=1)
        assert same(got[b], ref), b


@pytest.mark.parametrize("case", [
    ((2, 16, 16, 32, 2), (2, 32, 32, 3, 3), 8, True),     # V1 -> Residual (config[0] shape)
    ((1, 16, 16, 24, 2), (1, 64, 64, 1, 6), 16, False),   # config[2] geometry
    ((1, 8, 8, 5, 1), (1, 8, 8, 2, 1), 3, True),          # one-to-one
    ((1, 32, 32, 4, 5), (1, 16, 16, 3, 2), 6, True),      # many-to-one (4 kernels)
])
def test_hebbian_dw_bit_exact(case):
    rng = np.random.default_rng(28)
    (hp, op), (hq, oq), p, norm = both(case[0]), both(case[1]), case[2], case[3]
    a = ext_random(hp, rng, 0.2)
    r = ext_random(hq, rng)
    dw, cnt = O.hebbian_dw(op, oq, p, p, oq.nf, 0.25, a, r, counts=norm)
    hb = R.HebbianUpdater(hp, hq, p, p, 0.25, normalize_dw=norm, device=DEV)
    hb.update_dw(dev(a), dev(r))
    assert same(host(hb.dw), dw)
    if norm:
        assert (host(hb.counts) == cnt).all()
    # normalize_dW + updateWeights
    w = rng.standard_normal(dw.size).astype(np.float32)
    wr, dwr = w.copy(), dw.copy()
    O.lib().or_hebbian_normalize_apply(wr, dwr, None if cnt is None else cnt.ctypes.data, dw.size,
                                       int(norm))
    wd = dev(w)
    hb.apply(wd)
    assert same(host(wd), wr) and same(host(hb.dw), dwr)


def test_fp64_probes_and_normalize():
    rng = np.random.default_rng(29)
    hl, ol = both((3, 10, 9, 7, 2))
    buf = ext_random(hl, rng)
    for kind, p0, p1 in [(0, 0, 0), (1, 0, 0), (2, 0.025, 0.025), (3, 0.5, 0)]:
        out = host(R.cost_sum_f64(hl, dev(buf), True, kind, p0, p1))
        for b in range(3):
            assert out[b] == pytest.approx(O.cost_sum(ol, buf, True, b, kind, p0, p1), rel=1e-12)
    w = rng.standard_normal(40 * 256).astype(np.float32)
    wr = w.copy()
    O.lib().or_normalize_l2_f64(wr, 40, 256, 1.0, 0.0)
    wd = dev(w)
    R.normalize_l2_f64(wd, 40, 256)
    assert np.abs(host(wd) - wr).max() <= 2 ** -23


def c0_problem(nb=1, seed=0, n=32, nv=16, nfv=128, p=8, nf=3):
    rng = np.random.default_rng(seed)
    img = rng.standard_normal((nb, n, n, nf)).astype(np.float32)
    img -= img.reshape(nb, -1).mean(1).reshape(nb, 1, 1, 1).astype(np.float32)
    W = rng.standard_normal((nfv, p, p, nf))
    W = (W / np.sqrt((W.reshape(nfv, -1) ** 2).sum(1)).reshape(-1, 1, 1, 1)).astype(np.float32)
    return img, W


def oracle_cfg(col, p, vt=0.1):
    cfg = O.LCAConfig()
    il, rl, vl = col.in_loc, col.res_loc, col.v1_loc
    cfg.input = O.Loc.make(il.nbatch, il.nx, il.ny, il.nf, 0)
    cfg.res = O.Loc.make(rl.nbatch, rl.nx, rl.ny, rl.nf, rl.halo_tuple())
    cfg.v1 = O.Loc.make(vl.nbatch, vl.nx, vl.ny, vl.nf, vl.halo_tuple())
    cfg.nxp = cfg.nyp = p
    cfg.VThresh, cfg.AMin, cfg.AMax, cfg.AShift, cfg.VWidth = vt, 0.0, FMAX, vt, 0.0
    cfg.transfer_vertices = 0
    cfg.tau, cfg.dt = 100.0, 1.0
    cfg.selfInteract = 1
    cfg.errScale = 1.0
    cfg.num_threads = 1
    return cfg


@pytest.mark.parametrize("fused", [False, True])
def test_lca_column_config0_bit_exact_200_steps(fused):
    """config[0]: 32x32x3 -> 16x16x128, 8x8x3 stride 2, VThresh 0.1, tau 100, 200 steps."""
    img, W = c0_problem()
    col = LCAColumn(img, W, 16, 16, 128, 8, 8, VThresh=0.1, arith=abi.ARITH_EXACT, fused=fused,
                    device=DEV)
    nact = []
    for _ in range(200):
        col.step()
        nact.append(int(col.num_active()[0]))
    V, A = col.v1_state()
    ref = O.lca_run(oracle_cfg(col, 8), W, img, None, 200)
    assert nact == ref["num_active"][:, 0].tolist()
    assert same(V, ref["V"]) and same(A, ref["A"])
    assert same(host(col.rec_A), ref["recon"])
    e = host(col.probe_energy())
    assert e[0] == pytest.approx(ref["energy"][-1, 0], rel=1e-12)


def test_lca_column_batched_graph_replay():
    """hipGraph-captured step replays to the same bits as eager steps."""
    img, W = c0_problem(nb=4, seed=3)
    eager = LCAColumn(img, W, 16, 16, 128, 8, 8, VThresh=0.1, arith=abi.ARITH_EXACT, fused=True,
                      device=DEV)
    graph = LCAColumn(img, W, 16, 16, 128, 8, 8, VThresh=0.1, arith=abi.ARITH_EXACT, fused=True,
                      device=DEV)
    s = torch.cuda.Stream()
    torch.cuda.synchronize()
    graph.capture(s)           # capture records, it does not execute
    for i in range(12):
        eager.step()
        graph.replay(s)
        s.synchronize()
        torch.cuda.synchronize()
        Ve, Ae = eager.v1_state()
        Vg, Ag = graph.v1_state()
        assert same(Ve, Vg) and same(Ae, Ag), "replay %d differs" % i
        assert (eager.num_active() == graph.num_active()).all()
    graph.release_graph()


@pytest.mark.parametrize("arith", [abi.ARITH_EXACT, abi.ARITH_BF16])
def test_lca_column_multi_step_graph_replay(arith):
    """A hipGraph of four captured steps: each replay advances the column by
    four steps, to the same bits as four eager steps."""
    img, W = c0_problem(nb=4, seed=4)
    eager = LCAColumn(img, W, 16, 16, 128, 8, 8, VThresh=0.1, arith=arith, fused=True, device=DEV)
    graph = LCAColumn(img, W, 16, 16, 128, 8, 8, VThresh=0.1, arith=arith, fused=True, device=DEV)
    s = torch.cuda.Stream()
    torch.cuda.synchronize()
    graph.capture(s, steps=4)
    assert graph.graph_steps == 4
    for i in range(3):
        for _ in range(4):
            eager.step()
        graph.replay(s)
        s.synchronize()
        torch.cuda.synchronize()
        for x, y in ((eager.v1_V, graph.v1_V), (eager.v1_activity(), graph.v1_activity()),
                     (eager.rec_A, graph.rec_A), (eager.res_A, graph.res_A)):
            assert np.array_equal(host(x).view(np.uint32), host(y).view(np.uint32)), "replay %d" % i
        assert (eager.num_active() == graph.num_active()).all()
    graph.release_graph()
    with pytest.raises(ValueError):
        graph.capture(s, steps=0)


def test_lca_column_bf16_energy_and_sparsity():
    """BF16 MFMA delivery with FP32 state vs the EXACT column (2 x 32x32x3 ->
    16x16x128, 8x8 patches at stride 2), 200 steps: energy within 1e-3
    relative and active counts within 0.5 %."""
    img, W = c0_problem(nb=2, seed=5)
    ex = LCAColumn(img, W, 16, 16, 128, 8, 8, VThresh=0.1, arith=abi.ARITH_EXACT, fused=True,
                   device=DEV)
    bf = LCAColumn(img, W, 16, 16, 128, 8, 8, VThresh=0.1, arith=abi.ARITH_BF16, fused=True,
                   device=DEV)
    for _ in range(200):
        ex.step()
        bf.step()
    e_ex = host(ex.probe_energy())
    e_bf = host(bf.probe_energy())
    n_ex, n_bf = ex.num_active(), bf.num_active()
    rel = np.abs(e_bf - e_ex) / np.abs(e_ex)
    print("bf16 energy rel err", rel, "active", n_ex, n_bf, flush=True)
    assert rel.max() <= 1e-3
    assert (np.abs(n_bf - n_ex) <= 0.005 * n_ex).all()


@pytest.mark.parametrize("case", [
    # source, target, post-perspective patch (target larger: 16 / 4 data patches per feature)
    ((2, 16, 16, 64, 2), (2, 64, 64, 2, 0), 4),       # transpose of a stride-4 one-to-many
    ((1, 16, 16, 128, 2), (1, 32, 32, 3, 0), 4),      # config[0] Recon <- V1 transposed
    ((1, 8, 12, 32, 2), (1, 16, 24, 5, 1), 4),        # stride 2, ragged tiles
    ((1, 80, 20, 32, 2), (1, 160, 40, 3, 0), 4),      # window grid wider than one strip
    ((2, 20, 20, 32, 2), (2, 40, 40, 16, 0), 4),      # 64 kernels: four MFMA row blocks
])
@pytest.mark.parametrize("arith", [abi.ARITH_FP32, abi.ARITH_BF16])
def test_recv_post_several_patches_mfma(case, arith):
    rng = np.random.default_rng(30)
    hs, os_, ht, ot, p, W, A = gather_setup(case, rng, density=0.3)
    ref = O.recv_post(os_, ot, p, p, A, W).astype(np.float64)
    d = R.PostsynapticPerspectiveGPUDelivery(hs, ht, p, p, dev(W), arith=arith)
    g = torch.full((ht.nbatch * ht.num_res,), 5.0, device=DEV)
    d.deliver(dev(A), g)
    err = np.abs(host(g) - ref)
    scale = np.abs(ref).max()
    if arith == abi.ARITH_FP32:
        assert err.max() <= 1e-5 * scale
    else:
        assert np.sqrt((err ** 2).mean()) <= 1e-2 * np.sqrt((ref ** 2).mean())
    # accumulate adds onto the existing GSyn
    d.deliver(dev(A), g, accumulate=True)
    assert np.abs(host(g) - 2 * ref).max() <= (2e-5 if arith == abi.ARITH_FP32 else 4e-2) * scale


@pytest.mark.parametrize("case", [GATHER_CASES[0], GATHER_CASES[1], GATHER_CASES[2], GATHER_CASES[5]])
def test_recv_post_fp8_exact_on_representable_values(case):
    """FP8 (e4m3) gather on values that survive the conversion exactly: weights in
    {0, +-1/32, +-1/16} (per-kernel scale 2^11 -> {0, 64, 128}), sources in
    {0, +-1/4, +-1, +-2} (source scale 2^6 -> {0, 16, 64, 128}).  Every product
    and partial sum is exact in FP32, so the result equals the oracle's float64
    sum bit for bit: the FP8 path's fragment layout, conversion and scaling."""
    rng = np.random.default_rng(40)
    hs, os_, ht, ot, p, W, A = gather_setup(case, rng)
    W = (rng.choice([-2, -1, 0, 1, 2], size=W.shape) / 32.0).astype(np.float32)
    ext = A.reshape(hs.nbatch, hs.ny_ext, hs.nx_ext, hs.nf) != 0
    vals = rng.choice([-2.0, -1.0, -0.25, 0.0, 0.25, 1.0, 2.0], size=ext.shape).astype(np.float32)
    A = np.where(ext, vals, 0.0).astype(np.float32).ravel()
    ref = O.recv_post(os_, ot, p, p, A, W).astype(np.float64)
    # the float64 sum of exact products is itself exact here (small integers x 2^-7)
    d = R.PostsynapticPerspectiveGPUDelivery(hs, ht, p, p, dev(W), arith=abi.ARITH_FP8)
    d.set_source_scale(R.fp8_scale_for(2.0))
    assert R.fp8_scale_for(2.0) == 64.0
    g = torch.full((ht.nbatch * ht.num_res,), 9.0, device=DEV)
    d.deliver(dev(A), g)
    assert np.array_equal(host(g).astype(np.float64), ref)
    d.deliver(dev(A), g, accumulate=True)
    assert np.array_equal(host(g).astype(np.float64), 2 * ref)


@pytest.mark.parametrize("case", [GATHER_CASES[0], GATHER_CASES[1]])
def test_recv_post_fp8_error_bound(case):
    """FP8 gather on random data against the oracle: e4m3 keeps 3 mantissa bits
    (relative rounding <= 2^-4 per operand), so the RMS error of a dot product
    is a few percent of the RMS result; the largest error stays within 0.15 of
    the largest magnitude."""
    rng = np.random.default_rng(41)
    hs, os_, ht, ot, p, W, A = gather_setup(case, rng, density=0.5)
    ref = O.recv_post(os_, ot, p, p, A, W).astype(np.float64)
    d = R.PostsynapticPerspectiveGPUDelivery(hs, ht, p, p, dev(W), arith=abi.ARITH_FP8)
    d.set_source_scale(R.fp8_scale_for(np.abs(A).max()))
    g = torch.zeros(ht.nbatch * ht.num_res, device=DEV)
    d.deliver(dev(A), g)
    err = host(g) - ref
    rms = np.sqrt((err ** 2).mean()) / np.sqrt((ref ** 2).mean())
    print("fp8 rms rel err", rms, "max", np.abs(err).max() / np.abs(ref).max(), flush=True)
    assert rms <= 5e-2
    assert np.abs(err).max() <= 0.15 * np.abs(ref).max()


def test_lca_column_fp8_energy_and_sparsity():
    """The same column with the FP8 gather (the scatter runs its BF16 route),
    200 steps against the EXACT column: energy within 1e-3 relative and active
    counts within 0.5 % (measured values printed)."""
    img, W = c0_problem(nb=2, seed=5)
    cols = {}
    for name, ar in (("exact", abi.ARITH_EXACT), ("fp8", abi.ARITH_FP8)):
        c = LCAColumn(img, W, 16, 16, 128, 8, 8, VThresh=0.1, arith=ar, fused=True, device=DEV)
        for _ in range(200):
            c.step()
        cols[name] = c
    e_ex, e_lp = host(cols["exact"].probe_energy()), host(cols["fp8"].probe_energy())
    n_ex, n_lp = cols["exact"].num_active(), cols["fp8"].num_active()
    rel = np.abs(e_lp - e_ex) / np.abs(e_ex)
    print("fp8 column energy rel err", rel, "active", n_ex, n_lp, flush=True)
    assert rel.max() <= 1e-3
    assert (np.abs(n_lp - n_ex) <= 0.005 * n_ex).all()


@pytest.mark.parametrize("arith", [abi.ARITH_FP32, abi.ARITH_BF16])
def test_channel_gemm_split_equals_unsplit(arith):
    """The channel GEMM with its channels split over workgroups (partials in the
    caller's scratch, added in split order) against the same GEMM unsplit and
    the oracle, at the transposed stride-4 scatter of the LCA bench (4 images
    32x32x512 -> 128x128x1, 4x4 post patch); accumulate adds onto GSyn."""
    rng = np.random.default_rng(33)
    case = ((4, 32, 32, 512, 2), (4, 128, 128, 1, 0), 4)
    hs, os_, ht, ot, p, W, A = gather_setup(case, rng, density=0.1)
    ref = O.recv_post(os_, ot, p, p, A, W).astype(np.float64)
    split = R.PostsynapticPerspectiveGPUDelivery(hs, ht, p, p, dev(W), arith=arith)
    split.set_launch(cgemm_tile=0, cgemm_splits=4)
    whole = R.PostsynapticPerspectiveGPUDelivery(hs, ht, p, p, dev(W), arith=arith, split=False)
    assert split.scratch is not None and whole.scratch is None
    nbytes = split.scratch.numel() * 4
    assert nbytes % (ht.nbatch * ht.num_res * 4) == 0 and nbytes // (ht.nbatch * ht.num_res * 4) > 1
    gs = torch.full((ht.nbatch * ht.num_res,), 3.0, device=DEV)
    gw = torch.full((ht.nbatch * ht.num_res,), 3.0, device=DEV)
    split.deliver(dev(A), gs, accumulate=True)
    whole.deliver(dev(A), gw, accumulate=True)
    a, b = host(gs) - 3.0, host(gw) - 3.0
    scale = np.abs(ref).max()
    assert np.abs(a - b).max() <= 1e-5 * scale          # only the order of the FP32 sums differs
    tol = 1e-5 if arith == abi.ARITH_FP32 else 1e-2
    assert np.abs(a - ref).max() <= tol * scale
    # the EXACT form has no channel GEMM: no scratch
    ex = R.PostsynapticPerspectiveGPUDelivery(hs, ht, p, p, dev(W), arith=abi.ARITH_EXACT)
    assert ex.scratch is None


@pytest.mark.parametrize("case", SCATTER_CASES[:2])
@pytest.mark.parametrize("arith", [abi.ARITH_FP32, abi.ARITH_BF16])
def test_recv_pre_transposed_mfma_route(case, arith):
    """BF16 / FP32 scatter as a post-perspective gather with transposed weights."""
    rng = np.random.default_rng(31)
    (hp, op), (hq, oq), p, dens = both(case[0]), both(case[1]), case[2], case[3]
    kx, ky, kf = O.num_kernels(op, oq)
    W = (rng.standard_normal((kx * ky * kf, p, p, oq.nf)) * 0.1).astype(np.float32)
    A = ext_random(hp, rng, dens)
    ent, cnt = O.active_indices(op, A)
    ref = O.recv_pre(op, oq, p, p, W, entries=ent, num_active=cnt, num_threads=1).astype(np.float64)
    d = R.PresynapticPerspectiveGPUDelivery(hp, hq, p, p, dev(W), arith=arith)
    assert d.transposed is not None
    g = torch.zeros(hq.nbatch * hq.num_res, device=DEV)
    d.deliver(g, pre_activity=dev(A))
    err = np.abs(host(g) - ref)
    if arith == abi.ARITH_FP32:
        assert err.max() <= 1e-5 * np.abs(ref).max()
    else:
        assert np.sqrt((err ** 2).mean()) <= 1e-2 * np.sqrt((ref ** 2).mean())


@pytest.mark.parametrize("case,arith,halo_zero", [
    (GATHER_CASES[0], abi.ARITH_BF16, True),    # full tiles: active list emitted by the epilogue
    (((3, 64, 32, 1, 6), (3, 32, 16, 64, 1), 8), abi.ARITH_BF16, True),
    (GATHER_CASES[0], abi.ARITH_BF16, False),   # halo not guaranteed zero: ordered compaction
    (GATHER_CASES[5], abi.ARITH_BF16, True),    # ragged tiles: ordered compaction
    (GATHER_CASES[0], abi.ARITH_EXACT, True),   # tiled EXACT kernel: ordered compaction
])
def test_recv_post_fused_active_list(case, arith, halo_zero):
    """Gather + LCA update + published active list in one call == the same
    update followed by pvhip_active_compact, bit for bit (V, A, entries, counts,
    position segments)."""
    rng = np.random.default_rng(27)
    hs, os_, ht, ot, p, W, A = gather_setup(case, rng)
    n = ht.nbatch * ht.num_res
    V0 = (rng.standard_normal(n) * 0.1).astype(np.float32)
    tr = R.make_transfer(0.05, 0.0, FMAX, 0.05, 0.0)
    A0 = O.transfer(ot, V0, 0.05, 0.0, FMAX, 0.05, 0.0)
    d = R.PostsynapticPerspectiveGPUDelivery(hs, ht, p, p, dev(W), arith=arith)
    upd = R.HyPerLCAUpdate(ht, tr, 100.0, 1.0, True, activity_from_v=True, device=DEV)
    Va, Aa = dev(V0), dev(A0)
    d.deliver(dev(A), None, fused=upd.args(None, Va, Aa, num_channels=1))
    ref = R.ActiveList(ht, DEV)
    ref.compact(Aa)
    Vb, Ab = dev(V0), dev(A0)
    act = R.ActiveList(ht, DEV)
    act.entries.fill_(-7)
    d.deliver(dev(A), None, fused=upd.args(None, Vb, Ab, num_channels=1), active=act,
              halo_is_zero=halo_zero)
    assert same(host(Va), host(Vb)) and same(host(Aa), host(Ab))
    e_ref, n_ref = ref.host_entries()
    e_got, n_got = act.host_entries()
    assert (n_ref == n_got).all() and n_ref.sum() > 0
    for b in range(ht.nbatch):
        assert np.array_equal(e_ref[b, :n_ref[b]], e_got[b, :n_got[b]])
    assert np.array_equal(host(ref.pos_start), host(act.pos_start))


@pytest.mark.parametrize("case", [
    ((2, 16, 16, 40, 2), (2, 32, 32, 3, 3), 8, True),      # V1 -> Residual (config[0] shape)
    ((2, 16, 16, 96, 2), (2, 64, 64, 1, 6), 16, True),     # config[2] / config[4] geometry
    ((1, 8, 8, 70, 1), (1, 8, 8, 2, 1), 3, False),         # one-to-one, ragged feature tile
    ((1, 12, 10, 16, 2), (1, 24, 20, 5, 3), 6, True),      # non-square, even patch
])
@pytest.mark.parametrize("arith", [abi.ARITH_FP32, abi.ARITH_BF16])
def test_hebbian_dw_mfma(case, arith):
    """MFMA correlation form of the Hebbian dW against the oracle: dW to
    rounding (FP32) or to the BF16 input bound, activation counts exactly.
    Both halos hold nonzero values: the shrunken patches must ignore the post
    halo and include every pre extended neuron."""
    rng = np.random.default_rng(32)
    (hp, op), (hq, oq), p, norm = both(case[0]), both(case[1]), case[2], case[3]
    a = rng.standard_normal(hp.nbatch * hp.num_ext).astype(np.float32)
    a[rng.random(a.shape) < 0.7] = 0
    r = rng.standard_normal(hq.nbatch * hq.num_ext).astype(np.float32)
    dw, cnt = O.hebbian_dw(op, oq, p, p, oq.nf, 0.25, a, r, counts=norm)
    hb = R.HebbianUpdater(hp, hq, p, p, 0.25, normalize_dw=norm, device=DEV)
    assert hb.windows is not None
    hb.update_dw(dev(a), dev(r), arith=arith)
    got = host(hb.dw)
    scale = np.abs(dw).max()
    if arith == abi.ARITH_FP32:
        assert np.abs(got - dw).max() <= 2e-5 * scale
    else:
        assert np.sqrt(((got - dw) ** 2).mean()) <= 1e-2 * np.sqrt((dw ** 2).mean())
        assert np.abs(got - dw).max() <= 3e-2 * scale
    if norm:
        assert (host(hb.counts) == cnt).all()
    # dw accumulates across calls when not cleared
    hb.update_dw(dev(a), dev(r), clear=False, arith=arith)
    assert np.abs(host(hb.dw) - 2 * got).max() <= 1e-6 * scale


def test_autotuner_shapes_keep_results(tmp_path):
    """The launch-shape autotuner: for the channel GEMM (FP32 and BF16) and the
    Hebbian MFMA split-K it measures every candidate, applies the fastest and
    caches it; the tuned launch gives the oracle's result within each form's
    bar and a second tuner reads the cache without measuring."""
    from openpv_amd.autotune import LaunchTuner
    rng = np.random.default_rng(34)
    cache = str(tmp_path / "shapes.json")
    case = ((2, 32, 32, 512, 2), (2, 128, 128, 1, 0), 4)
    hs, os_, ht, ot, p, W, A = gather_setup(case, rng, density=0.1)
    ref = O.recv_post(os_, ot, p, p, A, W).astype(np.float64)
    scale = np.abs(ref).max()
    for arith, tol in ((abi.ARITH_FP32, 1e-5), (abi.ARITH_BF16, 1e-2)):
        d = R.PostsynapticPerspectiveGPUDelivery(hs, ht, p, p, dev(W), arith=arith)
        g = torch.zeros(ht.nbatch * ht.num_res, device=DEV)
        tuner = LaunchTuner(cache=cache)
        shape = tuner.tune_recv_post(d, dev(A), g)
        chunks = 512 // (32 if arith == abi.ARITH_BF16 else 16)
        assert shape[0] in (64, 128, 256) and 1 <= shape[1] <= min(8, chunks)
        assert (d.cgemm_tile, d.cgemm_splits) == tuple(shape) and len(tuner.log) == 1
        print("arith", arith, "tuned", shape, {k: round(v, 4) for k, v in tuner.log[0][1].items()}, flush=True)
        g.zero_()
        d.deliver(dev(A), g)
        assert np.abs(host(g) - ref).max() <= tol * scale
        d2 = R.PostsynapticPerspectiveGPUDelivery(hs, ht, p, p, dev(W), arith=arith)
        t2 = LaunchTuner(cache=cache)
        assert tuple(t2.tune_recv_post(d2, dev(A), g)) == tuple(shape) and t2.log == []
    # a pre-perspective delivery: dense MFMA route or ordered sparse scatter, both to the FP32 bar
    (hp, op), (hq, oq), p, dens = both(SCATTER_CASES[2][0]), both(SCATTER_CASES[2][1]), SCATTER_CASES[2][2], \
        SCATTER_CASES[2][3]
    kx, ky, kf = O.num_kernels(op, oq)
    Ws = (rng.standard_normal((kx * ky * kf, p, p, oq.nf)) * 0.1).astype(np.float32)
    As = ext_random(hp, rng, dens)
    ent, cnt = O.active_indices(op, As)
    sref = O.recv_pre(op, oq, p, p, Ws, entries=ent, num_active=cnt, num_threads=1).astype(np.float64)
    sd = R.PresynapticPerspectiveGPUDelivery(hp, hq, p, p, dev(Ws), arith=abi.ARITH_FP32)
    act = R.ActiveList(hp, DEV)
    act.compact(dev(As))
    gs = torch.zeros(hq.nbatch * hq.num_res, device=DEV)
    mode = LaunchTuner(cache=None).tune_recv_pre(sd, gs, act, dev(As))
    print("scatter mode", mode, flush=True)
    assert mode in ("route", "ordered") and sd.use_route == (mode == "route")
    gs.zero_()
    sd.deliver(gs, active=act, pre_activity=dev(As))
    assert np.abs(host(gs) - sref).max() <= 1e-5 * np.abs(sref).max()
    # Hebbian split-K rows
    (hp, op), (hq, oq), p = both((2, 16, 16, 96, 2)), both((2, 64, 64, 1, 6)), 16
    a = rng.standard_normal(hp.nbatch * hp.num_ext).astype(np.float32)
    a[rng.random(a.shape) < 0.7] = 0
    r = rng.standard_normal(hq.nbatch * hq.num_ext).astype(np.float32)
    dw, _ = O.hebbian_dw(op, oq, p, p, oq.nf, 0.25, a, r, counts=False)
    hb = R.HebbianUpdater(hp, hq, p, p, 0.25, normalize_dw=False, device=DEV)
    shape = LaunchTuner(cache=None).tune_hebbian(hb, dev(a), dev(r), abi.ARITH_FP32)
    assert hb.rows_per_chunk == shape[0]
    hb.update_dw(dev(a), dev(r), arith=abi.ARITH_FP32)
    assert np.abs(host(hb.dw) - dw).max() <= 2e-5 * np.abs(dw).max()


def test_config1_delivery_full_size():
    """Delivery drop-in at its full size (pre 256x256x16 halo 6 -> post 128x128x256,
    6x6x256 pre-perspective / 12x12x16 post-perspective patches): the EXACT gather
    over dense N(0,1) activity and the EXACT scatter over a 2 % Exp(1) active list
    are bit-identical to the oracle; the BF16 forms stay within their bounds."""
    import ctypes as C
    rng = np.random.default_rng(33)
    (hp, op), (hq, oq) = both((1, 256, 256, 16, 6)), both((1, 128, 128, 256, 0))
    nkx, nky, nkf = O.num_kernels(op, oq)
    Wpre = (rng.standard_normal((nkx * nky * nkf, 6, 6, 256)) * 0.05).astype(np.float32)
    Wpost = np.zeros((256, 12, 12, 16), np.float32)
    assert abi.host().pvhost_transpose_weights(C.byref(hp), C.byref(hq), 6, 6, 256,
                                               Wpre.ctypes.data_as(C.c_void_p),
                                               Wpost.ctypes.data_as(C.c_void_p)) == 0
    A = ext_random(hp, rng)
    print("config1 gather: oracle", flush=True)
    ref = O.recv_post(op, oq, 12, 12, A, Wpost)
    g = torch.zeros(hq.num_res, device=DEV)
    R.PostsynapticPerspectiveGPUDelivery(hp, hq, 12, 12, dev(Wpost)).deliver(dev(A), g)
    assert same(host(g), ref)
    R.PostsynapticPerspectiveGPUDelivery(hp, hq, 12, 12, dev(Wpost), arith=abi.ARITH_BF16).deliver(dev(A), g)
    err = host(g).astype(np.float64) - ref
    assert np.sqrt((err ** 2).mean()) <= 1e-2 * np.sqrt((ref.astype(np.float64) ** 2).mean())
    # sparse scatter
    As = np.zeros((1, hp.ny_ext, hp.nx_ext, 16), np.float32)
    v = np.zeros(256 * 256 * 16, np.float32)
    idx = rng.choice(v.size, size=v.size // 50, replace=False)
    v[idx] = rng.exponential(1.0, idx.size).astype(np.float32)
    As[:, 6:262, 6:262] = v.reshape(1, 256, 256, 16)
    As = As.ravel()
    ent, cnt = O.active_indices(op, As)
    print("config1 scatter: oracle", flush=True)
    ref = O.recv_pre(op, oq, 6, 6, Wpre, entries=ent, num_active=cnt, num_threads=1)
    act = R.ActiveList(hp, DEV)
    dA = dev(As)
    act.compact(dA)
    d = R.PresynapticPerspectiveGPUDelivery(hp, hq, 6, 6, dev(Wpre))
    g.zero_()
    d.deliver(g, active=act)
    assert same(host(g), ref)
    db = R.PresynapticPerspectiveGPUDelivery(hp, hq, 6, 6, dev(Wpre), arith=abi.ARITH_BF16)
    db.deliver(g, active=act, pre_activity=dA)
    err = host(g).astype(np.float64) - ref
    assert np.sqrt((err ** 2).mean()) <= 1e-2 * np.sqrt((ref.astype(np.float64) ** 2).mean())


# ---------------------------------------------------------------- adjoint
def _adjoint_setup(T=12, checkpoint=5, arith=abi.ARITH_FP32, seed=40):
    from openpv_amd.adjoint import LCAAdjoint
    rng = np.random.default_rng(seed)
    nb, n, nf, nv, nfv, p = 2, 16, 3, 8, 16, 8
    x = rng.standard_normal((nb, n, n, nf)).astype(np.float32)
    x -= x.mean(axis=(1, 2, 3), keepdims=True)
    W = rng.standard_normal((nfv, p, p, nf))
    W /= np.sqrt((W.reshape(nfv, -1) ** 2).sum(1)).reshape(-1, 1, 1, 1)
    W = W.astype(np.float32)
    adj = LCAAdjoint(nb, n, n, nf, nv, nv, nfv, p, p, VThresh=0.1, timeConstantTau=10.0,
                     arith=arith, checkpoint=checkpoint, device=DEV)
    return adj, x, W, T


def _energy_f64(adj, x, W, T):
    """The same unrolled loop in float64 torch (the FD checker): gather as conv2d on
    the residual's extended grid, scatter as its adjoint, soft threshold."""
    S, Tl = adj.res_loc, adj.v1_loc
    oS = O.Loc.make(S.nbatch, S.nx, S.ny, S.nf, S.halo_tuple())
    oT = O.Loc.make(Tl.nbatch, Tl.nx, Tl.ny, Tl.nf, Tl.halo_tuple())
    p = adj.nxp
    ux = O.patch_axis(oT, oS, p, p, 0)["unshrunken"][Tl.halo.lt:Tl.halo.lt + Tl.nx]
    uy = O.patch_axis(oT, oS, p, p, 1)["unshrunken"][Tl.halo.up:Tl.halo.up + Tl.ny]
    stride = int(ux[1] - ux[0])
    w = torch.tensor(W, dtype=torch.float64).permute(0, 3, 1, 2)
    nb, nf = S.nbatch, S.nf

    def gather(r):
        ext = torch.zeros((nb, nf, S.ny_ext, S.nx_ext), dtype=torch.float64)
        ext[:, :, S.halo.up:S.halo.up + S.ny, S.halo.lt:S.halo.lt + S.nx] = r
        return torch.nn.functional.conv2d(ext[:, :, uy[0]:, ux[0]:], w, stride=stride)[:, :, :Tl.ny, :Tl.nx]

    def scatter(a):
        r0 = torch.zeros((nb, nf, S.ny, S.nx), dtype=torch.float64, requires_grad=True)
        (g,) = torch.autograd.grad((gather(r0) * a).sum(), r0)
        return g

    e = float(adj.lca.exp_tau[0].item())
    vt = 0.1
    xx = torch.tensor(x, dtype=torch.float64).permute(0, 3, 1, 2)
    V = torch.zeros((nb, Tl.nf, Tl.ny, Tl.nx), dtype=torch.float64)
    A = torch.zeros_like(V)
    rec = torch.zeros_like(xx)
    for _ in range(T):
        G = gather(xx - rec)
        V = e * V + (1 - e) * (G + A)
        A = torch.where(V >= vt, V - vt, torch.zeros_like(V))
        rec = scatter(A)
    return float(0.5 * ((xx - rec) ** 2).sum() + float(np.float32(vt)) * A.abs().sum())


def test_adjoint_matches_fp64_finite_differences():
    """dE/dW and dE/dx of 12 unrolled LCA steps (checkpoint every 5) against
    central finite differences of the float64 loop along random directions."""
    adj, x, W, T = _adjoint_setup()
    E, ck = adj.forward(x, W, T)
    gW, gx = adj.backward(x, T, ck)
    gW, gx = host(gW).astype(np.float64), host(gx).astype(np.float64)
    e32 = float(E.sum().item())
    e64 = _energy_f64(adj, x, W, T)
    assert abs(e32 - e64) <= 1e-4 * abs(e64)
    rng = np.random.default_rng(41)
    eps = 1e-4
    for _ in range(2):
        d = rng.standard_normal(W.shape)
        fd = (_energy_f64(adj, x, W + eps * d, T) - _energy_f64(adj, x, W - eps * d, T)) / (2 * eps)
        ad = float((gW * d.ravel()).sum())
        print("dW direction: fd", fd, "adjoint", ad, flush=True)
        assert abs(fd - ad) <= 2e-3 * abs(fd) + 1e-6
        dx = rng.standard_normal(x.shape)
        fdx = (_energy_f64(adj, x + eps * dx, W, T) - _energy_f64(adj, x - eps * dx, W, T)) / (2 * eps)
        adx = float((gx * dx.ravel()).sum())
        print("dx direction: fd", fdx, "adjoint", adx, flush=True)
        assert abs(fdx - adx) <= 2e-3 * abs(fdx) + 1e-6


def test_adjoint_checkpoint_interval_does_not_change_gradients():
    outs = []
    for k in (1, 5, 100):
        adj, x, W, T = _adjoint_setup(checkpoint=k)
        E, ck = adj.forward(x, W, T)
        gW, gx = adj.backward(x, T, ck)
        outs.append((host(gW), host(gx)))
    for gW, gx in outs[1:]:
        assert same(gW, outs[0][0]) and same(gx, outs[0][1])


def test_adjoint_fixed_activity_equals_hebbian_dw():
    """With A_T held fixed the weight gradient is the reconstruction term,
    corr(A_T, -(x - Rec_T)) = minus HebbianUpdater's dW (dWMax 1, normalizeDw off),
    checked against the oracle's Hebbian restatement (EXACT forms: bit for bit)."""
    adj, x, W, T = _adjoint_setup(T=1, arith=abi.ARITH_EXACT)
    E, ck = adj.forward(x, W, T)
    gW, _ = adj.backward(x, T, ck, fixed_activity=True)
    S, Tl = adj.res_loc, adj.v1_loc
    oS = O.Loc.make(S.nbatch, S.nx, S.ny, S.nf, S.halo_tuple())
    oT = O.Loc.make(Tl.nbatch, Tl.nx, Tl.ny, Tl.nf, Tl.halo_tuple())
    A_T = host(adj.A)
    negR = -host(adj.R_ext)
    dw, _ = O.hebbian_dw(oT, oS, adj.nxp, adj.nyp, S.nf, 1.0, A_T, negR, counts=False)
    assert np.abs(A_T).sum() > 0
    assert same(host(gW), dw)


def test_adjoint_autograd_function():
    from openpv_amd.adjoint import LCAEnergyFunction
    adj, x, W, T = _adjoint_setup()
    xt = torch.tensor(x, device=DEV, requires_grad=True)
    Wt = torch.tensor(W, device=DEV, requires_grad=True)
    E = LCAEnergyFunction.apply(xt, Wt, adj, T)
    (2.0 * E).backward()
    E2, ck = adj.forward(x, W, T)
    gW, gx = adj.backward(x, T, ck)
    assert np.allclose(host(Wt.grad).ravel(), 2.0 * host(gW), rtol=1e-6, atol=1e-7)
    assert np.allclose(host(xt.grad).ravel(), 2.0 * host(gx), rtol=1e-6, atol=1e-7)


# ------------------------------------------------- end-to-end: reference fixture
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "gpulca")


def _gpulca_run(arith, fused=True, steps=40):
    """The CPU column of the reference's GPULCATest (tests/GPULCATest): Input
    128x64x1 (two frames, displayPeriod 20) -> Residual <-> V1 64x32x24 (18x18
    patches, firm threshold VThresh 0.025 / VWidth 0.025, tau 100, V from file)
    -> Reconstruction, weights from file L2-normalised on initialisation,
    adaptive time scales (baseMax 0.55, baseMin 0.5, tauFactor 0.1, growth
    0.01) on the total energy 0.5|r|^2 + VThresh*firm(a); 40 steps."""
    from openpv_amd import pvp
    from openpv_amd.lca import LCAColumn, run_column
    from openpv_amd.timescale import AdaptiveTimeScaleController
    _, _, frames = pvp.read_activity(os.path.join(GOLDEN, "inputImages.pvp"))
    _, _, V0 = pvp.read_activity(os.path.join(GOLDEN, "initialV1.pvp"))
    _, _, W = pvp.read_weights(os.path.join(GOLDEN, "weights.pvp"))
    col = LCAColumn(frames[:1], W[0], 64, 32, 24, 18, 18, VThresh=0.025, AMin=0.0, AMax=FMAX,
                    AShift=0.0, VWidth=0.025, timeConstantTau=100.0, dt=1.0, arith=arith,
                    fused=fused, V0=V0[0], normalize="exact", device=DEV)
    ctrl = AdaptiveTimeScaleController(1, 0.55, 0.5, 0.1, 0.01)
    recon = []
    run_column(col, steps, frames=[f[None] for f in frames], display_period=20, adaptive=ctrl,
               energy_cost="firm", on_step=lambda t, c, e: recon.append(c.rec_A.cpu().numpy().copy()))
    return np.stack(recon).reshape(steps, 64, 128, 1)


def _gpulca_oracle(steps, momentum=None):
    """Reconstruction of each step of the column on the oracle."""
    return _gpulca_oracle_run(steps, momentum)[0]


def _gpulca_oracle_run(steps, momentum=None):
    """The same column on the oracle alone (its deliveries, updates, cost sums,
    time-scale controller and input trigger; lcatest_column.oracle_run):
    (Reconstruction of each step, total energy of each step).  momentum: V1 as
    a MomentumLCALayer with that LCAMomentumRate.  Cached: callers must not
    modify the arrays."""
    import lcatest_column
    return lcatest_column.oracle_run(steps, momentum)


def test_gpulca_inputs_adaptive_timescales_exact():
    """The reference's GPULCATest inputs (images, initial V1, weights from its
    input/ folder) through the whole column with adaptive time scales: EXACT
    forms on the GPU == the oracle, bit for bit, for 25 steps (an input change
    and its timescale reset at step 21 included)."""
    steps = 25
    ref = _gpulca_oracle(steps)
    got = _gpulca_run(abi.ARITH_EXACT, steps=steps)
    for t in range(steps):
        assert same(got[t], ref[t]), "step %d differs (max %g)" % (t + 1, np.abs(got[t] - ref[t]).max())


@pytest.mark.parametrize("arith,tol", [(abi.ARITH_FP32, 1e-5), (abi.ARITH_BF16, 5e-3)])
def test_gpulca_inputs_adaptive_timescales_fast_forms(arith, tol):
    steps = 25
    ref = _gpulca_oracle(steps)
    got = _gpulca_run(arith, steps=steps)
    err = np.abs(got - ref).max()
    print("max abs error", err, "scale", np.abs(ref).max(), flush=True)
    assert err <= tol * np.abs(ref).max()


# ------------------------------------------------- the C++ host: pvrun on a .params file
def _pvrun(params, outdir, steps):
    """Runs pvrun on `params` in the fixture directory (the params' relative
    input paths resolve there); returns the completed process."""
    print("pvrun", params, "steps", steps, flush=True)
    r = subprocess.run([abi.pvrun_path(), "-p", str(params), "-o", str(outdir), "--steps", str(steps)],
                       cwd=GOLDEN, capture_output=True, text=True, timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    return r


def _energy_file(path):
    return np.array([float(line.split(",")[2]) for line in open(path) if line.strip()])


def test_pvrun_two_columns_from_params(tmp_path):
    """pvrun builds both columns of the committed params file (a host column with
    receiveGpu = false and a device column with receiveGpu = true on the gather)
    and runs 25 steps: the host column's reconstructions equal the oracle's bit
    for bit and its energy file matches the oracle's energies; the device column
    (FP32) is within 1e-5 of the oracle's scale; the Difference probe's verdict
    agrees with the two reconstructions."""
    from openpv_amd import pvp
    steps = 25
    r = _pvrun("LCAColumn.params", tmp_path, steps)
    ref, energy = _gpulca_oracle_run(steps)
    _, times, host = pvp.read_activity(str(tmp_path / "ReconHost.pvp"))
    _, _, dev = pvp.read_activity(str(tmp_path / "ReconDevice.pvp"))
    assert host.shape == (steps, 64, 128, 1) and list(times) == list(range(1, steps + 1))
    for t in range(steps):
        assert same(host[t], ref[t]), "host column, step %d (max %g)" % (t + 1, np.abs(host[t] - ref[t]).max())
    e = _energy_file(tmp_path / "EnergyHost.txt")
    assert np.allclose(e, energy, rtol=1e-8, atol=1e-9), (e[:4], energy[:4])
    err = np.abs(dev - ref).max()
    diff = np.abs(dev - host).max()
    print("device column: max |dev - oracle| %g (scale %g); max |dev - host| %g" % (err, np.abs(ref).max(), diff))
    assert err <= 1e-5 * np.abs(ref).max()
    # RequireAllZeroActivityProbe at nnzThreshold 2e-7 on Difference = ReconDevice - ReconHost
    assert (r.returncode == 0) == (diff <= np.float32(2e-7)), r.stderr[-500:]


def test_pvrun_momentum_lca_column(tmp_path):
    """The host column's V1 as a MomentumLCALayer (LCAMomentumRate 0.4, the
    rate of the reference's MomentumLCATest): pvrun's reconstructions equal the
    oracle's momentum column bit for bit over 25 steps, and differ from the
    plain LCA column's."""
    from openpv_amd import pvp
    steps = 25
    text = open(os.path.join(GOLDEN, "LCAColumn.params")).read()
    old = 'HyPerLCALayer "SparseCodeHost" = {'
    assert text.count(old) == 1
    params = tmp_path / "momentum.params"
    params.write_text(text.replace(old, 'MomentumLCALayer "SparseCodeHost" = {\n    LCAMomentumRate = 0.4;'))
    r = _pvrun(params, tmp_path, steps)
    assert r.returncode in (0, 1), r.stderr[-500:]   # 1: the host and device columns now differ
    ref = _gpulca_oracle(steps, 0.4)
    _, _, host = pvp.read_activity(str(tmp_path / "ReconHost.pvp"))
    for t in range(steps):
        assert same(host[t], ref[t]), "step %d (max %g)" % (t + 1, np.abs(host[t] - ref[t]).max())
    assert not same(host, _gpulca_oracle(steps))


@pytest.mark.parametrize("form,tol", [("bf16", 5e-3), ("fp8", 5e-2)])
def test_pvrun_low_precision_device_column(tmp_path, form, tol):
    """gpuArithmetic = "bf16" / "fp8": the device column's gather runs on the
    BF16 or FP8 MFMA path (FP32 state; the FP8 source scale comes from the
    input range); the host column is unchanged (bit-exact)."""
    from openpv_amd import pvp
    steps = 25
    text = open(os.path.join(GOLDEN, "LCAColumn.params")).read()
    assert 'gpuArithmetic = "fp32";' in text
    params = tmp_path / (form + ".params")
    params.write_text(text.replace('gpuArithmetic = "fp32";', 'gpuArithmetic = "%s";' % form))
    r = _pvrun(params, tmp_path, steps)
    assert r.returncode in (0, 1), r.stderr[-500:]   # 1: the 2e-7 comparison probe fails, as expected
    ref = _gpulca_oracle(steps)
    _, _, host = pvp.read_activity(str(tmp_path / "ReconHost.pvp"))
    _, _, dev = pvp.read_activity(str(tmp_path / "ReconDevice.pvp"))
    assert same(host, ref)
    err = np.abs(dev - ref).max()
    print("%s device column: max |dev - oracle| %g (scale %g)" % (form, err, np.abs(ref).max()))
    assert err <= tol * np.abs(ref).max()


def test_cost_sum_firm_threshold_matches_reference_fixture():
    """pvhip_cost_sum_f64 (firm threshold, VThresh 0.5, VWidth 0.3) over the
    reference FirmThresholdCostTest inputs placed in an extended buffer whose
    halo holds junk: per batch element equal to the sum of the reference's
    expected per-neuron costs (FP64 vs the reference layer's float values)."""
    from openpv_amd import pvp
    g = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "firmthresh")
    _, _, x = pvp.read_activity(os.path.join(g, "firmthreshinput.pvp"))
    _, _, y = pvp.read_activity(os.path.join(g, "correctoutput.pvp"))
    h = 2
    ext = np.full((2, 16 + 2 * h, 16 + 2 * h, 3), 0.7, np.float32)   # junk in the halo
    ext[:, h:h + 16, h:h + 16, :] = x[:2]
    loc = abi.Loc.make(2, 16, 16, 3, halo=h)
    out = R.cost_sum_f64(loc, dev(ext.ravel()), True, 2, 0.5, 0.3)
    got = out.cpu().numpy()
    want = y[:2].astype(np.float64).reshape(2, -1).sum(1)
    np.testing.assert_allclose(got, want, rtol=1e-6)
    # L0 count above a threshold (kind 3), the all-zero probe's test
    cnt = R.cost_sum_f64(loc, dev(ext.ravel()), True, 3, 0.5).cpu().numpy()
    assert list(cnt) == [float((np.abs(x[b]) > 0.5).sum()) for b in range(2)]


@pytest.mark.parametrize("zero_patch", [None, 5])
def test_normalize_l2_exact_bit_exact(zero_patch):
    """pvhip_normalize_l2 (the reference's float sum of squares, sqrtf and
    multiply per data patch) equals the oracle bit for bit on 24 patches of
    18x18 and on 512 patches of 256, including a degenerate patch that ends
    the loop."""
    rng = np.random.default_rng(31)
    for npatch, psize in ((24, 324), (512, 256)):
        w = rng.standard_normal((npatch, psize)).astype(np.float32)
        if zero_patch is not None:
            w[zero_patch] = 0.0
        ref = w.copy()
        O.lib().or_normalize_l2(ref.ravel(), npatch, psize, 0.75, 0.0)
        d = dev(w.ravel())
        abi.check(abi.hip().pvhip_normalize_l2(C.c_void_p(d.data_ptr()), npatch, psize, 0.75, 0.0, None),
                  "pvhip_normalize_l2")
        torch.cuda.synchronize()
        assert same(d.cpu().numpy(), ref.ravel())
