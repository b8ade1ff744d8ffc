"""The chunked BF16 activity copy (pvhip_activity_to_bf16c layout) and the BF16
channel GEMM that reads it.

The copy holds each activity rounded to BF16 (round to nearest even) in
[nbatch][nf/32][nyExt][nxExt][32] order.  Layer updates write it beside A, and
the BF16 channel GEMM reads it instead of converting the FP32 activity while
staging; both convert the same values the same way, so every result must be
bit-identical to the path without the copy.  The rounding itself is pinned by
a numpy restatement of round-to-nearest-even on the float bits.
"""
import numpy as np
import pytest
import torch

from openpv_amd import abi
from openpv_amd import runtime as R
from openpv_amd.lca import LCAColumn

import oracle as O

pytestmark = pytest.mark.gpu

DEV = "cuda"
FMAX = float(np.finfo(np.float32).max)


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a)).to(DEV, dtype=torch.float32).contiguous()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def bf16_bits(t):
    torch.cuda.synchronize()
    return t.view(torch.int16).cpu().numpy().view(np.uint16)


def bf16_rne(x):
    """float32 -> BF16 bits, round to nearest even (finite inputs)."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def chunked(loc, A):
    """numpy layout of the copy: [nb][nyE][nxE][nf] -> [nb][nf/32][nyE][nxE][32]."""
    a = np.asarray(A).reshape(loc.nbatch, loc.ny_ext, loc.nx_ext, loc.nf // 32, 32)
    return np.ascontiguousarray(a.transpose(0, 3, 1, 2, 4)).ravel()


def ext_random(loc, rng, density=1.0, scale=1.0):
    A = np.zeros((loc.nbatch, loc.ny_ext, loc.nx_ext, loc.nf), np.float32)
    v = (rng.standard_normal((loc.nbatch, loc.ny, loc.nx, loc.nf)) * scale).astype(np.float32)
    if density < 1.0:
        v[rng.random(v.shape) >= density] = 0
    A[:, loc.halo.up:loc.halo.up + loc.ny, loc.halo.lt:loc.halo.lt + loc.nx] = v
    return A.ravel()


def test_round_to_nearest_even_restatement():
    """The numpy rounding used as the check: ties go to the even BF16, and the
    values nearest a BF16 round to it (CPU only, no device call)."""
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -8 + 2 ** -20, -2.5, 0.0],
                 np.float32)
    got = bf16_rne(x)
    want = [0x3F80, 0x3F80, 0x3F82, 0x3F81, 0xC020, 0x0000]
    assert got.tolist() == want


def test_activity_to_bf16c_layout_and_rounding():
    rng = np.random.default_rng(61)
    loc = abi.Loc.make(2, 7, 5, 64, halo=(2, 1, 3, 2))
    A = ext_random(loc, rng).reshape(-1)
    # exact ties and the halo's values (the whole buffer is converted)
    A[:64] = (1.0 + np.arange(64) * 2 ** -8).astype(np.float32)
    A[-1] = np.float32(-3.0 - 2 ** -7)
    out = R.bf16c_buffer(loc, DEV)
    out.fill_(7.0)
    R.activity_to_bf16c(loc, dev(A), out)
    assert np.array_equal(bf16_bits(out), bf16_rne(chunked(loc, A)))


def test_activity_to_bf16c_refuses_bad_nf():
    loc = abi.Loc.make(1, 4, 4, 48, halo=0)
    out = torch.zeros(loc.nbatch * loc.num_ext, dtype=torch.bfloat16, device=DEV)
    with pytest.raises(abi.PVHIPError, match="multiple of 32"):
        R.activity_to_bf16c(loc, torch.zeros(loc.num_ext, device=DEV), out)


@pytest.mark.parametrize("vertices", [False, True])
def test_lca_update_writes_bf16c(vertices):
    """pvhip_lca_update with A_bf16c: V and A as without it, the copy equals the
    conversion of the new A, and the halo of the copy is left alone."""
    rng = np.random.default_rng(62)
    loc = abi.Loc.make(2, 6, 5, 64, halo=2)
    n = loc.nbatch * loc.num_res
    G = (rng.standard_normal(n) * 0.2).astype(np.float32)
    V0 = (rng.standard_normal(n) * 0.2).astype(np.float32)
    tr = R.make_transfer(0.1, 0.0, FMAX, 0.1, 0.05, vertices)
    upd = R.HyPerLCAUpdate(loc, tr, 50.0, 1.0, True, device=DEV)
    res = []
    for with_copy in (False, True):
        Va = dev(V0)
        Aa = torch.zeros(loc.nbatch * loc.num_ext, device=DEV)
        A16 = R.bf16c_buffer(loc, DEV) if with_copy else None
        upd.update(dev(G), Va, Aa, A_bf16c=A16)
        res.append((host(Va), host(Aa), A16))
    (V1, A1, _), (V2, A2, A16) = res
    assert np.array_equal(V1.view(np.uint32), V2.view(np.uint32))
    assert np.array_equal(A1.view(np.uint32), A2.view(np.uint32))
    assert np.array_equal(bf16_bits(A16), bf16_rne(chunked(loc, A2)))


def test_lca_update_without_fp32_activity():
    """A = NULL with A_bf16c and activity_from_v: V and the copy as with A;
    without the copy (or without activity_from_v) the call is refused."""
    rng = np.random.default_rng(66)
    loc = abi.Loc.make(2, 6, 5, 64, halo=2)
    n = loc.nbatch * loc.num_res
    G = (rng.standard_normal(n) * 0.2).astype(np.float32)
    V0 = (rng.standard_normal(n) * 0.2).astype(np.float32)
    tr = R.make_transfer(0.1, 0.0, FMAX, 0.1, 0.0)
    upd = R.HyPerLCAUpdate(loc, tr, 50.0, 1.0, True, activity_from_v=True, device=DEV)
    out = []
    for keep_a in (True, False):
        Va = dev(V0)
        Aa = torch.zeros(loc.nbatch * loc.num_ext, device=DEV)
        R.ann_activity(loc, tr, Va, Aa)
        A16 = R.bf16c_buffer(loc, DEV)
        upd.update(dev(G), Va, Aa if keep_a else None, A_bf16c=A16)
        out.append((host(Va), bf16_bits(A16), host(Aa)))
    assert np.array_equal(out[0][0].view(np.uint32), out[1][0].view(np.uint32))
    assert np.array_equal(out[0][1], out[1][1])
    # the FP32 activity derived afterwards from V equals the one the update wrote
    Ad = torch.zeros(loc.nbatch * loc.num_ext, device=DEV)
    R.ann_activity(loc, tr, dev(out[1][0]), Ad)
    assert np.array_equal(host(Ad).view(np.uint32), out[0][2].view(np.uint32))
    with pytest.raises(abi.PVHIPError, match="NULL A"):
        upd.update(dev(G), dev(V0), None)
    upd2 = R.HyPerLCAUpdate(loc, tr, 50.0, 1.0, True, activity_from_v=False, device=DEV)
    with pytest.raises(abi.PVHIPError, match="NULL A"):
        upd2.update(dev(G), dev(V0), None, A_bf16c=R.bf16c_buffer(loc, DEV))


GATHER_CASES = [
    # source, target, post-perspective patch, arith
    ((2, 32, 32, 3, 3), (2, 16, 16, 128, 2), 8, abi.ARITH_BF16),    # fused MFMA epilogue
    ((2, 32, 32, 3, 3), (2, 16, 16, 128, 2), 8, abi.ARITH_FP8),
    ((2, 32, 32, 3, 3), (2, 16, 16, 128, 2), 8, abi.ARITH_EXACT),   # tiled kernel, per-neuron update
    ((1, 20, 12, 2, (2, 3, 3, 2)), (1, 10, 6, 64, 1), 6, abi.ARITH_BF16),  # ragged tiles
]


@pytest.mark.parametrize("case", GATHER_CASES)
def test_fused_gather_writes_bf16c(case):
    """Gather + fused LCA update with A_bf16c: V and A bit-identical to the
    launch without it; the copy equals the conversion of the new A."""
    rng = np.random.default_rng(63)
    src, tgt, p, arith = case
    hs, ht = abi.Loc.make(*src[:4], halo=src[4]), abi.Loc.make(*tgt[:4], halo=tgt[4])
    W = (rng.standard_normal((ht.nf, p, p, hs.nf)) * 0.05).astype(np.float32)
    A = ext_random(hs, rng)
    n = ht.nbatch * ht.num_res
    V0 = (rng.standard_normal(n) * 0.1).astype(np.float32)
    tr = R.make_transfer(0.05, 0.0, FMAX, 0.05, 0.0)
    d = R.PostsynapticPerspectiveGPUDelivery(hs, ht, p, p, dev(W), arith=arith)
    if arith == abi.ARITH_FP8:
        d.set_source_scale(R.fp8_scale_for(np.abs(A).max()))
    upd = R.HyPerLCAUpdate(ht, tr, 100.0, 1.0, True, activity_from_v=True, device=DEV)
    out = []
    for with_copy in (False, True):
        Va = dev(V0)
        Aa = torch.zeros(ht.nbatch * ht.num_ext, device=DEV)
        R.ann_activity(ht, tr, Va, Aa)
        A16 = R.bf16c_buffer(ht, DEV) if with_copy else None
        d.deliver(dev(A), None, fused=upd.args(None, Va, Aa, num_channels=1, A_bf16c=A16))
        out.append((host(Va), host(Aa), A16))
    (V1, A1, _), (V2, A2, A16) = out
    assert np.array_equal(V1.view(np.uint32), V2.view(np.uint32))
    assert np.array_equal(A1.view(np.uint32), A2.view(np.uint32))
    assert np.array_equal(bf16_bits(A16), bf16_rne(chunked(ht, A2)))
    # only the copy written (A = NULL): the same V and copy
    Va = dev(V0)
    Aa = torch.zeros(ht.nbatch * ht.num_ext, device=DEV)
    R.ann_activity(ht, tr, Va, Aa)
    B16 = R.bf16c_buffer(ht, DEV)
    d.deliver(dev(A), None, fused=upd.args(None, Va, None, num_channels=1, A_bf16c=B16))
    assert np.array_equal(host(Va).view(np.uint32), V2.view(np.uint32))
    assert np.array_equal(bf16_bits(B16), bf16_bits(A16))


SCATTER_CASES = [
    # pre, post, pre-perspective patch, density
    ((2, 16, 16, 128, 2), (2, 32, 32, 3, 0), 8, 0.05),     # config[0] V1 -> Recon
    ((2, 16, 16, 96, 2), (2, 64, 64, 1, 0), 16, 0.03),     # config[2] geometry
    ((2, 10, 14, 64, 3), (2, 20, 28, 2, 0), 6, 0.2),       # halo wider than the margin
]


@pytest.mark.parametrize("case", SCATTER_CASES)
@pytest.mark.parametrize("shape", [(0, 0), (64, 1), (128, 2), (256, 4)])
def test_scatter_route_reads_bf16c(case, shape):
    """The BF16 dense route of a pre-perspective delivery (transposed channel
    GEMM) reading the chunked copy gives the same GSyn, bit for bit, as the
    route converting the FP32 activity, at every launch shape."""
    rng = np.random.default_rng(64)
    pre, post, p, dens = case
    hp, hq = abi.Loc.make(*pre[:4], halo=pre[4]), abi.Loc.make(*post[:4], halo=post[4])
    op, oq = O.Loc.make(*pre[:4], halo=pre[4]), O.Loc.make(*post[:4], halo=post[4])
    kx, ky, kf = O.num_kernels(op, oq)
    W = (rng.standard_normal((kx * ky * kf, p, p, hq.nf)) * 0.1).astype(np.float32)
    A = ext_random(hp, rng, dens)
    res = []
    for with_copy in (False, True):
        d = R.PresynapticPerspectiveGPUDelivery(hp, hq, p, p, dev(W), arith=abi.ARITH_BF16)
        assert d.transposed is not None and d.transposed.inverse is not None
        d.transposed.set_launch(*shape)
        Ad = dev(A)
        if with_copy:
            A16 = R.bf16c_buffer(hp, DEV)
            R.activity_to_bf16c(hp, Ad, A16)
            assert d.bind_pre_bf16c(A16)
        g = torch.full((hq.nbatch * hq.num_res,), 2.0, device=DEV)
        d.deliver(g, pre_activity=Ad, accumulate=True)
        res.append(host(g))
    assert np.array_equal(res[0].view(np.uint32), res[1].view(np.uint32))
    ent, cnt = O.active_indices(op, A)
    ref = O.recv_pre(op, oq, p, p, W, entries=ent, num_active=cnt, num_threads=1).astype(np.float64)
    err = res[1] - 2.0 - ref
    assert np.sqrt((err ** 2).mean()) <= 1e-2 * np.sqrt((ref ** 2).mean())


@pytest.mark.parametrize("halo", [(3, 2, 3, 2), (3, 3, 2, 2), (2, 2, 3, 2)])
def test_scatter_keeps_the_ordered_form_for_unequal_halos(halo):
    """A pre layer whose halo differs between sides has no transposed route
    (for (3, 3, 2, 2) and (2, 2, 3, 2) its gather would not equal the scatter,
    whose result is the one checked here): the BF16 delivery runs the ordered
    scatter in FP32.  One batch element: for lt != rt the reference's CPU
    scatter steps batches by nx + rt + rt (PresynapticPerspectiveConvolveDelivery.cpp:73),
    which the oracle keeps and the library does not."""
    rng = np.random.default_rng(65)
    pre, post, p = (1, 10, 14, 64, halo), (1, 20, 28, 2, 0), 6
    hp, hq = abi.Loc.make(*pre[:4], halo=pre[4]), abi.Loc.make(*post[:4], halo=post[4])
    op, oq = O.Loc.make(*pre[:4], halo=pre[4]), O.Loc.make(*post[:4], halo=post[4])
    kx, ky, kf = O.num_kernels(op, oq)
    W = (rng.standard_normal((kx * ky * kf, p, p, hq.nf)) * 0.1).astype(np.float32)
    A = ext_random(hp, rng, 0.2)
    d = R.PresynapticPerspectiveGPUDelivery(hp, hq, p, p, dev(W), arith=abi.ARITH_BF16)
    assert d.transposed is None
    active = R.ActiveList(hp, DEV)
    Ad = dev(A)
    active.compact(Ad)
    g = torch.zeros(hq.nbatch * hq.num_res, device=DEV)
    d.deliver(g, active=active, pre_activity=Ad)
    ent, cnt = O.active_indices(op, A)
    ref = O.recv_pre(op, oq, p, p, W, entries=ent, num_active=cnt, num_threads=1).astype(np.float64)
    assert np.abs(host(g) - ref).max() <= 1e-5 * np.abs(ref).max()


def c0_problem(nb=2, seed=5, n=32, nfv=128, p=8):
    rng = np.random.default_rng(seed)
    img = rng.standard_normal((nb, n, n, 3)).astype(np.float32)
    W = rng.standard_normal((nfv, p, p, 3)).astype(np.float32)
    W /= np.sqrt((W.reshape(nfv, -1) ** 2).sum(1)).reshape(nfv, 1, 1, 1)
    return img, W


@pytest.mark.parametrize("arith", [abi.ARITH_BF16, abi.ARITH_FP8])
def test_lca_column_with_bf16c_is_bit_identical(arith):
    """A BF16 / FP8 column whose scatter reads V1's chunked copy (the update
    writes it) against one whose scatter converts V1's FP32 activity: the same
    state after 40 steps, bit for bit, eager and replayed from a hipGraph."""
    img, W = c0_problem()
    cols = []
    for with_copy in (True, False):
        c = LCAColumn(img, W, 16, 16, 128, 8, 8, VThresh=0.1, arith=arith, fused=True,
                      bf16_copy=with_copy, device=DEV)
        assert (c.v1_A16 is not None) == with_copy and c.dense_scatter()
        for _ in range(20):
            c.step()
        s = torch.cuda.Stream()
        torch.cuda.synchronize()
        c.capture(s)
        for _ in range(20):
            c.replay(s)
        s.synchronize()
        cols.append(c)
    a, b = cols
    assert a._a_stale and not b._a_stale   # the copy column wrote no FP32 activity
    for x, y in ((a.v1_V, b.v1_V), (a.v1_activity(), b.v1_activity()), (a.rec_A, b.rec_A),
                 (a.res_A, b.res_A)):
        assert np.array_equal(host(x).view(np.uint32), host(y).view(np.uint32))
    assert np.array_equal(bf16_bits(a.v1_A16), bf16_rne(chunked(a.v1_loc, host(a.v1_activity()))))
    assert np.array_equal(a.num_active(), b.num_active()) and a.num_active().sum() > 0
    ea, eb = host(a.probe_energy()), host(b.probe_energy())
    assert np.array_equal(ea, eb)


@pytest.mark.parametrize("arith", [abi.ARITH_FP32, abi.ARITH_BF16])
@pytest.mark.parametrize("shape", [(0, 0), (128, 1), (128, 4)])
def test_channel_gemm_fused_layer_update(arith, shape):
    """The target layer's update applied in the channel GEMM (fused_layer) gives
    V and A bit-identical to the delivery into GSyn followed by
    pvhip_layer_update, split or not, with an error scale, a threshold transfer
    and a target halo."""
    rng = np.random.default_rng(67)
    pre, post, p = (2, 16, 16, 128, 2), (2, 32, 32, 3, 1), 8
    hp, hq = abi.Loc.make(*pre[:4], halo=pre[4]), abi.Loc.make(*post[:4], halo=post[4])
    op, oq = O.Loc.make(*pre[:4], halo=pre[4]), O.Loc.make(*post[:4], halo=post[4])
    kx, ky, kf = O.num_kernels(op, oq)
    W = (rng.standard_normal((kx * ky * kf, p, p, hq.nf)) * 0.1).astype(np.float32)
    A = dev(ext_random(hp, rng, 0.1))
    tr = R.make_transfer(0.05, 0.0, FMAX, 0.02, 0.0)
    d = R.PresynapticPerspectiveGPUDelivery(hp, hq, p, p, dev(W), arith=arith)
    assert d.transposed is not None and d.transposed.inverse is not None
    d.transposed.set_launch(*shape)
    n, ne = hq.nbatch * hq.num_res, hq.nbatch * hq.num_ext
    out = []
    for fused in (False, True):
        g = torch.full((n,), 5.0, device=DEV)
        V = torch.zeros(n, device=DEV)
        Ae = torch.zeros(ne, device=DEV)
        la = R.layer_update_args(hq, g, V, Ae, tr, channel_coeffs=(1.0,), err_scale=0.5)
        if fused:
            d.deliver(g, pre_activity=A, fused_layer=la)
            assert np.all(host(g) == 5.0)   # GSyn not written
        else:
            d.deliver(g, pre_activity=A)
            R.launch_layer_update(la)
        out.append((host(V), host(Ae)))
    assert np.array_equal(out[0][0].view(np.uint32), out[1][0].view(np.uint32))
    assert np.array_equal(out[0][1].view(np.uint32), out[1][1].view(np.uint32))
    assert np.abs(out[1][0]).max() > 0


def test_fused_layer_refused_outside_the_channel_gemm():
    """fused_layer on a one-patch gather (no channel GEMM) is an error."""
    rng = np.random.default_rng(68)
    hs, ht = abi.Loc.make(2, 32, 32, 3, halo=3), abi.Loc.make(2, 16, 16, 64, halo=0)
    W = (rng.standard_normal((64, 8, 8, 3)) * 0.05).astype(np.float32)
    d = R.PostsynapticPerspectiveGPUDelivery(hs, ht, 8, 8, dev(W), arith=abi.ARITH_BF16)
    V = torch.zeros(ht.nbatch * ht.num_res, device=DEV)
    la = R.layer_update_args(ht, None, V, None, R.make_transfer(-FMAX, -FMAX, FMAX, 0.0, 0.0))
    with pytest.raises(abi.PVHIPError, match="fused_layer"):
        d.deliver(dev(ext_random(hs, rng)), None, fused_layer=la)
