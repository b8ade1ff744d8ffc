"""The dense scatter (pvhip_recv_pre_dense, BF16 and FP32 forms): the
pre-perspective delivery of a dense source, one data patch per source feature,
into a target 4x finer, as a GEMM of source positions by patch offsets on the
matrix cores followed by the overlap-add.

Checks, against the oracle's scatter (the reference CPU loop order):
  * on operands already rounded to BF16 the products are exact in FP32, so
    the kernel differs from the oracle only by the order of FP32 sums:
    within 1e-5 of the largest output;
  * on FP32 operands, within the BF16 bound of the other BF16 forms;
  * the FP32 form within 1e-5 of the largest output (FP32 products; only the
    order of the sums differs from the oracle's);
  * halo activity (every extended position) and the restricted-only mode,
    accumulate, a batch that is not a multiple of 8, several target
    features, the fused target-layer update (the same bits as the delivery
    followed by pvhip_layer_update), and the same bits on every call and in a
    hipGraph replay.
"""
import numpy as np
import pytest
import torch

import oracle as O
from openpv_amd import abi, runtime as R

pytestmark = pytest.mark.gpu
DEV = "cuda"
FMAX = float(np.finfo(np.float32).max)


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a, np.float32)).to(DEV).contiguous()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def bf16_round(x):
    """float32 values rounded to the nearest BF16 (ties to even), as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def activity(loc, rng, density, halo_too):
    A = np.zeros((loc.nbatch, loc.ny_ext, loc.nx_ext, loc.nf), np.float32)
    v = rng.exponential(1.0, A.shape).astype(np.float32)
    v[rng.random(v.shape) >= density] = 0
    if halo_too:
        A[:] = v
    else:
        A[:, loc.halo.up:loc.halo.up + loc.ny, loc.halo.lt:loc.halo.lt + loc.nx] = \
            v[:, loc.halo.up:loc.halo.up + loc.ny, loc.halo.lt:loc.halo.lt + loc.nx]
    return A.ravel()


CASES = [
    # pre, post, pre-perspective patch, density
    ((2, 32, 32, 512, 2), (2, 128, 128, 1, 0), 16, 0.08),   # the batched LCA's reconstruction
    ((3, 8, 12, 128, 1), (3, 32, 48, 1, 0), 16, 0.3),       # batch not a multiple of 8, not square
    ((2, 16, 16, 256, 2), (2, 64, 64, 4, 0), 8, 0.2),       # four target features
    ((9, 6, 5, 128, 3), (9, 24, 20, 1, 0), 16, 0.5),        # halo wider than the margin
]


def _setup(case, seed):
    rng = np.random.default_rng(seed)
    pre, post, p, dens = case
    hp, hq = abi.Loc.make(*pre[:4], halo=pre[4]), abi.Loc.make(*post[:4], halo=post[4])
    op, oq = O.Loc.make(*pre[:4], halo=pre[4]), O.Loc.make(*post[:4], halo=post[4])
    W = (rng.standard_normal((pre[3], p, p, post[3])) * 0.1).astype(np.float32)
    return rng, hp, hq, op, oq, p, dens, W


def _delivery(hp, hq, p, W, A, halo_is_zero, arith=abi.ARITH_BF16):
    d = R.PresynapticPerspectiveGPUDelivery(hp, hq, p, p, dev(W), arith=arith)
    assert d.dense is not None
    if arith == abi.ARITH_FP32:
        d.pre_halo_is_zero = halo_is_zero
        assert d.mode == "dense"
        return d, None
    A16 = R.bf16c_buffer(hp, DEV)
    R.activity_to_bf16c(hp, dev(A), A16)
    assert d.bind_pre_bf16c(A16, halo_is_zero=halo_is_zero)
    assert d.mode == "dense"
    return d, A16


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("halo_too", [False, True])
def test_dense_scatter_exact_products(case, halo_too):
    """Operands representable in BF16: the oracle's FP32 scatter of the same
    values (exact products, the reference order) within 1e-5 of the largest
    output; the restricted-only mode when the halo is zero."""
    rng, hp, hq, op, oq, p, dens, W = _setup(case, 71)
    Wb = bf16_round(W)
    A = bf16_round(activity(hp, rng, dens, halo_too))
    print("dense scatter", case, "halo activity" if halo_too else "zero halo", flush=True)
    d, _ = _delivery(hp, hq, p, Wb, A, halo_is_zero=not halo_too)
    g = torch.full((hq.nbatch * hq.num_res,), 7.0, device=DEV)
    d.deliver(g, pre_activity=dev(A), accumulate=False)
    ref = O.recv_pre(op, oq, p, p, Wb, pre_activity=A, num_threads=1).astype(np.float64)
    got = host(g)
    assert np.abs(got - ref).max() <= 1e-5 * np.abs(ref).max()
    assert np.abs(ref).max() > 0


@pytest.mark.parametrize("case", CASES[:2])
def test_dense_scatter_fp32_operands_and_accumulate(case):
    """FP32 operands (rounded to BF16 inside): RMS error within 1e-2 of the
    oracle's RMS output, as the other BF16 forms; accumulate adds to GSyn."""
    rng, hp, hq, op, oq, p, dens, W = _setup(case, 72)
    A = activity(hp, rng, dens, True)
    d, _ = _delivery(hp, hq, p, W, A, halo_is_zero=False)
    g0 = rng.standard_normal(hq.nbatch * hq.num_res).astype(np.float32)
    g = dev(g0)
    d.deliver(g, pre_activity=dev(A), accumulate=True)
    ref = O.recv_pre(op, oq, p, p, W, pre_activity=A, num_threads=1).astype(np.float64)
    err = host(g) - g0 - ref
    assert np.sqrt((err ** 2).mean()) <= 1e-2 * np.sqrt((ref ** 2).mean())


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("halo_too", [False, True])
def test_dense_scatter_fp32_form(case, halo_too):
    """The FP32 form (FP32 operands on v_mfma_f32_16x16x4f32): the oracle's
    FP32 scatter within 1e-5 of the largest output (only the order of the
    FP32 sums differs); accumulate adds to GSyn."""
    rng, hp, hq, op, oq, p, dens, W = _setup(case, 75)
    A = activity(hp, rng, dens, halo_too)
    print("FP32 dense scatter", case, "halo activity" if halo_too else "zero halo", flush=True)
    d, _ = _delivery(hp, hq, p, W, A, halo_is_zero=not halo_too, arith=abi.ARITH_FP32)
    ref = O.recv_pre(op, oq, p, p, W, pre_activity=A, num_threads=1).astype(np.float64)
    g = torch.full((hq.nbatch * hq.num_res,), 7.0, device=DEV)
    d.deliver(g, pre_activity=dev(A), accumulate=False)
    assert np.abs(host(g) - ref).max() <= 1e-5 * np.abs(ref).max()
    g0 = rng.standard_normal(hq.nbatch * hq.num_res).astype(np.float32)
    g = dev(g0)
    d.deliver(g, pre_activity=dev(A), accumulate=True)
    assert np.abs(host(g) - g0 - ref).max() <= 1e-5 * (np.abs(ref).max() + np.abs(g0).max())


@pytest.mark.parametrize("arith", [abi.ARITH_BF16, abi.ARITH_FP32])
def test_dense_scatter_fused_layer_deterministic_and_graph(arith):
    """The fused target-layer update equals the delivery followed by
    pvhip_layer_update bit for bit; repeated calls and hipGraph replays give
    the same bits."""
    rng, hp, hq, op, oq, p, dens, W = _setup(CASES[0], 73)
    hq = abi.Loc.make(2, 128, 128, 1, halo=1)      # the target layer's own halo for A
    A = activity(hp, rng, dens, False)
    d, _ = _delivery(hp, hq, p, W, A, halo_is_zero=True, arith=arith)
    tr = R.make_transfer(0.05, 0.0, FMAX, 0.02, 0.0)
    n, ne = hq.nbatch * hq.num_res, hq.nbatch * hq.num_ext
    Ad = dev(A)
    g = torch.zeros(n, device=DEV)
    V1, A1 = torch.zeros(n, device=DEV), torch.zeros(ne, device=DEV)
    la1 = R.layer_update_args(hq, g, V1, A1, tr, channel_coeffs=(1.0,), err_scale=0.5)
    d.deliver(g, pre_activity=Ad)
    R.launch_layer_update(la1, None)
    V2, A2 = torch.zeros(n, device=DEV), torch.zeros(ne, device=DEV)
    g2 = torch.full((n,), 3.0, device=DEV)
    la2 = R.layer_update_args(hq, g2, V2, A2, tr, channel_coeffs=(1.0,), err_scale=0.5)
    d.deliver(g2, pre_activity=Ad, fused_layer=la2)
    assert np.all(host(g2) == 3.0)                 # GSyn not written
    assert np.array_equal(host(V1).view(np.uint32), host(V2).view(np.uint32))
    assert np.array_equal(host(A1).view(np.uint32), host(A2).view(np.uint32))
    first = host(g).copy()
    for _ in range(3):
        d.deliver(g, pre_activity=Ad)
        assert np.array_equal(host(g).view(np.uint32), first.view(np.uint32))
    s = torch.cuda.Stream()
    graph = torch.cuda.CUDAGraph()
    g.zero_()
    torch.cuda.synchronize()
    with torch.cuda.graph(graph, stream=s):
        d.deliver(g, pre_activity=Ad, stream=s)
    for _ in range(3):
        g.fill_(-1.0)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert np.array_equal(host(g).view(np.uint32), first.view(np.uint32))


@pytest.mark.parametrize("arith", [abi.ARITH_BF16, abi.ARITH_FP32])
def test_dense_weights_follow_refresh(arith):
    """refresh_weights re-derives the dense mode's BF16 (FP32) weights after a
    weight change (the delivery then matches the oracle on the new weights)."""
    rng, hp, hq, op, oq, p, dens, W = _setup(CASES[1], 74)
    A = bf16_round(activity(hp, rng, dens, False))
    d, _ = _delivery(hp, hq, p, bf16_round(W), A, halo_is_zero=True, arith=arith)
    W2 = bf16_round(W[::-1].copy() * 0.5)
    d.weights.copy_(dev(W2).reshape(d.weights.shape))
    d.refresh_weights()
    g = torch.zeros(hq.nbatch * hq.num_res, device=DEV)
    d.deliver(g, pre_activity=dev(A))
    ref = O.recv_pre(op, oq, p, p, W2, pre_activity=A, num_threads=1).astype(np.float64)
    assert np.abs(host(g) - ref).max() <= 1e-5 * np.abs(ref).max()
