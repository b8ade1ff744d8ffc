"""The gather + fused LCA update over a large batch (the FP32, BF16 and FP8
forms run the persistent kernel there, whose workgroups walk several tiles
and prefetch the next tile while computing the current one): every batch
element equals, bit for bit, the same delivery
run on that element alone.  Per neuron the arithmetic does not depend on how tiles
are assigned to workgroups, so any difference is a pipelining or indexing
mistake."""
import numpy as np
import pytest
import torch

from openpv_amd import abi
from openpv_amd import runtime as R

pytestmark = pytest.mark.gpu

DEV = "cuda"
FMAX = float(np.finfo(np.float32).max)


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a)).to(DEV, dtype=torch.float32).contiguous()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def run(hs, ht, p, W, A, V0, arith, src_scale, with_copy):
    tr = R.make_transfer(0.05, 0.0, FMAX, 0.05, 0.0)
    d = R.PostsynapticPerspectiveGPUDelivery(hs, ht, p, p, dev(W), arith=arith)
    if arith == abi.ARITH_FP8:
        d.set_source_scale(src_scale)
    upd = R.HyPerLCAUpdate(ht, tr, 100.0, 1.0, True, activity_from_v=True, device=DEV)
    Va = dev(V0)
    Aa = torch.zeros(ht.nbatch * ht.num_ext, device=DEV)
    R.ann_activity(ht, tr, Va, Aa)
    A16 = R.bf16c_buffer(ht, DEV) if with_copy else None
    d.deliver(dev(A), None, fused=upd.args(None, Va, Aa, num_channels=1, A_bf16c=A16))
    return host(Va), host(Aa), (None if A16 is None else A16.view(torch.int16).cpu().numpy())


@pytest.mark.parametrize("arith", [abi.ARITH_BF16, abi.ARITH_FP8, abi.ARITH_FP32])
def test_persistent_gather_equals_one_element_runs(arith):
    rng = np.random.default_rng(71)
    nb, p = 160, 8                      # 640 tiles x 2 feature blocks: three tiles per workgroup
    src, tgt = (32, 32, 3, 3), (16, 16, 128, 2)
    hs, ht = abi.Loc.make(nb, *src[:3], halo=src[3]), abi.Loc.make(nb, *tgt[:3], halo=tgt[3])
    W = (rng.standard_normal((tgt[2], p, p, src[2])) * 0.05).astype(np.float32)
    A = np.zeros((nb, hs.ny_ext, hs.nx_ext, src[2]), np.float32)
    A[:, 3:3 + 32, 3:3 + 32] = rng.standard_normal((nb, 32, 32, src[2]))
    V0 = (rng.standard_normal((nb, 16, 16, tgt[2])) * 0.1).astype(np.float32)
    scale = R.fp8_scale_for(np.abs(A).max())
    with_copy = arith != abi.ARITH_FP32
    V, Ae, C16 = run(hs, ht, p, W, A.ravel(), V0.ravel(), arith, scale, with_copy)
    V, Ae = V.reshape(nb, -1), Ae.reshape(nb, -1)
    hs1, ht1 = abi.Loc.make(1, *src[:3], halo=src[3]), abi.Loc.make(1, *tgt[:3], halo=tgt[3])
    for b in (0, 1, 2, 79, 158, 159):
        v1, a1, c1 = run(hs1, ht1, p, W, A[b].ravel(), V0[b].ravel(), arith, scale, with_copy)
        assert np.array_equal(V[b].view(np.uint32), v1.view(np.uint32)), b
        assert np.array_equal(Ae[b].view(np.uint32), a1.view(np.uint32)), b
        if with_copy:
            assert np.array_equal(C16.reshape(nb, -1)[b], c1), b
