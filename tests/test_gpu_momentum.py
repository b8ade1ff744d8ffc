"""GPU parity of the momentum forms against the CPU oracle, through the C ABI:
the MomentumLCA V update (MomentumLCAInternalStateBuffer.kpp) and the
MomentumUpdater weight step (MomentumUpdater.cpp).  Both are EXACT forms:
V, A, prevDrive, dW, prev_dW and the weights bit-identical to the oracle.
"""
import numpy as np
import pytest
import torch

import oracle as O
from openpv_amd import abi, runtime as R

pytestmark = pytest.mark.gpu
FMAX = float(np.finfo(np.float32).max)
DEV = "cuda"


def same(a, b):
    return np.array_equal(np.ascontiguousarray(a, np.float32).view(np.uint32),
                          np.ascontiguousarray(b, np.float32).view(np.uint32))


def dev(a, dtype=torch.float32):
    return torch.as_tensor(np.ascontiguousarray(a)).to(DEV, dtype=dtype).contiguous()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


@pytest.mark.parametrize("params", [(0.025, 0.0, FMAX, 0.0, 0.025), (0.1, 0.0, 0.4, 0.05, 0.1)])
@pytest.mark.parametrize("self_interact,from_v", [(True, False), (False, False), (True, True)])
def test_momentum_lca_update_bit_exact(params, self_interact, from_v):
    """Four steps carrying V, A and prevDrive (MomentumLCATest's firm threshold
    and a clamped ramp), two GSyn channels, per-batch dt, a ragged halo."""
    rng = np.random.default_rng(51)
    hl = abi.Loc.make(3, 11, 7, 5, (2, 1, 1, 2))
    ol = O.Loc.make(3, 11, 7, 5, (2, 1, 1, 2))
    n = hl.nbatch * hl.num_res
    dt = [1.0, 0.5, 2.0]
    V = (rng.standard_normal(n) * 0.2).astype(np.float32)
    A = O.transfer(ol, V, *params)
    prev = np.zeros(n, np.float32)
    upd = R.MomentumLCAUpdate(hl, R.make_transfer(*params), 100.0, 1.0, self_interact,
                              LCAMomentumRate=0.4, channel_coeffs=(1.0, -0.5),
                              activity_from_v=from_v, device=DEV)
    upd.set_dt(dt)
    Vd, Ad = dev(V), dev(A)
    for step in range(4):
        G = (rng.standard_normal(2 * n) * 0.3).astype(np.float32)
        acc = np.zeros(n, np.float32)
        O.lib().or_gsyn_accumulate(n, 2, np.array([1.0, -0.5], np.float32), G, acc)
        V, prev = O.update_momentum_lca(ol, self_interact, 0.4, np.array(dt), 100.0, acc, A, prev, V)
        A = O.transfer(ol, V, *params, A=A)
        upd.update(dev(G), Vd, Ad)
        print("momentum lca step", step, flush=True)
        assert same(host(Vd), V) and same(host(Ad), A) and same(host(upd.prev_drive), prev)


def test_momentum_lca_rejects_missing_prev_drive():
    hl = abi.Loc.make(1, 4, 4, 2, 0)
    upd = R.MomentumLCAUpdate(hl, R.make_transfer(0.1, 0.0, FMAX, 0.1, 0.0), device=DEV)
    V = torch.zeros(hl.num_res, device=DEV)
    A = torch.zeros(hl.num_ext, device=DEV)
    a = upd.args(torch.zeros(hl.num_res, device=DEV), V, A)
    import ctypes as C
    assert abi.hip().pvhip_momentum_lca_update(C.byref(a), 0.4, None, None) == 1  # PVHIP_ERR_INVALID


@pytest.mark.parametrize("method,tau,decay,deprecated", [
    ("viscosity", 100.0, 0.0, False), ("simple", 0.75, 0.01, False), ("alex", 0.9, 0.5, False),
    ("simple", 0.25, 0.01, True),
])
@pytest.mark.parametrize("normalize", [True, False])
def test_momentum_apply_bit_exact(method, tau, decay, deprecated, normalize):
    """Three weight updates in a row (prev_dW carried), counts with zeros."""
    rng = np.random.default_rng(52)
    n = 5 * 1024 + 37
    w = rng.standard_normal(n).astype(np.float32)
    prev = np.zeros(n, np.float32)
    code = {"viscosity": 0, "simple": 1, "alex": 2}[method]
    f, wf = O.momentum_factors(code, tau, decay, 0.5)
    assert (f, wf) == R.momentum_factors(method, tau, decay, 0.5)
    wd, prevd = dev(w), dev(prev)
    for rep in range(3):
        dw = (rng.standard_normal(n) * 3).astype(np.float32)
        cnt = rng.integers(0, 4, n).astype(np.int64) if normalize else None
        w, dwr, prev = O.momentum_apply(w, dw, prev, cnt, f, wf, deprecated)
        dwd = dev(dw)
        cd = dev(cnt, torch.int64) if normalize else None
        import ctypes as C
        abi.check(abi.hip().pvhip_momentum_apply(
            R._ptr(wd), R._ptr(dwd), R._ptr(prevd), R._ptr(cd), n, int(normalize), f, wf,
            int(deprecated), None), "pvhip_momentum_apply")
        print("momentum apply", method, rep, flush=True)
        assert same(host(wd), w) and same(host(dwd), dwr) and same(host(prevd), prev)


def test_momentum_updater_end_to_end():
    """MomentumUpdater (viscosity, decay) on a shared-weight connection: the
    Hebbian dW then two momentum weight steps, against the oracle's
    hebbian_dw + momentum_apply."""
    rng = np.random.default_rng(53)
    hp, op = abi.Loc.make(2, 16, 16, 3, 2), O.Loc.make(2, 16, 16, 3, 2)
    hq, oq = abi.Loc.make(2, 8, 8, 4, 0), O.Loc.make(2, 8, 8, 4, 0)
    mu = R.MomentumUpdater(hp, hq, 5, 5, 0.25, momentumMethod="viscosity", timeConstantTau=10.0,
                           momentumDecay=0.05, device=DEV)
    assert (mu.dw_factor, mu.w_factor) == O.momentum_factors(0, 10.0, 0.05, 0.25)
    w = rng.standard_normal(mu.size).astype(np.float32)
    prev = np.zeros(mu.size, np.float32)
    wd = dev(w)
    for rep in range(2):
        a = rng.standard_normal(hp.nbatch * hp.num_ext).astype(np.float32)
        a[rng.random(a.shape) < 0.6] = 0
        r = rng.standard_normal(hq.nbatch * hq.num_ext).astype(np.float32)
        dw, cnt = O.hebbian_dw(op, oq, 5, 5, oq.nf, 0.25, a, r, counts=True)
        w, dw, prev = O.momentum_apply(w, dw, prev, cnt, mu.dw_factor, mu.w_factor)
        mu.update_dw(dev(a), dev(r))
        mu.apply(wd)
        assert same(host(wd), w) and same(host(mu.dw), dw) and same(host(mu.prev_dw), prev)
