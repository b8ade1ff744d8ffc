"""GPU parity of the ISTA V update (ISTAInternalStateBuffer.kpp, the
reference's K9) against the oracle through the C ABI: V and A bit-identical
over several steps, with the GSyn accumulation and the transfer fused."""
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


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a, np.float32)).to(DEV).contiguous()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


@pytest.mark.parametrize("params", [(0.1, 0.0, FMAX, 0.1, 0.0), (0.025, 0.0, FMAX, 0.0, 0.025),
                                    (0.1, -FMAX, FMAX, 0.1, 0.0)])
@pytest.mark.parametrize("from_v", [False, True])
def test_ista_update_bit_exact(params, from_v):
    """Five steps carrying V and A: soft threshold, firm threshold, and a
    two-sided soft threshold (AMin -inf: negative activity, sign -1); two GSyn
    channels, per-batch dt, a ragged halo."""
    rng = np.random.default_rng(61)
    hl = abi.Loc.make(3, 11, 7, 5, (2, 1, 1, 2))
    ol = O.Loc.make(3, 11, 7, 5, (2, 1, 1, 2))
    n = hl.nbatch * hl.num_res
    dt = [1.0, 0.5, 2.0]
    tau = 10.0
    V = (rng.standard_normal(n) * 0.3).astype(np.float32)
    A = O.transfer(ol, V, *params)
    upd = R.ISTAUpdate(hl, R.make_transfer(*params), params[0], tau, 1.0, channel_coeffs=(1.0, -0.5),
                       activity_from_v=from_v, device=DEV)
    upd.set_dt(dt)
    Vd, Ad = dev(V), dev(A)
    for step in range(5):
        G = (rng.standard_normal(2 * n) * 0.3).astype(np.float32)
        acc = np.zeros(n, np.float32)
        O.lib().or_gsyn_accumulate(n, 2, np.array([1.0, -0.5], np.float32), G, acc)
        V = O.update_ista(ol, params[0], np.array(dt), np.float32(tau), acc, A, V)
        A = O.transfer(ol, V, *params, A=A)
        upd.update(dev(G), Vd, Ad)
        print("ista step", step, "active", int((A != 0).sum()), flush=True)
        assert same(host(Vd), V) and same(host(Ad), A)
    assert (A != 0).any()


def test_pvrun_ista_column_equals_oracle(tmp_path):
    """LCATest's network with V1 as an ISTALayer: pvrun's host column (EXACT
    deliveries) equals the oracle's ISTA column bit for bit over 25 steps, and
    differs from the HyPerLCA column."""
    import os
    import subprocess
    import lcatest_column as LC
    from openpv_amd import pvp
    steps = 25
    text = open(os.path.join(LC.LCATEST, "LCATest.params")).read()
    assert text.count('HyPerLCALayer "V1" = {') == 1
    text = text.replace('HyPerLCALayer "V1" = {', 'ISTALayer "V1" = {')
    text = text.replace("exitOnFailure = true;", "exitOnFailure = false;")
    params = tmp_path / "ista.params"
    params.write_text(text)
    r = subprocess.run([abi.pvrun_path(), "-p", str(params), "-o", str(tmp_path), "--steps", str(steps)],
                       cwd=LC.LCATEST, capture_output=True, text=True, timeout=300)
    print(r.stdout[-800:], r.stderr[-800:], flush=True)
    assert r.returncode == 0
    ref, _ = LC.oracle_run(steps, ista=True)
    _, _, rec = pvp.read_activity(str(tmp_path / "Reconstruction.pvp"))
    for t in range(steps):
        assert same(rec[t], ref[t]), "step %d (max %g)" % (t + 1, np.abs(rec[t] - ref[t]).max())
    assert not same(rec, LC.oracle_run(steps)[0])
