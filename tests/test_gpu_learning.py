"""Dictionary learning on the device: the LCA column whose dictionary is
updated by HebbianUpdater at each new input (immediateWeightUpdate, the trigger
at the last step of a display period), renormalised per feature and
re-derived for the deliveries.

  * EXACT column, EXACT dW, the reference's float NormalizeL2: weights, V1
    state and reconstructions bit-identical to the oracle over three display
    periods;
  * BF16 column, BF16 MFMA dW, FP64 renormalisation, the update replayed from a
    hipGraph: within stated bounds of the oracle (EXACT dW, FP64
    renormalisation), graph replays bit-identical to eager updates;
  * the large workload (8 x 512x512x1 -> 128x128x1024, 16x16 stride 4, lr 1e-3
    with HebbianUpdater's default normalizeDw, 300 steps per image) in BF16
    and FP8: FP64 norms of every feature 1 after each update, energy falling
    within each period.
"""
import functools

import numpy as np
import pytest
import torch

import oracle_column as OC
from openpv_amd import abi

pytestmark = pytest.mark.gpu


def _frames(n, nb, size, seed):
    import bench
    return [bench.one_over_f_images(nb, size, size, seed + i) for i in range(n)]


def _dictionary(nf, p, seed):
    import bench
    return bench.dictionary(nf, p, 1, seed)


SMALL = dict(nb=2, size=64, nf=32, nv=16, p=16, vt=0.15, steps=20, periods=3, lr=1e-2)


def _run_device(arith, learn_arith, renorm, graph=False):
    from openpv_amd.lca import LCAColumn
    c = SMALL
    frames = _frames(c["periods"], c["nb"], c["size"], 11)
    W = _dictionary(c["nf"], c["p"], 12)
    col = LCAColumn(frames[0], W, c["nv"], c["nv"], c["nf"], c["p"], c["p"], VThresh=c["vt"],
                    arith=arith, fused=True, device="cuda")
    col.enable_learning(c["lr"], arith=learn_arith, renormalize=renorm)
    s = torch.cuda.Stream()
    out = []
    for k in range(c["periods"]):
        col.set_input(frames[k])
        torch.cuda.synchronize()
        for _ in range(c["steps"]):
            col.step(stream=s)
        s.synchronize()
        energy = col.probe_energy(stream=s).cpu().numpy()
        if graph and k == 0:
            saved = col.weights.clone()
            col.update_weights(s)          # eager (also creates the updater's scratch)
            s.synchronize()
            w_eager = col.weights.clone()
            col.weights.copy_(saved)       # the same state again, then the update from its graph
            col.refresh_weights(s)
            col.capture_learning(s)
            col.replay_learning(s)
            s.synchronize()
            assert torch.equal(col.weights.view(torch.int32), w_eager.view(torch.int32))
        elif graph:
            col.replay_learning(s)
        else:
            col.update_weights(s)
        s.synchronize()
        V, A = col.v1_state()
        out.append(dict(energy=energy, W=col.weights.cpu().numpy().copy(), V=V, A=A,
                        rec=col.rec_A.cpu().numpy().copy()))
    col.release_graph()
    return out


@functools.lru_cache(maxsize=None)
def _run_oracle(renorm):
    c = SMALL
    frames = _frames(c["periods"], c["nb"], c["size"], 11)
    W = _dictionary(c["nf"], c["p"], 12)
    col = OC.OracleColumn(frames[0], W, c["nv"], c["nv"], c["nf"], c["p"], c["vt"])
    out = []
    for k in range(c["periods"]):
        col.set_input(frames[k])
        for _ in range(c["steps"]):
            col.step()
        energy = col.energy()
        col.update_weights(c["lr"], renorm)
        out.append(dict(energy=energy, W=col.W.copy(), V=col.V.copy(), A=col.A.copy(), rec=col.rc.copy()))
    return out


def test_learning_exact_bit_exact():
    print("learning column EXACT, %d periods of %d steps" % (SMALL["periods"], SMALL["steps"]), flush=True)
    got = _run_device(abi.ARITH_EXACT, abi.ARITH_EXACT, "exact")
    ref = _run_oracle("exact")
    for k, (g, r) in enumerate(zip(got, ref)):
        for key in ("W", "V", "A", "rec"):
            assert np.array_equal(g[key].view(np.uint32), np.ascontiguousarray(r[key]).view(np.uint32)), \
                "period %d: %s differs (max %g)" % (k, key, np.abs(g[key] - r[key]).max())
        assert np.allclose(g["energy"], r["energy"], rtol=1e-12)
    assert not np.array_equal(ref[0]["W"], ref[-1]["W"])


def test_learning_bf16_graph_against_oracle():
    """BF16 deliveries and BF16 MFMA dW with FP64 renormalisation, the weight
    update replayed from its hipGraph (the first replay checked bit for bit
    against the eager update of the same state): after each period the
    weights are within 2e-3 of the oracle's (EXACT dW, FP64 renormalisation;
    norm 1 per feature, so this is relative) and the energy within 1e-3."""
    got = _run_device(abi.ARITH_BF16, abi.ARITH_BF16, "fp64", graph=True)
    ref = _run_oracle("fp64")
    for k, (g, r) in enumerate(zip(got, ref)):
        werr = np.abs(g["W"] - r["W"]).max()
        eerr = np.abs(g["energy"] - r["energy"]).max() / np.abs(r["energy"]).max()
        print("period %d: max |W - oracle| %.3e, energy rel err %.3e" % (k, werr, eerr), flush=True)
        assert werr <= 2e-3
        assert eerr <= 1e-3
        norms = np.sqrt((g["W"].astype(np.float64).reshape(SMALL["nf"], -1) ** 2).sum(1))
        assert np.allclose(norms, 1.0, atol=1e-6)


@pytest.mark.parametrize("form", ["bf16", "fp8"])
def test_learning_large_workload(form):
    """8 x 512x512x1 (1/f^2) -> 128x128x1024, 16x16 stride 4, VThresh 0.15,
    dWMax 1e-3 with normalizeDw (HebbianUpdater's default: each kernel's dW
    divided by its activation count), FP64 renormalisation, 300 steps per
    image, two images per image slot: every feature's FP64 norm is 1 after
    each update, the update changes the weights, and the energy falls within
    each period.  (With normalizeDw off the step is the gradient summed over
    131,072 positions: the dictionary turns coherent and the next image's LCA
    iteration diverges.)"""
    from openpv_amd.lca import LCAColumn
    arith = abi.ARITH_BF16 if form == "bf16" else abi.ARITH_FP8
    frames = _frames(2, 8, 512, 21)
    W = _dictionary(1024, 16, 22)
    print("C4 learning column", form, flush=True)
    col = LCAColumn(frames[0], W, 128, 128, 1024, 16, 16, VThresh=0.15, arith=arith, fused=True,
                    device="cuda")
    col.enable_learning(1e-3, normalize_dw=True, renormalize="fp64")
    s = torch.cuda.Stream()
    torch.cuda.synchronize()
    col.capture(s)
    for k in range(2):
        col.set_input(frames[k])
        torch.cuda.synchronize()
        e = []
        for t in range(300):
            col.replay(s)
            if t in (9, 99, 299):
                energy = col.probe_energy(stream=s)
                s.synchronize()
                e.append(energy.cpu().numpy())
        print("period %d energies (steps 10, 100, 300):" % k, [float(x.sum()) for x in e], flush=True)
        assert np.isfinite(e).all()
        assert (e[1] < e[0]).all() and (e[2] < e[1]).all()
        w_before = col.weights.clone()
        col.update_weights(s)
        s.synchronize()
        norms = col.weights.double().view(1024, -1).pow(2).sum(1).sqrt().cpu().numpy()
        assert np.allclose(norms, 1.0, atol=1e-6), (norms.min(), norms.max())
        assert not torch.equal(w_before, col.weights)
    col.release_graph()
