"""The reference's LCATest / GPULCATest column on the oracle alone (a helper of
the tests, not collected by pytest).

Network (tests/LCATest/input/LCATest.params): Input 128x64x1 (two frames from
inputImages.pvp, displayPeriod 20) -> Residual (ANNErrorLayer, errScale 1)
<-> V1 (HyPerLCALayer 64x32x24, 18x18 patches, firm threshold VThresh 0.025 /
VWidth 0.025, tau 100, adaptive time scales on the total energy
0.5|r|^2 + VThresh*firm(a)) -> Reconstruction.  LCATest draws V1's initial V
and the weights from the reference's RNG; tests/GPULCATest/input holds that
state as files (initialV1.pvp; weights.pvp, 90 % zeros = sparseFraction 0.9,
L2-normalised on initialisation), which is how this column feeds them.
LCATest compares every step's Reconstruction with input/correct.pvp at
nnzThreshold 1e-6 (RequireAllZeroActivityProbe "ComparisonTest").

Every piece here is the oracle's: deliveries, updates, the active list, the
cost sums, the time-scale controller and the input trigger.
"""
import functools
import os

import numpy as np

import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
GPULCA = os.path.join(GOLDEN, "gpulca")
LCATEST = os.path.join(GOLDEN, "lcatest")
FMAX = float(np.finfo(np.float32).max)
# LCATest.params: RequireAllZeroActivityProbe "ComparisonTest" nnzThreshold
NNZ_THRESHOLD = 1e-6


@functools.lru_cache(maxsize=None)
def inputs():
    """(frames [2][64][128][1], V0 [64*32*24 restricted], raw weights [24][18][18][1])."""
    from openpv_amd import pvp
    _, _, frames = pvp.read_activity(os.path.join(GPULCA, "inputImages.pvp"))
    _, _, V0 = pvp.read_activity(os.path.join(GPULCA, "initialV1.pvp"))
    _, _, W = pvp.read_weights(os.path.join(GPULCA, "weights.pvp"))
    return frames, V0[0].ravel().copy(), W[0].copy()


@functools.lru_cache(maxsize=None)
def correct():
    """LCATest's expected Reconstruction, [40][64][128][1] (timestamps 1..40)."""
    from openpv_amd import pvp
    _, times, c = pvp.read_activity(os.path.join(LCATEST, "correct.pvp"))
    assert list(times) == list(range(1, 41))
    return c


@functools.lru_cache(maxsize=None)
def oracle_run(steps, momentum=None, adaptive=True):
    """(Reconstruction of each step [steps][64][128][1], total energy of each
    step [steps]).  momentum: V1 as a MomentumLCALayer with that
    LCAMomentumRate.  adaptive=False holds V1's time scale at baseMin (a
    wrong column, for tests that show the fixture tells it apart).  Cached:
    callers must not modify the arrays."""
    frames, V0, W = inputs()
    W = W.copy()
    O.lib().or_normalize_l2(W.ravel(), 24, 18 * 18, 1.0, 0.0)  # normalizeOnInitialize
    nxpT = O.lib().or_calcPostPatchSize(18, 64, 128)
    mres = O.lib().or_requiredConvolveMargin(128, 64, nxpT)
    mv1 = O.lib().or_requiredConvolveMargin(64, 128, 18)
    res = O.Loc.make(1, 128, 64, 1, mres)
    v1 = O.Loc.make(1, 64, 32, 24, mv1)
    rec = O.Loc.make(1, 128, 64, 1, 0)
    tr = (0.025, 0.0, FMAX, 0.0, 0.025)
    V = V0.copy()
    Av = O.transfer(v1, V, *tr)
    prev = np.zeros(v1.num_res, np.float32)
    rc = np.zeros(rec.num_res, np.float32)
    ctrl = O.AdaptiveTimeScales(1, 0.55, 0.5, 0.1, 0.01)
    last_input = 1.0   # LayerUpdateController: mLastUpdateTime starts at dt
    x, idx, e_prev, out, energies = frames[0], 0, np.zeros(1), [], []
    vt = float(np.float32(0.025))
    for t in range(1, steps + 1):
        up = O.need_update(float(t), 1.0, last_input, 20.0)
        ts = ctrl.calc(np.full(1, -1.0) if up else e_prev)
        if not adaptive:
            ts = np.full(1, 0.5)
        if up:
            x, idx = frames[idx], (idx + 1) % len(frames)
            last_input = float(t)
        Rres = np.zeros(res.num_res, np.float32)
        O.lib().or_gsyn_accumulate(res.num_res, 2, np.array([1.0, -1.0], np.float32),
                                   np.concatenate([x.ravel(), rc]), Rres)
        Rext = np.zeros((res.ny_ext, res.nx_ext), np.float32)
        Rext[mres:mres + 64, mres:mres + 128] = Rres.reshape(64, 128)
        G = O.recv_post(res, v1, 18, 18, Rext.ravel(), W, num_threads=1)
        if momentum is None:
            V = O.update_lca(v1, True, ts, 100.0, G, Av, V)
        else:
            V, prev = O.update_momentum_lca(v1, True, momentum, ts, 100.0, G, Av, prev, V)
        Av = O.transfer(v1, V, *tr)
        ent, cnt = O.active_indices(v1, Av)
        rc = O.recv_pre(v1, rec, 18, 18, W, entries=ent, num_active=cnt, num_threads=1)
        e_prev = np.array([0.5 * O.cost_sum(res, Rext.ravel(), True, 0, 0)
                           + vt * O.cost_sum(v1, Av, True, 0, 2, vt, vt)])
        out.append(rc.copy())
        energies.append(e_prev[0])
    return np.stack(out).reshape(steps, 64, 128, 1), np.array(energies)
