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
def oracle_run(steps, momentum=None, adaptive=True, learn=None, input_delay=0, recon_delay=0,
               ista=False):
    """(Reconstruction of each step [steps][64][128][1], total energy of each
    step [steps]).  momentum: V1 as a MomentumLCALayer with that
    LCAMomentumRate.  adaptive=False holds V1's time scale at baseMin (a
    wrong column, for tests that show the fixture tells it apart).
    learn: the dictionary plastic, as (dWMax, normalizeDw, weight momentum)
    with weight momentum None or (method code, timeConstantTau, decay): at the
    end of each step whose next step shows a new input (triggerLayerName Input,
    triggerOffset 1) HebbianUpdater's immediate update, then NormalizeL2
    (normalizeOnWeightUpdate).  With learn, a third value holds the weights
    after each step [steps][24*18*18].
    input_delay / recon_delay: the delays (timesteps) of the Input -> Residual
    IdentConn and of the V1 -> Reconstruction CloneConn: the reader gets the
    activity published that many steps earlier (the DataStore levels of
    src/columns/DataStore.cpp and Publisher.cpp: the initial publish at time 0
    -- the input's first frame, V1's T(V0) -- and zeros before it).
    ista: V1 as an ISTALayer (ISTAInternalStateBuffer.kpp) with the same
    parameters.
    Cached: callers must not modify the arrays."""
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
    weights, prev_dw = [], None
    shown = {0: frames[0]}                # the input published at each time
    published = {0: Av.copy()}            # V1's activity published at each time
    zero_in = np.zeros_like(frames[0])
    zero_v1 = np.zeros_like(Av)
    vt = float(np.float32(0.025))
    for t in range(1, steps + 1):
        up = O.need_update(float(t), 1.0, last_input, 20.0)
        ts = ctrl.calc(np.full(1, -1.0) if up else e_prev)
        if not adaptive:
            ts = np.full(1, 0.5)
        if up:
            x, idx = frames[idx], (idx + 1) % len(frames)
            last_input = float(t)
        shown[t] = x
        x_in = shown.get(t - input_delay, zero_in)
        Rres = np.zeros(res.num_res, np.float32)
        O.lib().or_gsyn_accumulate(res.num_res, 2, np.array([1.0, -1.0], np.float32),
                                   np.concatenate([x_in.ravel(), rc]), Rres)
        Rext = np.zeros((res.ny_ext, res.nx_ext), np.float32)
        Rext[mres:mres + 64, mres:mres + 128] = Rres.reshape(64, 128)
        G = O.recv_post(res, v1, 18, 18, Rext.ravel(), W, num_threads=1)
        if ista:
            V = O.update_ista(v1, 0.025, ts, 100.0, G, Av, V)
        elif momentum is None:
            V = O.update_lca(v1, True, ts, 100.0, G, Av, V)
        else:
            V, prev = O.update_momentum_lca(v1, True, momentum, ts, 100.0, G, Av, prev, V)
        Av = O.transfer(v1, V, *tr)
        published[t] = Av.copy()
        ent, cnt = O.active_indices(v1, published.get(t - recon_delay, zero_v1))
        rc = O.recv_pre(v1, rec, 18, 18, W, entries=ent, num_active=cnt, num_threads=1)
        e_prev = np.array([0.5 * O.cost_sum(res, Rext.ravel(), True, 0, 0)
                           + vt * O.cost_sum(v1, Av, True, 0, 2, vt, vt)])
        out.append(rc.copy())
        energies.append(e_prev[0])
        if learn is not None:
            if O.need_update(float(t + 1), 1.0, last_input, 20.0):
                dw_max, normalize_dw, wmom = learn
                dw, cnt = O.hebbian_dw(v1, res, 18, 18, 1, dw_max, Av, Rext.ravel(), counts=normalize_dw)
                if wmom is None:
                    O.lib().or_hebbian_normalize_apply(W.ravel(), dw, None if cnt is None else cnt.ctypes.data,
                                                       W.size, int(normalize_dw))
                else:
                    f, wf = O.momentum_factors(wmom[0], wmom[1], wmom[2], dw_max)
                    if prev_dw is None:
                        prev_dw = np.zeros(W.size, np.float32)
                    w2, _, prev_dw = O.momentum_apply(W.ravel(), dw, prev_dw, cnt, f, wf)
                    W.ravel()[:] = w2
                O.lib().or_normalize_l2(W.ravel(), 24, 18 * 18, 1.0, 0.0)
            weights.append(W.ravel().copy())
    if learn is not None:
        return np.stack(out).reshape(steps, 64, 128, 1), np.array(energies), np.stack(weights)
    return np.stack(out).reshape(steps, 64, 128, 1), np.array(energies)
