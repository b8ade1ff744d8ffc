"""The whole LCA loop on the oracle against the reference's own acceptance
fixture: tests/LCATest/input/correct.pvp (kept as tests/golden/lcatest/),
the Reconstruction of each of LCATest's 40 steps, which LCATest compares at
nnzThreshold 1e-6 (RequireAllZeroActivityProbe "ComparisonTest").

These run on the CPU.  The device columns meet the same fixture in
test_gpu_lcatest.py.
"""
import numpy as np
import pytest

import oracle as O
import lcatest_column as LC
from openpv_amd.timescale import AdaptiveTimeScaleController, PeriodicUpdate


def test_oracle_column_reproduces_lcatest_correct_pvp():
    """Every one of the 40 steps (the input switch and time-scale reset at
    step 21 included) is within LCATest's nnzThreshold of correct.pvp."""
    recon, _ = LC.oracle_run(40)
    want = LC.correct()
    err = np.abs(recon - want).reshape(40, -1).max(1)
    print("max |oracle - correct.pvp| per step:", err)
    assert (err <= LC.NNZ_THRESHOLD).all(), np.nonzero(err > LC.NNZ_THRESHOLD)[0] + 1


def test_lcatest_fixture_detects_a_wrong_column():
    """The fixture is discriminating: the same column with V1 as a
    MomentumLCALayer (rate 0.4), or with the time scale held at baseMin, fails
    LCATest's comparison."""
    want = LC.correct()[:25]
    mom, _ = LC.oracle_run(25, 0.4)
    assert np.abs(mom - want).max() > 100 * LC.NNZ_THRESHOLD
    fixed, _ = LC.oracle_run(25, None, False)
    assert np.abs(fixed - want).max() > 100 * LC.NNZ_THRESHOLD


def test_timescale_controller_matches_the_oracles():
    """The host controller of the product (openpv_amd.timescale, also restated
    in pvrun) against the oracle's restatement of
    AdaptiveTimeScaleController::calcTimesteps, on LCATest's own energy
    trajectory with its trigger steps, and on a trajectory that exercises the
    cap, the growth of the cap and equal successive values."""
    _, energy = LC.oracle_run(40)
    prod = AdaptiveTimeScaleController(1, 0.55, 0.5, 0.1, 0.01)
    orc = O.AdaptiveTimeScales(1, 0.55, 0.5, 0.1, 0.01)
    trig = PeriodicUpdate(20)
    last, prev = 1.0, np.zeros(1)
    for t in range(1, 41):
        up = trig.need_update(t)
        assert up == O.need_update(float(t), 1.0, last, 20.0)
        vals = np.full(1, -1.0) if up else prev
        assert np.array_equal(prod.calc_timesteps(vals), orc.calc(vals))
        if up:
            trig.updated(t)
            last = float(t)
        prev = energy[t - 1:t]
    rng = np.random.default_rng(5)
    prod = AdaptiveTimeScaleController(3, 2.0, 0.1, 1.0, 0.5)
    orc = O.AdaptiveTimeScales(3, 2.0, 0.1, 1.0, 0.5)
    e = np.array([10.0, 5.0, -1.0])
    for _ in range(60):
        kind = rng.integers(4)
        if kind == 0:
            e = e * 0.5                     # fast fall: cap, then cap growth
        elif kind == 1:
            e = e * 1.1                     # rise: reset to baseMin
        elif kind == 2:
            e = e.copy()                    # equal values: unchanged
        else:
            e = e * rng.uniform(0.9, 1.0, 3)
        assert np.array_equal(prod.calc_timesteps(e), orc.calc(e))


@pytest.mark.parametrize("period,last,t,want", [(20, 1.0, 1, True), (20, 1.0, 20, False),
                                                (20, 1.0, 21, True), (1, 5.0, 6, True),
                                                (0, 1.0, 1, False)])
def test_oracle_need_update_cases(period, last, t, want):
    """LayerUpdateController::needUpdate closed cases: an input shown every
    `period` steps from its last update; period 0 never updates."""
    assert O.need_update(float(t), 1.0, last, float(period)) == want
