"""Plasticity through the .params drop-in: pvrun reads plasticityFlag, dWMax,
normalizeDw, immediateWeightUpdate, triggerLayerName / triggerOffset (or
weightUpdatePeriod), and MomentumConn's momentumMethod / timeConstantTau /
momentumDecay, and runs HebbianUpdater's dW, normalize_dW, the momentum blend,
W += dW, NormalizeL2 and the re-derivation of every delivery's weights.

  * LCATest's network with its dictionary plastic (HyPerConn, and MomentumConn
    "viscosity" as in the reference's LCA-CIFAR tutorial): the host column's
    reconstructions and the written weights equal the oracle's bit for bit,
    over 25 steps (one update at step 20, the input switch at step 21);
  * the network of the reference's MomentumConnSimpleCheckpointerTest: the
    weight, input and output of every step equal that test's own expected
    state (CorrectState.cpp).
"""
import os
import subprocess

import numpy as np
import pytest

import lcatest_column as LC
from openpv_amd import abi, pvp

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _pvrun(params, cwd, outdir, steps):
    print("pvrun", params, "steps", steps, flush=True)
    r = subprocess.run([abi.pvrun_path(), "-p", str(params), "-o", str(outdir), "--steps", str(steps)],
                       cwd=cwd, capture_output=True, text=True, timeout=300)
    print(r.stdout[-1500:], r.stderr[-1500:], flush=True)
    return r


def _plastic_lcatest(tmp_path, keyword, extra):
    text = open(os.path.join(LC.LCATEST, "LCATest.params")).read()
    old = 'HyPerConn "V1ToResidual" = {'
    assert text.count(old) == 1 and text.count("plasticityFlag = false;") == 1
    text = text.replace(old, '%s "V1ToResidual" = {' % keyword)
    text = text.replace("plasticityFlag = false;",
                        "plasticityFlag = true;\n    triggerLayerName = \"Input\";\n    triggerOffset = 1;\n"
                        "    immediateWeightUpdate = true;\n    writeStep = 1;\n    initialWriteTime = 1;\n"
                        + extra)
    # the fixed-dictionary comparison with correct.pvp no longer applies
    assert text.count("exitOnFailure = true;") == 1
    text = text.replace("exitOnFailure = true;", "exitOnFailure = false;")
    path = tmp_path / "plastic.params"
    path.write_text(text)
    return path


@pytest.mark.parametrize("case", ["hebbian_normalize_dw", "hebbian_plain", "momentum_viscosity"])
def test_pvrun_plastic_lcatest_column_equals_oracle(tmp_path, case):
    steps = 25
    if case == "hebbian_normalize_dw":
        params = _plastic_lcatest(tmp_path, "HyPerConn", "    dWMax = 0.05;\n    normalizeDw = true;\n")
        learn = (0.05, True, None)
    elif case == "hebbian_plain":
        params = _plastic_lcatest(tmp_path, "HyPerConn", "    dWMax = 0.001;\n    normalizeDw = false;\n")
        learn = (0.001, False, None)
    else:
        params = _plastic_lcatest(tmp_path, "MomentumConn",
                                  "    dWMax = 1.0;\n    normalizeDw = true;\n    momentumMethod = \"viscosity\";\n"
                                  "    timeConstantTau = 5;\n    momentumDecay = 0;\n")
        learn = (1.0, True, (0, 5.0, 0.0))
    r = _pvrun(params, LC.LCATEST, tmp_path, steps)
    assert r.returncode == 0, r.stderr[-800:]
    ref, _, wref = LC.oracle_run(steps, learn=learn)
    _, times, rec = pvp.read_activity(str(tmp_path / "Reconstruction.pvp"))
    assert list(times) == list(range(1, steps + 1))
    for t in range(steps):
        assert np.array_equal(rec[t].view(np.uint32), ref[t].view(np.uint32)), \
            "step %d (max %g)" % (t + 1, np.abs(rec[t] - ref[t]).max())
    _, wt, w = pvp.read_weights(str(tmp_path / "V1ToResidual.pvp"), all_frames=True)
    assert list(wt) == list(range(1, steps + 1))
    got = w[:, 0].reshape(steps, -1)
    for t in range(steps):
        assert np.array_equal(got[t].view(np.uint32), wref[t].view(np.uint32)), "weights after step %d" % (t + 1)
    assert not np.array_equal(got[18], got[19])      # the update at step 20
    assert np.array_equal(got[19], got[24])          # none after it within 25 steps


def _correct_state(steps, tau=0.75):
    """The expected state of the reference's MomentumConnSimpleCheckpointerTest
    (tests/MomentumConnSimpleCheckpointerTest/src/CorrectState.cpp): input
    ceil(t/4), output input*weight, and at t with (t + 1) mod 4 == 0
    dw = (1 - tau)*input*output + tau*dw, weight += dw (float arithmetic)."""
    f = np.float32
    w, dw = f(1.0), f(0.0)
    out = []
    for t in range(1, steps + 1):
        inp = f(np.ceil(f(t) / f(4.0)))
        o = inp * w
        if abs(np.fmod(t + 1, 4.0) - 1) < 0.5:
            base = inp * o
            dw = (f(1) - f(tau)) * base + f(tau) * dw
            w = w + dw
        out.append((inp, o, w))
    return out


def test_pvrun_momentum_conn_simple_expected_state(tmp_path):
    steps = 20
    d = os.path.join(GOLDEN, "momentumconn")
    r = _pvrun(os.path.join(d, "MomentumConnSimple.params"), d, tmp_path, steps)
    assert r.returncode == 0, r.stderr[-800:]
    _, _, inp = pvp.read_activity(str(tmp_path / "Input.pvp"))
    _, _, out = pvp.read_activity(str(tmp_path / "Output.pvp"))
    _, _, w = pvp.read_weights(str(tmp_path / "InputToOutput.pvp"), all_frames=True)
    want = _correct_state(steps)
    for t in range(steps):
        ci, co, cw = want[t]
        assert (inp[t] == ci).all(), "input at t = %d" % (t + 1)
        assert (out[t] == co).all(), "output at t = %d: %s vs %s" % (t + 1, out[t].ravel()[:2], co)
        assert w[t, 0].ravel()[0] == cw, "weight at t = %d: %r vs %r" % (t + 1, w[t, 0].ravel()[0], cw)


@pytest.mark.parametrize("conn,delay,kw", [("InputToResidual", 2, "input_delay"),
                                           ("V1ToReconstruction", 1, "recon_delay"),
                                           ("V1ToReconstruction", 3, "recon_delay")])
def test_pvrun_delays_equal_oracle(tmp_path, conn, delay, kw):
    """Connection delays (the DataStore levels of src/columns/DataStore.cpp,
    Publisher.cpp; ArborList::convertDelay): LCATest's network with a delay on
    the Input -> Residual IdentConn, or on the sparse V1 -> Reconstruction
    CloneConn (its active list then comes from the delayed level), equals the
    oracle column with the same delay bit for bit over 25 steps, and differs
    from the undelayed column."""
    steps = 25
    text = open(os.path.join(LC.LCATEST, "LCATest.params")).read()
    head = '"%s" = {' % conn
    i = text.index(head)
    text = text[:i + len(head)] + "\n    delay = [%d.0];" % delay + text[i + len(head):]
    text = text.replace("exitOnFailure = true;", "exitOnFailure = false;")
    params = tmp_path / "delay.params"
    params.write_text(text)
    r = _pvrun(params, LC.LCATEST, tmp_path, steps)
    assert r.returncode == 0, r.stderr[-800:]
    ref, _ = LC.oracle_run(steps, **{kw: delay})
    _, _, rec = pvp.read_activity(str(tmp_path / "Reconstruction.pvp"))
    for t in range(steps):
        assert np.array_equal(rec[t].view(np.uint32), ref[t].view(np.uint32)), \
            "step %d (max %g)" % (t + 1, np.abs(rec[t] - ref[t]).max())
    plain, _ = LC.oracle_run(steps)
    assert not np.array_equal(rec, plain)


def test_pvrun_checkpoint_and_restart(tmp_path):
    """Checkpointer directories (checkpointWrite, checkpointWriteStepInterval,
    initializeFromCheckpointDir): LCATest's network with a plastic MomentumConn
    dictionary and a delayed reconstruction, written every 10 steps over 25;
    a second run started from Checkpoint10 gives the same reconstructions and
    weights for steps 11..25, bit for bit.  The directories hold the
    reference's file names (timeinfo.bin, <layer>_V.pvp, <layer>_A.pvp,
    <layer>_Delays.pvp, <conn>_W.pvp, <conn>_prev_dW.pvp, <probe>_timescaleinfo.bin)."""
    steps = 25
    params = _plastic_lcatest(tmp_path, "MomentumConn",
                              "    dWMax = 1.0;\n    normalizeDw = true;\n    momentumMethod = \"viscosity\";\n"
                              "    timeConstantTau = 5;\n    momentumDecay = 0;\n")
    text = open(params).read()
    head = '"V1ToReconstruction" = {'
    i = text.index(head)
    text = text[:i + len(head)] + "\n    delay = [2.0];" + text[i + len(head):]
    ck = tmp_path / "ck"
    text = text.replace("    outputPath = \"output/\";",
                        "    outputPath = \"output/\";\n    checkpointWrite = true;\n"
                        "    checkpointWriteDir = \"%s\";\n    checkpointWriteTriggerMode = \"step\";\n"
                        "    checkpointWriteStepInterval = 10;" % ck)
    full = tmp_path / "full.params"
    full.write_text(text)
    out_a = tmp_path / "a"
    r = _pvrun(full, LC.LCATEST, out_a, steps)
    assert r.returncode == 0, r.stderr[-800:]
    names = sorted(os.listdir(ck))
    assert names == ["Checkpoint00", "Checkpoint10", "Checkpoint20", "Checkpoint25"], names
    c10 = ck / "Checkpoint10"
    for f in ("timeinfo.bin", "V1_V.pvp", "V1_A.pvp", "V1_Delays.pvp", "Residual_V.pvp",
              "V1ToResidual_W.pvp", "V1ToResidual_prev_dW.pvp", "AdaptiveTimeScales_timescaleinfo.bin",
              "Input_FrameNumbers.bin"):
        assert (c10 / f).exists(), f
    t, step = np.frombuffer(open(c10 / "timeinfo.bin", "rb").read()[:16], dtype=[("t", "<f8"), ("s", "<i8")])[0]
    assert t == 10.0 and step == 10
    _, _, v1 = pvp.read_activity(str(c10 / "V1_Delays.pvp"))
    assert v1.shape[0] == 3      # levels 0..2 of the delayed V1, one batch element
    restart = tmp_path / "restart.params"
    restart.write_text(text.replace("    checkpointWrite = true;",
                                    "    checkpointWrite = false;\n    initializeFromCheckpointDir = \"%s\";" % c10))
    out_b = tmp_path / "b"
    r = _pvrun(restart, LC.LCATEST, out_b, steps)
    assert r.returncode == 0, r.stderr[-800:]
    _, ta, ra = pvp.read_activity(str(out_a / "Reconstruction.pvp"))
    _, tb, rb = pvp.read_activity(str(out_b / "Reconstruction.pvp"))
    assert list(tb) == list(range(11, steps + 1))
    assert np.array_equal(ra[10:].view(np.uint32), rb.view(np.uint32))
    _, _, wa = pvp.read_weights(str(out_a / "V1ToResidual.pvp"), all_frames=True)
    _, _, wb = pvp.read_weights(str(out_b / "V1ToResidual.pvp"), all_frames=True)
    assert np.array_equal(wa[10:].view(np.uint32), wb.view(np.uint32))
    assert not np.array_equal(wa[10], wa[-1])     # the dictionary learned in between
