"""The device columns against the reference's own acceptance fixture,
tests/LCATest/input/correct.pvp: LCATest's Reconstruction for each of 40
steps, compared by its RequireAllZeroActivityProbe at nnzThreshold 1e-6.

The Python column (openpv_amd.lca) and the C++ host (pvrun on a params file
with LCATest's network) run the EXACT and FP32 forms to that threshold;
BF16 to the bound written at its test.
"""
import os
import subprocess

import numpy as np
import pytest

import lcatest_column as LC
from openpv_amd import abi, pvp

pytestmark = pytest.mark.gpu
FMAX = float(np.finfo(np.float32).max)


def _python_column(arith, fused, steps=40):
    """LCATest's column on the device through the Python host, with the
    product's own time-scale controller; the Reconstruction of each step."""
    from openpv_amd.lca import LCAColumn, run_column
    from openpv_amd.timescale import AdaptiveTimeScaleController
    frames, V0, W = LC.inputs()
    col = LCAColumn(frames[:1], W, 64, 32, 24, 18, 18, VThresh=0.025, AMin=0.0, AMax=FMAX,
                    AShift=0.0, VWidth=0.025, timeConstantTau=100.0, dt=1.0, arith=arith,
                    fused=fused, V0=V0, normalize="exact", device="cuda")
    ctrl = AdaptiveTimeScaleController(1, 0.55, 0.5, 0.1, 0.01)
    recon = []
    run_column(col, steps, frames=[f[None] for f in frames], display_period=20, adaptive=ctrl,
               energy_cost="firm", on_step=lambda t, c, e: recon.append(c.rec_A.cpu().numpy().copy()))
    return np.stack(recon).reshape(steps, 64, 128, 1)


@pytest.mark.parametrize("arith,fused", [(abi.ARITH_EXACT, False), (abi.ARITH_EXACT, True),
                                         (abi.ARITH_FP32, False), (abi.ARITH_FP32, True)])
def test_python_column_meets_lcatest_threshold(arith, fused):
    print("LCATest column, arith", arith, "fused", fused, flush=True)
    got = _python_column(arith, fused)
    err = np.abs(got - LC.correct()).reshape(40, -1).max(1)
    print("max |device - correct.pvp| per step:", err, flush=True)
    assert (err <= LC.NNZ_THRESHOLD).all(), np.nonzero(err > LC.NNZ_THRESHOLD)[0] + 1
    if arith == abi.ARITH_EXACT:
        ref, _ = LC.oracle_run(40)
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


# BF16 delivery (weights and the residual rounded to BF16, FP32 accumulation
# and state): the largest error of any step's Reconstruction, relative to the
# largest |value| of correct.pvp.
BF16_BOUND = 2e-2


def test_python_column_bf16_bound():
    got = _python_column(abi.ARITH_BF16, True)
    want = LC.correct()
    err = np.abs(got - want).max() / np.abs(want).max()
    print("BF16: max |device - correct.pvp| / max|correct| = %.3e" % err, flush=True)
    assert err <= BF16_BOUND


def _pvrun(params, cwd, outdir, steps=40):
    print("pvrun", params, flush=True)
    r = subprocess.run([abi.pvrun_path(), "-p", str(params), "-o", str(outdir), "--steps", str(steps)],
                       cwd=cwd, capture_output=True, text=True, timeout=300)
    print(r.stdout[-1500:], r.stderr[-1500:], flush=True)
    return r


@pytest.mark.parametrize("form", ["host", "fp32", "fp32_vertices"])
def test_pvrun_lcatest_params(tmp_path, form):
    """pvrun on LCATest's network (tests/golden/lcatest/LCATest.params): the
    RequireAllZeroActivityProbe against correct.pvp passes (exit status 0) and
    the written Reconstruction is within 1e-6 of correct.pvp at every step.
    "host": receiveGpu = false everywhere (the EXACT forms; bit-identical to
    the oracle as well); "fp32": receiveGpu = true on the gather and the
    scatter; "fp32_vertices": also updateGpu = true on V1 (the transfer in
    the reference GPU kernel's vertex form)."""
    params = os.path.join(LC.LCATEST, "LCATest.params")
    if form != "host":
        text = open(params).read()
        n = text.count("receiveGpu = false;")
        assert n == 3
        # the TransposeConn's and CloneConn's deliveries (the HyPerConn delivers nothing)
        text = text.replace("receiveGpu = false;\n    updateGSynFromPostPerspective = true;",
                            "receiveGpu = true;\n    updateGSynFromPostPerspective = true;")
        text = text.replace('originalConnName = "V1ToResidual";\n    channelCode = 0;\n    receiveGpu = false;\n'
                            '    updateGSynFromPostPerspective = false;',
                            'originalConnName = "V1ToResidual";\n    channelCode = 0;\n    receiveGpu = true;\n'
                            '    updateGSynFromPostPerspective = false;')
        assert text.count("receiveGpu = true;") == 2
        if form == "fp32_vertices":
            assert text.count("updateGpu = false;") == 1
            text = text.replace("updateGpu = false;", "updateGpu = true;")
        params = str(tmp_path / ("lcatest_%s.params" % form))
        open(params, "w").write(text)
    r = _pvrun(params, LC.LCATEST, tmp_path)
    assert r.returncode == 0, r.stderr[-800:]
    _, times, rec = pvp.read_activity(str(tmp_path / "Reconstruction.pvp"))
    assert list(times) == list(range(1, 41))
    err = np.abs(rec - LC.correct()).reshape(40, -1).max(1)
    print(form, "max |pvrun - correct.pvp| per step:", err, flush=True)
    assert (err <= LC.NNZ_THRESHOLD).all()
    lines = open(tmp_path / "ComparisonTest.txt").read().split("\n")
    assert sum("numNonzero = 0" in ln for ln in lines) == 40
    if form == "host":
        ref, _ = LC.oracle_run(40)
        assert np.array_equal(rec.view(np.uint32), ref.view(np.uint32))
