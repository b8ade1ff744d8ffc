"""The dense scatter through the .params drop-in: pvrun builds the LCA
column of the batched workload's geometry (128x128x1 images, V1 32x32x128,
16x16 patches at stride 4, VThresh 0.15, tau 100, fixed dt) with
gpuArithmetic "bf16", "fp8" or "fp32"; the CloneConn V1 -> Reconstruction then
runs pvhip_recv_pre_dense (BF16: V1's update keeps its chunked BF16 activity;
FP32: it reads V1's FP32 activity), and the reconstructions stay within the
BF16 / FP8 / FP32 bounds of the oracle column's (EXACT) over 30 steps.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle_column as OC
from openpv_amd import abi, pvp

pytestmark = pytest.mark.gpu

PARAMS = """
HyPerCol "column" = {
    nx = 128;
    ny = 128;
    nbatch = 2;
    dt = 1;
    stopTime = %(steps)d;
    outputPath = "output/";
    gpuArithmetic = "%(form)s";
};

PvpLayer "Input" = {
    nxScale = 1;   nyScale = 1;   nf = 1;
    phase = 0;
    inputPath = "images.pvp";
    displayPeriod = 0;
    writeStep = -1;
};

ANNErrorLayer "Residual" = {
    nxScale = 1;   nyScale = 1;   nf = 1;
    phase = 1;
    errScale = 1;
    VThresh = -infinity;
    writeStep = -1;
};

HyPerLCALayer "V1" = {
    nxScale = 0.25; nyScale = 0.25; nf = 128;
    phase = 2;
    sparseLayer = true;
    VThresh = 0.15;  AMin = 0;  AMax = infinity;  AShift = 0.15;  VWidth = 0;
    timeConstantTau = 100;
    selfInteract = true;
    writeStep = -1;
};

ANNLayer "Reconstruction" = {
    nxScale = 1;   nyScale = 1;   nf = 1;
    phase = 3;
    VThresh = -infinity;  AMin = -infinity;  AMax = infinity;  AShift = 0;  VWidth = 0;
    writeStep = 1;
    initialWriteTime = 1;
};

IdentConn "InputToResidual" = {
    preLayerName = "Input";   postLayerName = "Residual";
    channelCode = 0;
};

HyPerConn "V1ToResidual" = {
    preLayerName = "V1";   postLayerName = "Residual";
    channelCode = -1;
    plasticityFlag = false;
    sharedWeights = true;
    weightInitType = "FileWeight";
    initWeightsFile = "weights.pvp";
    nxp = 16;  nyp = 16;  nfp = 1;
    receiveGpu = true;
    updateGSynFromPostPerspective = false;
    pvpatchAccumulateType = "convolve";
};

TransposeConn "ResidualToV1" = {
    preLayerName = "Residual";   postLayerName = "V1";
    originalConnName = "V1ToResidual";
    channelCode = 0;
    receiveGpu = true;
    updateGSynFromPostPerspective = true;
    pvpatchAccumulateType = "convolve";
};

CloneConn "V1ToReconstruction" = {
    preLayerName = "V1";   postLayerName = "Reconstruction";
    originalConnName = "V1ToResidual";
    channelCode = 0;
    receiveGpu = true;
    updateGSynFromPostPerspective = false;
    pvpatchAccumulateType = "convolve";
};

IdentConn "ReconstructionToResidual" = {
    preLayerName = "Reconstruction";   postLayerName = "Residual";
    channelCode = 1;
};
"""


@pytest.mark.parametrize("form,tol", [("bf16", 5e-3), ("fp8", 5e-2), ("fp32", 1e-4)])
def test_pvrun_dense_scatter_column(tmp_path, form, tol):
    import bench
    steps = 30
    images = bench.one_over_f_images(2, 128, 128, seed=31)
    W = bench.dictionary(128, 16, 1, seed=32)
    pvp.write_activity(str(tmp_path / "images.pvp"), images, [0.0, 0.0])
    pvp.write_weights(str(tmp_path / "weights.pvp"), W[None], [0.0])
    (tmp_path / "column.params").write_text(PARAMS % {"steps": steps, "form": form})
    print("pvrun", form, "dense column", flush=True)
    r = subprocess.run([abi.pvrun_path(), "-p", "column.params", "-o", str(tmp_path / "out"), "--steps",
                        str(steps)], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    print(r.stdout[-1500:], r.stderr[-1500:], flush=True)
    assert r.returncode == 0, r.stderr[-800:]
    assert "V1ToReconstruction: dense %s scatter" % ("FP32" if form == "fp32" else "BF16") in r.stdout
    _, times, rec = pvp.read_activity(str(tmp_path / "out" / "Reconstruction.pvp"))
    assert list(times) == [float(t) for t in range(1, steps + 1) for _ in range(2)]   # two batch elements
    rec = rec.reshape(steps, 2, 128, 128, 1)
    col = OC.OracleColumn(images, W, 32, 32, 128, 16, 0.15)
    refs = []
    for t in range(steps):
        col.step()
        refs.append(col.rc.reshape(rec[t].shape).copy())
    refs = np.stack(refs)
    scale = np.abs(refs).max()   # the reconstruction grows from 0: one scale for all steps
    err = np.abs(rec - refs).reshape(steps, -1).max(1)
    print("%s: max |pvrun - oracle| per step %s (scale %g)" % (form, np.array2string(err, precision=2), scale),
          flush=True)
    assert (err <= tol * scale).all()


def test_pvrun_fused_column_with_activity_readers(tmp_path):
    """In the BF16 column pvrun fuses every layer's deliveries into its update
    (Residual: both IdentConn in one pass; V1: the gather; Reconstruction: the
    dense scatter).  Without readers V1's update writes only its BF16 copy;
    with V1's activity written every step the update writes the FP32 activity
    too.  Both runs give the same reconstructions bit for bit, and V1's
    written activity is the oracle column's within the BF16 bound."""
    import bench
    steps, nb = 12, 2
    images = bench.one_over_f_images(nb, 128, 128, seed=33)
    W = bench.dictionary(128, 16, 1, seed=34)
    pvp.write_activity(str(tmp_path / "images.pvp"), images, [0.0, 0.0])
    pvp.write_weights(str(tmp_path / "weights.pvp"), W[None], [0.0])
    base = PARAMS % {"steps": steps, "form": "bf16"}
    v1_block = base.split('HyPerLCALayer "V1"')[1].split("};")[0]
    reader = base.replace(v1_block, v1_block.replace("writeStep = -1;", "writeStep = 1;\n    initialWriteTime = 1;"))
    assert reader != base
    recs, outs = [], []
    for name, params in (("base", base), ("reader", reader)):
        d = tmp_path / name
        d.mkdir()
        for f in ("images.pvp", "weights.pvp"):
            (d / f).write_bytes((tmp_path / f).read_bytes())
        (d / "column.params").write_text(params)
        print("pvrun bf16 fused column", name, flush=True)
        r = subprocess.run([abi.pvrun_path(), "-p", "column.params", "-o", str(d / "out"), "--steps", str(steps)],
                           cwd=d, capture_output=True, text=True, timeout=300)
        print(r.stdout[-1500:], r.stderr[-1500:], flush=True)
        assert r.returncode == 0, r.stderr[-800:]
        for msg in ("Residual: deliveries InputToResidual, ReconstructionToResidual and the update in one pass",
                    "V1: gather ResidualToV1 with the LCA update fused",
                    "Reconstruction: dense scatter V1ToReconstruction with the layer update fused"):
            assert msg in r.stdout
        recs.append(pvp.read_activity(str(d / "out" / "Reconstruction.pvp"))[2])
        outs.append(d / "out")
    assert np.array_equal(recs[0].view(np.uint32), recs[1].view(np.uint32))
    _, times, a1 = pvp.read_activity(str(outs[1] / "V1.pvp"))
    assert list(times) == [float(t) for t in range(1, steps + 1) for _ in range(nb)]
    a1 = a1.reshape(steps, nb, 32, 32, 128)
    col = OC.OracleColumn(images, W, 32, 32, 128, 16, 0.15)
    m = (int(round(np.sqrt(col.v1.num_ext / 128))) - 32) // 2
    refs = []
    for _ in range(steps):
        col.step()
        refs.append(col.A.reshape(nb, 32 + 2 * m, 32 + 2 * m, 128)[:, m:m + 32, m:m + 32, :].copy())
    refs = np.stack(refs)
    scale = np.abs(refs).max()
    err = np.abs(a1 - refs).reshape(steps, -1).max(1)
    print("V1 activity: max |pvrun - oracle| per step %s (scale %g)"
          % (" ".join("%.2e" % e for e in err), scale), flush=True)
    assert scale > 0 and (err <= 5e-3 * scale).all()


def _run(d, params, steps):
    (d / "column.params").write_text(params)
    r = subprocess.run([abi.pvrun_path(), "-p", "column.params", "-o", str(d / "out"), "--steps", str(steps)],
                       cwd=d, capture_output=True, text=True, timeout=300)
    print(r.stdout[-1200:], r.stderr[-1200:], flush=True)
    assert r.returncode == 0, r.stderr[-800:]
    return r


def test_pvrun_fused_column_probes_and_restart(tmp_path):
    """The BF16 fused column with checkpoints every 6 steps, without readers of
    V1's FP32 activity (derived from V on request: the checkpoint asks for it)
    and with an energy probe (0.5*|Residual|^2 + VThresh*|V1|_1, read every
    step; V1's update then writes it): the energies follow the oracle
    column's within the BF16 bound, both runs give the same reconstructions
    bit for bit, and each run restarted from its Checkpoint06 reproduces its
    steps 7..12 bit for bit."""
    import bench
    steps, nb = 12, 2
    images = bench.one_over_f_images(nb, 128, 128, seed=35)
    W = bench.dictionary(128, 16, 1, seed=36)
    def checkpointed(text, ck):
        return text.replace('    outputPath = "output/";',
                            '    outputPath = "output/";\n    checkpointWrite = true;\n'
                            '    checkpointWriteDir = "%s";\n    checkpointWriteTriggerMode = "step";\n'
                            '    checkpointWriteStepInterval = 6;' % ck)

    def restarted(text, ck):
        return text.replace("    checkpointWrite = true;",
                            '    checkpointWrite = false;\n    initializeFromCheckpointDir = "%s";'
                            % (ck / "Checkpoint06"))

    plain = PARAMS % {"steps": steps, "form": "bf16"}
    ckb, ck = tmp_path / "ckb", tmp_path / "ck"
    base = checkpointed(plain, ckb)
    probes = checkpointed(plain, ck)
    probes += """
ColumnEnergyProbe "Energy" = {
    probeOutputFile = "Energy.txt";
};
L2NormProbe "ResidualNorm" = {
    targetLayer = "Residual";
    energyProbe = "Energy";
    coefficient = 0.5;
    exponent = 2;
};
L1NormLCAProbe "Sparsity" = {
    targetLayer = "V1";
    energyProbe = "Energy";
};
"""
    recs = {}
    for name, params in (("base", base), ("probes", probes), ("restart_base", restarted(base, ckb)),
                         ("restart", restarted(probes, ck))):
        d = tmp_path / name
        d.mkdir()
        pvp.write_activity(str(d / "images.pvp"), images, [0.0, 0.0])
        pvp.write_weights(str(d / "weights.pvp"), W[None], [0.0])
        print("pvrun bf16 fused column", name, flush=True)
        _run(d, params, steps)
        recs[name] = pvp.read_activity(str(d / "out" / "Reconstruction.pvp"))[1:]
    assert np.array_equal(recs["base"][1].view(np.uint32), recs["probes"][1].view(np.uint32))
    for name in ("restart_base", "restart"):
        tb, rb = recs[name]
        assert list(tb) == [float(t) for t in range(7, steps + 1) for _ in range(nb)]
        assert np.array_equal(recs["base"][1][6 * nb:].view(np.uint32), rb.view(np.uint32)), name
    rows = [l.split(",") for l in (tmp_path / "probes" / "out" / "Energy.txt").read_text().splitlines()
            if l.strip()]
    e = {(round(float(t)), int(b)): float(v) for t, b, v in rows}
    col = OC.OracleColumn(images, W, 32, 32, 128, 16, 0.15)
    worst = 0.0
    for t in range(1, steps + 1):
        col.step()
        ref = col.energy()
        for b in range(nb):
            worst = max(worst, abs(e[(t, b)] - ref[b]) / abs(ref[b]))
    print("energy: largest relative difference to the oracle %.2e" % worst, flush=True)
    assert worst <= 1e-3
