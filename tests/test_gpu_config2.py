"""The batched LCA workload at full size against the accuracy bar of its
low-precision forms.

Workload: 64 grayscale 128x128x1 images with a 1/f^2 power spectrum (random
phase, unit variance) -> V1 32x32x512, 16x16 patches at stride 4, VThresh
0.15, tau 100, dt 1, soft threshold, 500 timesteps, each form replayed from
one hipGraph per timestep.

Bar, against the EXACT form (the reference CPU order, bit-identical to the
oracle) on the same inputs:
  FP32  total energy within 1e-5 relative at every step;
  BF16  total energy within 1e-3 relative at every step, and the active
        count within 0.5 % after the 500 steps;
  FP8   (e4m3 weights and residual) the same bar as BF16 after the 500
        steps; its error over the first few tens of steps is above 1e-3
        (the e4m3 rounding of the residual while it is large), which the
        every-step check records as an expected failure.
"""
import functools

import numpy as np
import pytest
import torch

from openpv_amd import abi

pytestmark = pytest.mark.gpu
STEPS = 500


@functools.lru_cache(maxsize=None)
def _problem():
    import bench
    return bench.one_over_f_images(64, 128, 128, seed=1), bench.dictionary(512, 16, 1, seed=2)


@functools.lru_cache(maxsize=None)
def _trajectory(form):
    """Per-step total energy [STEPS][64] and active counts [STEPS][64] of one
    form; the column's step is one hipGraph replay."""
    from openpv_amd.lca import LCAColumn
    arith = {"exact": abi.ARITH_EXACT, "fp32": abi.ARITH_FP32, "bf16": abi.ARITH_BF16,
             "fp8": abi.ARITH_FP8}[form]
    images, W = _problem()
    print("C2 column", form, STEPS, "steps", flush=True)
    col = LCAColumn(images, W, 32, 32, 512, 16, 16, VThresh=0.15, timeConstantTau=100.0, dt=1.0,
                    arith=arith, fused=True, device="cuda")
    s = torch.cuda.Stream()
    torch.cuda.synchronize()
    col.capture(s)
    energy, active = [], []
    for t in range(STEPS):
        col.replay(s)
        s.synchronize()
        energy.append(col.probe_energy().cpu().numpy())
        active.append(col.num_active())
    col.release_graph()
    del col
    torch.cuda.empty_cache()
    return np.array(energy), np.array(active, np.int64)


def _errors(form):
    e_ref, n_ref = _trajectory("exact")
    e, n = _trajectory(form)
    rel_e = np.abs(e - e_ref).max(1) / np.abs(e_ref).max(1)     # per step, worst image
    rel_n = np.abs(n - n_ref) / np.maximum(n_ref, 1)             # [step][image]
    print("%s: energy rel err max over steps %.3e (step %d), at step %d %.3e; active rel diff "
          "at step %d %.3e (mean active %.0f)"
          % (form, rel_e.max(), rel_e.argmax() + 1, STEPS, rel_e[-1], STEPS, rel_n[-1].max(),
             n_ref[-1].mean()), flush=True)
    return rel_e, rel_n


def test_fp32_energy_every_step():
    rel_e, rel_n = _errors("fp32")
    assert rel_e.max() <= 1e-5
    assert rel_n[-1].max() <= 5e-3


def test_bf16_energy_every_step_and_final_sparsity():
    rel_e, rel_n = _errors("bf16")
    assert rel_e.max() <= 1e-3
    assert rel_n[-1].max() <= 5e-3


def test_fp8_energy_and_sparsity_after_500_steps():
    rel_e, rel_n = _errors("fp8")
    assert rel_e[-1] <= 1e-3
    assert rel_n[-1].max() <= 5e-3


@pytest.mark.xfail(reason="FP8 (e4m3) residual rounding: energy error above 1e-3 during the "
                          "first tens of steps", strict=False)
def test_fp8_energy_every_step():
    rel_e, _ = _errors("fp8")
    assert rel_e.max() <= 1e-3


def test_exact_energy_descends():
    """EXACT's total energy falls over the run for every image and its code is
    sparse (some, not all, of V1's 524,288 neurons per image active)."""
    e, n = _trajectory("exact")
    assert np.isfinite(e).all() and (e[-1] < 0.5 * e[0]).all()
    assert (n[-1] > 0).all() and (n[-1] < 32 * 32 * 512 // 10).all()
