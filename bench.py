"""LCA steps/s + synaptic ops/s on the batched-LCA workload, BF16-MFMA vs FP32.

Workload (the metric's configuration): nbatch 64 grayscale 128x128x1 images
with a 1/f^2 power spectrum (random phase, unit variance), V1 32x32x512 with
16x16 patches at stride 4, VThresh 0.15, tau 100, dt 1, soft threshold.  One
step is one LCA timestep of the Input -> Residual <-> V1 -> Recon column over
the whole batch: the Residual update, the V1 gather (post perspective) with the
LCA update fused, and the Recon scatter (pre perspective: the dense GEMM in the
FP32, BF16 and FP8 forms, the ordered sparse scatter after the ordered active
list in EXACT), with Recon's update fused.  Steps are replayed from a hipGraph
that holds several of them (graph_steps), so the gap between two graph
launches is paid once per graph, not once per step.

Forms timed: EXACT (bit-identical to the reference CPU order), FP32 (FMA) and
BF16 (MFMA gather, FP32 state).  `value` is the BF16 form's steps/s; the line
also carries the FP32 forms, synaptic ops/s, the BF16 energy error against
EXACT, the sparsity difference, a roofline record for the dominant kernel and
a CPU baseline (the oracle on the host cores, bounded sample).

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

NB, NX, NY, NF = 64, 128, 128, 1
NXV, NYV, NFV, NXP = 32, 32, 512, 16
VTHRESH, TAU, DT = 0.15, 100.0, 1.0
HBM_PEAK_GBS = 8000.0
MFMA_BF16_PEAK_TFLOPS = 2500.0


def one_over_f_images(nb, ny, nx, seed):
    """1/f^2 power spectrum (amplitude 1/f), random phase, zero mean, unit variance."""
    rng = np.random.default_rng(seed)
    fy = np.fft.fftfreq(ny)[:, None]
    fx = np.fft.fftfreq(nx)[None, :]
    f = np.sqrt(fx * fx + fy * fy)
    amp = np.zeros_like(f)
    amp[f > 0] = 1.0 / f[f > 0]
    out = np.empty((nb, ny, nx, 1), np.float32)
    for b in range(nb):
        phase = np.exp(2j * np.pi * rng.random((ny, nx)))
        img = np.real(np.fft.ifft2(amp * phase))
        img -= img.mean()
        img /= img.std()
        out[b, :, :, 0] = img
    return out


def dictionary(nf, nxp, nfp, seed):
    """~N(0,1) then L2-normalised per feature in FP64."""
    rng = np.random.default_rng(seed)
    W = rng.standard_normal((nf, nxp, nxp, nfp))
    W /= np.sqrt((W.reshape(nf, -1) ** 2).sum(1)).reshape(-1, 1, 1, 1)
    return W.astype(np.float32)


def scatter_macs(col):
    """Sum over active V1 neurons of their (shrunken) patch size: the pre-perspective
    scatter's synaptic operations for the current active lists."""
    import ctypes as C
    from openpv_amd import abi
    lp, lq = col.v1_loc, col.rec_loc
    dims = []
    for axis, n in ((0, lp.nx_ext), (1, lp.ny_ext)):
        d = np.zeros(n, np.int32)
        abi.host().pvhost_patch_axis(C.byref(lp), C.byref(lq), NXP, NXP, axis,
                                     d.ctypes.data_as(C.c_void_p), None, None, None, None)
        dims.append(d)
    e, n = col.active_list().host_entries()
    total = 0
    for b in range(lp.nbatch):
        idx = e[b, :n[b]]["index"].astype(np.int64) // lp.nf
        xe, ye = idx % lp.nx_ext, idx // lp.nx_ext
        total += int((dims[0][xe] * dims[1][ye]).sum()) * lq.nf
    return total


def graph_steps(steps):
    """Steps captured per hipGraph: the largest of 25, 20, 10, 5, 4, 2, 1 that
    divides the timed steps (each replay advances the column by that many)."""
    return next(k for k in (25, 20, 10, 5, 4, 2, 1) if steps % k == 0)


def time_form(arith, images, W, steps, warmup, torch, tuner=None):
    from openpv_amd.lca import LCAColumn
    col = LCAColumn(images, W, NXV, NYV, NFV, NXP, NXP, VThresh=VTHRESH, timeConstantTau=TAU,
                    dt=DT, arith=arith, fused=True, device="cuda")
    s = torch.cuda.Stream()
    torch.cuda.synchronize()
    shapes = col.autotune(tuner, stream=s) if tuner is not None else {}
    torch.cuda.synchronize()
    k = graph_steps(steps)
    col.capture(s, steps=k)
    # warm-up: whole graphs, then the remainder as eager steps (the same launches)
    for _ in range(warmup // k):
        col.replay(s)
    for _ in range(warmup % k):
        col.step(s)
    s.synchronize()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(s)
    t0 = time.perf_counter()
    for _ in range(steps // k):
        col.replay(s)
    e1.record(s)
    s.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = e0.elapsed_time(e1)
    return col, s, ms / steps, wall, shapes, k


ACCURACY_STEPS = 500   # the workload's run length


def accuracy(forms, images, W, torch, steps=ACCURACY_STEPS):
    """Each form against EXACT over the workload's own run (500 steps from the
    same start, one hipGraph replay per step): the relative error of the total
    energy (worst image) at the last step, its largest value over all steps
    and over steps 50..500, and the relative difference of the active counts
    at the last step (worst image).  Independent of --steps / --warmup."""
    from openpv_amd import abi
    from openpv_amd.lca import LCAColumn
    arith = {"exact": abi.ARITH_EXACT, "fp32": abi.ARITH_FP32, "bf16": abi.ARITH_BF16,
             "fp8": abi.ARITH_FP8}
    traj = {}
    for f in ["exact"] + [x for x in forms if x != "exact"]:
        col = LCAColumn(images, W, NXV, NYV, NFV, NXP, NXP, VThresh=VTHRESH, timeConstantTau=TAU,
                        dt=DT, arith=arith[f], fused=True, device="cuda")
        s = torch.cuda.Stream()
        torch.cuda.synchronize()
        col.capture(s)
        energy = []
        for _ in range(steps):
            col.replay(s)
            s.synchronize()
            energy.append(col.probe_energy(stream=s).cpu().numpy())
        traj[f] = (np.array(energy), col.num_active().astype(np.float64))
        col.release_graph()
        del col
        torch.cuda.empty_cache()
    e_ref, n_ref = traj["exact"]
    out = {"steps": steps}
    for f in forms:
        if f == "exact":
            continue
        e, n = traj[f]
        rel = np.abs(e - e_ref).max(1) / np.abs(e_ref).max(1)
        out[f] = {"energy_rel_err_final": float(rel[-1]), "energy_rel_err_max": float(rel.max()),
                  "energy_rel_err_max_step": int(rel.argmax()) + 1,
                  "energy_rel_err_max_after_50": float(rel[49:].max()),
                  "active_rel_diff_final": float(np.max(np.abs(n - n_ref) / np.maximum(n_ref, 1)))}
    out["active_mean_final"] = float(n_ref.mean())
    return out


def step_kernel_timing(col, s, n, torch):
    """Average and minimum durations of the fused gather + LCA-update kernel
    (recv_post_mfma_persist_kernel<4, GatherBF16, 8> for the BF16 form) and of
    the Recon scatter (recv_pre_dense_kernel for the BF16 form) as the step
    runs them: n steps enqueued eagerly on the column's stream, each kernel
    bracketed by HIP events recorded on that stream, so every kernel runs
    between the same neighbours as in the replayed graph (the Residual update
    before the gather, the scatter after it).  Returns ((avg, min) gather,
    (avg, min) scatter) in ms."""
    ev = [tuple(torch.cuda.Event(enable_timing=True) for _ in range(4)) for _ in range(n)]
    with torch.cuda.stream(s):
        for g0, g1, c0, c1 in ev:
            col.step(s, gather_events=(g0, g1), scatter_events=(c0, c1))
    s.synchronize()
    torch.cuda.synchronize()
    dg = [g0.elapsed_time(g1) for g0, g1, _, _ in ev]
    dc = [c0.elapsed_time(c1) for _, _, c0, c1 in ev]
    return (float(np.mean(dg)), float(np.min(dg))), (float(np.mean(dc)), float(np.min(dc)))


TRAFFIC_SUMMARY = os.path.join("profiles", "lca64_bf16_traffic.md")


def profiled_traffic():
    """HBM bytes per launch of the fused BF16 gather on this workload, measured
    by two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE) scaled by the same
    counters on a 256 MiB copy; the summary file records the commands."""
    try:
        with open(os.path.join(ROOT, TRAFFIC_SUMMARY)) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(line.split(":")[1])
    except OSError:
        pass
    return None


def host_cores():
    """Threads for the host baseline and what the machine reports: the cores
    this process may run on, capped by OMP_NUM_THREADS where the job sets it
    (the job's CPU share), plus nproc, the cgroup CPU quota and the CPU model."""
    affinity = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS")
    threads = min(affinity, int(omp)) if omp and omp.isdigit() and int(omp) > 0 else affinity
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, period = f.read().split()[:2]
            quota = None if q == "max" else round(int(q) / int(period), 2)
    except (OSError, ValueError):
        pass
    return threads, {"nproc": os.cpu_count(), "affinity_cores": affinity, "omp_num_threads": omp,
                     "cgroup_cpu_quota": quota, "cpu_model": model}


def cpu_baseline(images, W, budget_s=15.0):
    """The oracle (C restatement of the reference loops, OpenMP over the host
    cores) on a bounded sample of the workload: 4 of its images for as many
    steps as fit in about budget_s seconds."""
    import oracle as O
    cores, machine = host_cores()
    nb = 4
    cfg = O.LCAConfig()
    mres = O.lib().or_requiredConvolveMargin(NX, NXV, O.lib().or_calcPostPatchSize(NXP, NXV, NX))
    mv1 = O.lib().or_requiredConvolveMargin(NXV, NX, NXP)
    cfg.input = O.Loc.make(nb, NX, NY, NF, 0)
    cfg.res = O.Loc.make(nb, NX, NY, NF, mres)
    cfg.v1 = O.Loc.make(nb, NXV, NYV, NFV, mv1)
    cfg.nxp = cfg.nyp = NXP
    fmax = float(np.finfo(np.float32).max)
    cfg.VThresh, cfg.AMin, cfg.AMax, cfg.AShift, cfg.VWidth = VTHRESH, 0.0, fmax, VTHRESH, 0.0
    cfg.transfer_vertices = 0
    cfg.tau, cfg.dt = TAU, DT
    cfg.selfInteract = 1
    cfg.errScale = 1.0
    cfg.num_threads = cores
    img = np.ascontiguousarray(images[:nb])
    O.lca_run(cfg, W, img, None, 1)  # warm (page-in, thread pool)
    steps, elapsed = 1, 0.0
    while True:
        t0 = time.perf_counter()
        O.lca_run(cfg, W, img, None, steps)
        elapsed = time.perf_counter() - t0
        if elapsed >= budget_s or steps >= 1000:
            break
        steps = max(steps + 1, int(steps * min(4.0, budget_s / max(elapsed, 1e-3))))
    per_image_step = elapsed / (steps * nb)
    out = {"value": 1.0 / (per_image_step * NB), "unit": "LCA steps/s", "cores": cores,
           "kind": "port",
           "sample": "oracle LCA loop, %d images x %d steps of the 64-image workload (%.1f s), "
                     "scaled to the 64-image batch" % (nb, steps, elapsed)}
    out.update(machine)
    return out


def copy_bandwidth(nbytes, torch, reps=20):
    """GB/s of a plain device-to-device copy (torch's copy kernel) reading
    nbytes/2 and writing nbytes/2: what a streaming kernel of that size reaches
    on this device, beside the published peak."""
    n = int(nbytes) // 8
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del a, b
    torch.cuda.empty_cache()
    return 8.0 * n / (ms * 1e-3) / 1e9


def _time_events(fn, s, reps, torch):
    """Mean and min duration (ms) of fn() on stream s, each call bracketed by HIP events."""
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(reps)]
    with torch.cuda.stream(s):
        fn()
        for a, b in evs:
            a.record(s)
            fn()
            b.record(s)
    s.synchronize()
    d = [a.elapsed_time(b) for a, b in evs]
    return float(np.mean(d)), float(np.min(d))


def c1_delivery(torch, reps=10):
    """Full-size delivery case: pre 256x256x16 halo 6 -> post 128x128x256,
    pre-perspective patch 6x6x256 (post-perspective 12x12x16), W ~ N(0, 0.05^2).
    Dense gather (pre ~N(0,1)) and sparse scatter (2 % of pre neurons, Exp(1))."""
    import ctypes as C
    from openpv_amd import abi, runtime as R
    dev = "cuda"
    rng = np.random.default_rng(3)
    pre = R.loc(1, 256, 256, 16, 6)
    post = R.loc(1, 128, 128, 256, 0)
    nkx, nky = C.c_int(), C.c_int()
    abi.host().pvhost_num_kernels(C.byref(pre), C.byref(post), C.byref(nkx), C.byref(nky))
    nk = nkx.value * nky.value * pre.nf
    Wpre = (rng.standard_normal((nk, 6, 6, 256)) * 0.05).astype(np.float32)
    Wpost = np.zeros((256, 12, 12, 16), np.float32)
    abi.host().pvhost_transpose_weights(C.byref(pre), C.byref(post), 6, 6, 256,
                                        Wpre.ctypes.data_as(C.c_void_p), Wpost.ctypes.data_as(C.c_void_p))
    A = np.zeros((1, pre.ny_ext, pre.nx_ext, 16), np.float32)
    A[:, 6:262, 6:262] = rng.standard_normal((1, 256, 256, 16))
    As = np.zeros_like(A)
    v = np.zeros(256 * 256 * 16, np.float32)
    idx = rng.choice(v.size, size=v.size // 50, replace=False)
    v[idx] = rng.exponential(1.0, idx.size).astype(np.float32)
    As[:, 6:262, 6:262] = v.reshape(1, 256, 256, 16)
    dA = torch.as_tensor(A.ravel()).to(dev)
    dAs = torch.as_tensor(As.ravel()).to(dev)
    g = torch.zeros(post.num_res, device=dev)
    s = torch.cuda.Stream()
    out = {"workload": "pre 256x256x16 (halo 6) -> post 128x128x256, 12x12x16 post patch",
           "gather_macs": 128 * 128 * 256 * 12 * 12 * 16, "gather_ms": {}, "scatter_ms": {}}
    ref = None
    for name, ar in (("exact", abi.ARITH_EXACT), ("fp32", abi.ARITH_FP32), ("bf16", abi.ARITH_BF16),
                     ("fp8", abi.ARITH_FP8)):
        d = R.PostsynapticPerspectiveGPUDelivery(pre, post, 12, 12, torch.as_tensor(Wpost.ravel()).to(dev),
                                                 arith=ar)
        d.set_source_scale(R.fp8_scale_for(np.abs(A).max()))
        ms, _ = _time_events(lambda: d.deliver(dA, g, stream=s), s, reps, torch)
        out["gather_ms"][name] = round(ms, 4)
        gh = g.cpu().numpy().astype(np.float64)
        if ref is None:
            ref = gh
        else:
            out.setdefault("gather_rel_err_vs_exact", {})[name] = float(
                np.sqrt(((gh - ref) ** 2).mean() / (ref ** 2).mean()))
    act = R.ActiveList(pre, dev)
    act.compact(dAs)
    torch.cuda.synchronize()
    out["scatter_active"] = int(act.num_active.cpu().numpy()[0])
    out["scatter_macs"] = out["scatter_active"] * 6 * 6 * 256
    for name, ar in (("exact", abi.ARITH_EXACT), ("bf16", abi.ARITH_BF16)):
        d = R.PresynapticPerspectiveGPUDelivery(pre, post, 6, 6, torch.as_tensor(Wpre.ravel()).to(dev),
                                                arith=ar)
        ms, _ = _time_events(lambda: d.deliver(g, active=act, pre_activity=dAs, stream=s), s, reps, torch)
        out["scatter_ms"][name] = round(ms, 4)
    ms_b = out["gather_ms"]["bf16"]
    out["gather_bf16_tflops"] = round(2.0 * out["gather_macs"] / (ms_b * 1e-3) / 1e12, 1)
    out["gather_fp8_tflops"] = round(2.0 * out["gather_macs"] / (out["gather_ms"]["fp8"] * 1e-3) / 1e12, 1)
    # compulsory bytes of the sparse scatter: active list, weights, GSyn read + write
    sbytes = 8 * out["scatter_active"] + 4 * Wpre.size + 8 * post.num_res
    out["scatter_exact_gbs"] = round(sbytes / (out["scatter_ms"]["exact"] * 1e-3) / 1e9, 1)
    return out


def c4_hebbian(torch, reps=5, tuner=None):
    """HebbianUpdater dW at full size: V1 8 x 128x128x1024 (5 % active, Exp(1))
    x Residual 8 x 512x512x1 (~N(0,1)), 16x16 patches at stride 4, MFMA forms."""
    import ctypes as C
    from openpv_amd import abi, runtime as R
    dev = "cuda"
    H = abi.host()
    mv = H.pvhost_required_margin(128, 512, 16)
    mr = H.pvhost_required_margin(512, 128, H.pvhost_post_patch_size(16, 128, 512))
    pre = R.loc(8, 128, 128, 1024, mv)
    post = R.loc(8, 512, 512, 1, mr)
    g = torch.Generator(device=dev)
    g.manual_seed(4)
    a = torch.zeros(pre.nbatch * pre.num_ext, device=dev)
    a.view(8, pre.ny_ext, pre.nx_ext, 1024)[:, mv:mv + 128, mv:mv + 128] = torch.where(
        torch.rand(8, 128, 128, 1024, device=dev, generator=g) < 0.05,
        torch.empty(8, 128, 128, 1024, device=dev).exponential_(1.0, generator=g), 0.0)
    r = torch.zeros(post.nbatch * post.num_ext, device=dev)
    r.view(8, post.ny_ext, post.nx_ext)[:, mr:mr + 512, mr:mr + 512] = torch.randn(
        8, 512, 512, device=dev, generator=g)
    hb = R.HebbianUpdater(pre, post, 16, 16, 1e-3, normalize_dw=False, device=dev)
    s = torch.cuda.Stream()
    flops = 2.0 * 1024 * 256 * pre.nbatch * pre.nx_ext * pre.ny_ext
    out = {"workload": "V1 8x128x128x1024 (halo %d) x Residual 8x512x512x1, 16x16 stride 4" % mv,
           "gflop_per_update": round(flops / 1e9, 1)}
    dws = {}
    for name, ar in (("fp32", abi.ARITH_FP32), ("bf16", abi.ARITH_BF16)):
        if tuner is not None:
            shape = tuner.tune_hebbian(hb, a, r, ar, stream=s)
            out[name + "_rows_per_chunk"] = shape[0] if shape else None
        ms, _ = _time_events(lambda: hb.update_dw(a, r, arith=ar, stream=s), s, reps, torch)
        out[name + "_ms"] = round(ms, 4)
        out[name + "_tflops"] = round(flops / (ms * 1e-3) / 1e12, 1)
        dws[name] = hb.dw.cpu().numpy().astype(np.float64)
    out["bf16_rel_err_vs_fp32"] = float(np.sqrt(((dws["bf16"] - dws["fp32"]) ** 2).mean()
                                                / (dws["fp32"] ** 2).mean()))
    # activity read once per update in FP32: the HBM floor of this kernel
    out["pre_activity_bytes"] = int(4 * pre.nbatch * pre.num_ext)
    return out


def c4_learning(torch, steps=300, updates=3):
    """Large batched LCA + dictionary learning: 8 images 512x512x1 (1/f^2),
    1024 features 16x16 stride 4 (V1 128x128x1024), VThresh 0.15, BF16 then
    FP8 deliveries with FP32 state; every 300 steps one weight update (BF16
    MFMA Hebbian dW at dWMax 1e-3 with normalizeDw, HebbianUpdater's default,
    FP64 renormalisation per feature, the deliveries' weights re-derived), the
    step and the update each
    replayed from a hipGraph.  Timed: `updates` periods of `steps` steps plus
    their update, with HIP events on the column's stream."""
    from openpv_amd import abi
    from openpv_amd.lca import LCAColumn
    images = one_over_f_images(8, 512, 512, seed=7)
    W = dictionary(1024, 16, 1, seed=8)
    out = {"workload": "8 x 512x512x1 -> 128x128x1024, 16x16 stride 4, %d steps per weight update" % steps}
    for name, ar in (("bf16", abi.ARITH_BF16), ("fp8", abi.ARITH_FP8)):
        col = LCAColumn(images, W, 128, 128, 1024, 16, 16, VThresh=VTHRESH, timeConstantTau=TAU, dt=DT,
                        arith=ar, fused=True, device="cuda")
        col.enable_learning(1e-3, normalize_dw=True, renormalize="fp64")
        s = torch.cuda.Stream()
        col.update_weights(s)          # warm-up: the updater's scratch, outside any capture
        s.synchronize()
        torch.cuda.synchronize()
        k = graph_steps(steps)
        col.capture(s, steps=k)
        col.capture_learning(s)
        col.replay(s)
        col.replay_learning(s)
        s.synchronize()
        e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
        e0.record(s)
        for _ in range(updates):
            for _ in range(steps // k):
                col.replay(s)
            col.replay_learning(s)
        e1.record(s)
        s.synchronize()
        total = e0.elapsed_time(e1)
        # the update alone
        e0.record(s)
        for _ in range(5):
            col.replay_learning(s)
        e2.record(s)
        s.synchronize()
        upd = e0.elapsed_time(e2) / 5
        energy = col.probe_energy(stream=s)
        s.synchronize()
        energy = energy.cpu().numpy()
        norms = col.weights.double().view(1024, -1).pow(2).sum(1).sqrt()
        out[name] = {"ms_per_step_with_learning": round(total / (updates * steps), 5),
                     "steps_per_graph": k,
                     "steps_per_s_with_learning": round(updates * steps * 1000.0 / total, 2),
                     "weight_update_ms": round(upd, 4),
                     "energy_after": float(energy.sum()),
                     "max_norm_error_fp64": float((norms - 1).abs().max().item())}
        col.release_graph()
        del col
        torch.cuda.empty_cache()
    return out


def c2_momentum(torch, reps=20):
    """Momentum forms on the LCA workload's V1 (64 x 32x32x512): the
    MomentumLCA V update against the HyPerLCA one (both reading GSyn, reading
    and writing V, writing A, A(t-1) from V; momentum also reads and writes
    prevDrive: 16 vs 24 bytes per neuron), and the MomentumUpdater weight step
    at the Hebbian workload's 1024 x 16x16 weights (w, dW and prev_dW read and
    written, counts read: 32 bytes per weight)."""
    from openpv_amd import abi, runtime as R
    dev = "cuda"
    loc = R.loc(64, 32, 32, 512, 2)
    n = loc.nbatch * loc.num_res
    g = torch.Generator(device=dev)
    g.manual_seed(5)
    G = torch.randn(n, device=dev, generator=g) * 0.1
    V = torch.randn(n, device=dev, generator=g) * 0.1
    A = torch.zeros(loc.nbatch * loc.num_ext, device=dev)
    t = R.make_transfer(VTHRESH, 0.0, float(np.finfo(np.float32).max), VTHRESH, 0.0)
    R.ann_activity(loc, t, V, A)
    s = torch.cuda.Stream()
    out = {"workload": "V1 64x32x32x512 (halo 2); weights 1024x16x16"}
    plain = R.HyPerLCAUpdate(loc, t, TAU, DT, True, activity_from_v=True, device=dev)
    mom = R.MomentumLCAUpdate(loc, t, TAU, DT, True, LCAMomentumRate=0.4, activity_from_v=True,
                              device=dev)
    for name, upd, nbytes in (("lca", plain, 16), ("momentum_lca", mom, 24)):
        ms, _ = _time_events(lambda: upd.update(G, V, A, stream=s), s, reps, torch)
        out[name + "_us"] = round(1000 * ms, 2)
        out[name + "_gbs"] = round(nbytes * n / (ms * 1e-3) / 1e9, 1)
    nw = 1024 * 256
    w = torch.randn(nw, device=dev, generator=g)
    dw = torch.randn(nw, device=dev, generator=g)
    prev = torch.zeros(nw, device=dev)
    cnt = torch.ones(nw, dtype=torch.int64, device=dev)
    f, wf = R.momentum_factors("viscosity", 100.0, 0.0, 1.0)
    import ctypes as C
    ptr = lambda x: C.c_void_p(x.data_ptr())

    def step():
        abi.check(abi.hip().pvhip_momentum_apply(ptr(w), ptr(dw), ptr(prev), ptr(cnt), nw, 1, f, wf,
                                                 0, C.c_void_p(s.cuda_stream)), "momentum_apply")
    ms, _ = _time_events(step, s, reps, torch)
    out["momentum_apply_us"] = round(1000 * ms, 2)
    out["momentum_apply_gbs"] = round(32 * nw / (ms * 1e-3) / 1e9, 1)
    return out


def c3_adjoint(torch, T=100, checkpoint=10):
    """Adjoint workload: 16 images 64x64x3 ~N(0,1), 256 features
    8x8x3 stride 2 (V1 32x32x256), 100 unrolled steps, checkpoint every 10;
    E = 0.5||x - W^T a||^2 + 0.1||a||_1.  FP32 forms; forward and backward timed."""
    from openpv_amd import abi
    from openpv_amd.adjoint import LCAAdjoint
    rng = np.random.default_rng(5)
    x = rng.standard_normal((16, 64, 64, 3)).astype(np.float32)
    x -= x.mean(axis=(1, 2, 3), keepdims=True)
    W = rng.standard_normal((256, 8, 8, 3))
    W = (W / np.sqrt((W.reshape(256, -1) ** 2).sum(1)).reshape(-1, 1, 1, 1)).astype(np.float32)
    adj = LCAAdjoint(16, 64, 64, 3, 32, 32, 256, 8, 8, VThresh=0.1, timeConstantTau=100.0,
                     arith=abi.ARITH_FP32, checkpoint=checkpoint, device="cuda")
    adj.forward(x, W, 5)  # warm-up (tables, shadows, scratch)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    E, ck = adj.forward(x, W, T)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    gW, gx = adj.backward(x, T, ck)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    return {"workload": "16 x 64x64x3 -> 32x32x256, 8x8 stride 2, %d steps, checkpoint %d"
                        % (T, checkpoint),
            "forward_ms": round((t1 - t0) * 1e3, 2), "backward_ms": round((t2 - t1) * 1e3, 2),
            "energy": float(E.sum().item()), "grad_w_norm": float(gW.norm().item())}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--forms", default="exact,fp32,bf16,fp8")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the library's default launch shapes (no autotuner)")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the delivery (C1) and Hebbian (C4) measurements")
    ap.add_argument("--profile-copy", type=int, default=0, metavar="MIB",
                    help="also run a plain device copy of MIB MiB (profiler counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU: the library has no multi-GPU path (the batch would "
              "shard over GPUs with one dW all-reduce per update, not built here).",
              file=sys.stderr)
        sys.exit(2)
    import torch
    from openpv_amd import abi
    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(2)
    abi.hip()
    from openpv_amd.autotune import LaunchTuner
    images = one_over_f_images(NB, NY, NX, seed=1)
    W = dictionary(NFV, NXP, NF, seed=2)
    forms = [f for f in args.forms.split(",") if f]
    arith = {"exact": abi.ARITH_EXACT, "fp32": abi.ARITH_FP32, "bf16": abi.ARITH_BF16,
             "fp8": abi.ARITH_FP8}
    res = {}
    cols = {}
    # launch shapes measured afresh on this machine (no cached choices)
    tuner = None if args.no_autotune else LaunchTuner(cache=None)
    for f in forms:
        col, s, ms, wall, shapes, k = time_form(arith[f], images, W, args.steps, args.warmup, torch,
                                                tuner)
        res[f] = {"ms_per_step": ms, "steps_per_s": 1000.0 / ms, "wall_s": wall, "steps_per_graph": k}
        if shapes:
            res[f]["launch_shapes"] = shapes
        cols[f] = (col, s)
        print("# %s: %.4f ms/step (%.1f steps/s)" % (f, ms, 1000.0 / ms), file=sys.stderr,
              flush=True)
    head = "bf16" if "bf16" in res else forms[-1]
    col, s = cols[head]
    gather_macs = NB * NXV * NYV * NFV * NXP * NXP * NF
    smacs = scatter_macs(col)
    ms = res[head]["ms_per_step"]
    # energy and sparsity of each form against EXACT over the workload's 500 steps
    extra = {}
    if "exact" in cols and head != "exact":
        acc = accuracy([f for f in forms if f != "exact"], images, W, torch)
        for f in res:
            if f in acc:
                res[f]["energy_rel_err_vs_exact"] = acc[f]["energy_rel_err_final"]
                res[f]["energy_rel_err_vs_exact_max_over_steps"] = acc[f]["energy_rel_err_max"]
                res[f]["active_rel_diff_vs_exact"] = acc[f]["active_rel_diff_final"]
        extra["accuracy"] = acc
        extra["energy_rel_err_vs_exact"] = acc[head]["energy_rel_err_final"]
        extra["energy_rel_err_vs_exact_max_over_steps"] = acc[head]["energy_rel_err_max"]
        extra["active_rel_diff_vs_exact"] = acc[head]["active_rel_diff_final"]
        extra["accuracy_steps"] = acc["steps"]
        extra["active_mean"] = acc["active_mean_final"]
    for f in res:
        res[f]["steps_per_s"] = round(res[f]["steps_per_s"], 2)
        res[f]["ms_per_step"] = round(res[f]["ms_per_step"], 5)
        res[f]["wall_s"] = round(res[f]["wall_s"], 3)
    # dominant kernel: fused gather + LCA update of the headline form
    (kavg, kmin), (sms, smin) = step_kernel_timing(col, s, 50, torch)
    n_v1 = NB * NXV * NYV * NFV
    src_bytes = NB * col.res_loc.num_ext * 4
    w_bytes = NFV * NXP * NXP * NF * (2 if head == "bf16" else 4)
    # V read + write, then A written in FP32 (4 B) or, when the scatter reads
    # V1's chunked BF16 copy, only that copy (2 B); source; weights
    alg_bytes = n_v1 * (10 if col.v1_A16 is not None else 12) + src_bytes + w_bytes
    achieved = alg_bytes / (kavg * 1e-3) / 1e9
    copy_gbs = copy_bandwidth(alg_bytes, torch)
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            # a plain device copy moving the same bytes (half read, half written)
            "copy_gbs": round(copy_gbs, 1), "frac_of_copy": round(achieved / copy_gbs, 4),
            "traffic": profiled_traffic() if head == "bf16" else None,
            "kernel": ("recv_post_mfma_persist_kernel<4, GatherBF16, 8> (gather + fused LCA update)"
                       if head == "bf16" else "recv_post (gather + fused LCA update)"),
            "kernel_ms": round(kavg, 5), "kernel_ms_min": round(kmin, 5),
            "algorithmic_bytes": int(alg_bytes),
            "traffic_source": TRAFFIC_SUMMARY if head == "bf16" else None,
            "mfma_tflops": round(2.0 * gather_macs / (kavg * 1e-3) / 1e12, 1)}
    line = {
        "metric": "LCA steps/s + synaptic ops/s, BF16-MFMA vs FP32 GPU at <=1e-3 rel energy err",
        "value": round(1000.0 / ms, 2), "unit": "LCA steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": head if head != "exact" else "fp32",
        "data": "synthetic (1/f^2 images, random N(0,1) L2-normalised dictionary)",
        "config": {"workload": "batched LCA: 64 x 128x128x1 -> 32x32x512, 16x16 stride 4, "
                               "VThresh 0.15, one hipGraph per timestep",
                   "global_batch": NB, "parallelism": "1 GPU"},
        "synaptic_ops_per_s": round((gather_macs + smacs) * 1000.0 / ms, 1),
        "gather_macs_per_step": gather_macs, "scatter_macs_per_step": smacs,
        "forms": res,
    }
    if "fp32" in res and head == "bf16":
        line["bf16_vs_fp32"] = round(res["fp32"]["ms_per_step"] / res["bf16"]["ms_per_step"], 3)
    if "exact" in res and head == "bf16":
        line["bf16_vs_exact"] = round(res["exact"]["ms_per_step"] / res["bf16"]["ms_per_step"], 3)
    line.update(extra)
    line["roofline"] = roof
    if col.scatter.mode == "dense":
        # the dense scatter: a GEMM of V1's 64 x 32x32 positions by the 16x16
        # patch offsets over the 512 features, on the BF16 matrix cores
        sflop = 2.0 * NB * NXV * NYV * NXP * NXP * NFV
        # algorithmic bytes: V1's BF16 activity read once, the BF16 weights,
        # Recon's V and A written; at the peaks the bytes take longer than
        # the flops (9.5 us against 6.9 us), so HBM bounds the kernel
        sbytes = 2 * NB * NXV * NYV * NFV + 2 * NXP * NXP * NFV + 8 * NB * (4 * NXV) * (4 * NYV)
        line["scatter_kernel"] = {
            "kernel": "recv_pre_dense_kernel<false, 16, 8, 4, 2> (BF16 GEMM, overlap-add, Recon update)",
            "kernel_ms": round(sms, 5), "kernel_ms_min": round(smin, 5), "bound": "hbm",
            "achieved": round(sbytes / (sms * 1e-3) / 1e9, 1), "peak": 8000.0, "unit": "GB/s",
            "frac": round(sbytes / (sms * 1e-3) / 1e9 / 8000.0, 4), "algorithmic_bytes": sbytes,
            "mfma_tflops": round(sflop / (sms * 1e-3) / 1e12, 1),
            "mfma_frac": round(sflop / (sms * 1e-3) / 1e12 / 2500.0, 4)}
    if not args.no_extras:
        line["delivery_c1"] = c1_delivery(torch)
        line["hebbian_c4"] = c4_hebbian(torch, tuner=tuner)
        line["learning_c4"] = c4_learning(torch)
        line["momentum_c2"] = c2_momentum(torch)
        line["adjoint_c3"] = c3_adjoint(torch)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(images, W)
        line["gpu_vs_cpu_exact"] = (round(res["exact"]["steps_per_s"] / line["cpu_baseline"]["value"], 1)
                                    if "exact" in res else None)
    print(json.dumps(line), flush=True)
    if args.profile_copy:
        # a plain copy kernel of known size: FETCH_SIZE / WRITE_SIZE of this
        # dispatch calibrate the counters of the kernels above
        n = args.profile_copy * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print("# copy calibration: %d bytes read, %d bytes written per dispatch" % (4 * n, 4 * n),
              file=sys.stderr)


if __name__ == "__main__":
    main()
