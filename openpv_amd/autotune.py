"""Launch-shape autotuner for the delivery and plasticity kernels.

Some kernels take a launch shape that changes only their speed, not their
results beyond the order of FP32 sums:
- the channel GEMM of `pvhip_recv_post` (FP32/BF16 forms with several data
  patches per feature): `cgemm_tile` window starts per workgroup (64, 128 or
  256) and `cgemm_splits` channel ranges;
- the MFMA forms of `pvhip_hebbian_dw`: `rows_per_chunk` split-K rows;
- the FP32/BF16 forms of a pre-perspective delivery: its dense MFMA route
  (the transposed gather) or the ordered sparse scatter over the active list
  in FP32, whichever is faster on the caller's current activity (the choice
  depends on the density, so it is made on representative data).

The tuner times each candidate on the caller's own buffers with HIP events
(a few warm-up launches, then the median of the timed ones), keeps the
fastest and applies it.  Results are cached in a JSON file keyed by the
device, the ABI version and the geometry, so a shape is measured once per
machine.  Default cache: `$OPENPV_AMD_TUNE_CACHE`, else
`~/.cache/openpv_amd/launch_shapes.json`; `cache=None` keeps it in memory.
"""
import json
import os

import numpy as np
import torch

from . import abi

DEFAULT_CACHE = os.environ.get(
    "OPENPV_AMD_TUNE_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "openpv_amd",
                                          "launch_shapes.json"))


def _loc_key(l):
    return "%dx%dx%dx%d+%d.%d.%d.%d" % (l.nbatch, l.nx, l.ny, l.nf, l.halo.lt, l.halo.rt, l.halo.dn,
                                        l.halo.up)


class LaunchTuner:
    def __init__(self, cache=DEFAULT_CACHE, warmup=2, reps=7):
        self.cache_path = cache
        self.warmup, self.reps = warmup, reps
        self.table = {}
        if cache and os.path.exists(cache):
            try:
                with open(cache) as f:
                    self.table = json.load(f)
            except (OSError, ValueError):
                self.table = {}
        self.log = []   # (key, {candidate: ms}, best) of the shapes measured by this tuner

    # ------------------------------------------------------------------ helpers
    def _device(self):
        return torch.cuda.get_device_name(torch.cuda.current_device())

    def _time(self, launch, stream):
        """Median milliseconds of `reps` launches after `warmup` launches."""
        for _ in range(self.warmup):
            launch()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in range(self.reps)]
        for e0, e1 in ev:
            e0.record(stream)
            launch()
            e1.record(stream)
        stream.synchronize()
        return float(np.median([e0.elapsed_time(e1) for e0, e1 in ev]))

    def _save(self):
        if not self.cache_path:
            return
        try:
            os.makedirs(os.path.dirname(self.cache_path), exist_ok=True)
            tmp = self.cache_path + ".tmp%d" % os.getpid()
            with open(tmp, "w") as f:
                json.dump(self.table, f, indent=1, sort_keys=True)
            os.replace(tmp, self.cache_path)
        except OSError:
            pass   # a read-only home: keep the shapes in memory

    def _pick(self, key, candidates, apply, launch, stream):
        if key in self.table:
            best = self.table[key]["best"]
            apply(*best)
            return tuple(best)
        times = {}
        for c in candidates:
            apply(*c)
            times[c] = self._time(launch, stream)
        best = min(times, key=times.get)
        apply(*best)
        self.table[key] = {"best": list(best), "ms": {",".join(map(str, c)): t for c, t in times.items()}}
        self.log.append((key, times, best))
        self._save()
        return best

    # ------------------------------------------------------------------ kernels
    def tune_recv_post(self, delivery, pre_activity, gsyn, stream=None):
        """Channel GEMM launch shape of a PostsynapticPerspectiveGPUDelivery (or
        of a presynaptic delivery's transposed route).  Returns (tile, splits),
        or None when the delivery has no channel GEMM."""
        d = getattr(delivery, "transposed", None) or delivery
        if not hasattr(d, "set_launch") or d.inverse is None or d.arith == abi.ARITH_EXACT:
            return None
        ck = 32 if d.arith == abi.ARITH_BF16 else 16
        chunks = d.nfp // ck
        stream = stream or torch.cuda.current_stream()
        key = "recv_post|%s|abi%d|arith%d|%s|%s|%dx%d" % (
            self._device(), abi.hip().pvhip_abi_version(), d.arith, _loc_key(d.source),
            _loc_key(d.target), d.nxp, d.nyp)
        cands = [(t, k) for t in (64, 128, 256) for k in (1, 2, 4, 8) if k <= chunks]
        return self._pick(key, cands, d.set_launch,
                          lambda: d.deliver(pre_activity, gsyn, accumulate=False, stream=stream), stream)

    def tune_hebbian(self, updater, pre_activity, post_activity, arith, stream=None):
        """Split-K rows of a HebbianUpdater's MFMA form.  Returns (rows,) or None."""
        if updater.windows is None or arith == abi.ARITH_EXACT:
            return None
        stream = stream or torch.cuda.current_stream()
        rows = updater.pre.nbatch * updater.pre.ny_ext
        key = "hebbian|%s|abi%d|arith%d|%s|%s|%dx%d" % (
            self._device(), abi.hip().pvhip_abi_version(), arith, _loc_key(updater.pre),
            _loc_key(updater.post), updater.nxp, updater.nyp)
        # 8 to 128 chunks (the partials take chunks x dW of scratch)
        cands = [(0,)] + sorted({((rows + c - 1) // c,) for c in (8, 16, 32, 64, 128) if c <= rows})

        def apply(r):
            updater.rows_per_chunk = r
        best = self._pick(key, cands, apply,
                          lambda: updater.update_dw(pre_activity, post_activity, clear=True,
                                                    arith=arith, stream=stream), stream)
        return best

    def tune_recv_pre(self, delivery, gsyn, active, pre_activity, stream=None):
        """Mode of a PresynapticPerspectiveGPUDelivery (FP32/BF16 forms): the
        dense GEMM ("dense", when its geometry and, in BF16, the source's BF16
        copy are there), the transposed route's channel GEMM ("route") or the ordered
        sparse scatter ("ordered").  Returns the mode, or None when there is no
        choice."""
        if delivery.arith == abi.ARITH_EXACT or active is None:
            return None
        cands = []
        if delivery._dense_ready():
            cands.append(("dense",))
        if delivery.transposed is not None:
            cands.append(("route",))
            self.tune_recv_post(delivery.transposed, pre_activity, gsyn, stream=stream)
        if not cands:
            return None
        cands.append(("ordered",))
        stream = stream or torch.cuda.current_stream()
        key = "recv_pre|%s|abi%d|arith%d|%s|%s|%dx%d|%s" % (
            self._device(), abi.hip().pvhip_abi_version(), delivery.arith, _loc_key(delivery.pre),
            _loc_key(delivery.post), delivery.nxp, delivery.nyp, ",".join(c[0] for c in cands))

        def apply(mode):
            delivery.mode = mode
        best = self._pick(key, cands, apply,
                          lambda: delivery.deliver(gsyn, active=active, pre_activity=pre_activity,
                                                   stream=stream), stream)
        return best[0]
