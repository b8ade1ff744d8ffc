"""The LCA column of the reference's sparse-coding graphs on the device.

Graph (tests/GPULCATest/input/GPULCATest.params, tutorials/LCA-CIFAR):

    phase 0  Input            (images, constant over the run)
    phase 1  Residual         ANNErrorLayer   GSyn0 = Input (IdentConn),
                                              GSyn1 = Recon(t-1) (IdentConn)
    phase 2  V1               HyPerLCALayer   GSyn0 = TransposeConn gather of Residual
                                              (post perspective, receiveGpu)
    phase 3  Recon            ANNLayer        GSyn0 = CloneConn scatter of sparse V1
                                              (pre perspective)

One `step()` is HyPerCol::advanceTime (src/columns/HyPerCol.cpp:659-791) for
this graph with a fixed dt.  All state stays in HBM; a step is a fixed
sequence of libpvhip launches on one stream, so it can be captured once as a
hipGraph and replayed.
"""
import ctypes as C

import numpy as np
import torch

from . import abi
from .runtime import (ActiveList, HyPerLCAUpdate, PostsynapticPerspectiveGPUDelivery,
                      PresynapticPerspectiveGPUDelivery, activity_to_bf16c, bf16c_buffer,
                      cost_sum_f64, error_update, fill_zero, layer_update_args, loc,
                      make_transfer, _stream)

FLT_MAX = float(np.finfo(np.float32).max)


class LCAColumn:
    """Input -> Residual <-> V1 -> Recon on one device.

    weights: V1->Residual pre-perspective shared weights [nf_v1][nyp][nxp][nf_in]
    (the HyPerConn the TransposeConn and CloneConn share).
    arith: delivery arithmetic of the gather and the scatter (EXACT, FP32, BF16).
    fused: apply the V1 update in the gather epilogue.
    """

    def __init__(self, images, weights, nx_v1, ny_v1, nf_v1, nxp, nyp, VThresh=0.1, AMin=0.0,
                 AMax=FLT_MAX, AShift=None, VWidth=0.0, timeConstantTau=100.0, dt=1.0,
                 self_interact=True, err_scale=1.0, vertices=False, arith=abi.ARITH_EXACT,
                 fused=False, V0=None, res_halo=None, v1_halo=None, normalize=None,
                 bf16_copy=True, device="cuda"):
        """normalize: NormalizeL2 of the weights on initialisation (strength 1):
        "exact" (the reference's float arithmetic), "fp64", or None.
        bf16_copy: in the BF16/FP8 forms, let V1's update write the chunked BF16
        copy of its activity that the scatter's channel GEMM reads (where the
        geometry allows); False keeps the FP32 activity as the scatter's source."""
        images = torch.as_tensor(images, dtype=torch.float32)
        nb, ny, nx, nf = images.shape
        self.device = torch.device(device)
        H = abi.host()
        if res_halo is None:
            # TransposeConn Residual -> V1 with post-perspective patch: margin of Residual
            nxpT = H.pvhost_post_patch_size(nxp, nx_v1, nx)
            res_halo = H.pvhost_required_margin(nx, nx_v1, nxpT)
        if v1_halo is None:
            v1_halo = H.pvhost_required_margin(nx_v1, nx, nxp)
        self.in_loc = loc(nb, nx, ny, nf, 0)
        self.res_loc = loc(nb, nx, ny, nf, res_halo)
        self.rec_loc = loc(nb, nx, ny, nf, 0)
        self.v1_loc = loc(nb, nx_v1, ny_v1, nf_v1, v1_halo)
        d = self.device
        f = dict(dtype=torch.float32, device=d)
        self.inputA = images.reshape(-1).to(d).contiguous()
        self.res_V = torch.zeros(nb * self.res_loc.num_res, **f)
        self.res_A = torch.zeros(nb * self.res_loc.num_ext, **f)
        self.v1_gsyn = torch.zeros(nb * self.v1_loc.num_res, **f)
        self.v1_V = torch.zeros(nb * self.v1_loc.num_res, **f)
        self.v1_A = torch.zeros(nb * self.v1_loc.num_ext, **f)
        self.rec_gsyn = torch.zeros(nb * self.rec_loc.num_res, **f)
        self.rec_V = torch.zeros(nb * self.rec_loc.num_res, **f)
        self.rec_A = torch.zeros(nb * self.rec_loc.num_res, **f)
        self.weights = torch.as_tensor(weights, dtype=torch.float32).reshape(-1).to(d).contiguous()
        if normalize is not None:
            npatch = nf_v1
            psize = self.weights.numel() // npatch
            H2 = abi.hip()
            if normalize == "exact":
                abi.check(H2.pvhip_normalize_l2(C.c_void_p(self.weights.data_ptr()), npatch, psize,
                                                1.0, 0.0, None), "pvhip_normalize_l2")
            elif normalize == "fp64":
                abi.check(H2.pvhip_normalize_l2_f64(C.c_void_p(self.weights.data_ptr()), npatch,
                                                    psize, 1.0, 0.0, None), "pvhip_normalize_l2_f64")
            else:
                raise ValueError("normalize must be 'exact', 'fp64' or None")
            torch.cuda.synchronize()
        if AShift is None:
            AShift = VThresh
        # probe coefficient: the layer's float VThresh widened to double (L1NormLCAProbe)
        self.VThresh = float(np.float32(VThresh))
        self.VWidth = float(np.float32(VWidth))
        self.v1_transfer = make_transfer(VThresh, AMin, AMax, AShift, VWidth, vertices)
        self.none_transfer = make_transfer(-FLT_MAX, -FLT_MAX, FLT_MAX, 0.0, 0.0, False)
        self.err_scale = float(err_scale)
        self.arith = arith
        self.fused = fused
        # TransposeConn: post weights are the original's pre weights, post-perspective
        # patch = nxp x nyp x nf_in (TransposeWeightsPair.cpp:141-147).
        self.gather = PostsynapticPerspectiveGPUDelivery(self.res_loc, self.v1_loc, nxp, nyp,
                                                         self.weights, arith=arith)
        if arith == abi.ARITH_FP8:
            # the residual err_scale*(x - recon) stays within err_scale*2*max|x| once the
            # reconstruction approximates x; its FP8 scale is fixed from that bound
            from .runtime import fp8_scale_for
            bound = 2.0 * abs(self.err_scale) * float(images.abs().max())
            self.gather.set_source_scale(fp8_scale_for(bound))
        self.scatter = PresynapticPerspectiveGPUDelivery(self.v1_loc, self.rec_loc, nxp, nyp,
                                                         self.weights, arith=arith)
        self.active = ActiveList(self.v1_loc, d)
        # V1's activity is written only by its own update and starts as T(V0).
        self.lca = HyPerLCAUpdate(self.v1_loc, self.v1_transfer, timeConstantTau, dt,
                                  self_interact, activity_from_v=True, device=d)
        self.energy = torch.zeros(2, nb, dtype=torch.float64, device=d)
        # BF16 forms whose scatter reads the dense activity (the dense GEMM or
        # the transposed route's channel GEMM): V1's update writes its activity
        # as chunked BF16, which the scatter reads instead of converting the
        # FP32 activity (the same BF16 values, half the bytes); V1's halo is
        # never written, so the dense GEMM multiplies restricted positions only.
        # Such a step writes no FP32 activity: nothing in it reads one (the next
        # update recomputes A(t-1) = T(V) from V), so v1_A is derived from V, bit
        # for bit, when a reader asks for it (v1_activity()).
        self.v1_A16 = None
        self._a_stale = False
        if bf16_copy and arith in (abi.ARITH_BF16, abi.ARITH_FP8) and nf_v1 % 32 == 0:
            buf = bf16c_buffer(self.v1_loc, d)
            if self.scatter.bind_pre_bf16c(buf, halo_is_zero=True):
                self.v1_A16 = buf
        # the FP32 dense scatter reads v1_A itself: restricted positions only
        self.scatter.pre_halo_is_zero = True
        # Recon = ANNLayer with the scatter as its only channel: the dense
        # route's channel GEMM applies this update in its epilogue
        self._rec_update = layer_update_args(self.rec_loc, self.rec_gsyn, self.rec_V, self.rec_A,
                                             self.none_transfer, channel_coeffs=(1.0,))
        self._graph = None
        self._fused_args = self._gather_args = self._gather_args_plain = None
        # V1's active list is published when a consumer reads it in the step (the
        # ordered scatter), else marked stale and rebuilt on request -- the
        # reference's publisher also rebuilds it lazily (Publisher.cpp:85-134)
        self._active_stale = False
        if fused:
            # argument blocks (and the delivery's scratch) prepared outside any graph capture
            self._fused_args = self.lca.args(None, self.v1_V, self.v1_A, num_channels=1)
            self._fused_args_lp = self.lca.args(
                None, self.v1_V, self.v1_A if self.v1_A16 is None else None, num_channels=1,
                A_bf16c=self.v1_A16)
            self._gather_args = self.gather.args(self.res_A, None, False, fused=self._fused_args,
                                                 active=self.active, halo_is_zero=True)
            self._gather_args_plain = self.gather.args(self.res_A, None, False,
                                                       fused=self._fused_args_lp)
        # InitializeState: V1 V from V0, A = T(V); Residual and Recon start at 0.
        if V0 is not None:
            self.v1_V.copy_(torch.as_tensor(V0, dtype=torch.float32).reshape(-1).to(d))
        from .runtime import ann_activity
        ann_activity(self.v1_loc, self.v1_transfer, self.v1_V, self.v1_A)
        self.active.compact(self.v1_A)
        self._sync_bf16c()

    def _sync_bf16c(self, stream=None):
        """Rebuild V1's chunked BF16 activity from its FP32 activity (after V1's
        state was written other than by a step)."""
        if self.v1_A16 is not None:
            activity_to_bf16c(self.v1_loc, self.v1_activity(stream), self.v1_A16, stream)

    def autotune(self, tuner, stream=None, warm_steps=20):
        """Launch shapes of the column's tunable kernels (the FP32/BF16 scatter's
        channel GEMM, and its dense route against the ordered sparse scatter),
        measured on the column's own buffers by an autotune.LaunchTuner after
        `warm_steps` steps (representative sparsity); the state those steps
        changed is restored afterwards.  Call before capture().  Returns
        {kernel: shape}."""
        saved = [t.clone() for t in (self.v1_V, self.v1_activity(stream), self.rec_A)]
        for _ in range(warm_steps):
            self.step(stream)
        if stream is not None:
            stream.synchronize()
        self._sync_bf16c(stream)
        out = self._tune(tuner, stream)
        torch.cuda.synchronize()
        for t, v in zip((self.v1_V, self.v1_A, self.rec_A), saved):
            t.copy_(v)
        self._a_stale = False
        self.active.compact(self.v1_A)
        self._sync_bf16c()
        torch.cuda.synchronize()
        return out

    def _tune(self, tuner, stream):
        out = {}
        shape = tuner.tune_recv_post(self.scatter, self.v1_activity(stream), self.rec_gsyn,
                                     stream=stream)
        if shape is not None:
            out["scatter_cgemm(tile,splits)"] = list(shape)
        mode = tuner.tune_recv_pre(self.scatter, self.rec_gsyn, self.active_list(stream),
                                   self.v1_activity(stream), stream=stream)
        if mode is not None:
            out["scatter_mode"] = mode
        shape = tuner.tune_recv_post(self.gather, self.res_A, self.v1_gsyn, stream=stream)
        if shape is not None:
            out["gather_cgemm(tile,splits)"] = list(shape)
        return out

    # one timestep ---------------------------------------------------------
    def step(self, stream=None, gather_events=None, scatter_events=None):
        """One timestep.  gather_events=(start, stop) torch events are recorded on
        `stream` around the V1 gather (+ fused update) launch, scatter_events
        around the Recon scatter (+ Recon's update)."""
        s = stream
        nb = self.in_loc.nbatch
        # phase 1: Residual = err_scale * (Input - Recon(t-1)): both identity
        # deliveries and the layer update in one pass (GSyn not materialised)
        error_update(self.res_loc, self.in_loc, self.inputA, self.rec_loc, self.rec_A, self.res_V,
                     self.res_A, self.none_transfer, channel_coeffs=(1.0, -1.0),
                     err_scale=self.err_scale, stream=s)
        # phase 2: V1 gather + LCA update (+ threshold), then its active list
        if gather_events is not None:
            gather_events[0].record(s)
        needs_list = not self.dense_scatter()
        if self.fused:
            # gather + LCA update (+ V1's active list, its halo is never written) in one call
            self.gather.launch(self._gather_args if needs_list else self._gather_args_plain, stream=s)
            if gather_events is not None:
                gather_events[1].record(s)
        else:
            self.gather.deliver(self.res_A, self.v1_gsyn, accumulate=False, stream=s)
            self.lca.update(self.v1_gsyn, self.v1_V, self.v1_A, stream=s,
                            A_bf16c=None if needs_list else self.v1_A16)
            if gather_events is not None:
                gather_events[1].record(s)
            if needs_list:
                self.active.compact(self.v1_A, stream=s)
        self._active_stale = not needs_list
        self._a_stale = self.fused and not needs_list and self.v1_A16 is not None
        # phase 3: Recon = scatter of the sparse V1 activity
        if scatter_events is not None:
            scatter_events[0].record(s)
        self.scatter.deliver(self.rec_gsyn, active=self.active, pre_activity=self.v1_A,
                             accumulate=False, stream=s, fused_layer=self._rec_update)
        if scatter_events is not None:
            scatter_events[1].record(s)

    # dictionary learning -----------------------------------------------------
    def enable_learning(self, dw_max, normalize_dw=False, arith=None, renormalize="fp64",
                        strength=1.0, momentum=None):
        """Plasticity of the dictionary (the V1 -> Residual HyPerConn that the
        gather's TransposeConn and the scatter's CloneConn share): at each
        update_weights(), HebbianUpdater's dW = dw_max * sum a_V1 (x) r over the
        batch (pre: V1's activity, post: the Residual's; -dw_max * dE/dW of
        E = 0.5|r|^2 with the activity held), W += dW, NormalizeL2 per feature
        (normalizeOnWeightUpdate; "fp64": the sum of squares and scale in FP64,
        "exact": the reference's float arithmetic, None: off), then the
        delivery copies of the weights are re-derived (WeightsPair::finalizeUpdate).
        arith: the dW correlation's form (EXACT: the reference's order; FP32 /
        BF16: the MFMA correlation); default EXACT for the EXACT column, else
        the column's own MFMA form (BF16 for FP8).  momentum: dict of
        MomentumUpdater parameters (momentumMethod, timeConstantTau,
        momentumDecay, momentumTau) or None."""
        from .runtime import HebbianUpdater, MomentumUpdater
        if arith is None:
            arith = {abi.ARITH_EXACT: abi.ARITH_EXACT, abi.ARITH_FP32: abi.ARITH_FP32}.get(
                self.arith, abi.ARITH_BF16)
        if renormalize not in ("fp64", "exact", None):
            raise ValueError("renormalize must be 'fp64', 'exact' or None")
        nxp, nyp = self.gather.nxp, self.gather.nyp
        if momentum is None:
            self.learner = HebbianUpdater(self.v1_loc, self.res_loc, nxp, nyp, dw_max,
                                          normalize_dw=normalize_dw, device=self.device)
        else:
            self.learner = MomentumUpdater(self.v1_loc, self.res_loc, nxp, nyp, dw_max,
                                           normalize_dw=normalize_dw, device=self.device, **momentum)
        if self.learner.size != self.weights.numel():
            raise abi.PVHIPError("the dictionary has %d weights, its updater %d"
                                 % (self.weights.numel(), self.learner.size))
        self.learn_arith = arith
        self.renormalize = renormalize
        self.strength = float(strength)
        self._learn_graph = None

    def update_weights(self, stream=None):
        """One weight update of the dictionary (HebbianUpdater::updateState with
        immediateWeightUpdate at a trigger step, HebbianUpdater.cpp:343-382,
        822-833; NormalizeL2, NormalizeL2.cpp:36-120; the copies the deliveries
        read), from the activities of the last step.  Device work only, so it
        can be captured (capture_learning)."""
        L = self.learner
        L.update_dw(self.v1_activity(stream), self.res_A, clear=True, arith=self.learn_arith,
                    stream=stream)
        L.apply(self.weights, stream=stream)
        npatch = self.v1_loc.nf
        psize = self.weights.numel() // npatch
        H = abi.hip()
        if self.renormalize == "fp64":
            abi.check(H.pvhip_normalize_l2_f64(C.c_void_p(self.weights.data_ptr()), npatch, psize,
                                               self.strength, 0.0, _stream(stream)),
                      "pvhip_normalize_l2_f64")
        elif self.renormalize == "exact":
            abi.check(H.pvhip_normalize_l2(C.c_void_p(self.weights.data_ptr()), npatch, psize,
                                           self.strength, 0.0, _stream(stream)), "pvhip_normalize_l2")
        self.refresh_weights(stream)

    def refresh_weights(self, stream=None):
        """Re-derive what the deliveries read from the dictionary: the gather's
        BF16 / FP8 shadow, the scatter route's transposed weights and their
        shadow (all on the device)."""
        self.gather.refresh_weights(stream)
        self.scatter.refresh_weights(stream)

    def capture_learning(self, stream):
        """Capture update_weights() as a hipGraph on `stream` (call it once
        eagerly first, so the updater's scratch exists)."""
        H = abi.hip()
        with torch.cuda.stream(stream):
            abi.check(H.pvhip_graph_begin(_stream(stream)), "pvhip_graph_begin")
            try:
                self.update_weights(stream)
            finally:
                g = C.c_void_p()
                st = H.pvhip_graph_end(_stream(stream), C.byref(g))
            abi.check(st, "pvhip_graph_end")
        self._learn_graph = g
        return g

    def replay_learning(self, stream):
        abi.check(abi.hip().pvhip_graph_launch(self._learn_graph, _stream(stream)),
                  "pvhip_graph_launch")
        self._a_stale = False

    def probe_energy(self, stream=None, cost="l1"):
        """ColumnEnergyProbe in FP64: E_b = 0.5*sum r^2 (L2NormProbe, exponent 2) +
        VThresh*cost(A) with cost "l1" (L1NormLCAProbe: sum |a|) or "firm"
        (FirmThresholdCostFnLCAProbe with the layer's VThresh and VWidth)."""
        cost_sum_f64(self.res_loc, self.res_A, True, 0, out=self.energy[0], stream=stream)
        if cost == "firm":
            cost_sum_f64(self.v1_loc, self.v1_activity(stream), True, 2, self.VThresh, self.VWidth,
                         out=self.energy[1], stream=stream)
        else:
            cost_sum_f64(self.v1_loc, self.v1_activity(stream), True, 1, out=self.energy[1],
                         stream=stream)
        if stream is None:
            return 0.5 * self.energy[0] + self.VThresh * self.energy[1]
        with torch.cuda.stream(stream):     # ordered after the sums on the same stream
            e = 0.5 * self.energy[0] + self.VThresh * self.energy[1]
        stream.synchronize()                # a reader on another stream sees the result
        return e

    def set_input(self, images):
        """New input frame(s) [nb][ny][nx][nf] (an input layer's retrieveInput)."""
        self.inputA.copy_(torch.as_tensor(images, dtype=torch.float32).reshape(-1))

    def set_dt(self, dt):
        """Per-batch dt of the V1 update for the coming steps (adaptive time scales)."""
        self.lca.set_dt(dt)

    def capture(self, stream, steps=1):
        """Capture `steps` consecutive steps as one hipGraph on `stream` (a
        non-default torch stream); each replay() advances the column by that
        many steps.  Several steps per graph spare the gap between two graph
        launches, which costs about as much as a short kernel."""
        if steps < 1:
            raise ValueError("capture: steps must be at least 1")
        H = abi.hip()
        with torch.cuda.stream(stream):
            abi.check(H.pvhip_graph_begin(_stream(stream)), "pvhip_graph_begin")
            try:
                for _ in range(steps):
                    self.step(stream)
            finally:
                g = C.c_void_p()
                st = H.pvhip_graph_end(_stream(stream), C.byref(g))
            abi.check(st, "pvhip_graph_end")
        self._graph = g
        self.graph_steps = steps
        self._graph_leaves_stale = self._active_stale   # host flags set while capturing
        self._graph_leaves_a_stale = self._a_stale
        return g

    def replay(self, stream):
        abi.check(abi.hip().pvhip_graph_launch(self._graph, _stream(stream)), "pvhip_graph_launch")
        self._active_stale = self._graph_leaves_stale
        self._a_stale = self._graph_leaves_a_stale

    def release_graph(self):
        if self._graph is not None:
            abi.hip().pvhip_graph_destroy(self._graph)
            self._graph = None
        if getattr(self, "_learn_graph", None) is not None:
            abi.hip().pvhip_graph_destroy(self._learn_graph)
            self._learn_graph = None

    # host views -----------------------------------------------------------
    def v1_state(self):
        return self.v1_V.cpu().numpy(), self.v1_activity().cpu().numpy()

    def v1_activity(self, stream=None):
        """V1's FP32 activity after the last step (derived here as T(V) when the
        step wrote only its chunked BF16 copy; the same bits the update gives)."""
        if self._a_stale:
            from .runtime import ann_activity
            ann_activity(self.v1_loc, self.v1_transfer, self.v1_V, self.v1_A, stream=stream)
            self._a_stale = False
        return self.v1_A

    def dense_scatter(self):
        """True when the Recon scatter reads V1's dense activity (the MFMA route)
        rather than its active list."""
        return self.scatter.use_route

    def active_list(self, stream=None):
        """V1's active list after the last step (rebuilt here when the step did
        not need it)."""
        if self._active_stale:
            self.active.compact(self.v1_activity(stream), stream=stream)
            self._active_stale = False
        return self.active

    def num_active(self, stream=None):
        n = self.active_list(stream).num_active
        if stream is not None:
            stream.synchronize()
        return n.cpu().numpy()


def run_column(col, steps, frames=None, display_period=0, adaptive=None, energy_cost="l1",
               on_step=None, stream=None):
    """HyPerCol::advanceTime for the column (src/columns/HyPerCol.cpp:659-791), one
    step per timestep t = 1 .. steps (dt 1):
      1. AdaptTimestep: with `adaptive` (timescale.AdaptiveTimeScaleController),
         the per-batch dt of the V1 update from the energy of step t-1, or from
         -1 when the input layer updates this step (its trigger);
      2. the input layer: every `display_period` steps starting at t = 1, the
         next frame of `frames` [n][nb][ny][nx][nf] (retrieveInput, then the
         index advances; InputActivityBuffer.cpp:614-660);
      3. the layer phases (LCAColumn.step);
      4. the column energy probe (ColProbeOutputState), kept for step t+1.
    on_step(t, col, energy) is called after each step.  Returns the energies."""
    from .timescale import PeriodicUpdate
    nb = col.in_loc.nbatch
    trigger = PeriodicUpdate(display_period) if (frames is not None and display_period > 0) else None
    index = 0
    if frames is not None:
        col.set_input(frames[0])
    energy_prev = np.zeros(nb)
    energies = []
    for t in range(1, steps + 1):
        updating = trigger is not None and trigger.need_update(t)
        if adaptive is not None:
            vals = np.full(nb, -1.0) if updating else energy_prev
            col.set_dt(adaptive.calc_timesteps(vals))
        if updating:
            col.set_input(frames[index])
            index = (index + 1) % len(frames)
            trigger.updated(t)
        col.step(stream=stream)
        energy_prev = col.probe_energy(stream=stream, cost=energy_cost).cpu().numpy()
        energies.append(energy_prev)
        if on_step is not None:
            on_step(t, col, energy_prev)
    return np.array(energies)
