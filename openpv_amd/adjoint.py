"""Hand-written adjoint of the unrolled LCA loop (an extension: the reference
has no backward pass).

Forward (the column of lca.LCAColumn, one image batch, errScale 1):
    R_t   = x - Rec_{t-1}                                   (ANNErrorLayer)
    G_t   = D_W(R_t)                                        (TransposeConn gather)
    V_t   = e*V_{t-1} + (1-e)*(G_t + A_{t-1})               (HyPerLCAInternalStateBuffer)
    A_t   = T(V_t)                                          (ANNActivityBuffer)
    Rec_t = S_W(A_t)                                        (CloneConn scatter)
    E     = 0.5*||x - Rec_T||^2 + l1*sum|A_T|               (energy probes)

Backward, t = T .. 1 (S_W is the transpose of D_W; corr(p, q) is the Hebbian
correlation, dW[f][item] = sum_pos p[pos][f] * q[window(pos)+item]):
    gA_t  += D_W(gRec_t);          gW += corr(A_t, gRec_t)
    gVt    = gV + T'(V_t)*(gA_t + l1*sign(A_t) at t = T)
    gG_t   = (1-e)*gVt;  gA_{t-1} = (1-e)*gVt;  gV = e*gVt
    gR_t   = S_W(gG_t);            gW += corr(gG_t, R_t)
    gx    += gR_t;                 gRec_{t-1} = -gR_t
with gRec_T = -(x - Rec_T).  States are kept every `checkpoint` steps and each
segment is recomputed forward before it is walked backward.  Every operation
is a libpvhip kernel (deliveries, Hebbian correlation, the adjoint step);
torch only provides the device memory.
"""
import ctypes as C

import numpy as np
import torch

from . import abi
from .runtime import (FLT_MAX, HebbianUpdater, HyPerLCAUpdate, PostsynapticPerspectiveGPUDelivery,
                      PresynapticPerspectiveGPUDelivery, _ptr, _stream, cost_sum_f64, loc,
                      make_transfer)


def _s2e(loc_, scale, src, dst, accumulate, stream):
    abi.check(abi.hip().pvhip_scale_to_extended(C.byref(loc_), scale, _ptr(src), _ptr(dst),
                                                int(accumulate), _stream(stream)),
              "pvhip_scale_to_extended")


def _axpby(a, x, b, y, stream):
    abi.check(abi.hip().pvhip_axpby(y.numel(), a, _ptr(x), b, _ptr(y), _stream(stream)),
              "pvhip_axpby")


class LCAAdjoint:
    """Unrolled LCA (x: [nb][ny][nx][nf], V1 ny_v1 x nx_v1 x nf_v1, shared
    weights W [nf_v1][nyp][nxp][nf]) with energy and gradients dE/dW, dE/dx.

    arith: EXACT or FP32 for the deliveries and the weight correlation."""

    def __init__(self, nb, nx, ny, nf, nx_v1, ny_v1, nf_v1, nxp, nyp, VThresh=0.1, AMin=0.0,
                 AMax=FLT_MAX, AShift=None, VWidth=0.0, timeConstantTau=100.0, dt=1.0, l1=None,
                 self_interact=True, arith=abi.ARITH_FP32, checkpoint=10, device="cuda"):
        H = abi.host()
        nxpT = H.pvhost_post_patch_size(nxp, nx_v1, nx)
        mres = H.pvhost_required_margin(nx, nx_v1, nxpT)
        mv1 = H.pvhost_required_margin(nx_v1, nx, nxp)
        self.res_loc = loc(nb, nx, ny, nf, mres)
        self.v1_loc = loc(nb, nx_v1, ny_v1, nf_v1, mv1)
        self.rec_loc = loc(nb, nx, ny, nf, 0)
        self.nxp, self.nyp = nxp, nyp
        self.arith = arith
        self.checkpoint = max(1, int(checkpoint))
        self.device = torch.device(device)
        if AShift is None:
            AShift = VThresh
        self.transfer = make_transfer(VThresh, AMin, AMax, AShift, VWidth)
        # energy: L1 coefficient of the LCA probe (float VThresh widened to double)
        self.l1 = float(np.float32(VThresh)) if l1 is None else float(l1)
        self.lca = HyPerLCAUpdate(self.v1_loc, self.transfer, timeConstantTau, dt, self_interact,
                                  activity_from_v=True, device=device)
        self.self_interact = bool(self_interact)
        f = dict(dtype=torch.float32, device=self.device)
        nr, ne = nb * self.res_loc.num_res, nb * self.res_loc.num_ext
        nv, nve = nb * self.v1_loc.num_res, nb * self.v1_loc.num_ext
        self.R_ext = torch.zeros(ne, **f)
        self.rec = torch.zeros(nr, **f)
        self.G = torch.zeros(nv, **f)
        self.V = torch.zeros(nv, **f)
        self.A = torch.zeros(nve, **f)
        # backward buffers
        self.gA = torch.zeros(nv, **f)
        self.gAprev = torch.zeros(nv, **f)
        self.gVc = torch.zeros(nv, **f)
        self.gG_ext = torch.zeros(nve, **f)
        self.gR = torch.zeros(nr, **f)
        self.gRec_ext = torch.zeros(ne, **f)
        self.energy = torch.zeros(2, nb, dtype=torch.float64, device=self.device)
        self._W = None

    # weights ---------------------------------------------------------------
    def set_weights(self, W):
        W = torch.as_tensor(W, dtype=torch.float32, device=self.device).reshape(-1).contiguous()
        if self._W is None:
            self._W = W.clone()
            self.gather = PostsynapticPerspectiveGPUDelivery(self.res_loc, self.v1_loc, self.nxp,
                                                             self.nyp, self._W, arith=self.arith)
            self.scatter = PresynapticPerspectiveGPUDelivery(self.v1_loc, self.rec_loc, self.nxp,
                                                             self.nyp, self._W, arith=self.arith)
            self.corr = HebbianUpdater(self.v1_loc, self.res_loc, self.nxp, self.nyp, 1.0,
                                       normalize_dw=False, device=self.device)
        else:
            self._W.copy_(W)
            self.gather.refresh_weights()
            self.scatter.refresh_weights()

    # forward -----------------------------------------------------------------
    def _residual(self, x, stream):
        _s2e(self.res_loc, 1.0, x, self.R_ext, False, stream)
        _s2e(self.res_loc, -1.0, self.rec, self.R_ext, True, stream)

    def _step(self, x, stream):
        """One LCA timestep from (V, A, rec) of t-1 to t; R_ext holds R_t after it."""
        self._residual(x, stream)
        self.gather.deliver(self.R_ext, self.G, accumulate=False, stream=stream)
        self.lca.update(self.G, self.V, self.A, stream=stream)
        self._scatter(self.A, self.rec, stream)

    def _scatter(self, a_ext, out, stream):
        # dense pre activity: the transposed MFMA route (FP32) or the ordered
        # gather over every presynaptic neuron (EXACT)
        self.scatter.deliver(out, pre_activity=a_ext, stream=stream)

    def _init_state(self, V0, stream):
        self.V.zero_() if V0 is None else self.V.copy_(torch.as_tensor(V0).reshape(-1))
        from .runtime import ann_activity
        ann_activity(self.v1_loc, self.transfer, self.V, self.A, stream=stream)
        self.rec.zero_()

    def forward(self, x, W, T, V0=None, stream=None):
        """Runs T steps; returns (E [nb] float64, checkpoints) with checkpoints
        {t: (V_t, rec_t)} for t = 0, k, 2k, ..."""
        self.set_weights(W)
        x = torch.as_tensor(x, dtype=torch.float32, device=self.device).reshape(-1).contiguous()
        self._init_state(V0, stream)
        ck = {}
        for t in range(T):
            if t % self.checkpoint == 0:
                ck[t] = (self.V.clone(), self.rec.clone())
            self._step(x, stream)
        self._residual(x, stream)  # R_{T+1} = x - Rec_T
        cost_sum_f64(self.res_loc, self.R_ext, True, 0, out=self.energy[0], stream=stream)
        cost_sum_f64(self.v1_loc, self.A, True, 1, out=self.energy[1], stream=stream)
        E = 0.5 * self.energy[0] + self.l1 * self.energy[1]
        return E.clone(), ck

    # backward ----------------------------------------------------------------
    def backward(self, x, T, ck, fixed_activity=False, stream=None):
        """dE/dW (flat, like W) and dE/dx ([nb*ny*nx*nf]) of E summed over the batch.
        fixed_activity: the gradients of E with the final activity A_T held fixed
        (only the reconstruction term: dE/dW = corr(A_T, -(x - Rec_T)), which is
        minus the Hebbian dW with dWMax 1).  Leaves the forward state at step T."""
        x = torch.as_tensor(x, dtype=torch.float32, device=self.device).reshape(-1).contiguous()
        gW = self.corr.dw
        gW.zero_()
        self.gVc.zero_()
        self.gA.zero_()
        # gRec_T = -(x - Rec_T) and the direct dE/dx = x - Rec_T = R_{T+1}
        self._restore(ck, T, x, stream)
        self._residual(x, stream)
        gx = self._restricted_R().clone()
        self.gRec_ext.zero_()
        _s2e(self.res_loc, -1.0, gx, self.gRec_ext, False, stream)
        if fixed_activity:
            self.corr.update_dw(self.A, self.gRec_ext, clear=False, arith=self.arith, stream=stream)
            return gW.clone(), gx
        k = self.checkpoint
        seg_starts = sorted(ck)
        for s0 in reversed(seg_starts):
            s1 = min(s0 + k, T)
            # recompute the segment, keeping (V_t, A_t, R_t) for t in (s0, s1]
            self.V.copy_(ck[s0][0])
            self.rec.copy_(ck[s0][1])
            from .runtime import ann_activity
            ann_activity(self.v1_loc, self.transfer, self.V, self.A, stream=stream)
            states = []
            for t in range(s0 + 1, s1 + 1):
                self._step(x, stream)
                states.append((self.V.clone(), self.A.clone(), self.R_ext.clone()))
            for t in range(s1, s0, -1):
                Vt, At, Rt = states[t - s0 - 1]
                # Rec_t = S_W(A_t)
                self.gather.deliver(self.gRec_ext, self.gA, accumulate=True, stream=stream)
                self.corr.update_dw(At, self.gRec_ext, clear=False, arith=self.arith, stream=stream)
                # A_t = T(V_t); V_t = e V_{t-1} + (1-e)(G_t + A_{t-1})
                a = abi.LCAAdjointArgs()
                a.loc, a.transfer = self.v1_loc, self.transfer
                a.exp_tau = _ptr(self.lca.exp_tau)
                a.self_interact = int(self.self_interact)
                a.V, a.gA = _ptr(Vt), _ptr(self.gA)
                a.l1 = self.l1 if t == T else 0.0
                a.gV, a.gG, a.gA_prev = _ptr(self.gVc), _ptr(self.gG_ext), _ptr(self.gAprev)
                abi.check(abi.hip().pvhip_lca_adjoint_step(C.byref(a), _stream(stream)),
                          "pvhip_lca_adjoint_step")
                self.gA, self.gAprev = self.gAprev, self.gA
                # G_t = D_W(R_t)
                self._scatter(self.gG_ext, self.gR, stream)
                self.corr.update_dw(self.gG_ext, Rt, clear=False, arith=self.arith, stream=stream)
                # R_t = x - Rec_{t-1}
                _axpby(1.0, self.gR, 1.0, gx, stream)
                _s2e(self.res_loc, -1.0, self.gR, self.gRec_ext, False, stream)
        return gW.clone(), gx

    def _restricted_R(self):
        nb, L = self.res_loc.nbatch, self.res_loc
        R = self.R_ext.view(nb, L.ny_ext, L.nx_ext, L.nf)
        return R[:, L.halo.up:L.halo.up + L.ny, L.halo.lt:L.halo.lt + L.nx].contiguous().reshape(-1)

    def _restore(self, ck, T, x, stream):
        """Forward state at step T from the last checkpoint."""
        s0 = max(ck)
        self.V.copy_(ck[s0][0])
        self.rec.copy_(ck[s0][1])
        from .runtime import ann_activity
        ann_activity(self.v1_loc, self.transfer, self.V, self.A, stream=stream)
        for _ in range(s0, T):
            self._step(x, stream)


class LCAEnergyFunction(torch.autograd.Function):
    """E = sum_b E_b of T unrolled LCA steps as a torch.autograd.Function whose
    backward is the hand-written adjoint (checkpoint every `lca.checkpoint`)."""

    @staticmethod
    def forward(ctx, x, W, lca, T):
        E, ck = lca.forward(x, W, T)
        ctx.lca, ctx.T, ctx.ck, ctx.wshape = lca, T, ck, W.shape
        ctx.save_for_backward(x)
        return E.sum()

    @staticmethod
    def backward(ctx, gE):
        (x,) = ctx.saved_tensors
        gW, gx = ctx.lca.backward(x, ctx.T, ctx.ck)
        g = gE.to(torch.float32)
        return (gx * g).reshape(x.shape), (gW * g).reshape(ctx.wshape), None, None
