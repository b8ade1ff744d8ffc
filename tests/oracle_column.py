"""An LCA column with a learning dictionary on the oracle alone (a helper of
the tests, not collected by pytest).

Input -> Residual (ANNErrorLayer, errScale 1) <-> V1 (HyPerLCALayer, soft
threshold, fixed dt) -> Recon, as the device column openpv_amd.lca.LCAColumn
runs it, each step composed from the oracle's restatements of the
reference's functions; and the dictionary's weight update at a trigger step:
HebbianUpdater's dW (pre V1, post Residual, normalizeDw off), W += dW, then
NormalizeL2 in the reference's float arithmetic or in FP64.
"""
import numpy as np

import oracle as O

FMAX = float(np.finfo(np.float32).max)


class OracleColumn:
    def __init__(self, images, W, nx_v1, ny_v1, nf_v1, nxp, VThresh, tau=100.0, dt=1.0):
        images = np.ascontiguousarray(images, np.float32)
        nb, ny, nx, nf = images.shape
        nxpT = O.lib().or_calcPostPatchSize(nxp, nx_v1, nx)
        self.mres = O.lib().or_requiredConvolveMargin(nx, nx_v1, nxpT)
        mv1 = O.lib().or_requiredConvolveMargin(nx_v1, nx, nxp)
        self.res = O.Loc.make(nb, nx, ny, nf, self.mres)
        self.v1 = O.Loc.make(nb, nx_v1, ny_v1, nf_v1, mv1)
        self.rec = O.Loc.make(nb, nx, ny, nf, 0)
        self.nxp = nxp
        self.x = images
        self.W = np.ascontiguousarray(W, np.float32).ravel().copy()
        self.tr = (VThresh, 0.0, FMAX, VThresh, 0.0)
        self.VThresh = VThresh
        self.tau = float(np.float32(tau / dt))
        self.dt = np.full(nb, dt, np.float64)
        self.V = np.zeros(nb * self.v1.num_res, np.float32)
        self.A = O.transfer(self.v1, self.V, *self.tr)
        self.rc = np.zeros(nb * self.rec.num_res, np.float32)
        self.R_ext = np.zeros(nb * self.res.num_ext, np.float32)

    def set_input(self, images):
        self.x = np.ascontiguousarray(images, np.float32)

    def step(self):
        res, v1, rec, m = self.res, self.v1, self.rec, self.mres
        nb = res.nbatch
        Rres = np.zeros(nb * res.num_res, np.float32)
        O.lib().or_gsyn_accumulate(nb * res.num_res, 2, np.array([1.0, -1.0], np.float32),
                                   np.concatenate([self.x.ravel(), self.rc]), Rres)
        R = np.zeros((nb, res.ny_ext, res.nx_ext, res.nf), np.float32)
        R[:, m:m + res.ny, m:m + res.nx] = Rres.reshape(nb, res.ny, res.nx, res.nf)
        self.R_ext = R.ravel()
        G = O.recv_post(res, v1, self.nxp, self.nxp, self.R_ext, self.W, num_threads=1)
        self.V = O.update_lca(v1, True, self.dt, self.tau, G, self.A, self.V)
        self.A = O.transfer(v1, self.V, *self.tr)
        ent, cnt = O.active_indices(v1, self.A)
        self.rc = O.recv_pre(v1, rec, self.nxp, self.nxp, self.W, entries=ent, num_active=cnt,
                             num_threads=1)
        return cnt

    def energy(self):
        """Total energy per image: 0.5*|r|^2 + VThresh*|a|_1 (FP64)."""
        vt = float(np.float32(self.VThresh))
        return np.array([0.5 * O.cost_sum(self.res, self.R_ext, True, b, 0)
                         + vt * O.cost_sum(self.v1, self.A, True, b, 1)
                         for b in range(self.res.nbatch)])

    def update_weights(self, dw_max, renormalize="exact", strength=1.0):
        dw, _ = O.hebbian_dw(self.v1, self.res, self.nxp, self.nxp, self.res.nf, dw_max, self.A,
                             self.R_ext, counts=False)
        O.lib().or_hebbian_normalize_apply(self.W, dw, None, self.W.size, 0)
        npatch = self.v1.nf
        if renormalize == "exact":
            O.lib().or_normalize_l2(self.W, npatch, self.W.size // npatch, strength, 0.0)
        elif renormalize == "fp64":
            O.lib().or_normalize_l2_f64(self.W, npatch, self.W.size // npatch, strength, 0.0)
        return dw
