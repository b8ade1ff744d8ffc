"""Adaptive time scales of the LCA layer (host side, FP64).

AdaptiveTimeScaleController::calcTimesteps
(src/components/AdaptiveTimeScaleController.cpp:60-88) and the probe around it
(src/probes/AdaptiveTimeScaleProbe.cpp:185-206): at the start of each step
the probe reads the target energy probe's values from the end of the
previous step (or -1 for every batch element when its trigger layer updates
this step) and the controller turns them into per-batch time scales, which
the HyPerLCA update uses as its dt (HyPerLCAInternalStateBuffer.cpp:166-174).
"""
import numpy as np


class AdaptiveTimeScaleController:
    def __init__(self, batch_width, base_max, base_min, tau_factor, growth_factor):
        self.base_max = float(base_max)
        self.base_min = float(base_min)
        self.tau_factor = float(tau_factor)
        self.growth_factor = float(growth_factor)
        self.time_scale = np.full(batch_width, self.base_min)
        self.time_scale_max = np.full(batch_width, self.base_max)
        self.time_scale_true = np.full(batch_width, -1.0)

    def calc_timesteps(self, values):
        """values: the target probe's energy per batch element (float64)."""
        values = np.asarray(values, np.float64)
        old_true = self.time_scale_true.copy()
        for b in range(len(values)):
            self.time_scale_true[b] = values[b]
            E_dt = self.time_scale_true[b]
            E_0 = old_true[b]
            if E_dt == E_0:
                continue
            dE_dt_scaled = (E_0 - E_dt) / self.time_scale[b]
            if dE_dt_scaled < 0.0 or E_0 <= 0.0 or E_dt <= 0.0:
                self.time_scale[b] = self.base_min
                self.time_scale_max[b] = self.base_max
            else:
                tau_eff_scaled = E_0 / dE_dt_scaled
                self.time_scale[b] = self.tau_factor * tau_eff_scaled
                if self.time_scale[b] >= self.time_scale_max[b]:
                    self.time_scale[b] = self.time_scale_max[b]
                    self.time_scale_max[b] *= (1 + self.growth_factor)
        return self.time_scale.copy()


class PeriodicUpdate:
    """LayerUpdateController::needUpdate for a layer updating every `period`
    (src/components/LayerUpdateController.cpp:318-335): the last update time
    starts at dt, so the first update happens on the first step."""

    def __init__(self, period, dt=1.0):
        self.period = float(period)
        self.last = float(dt)

    def need_update(self, sim_time, dt=1.0):
        if self.period <= 0:
            return False
        n = (sim_time - self.last) / self.period
        return abs(n - np.rint(n)) * self.period < 0.5 * dt

    def updated(self, sim_time):
        self.last = float(sim_time)
