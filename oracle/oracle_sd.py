"""Steepest descent on the CPU oracle (test infrastructure only).

Restates src/sd.jl (compute_gk_kernel_1!/_2!, compute_tau, run_step_kernel!,
run_step) and the SD branch of relax (src/sim.jl:392-470) over the C oracle's
effective field (OracleSim.effective_field)."""
import numpy as np


def _cross(a, b):
    """cross_product: src/util.jl:30-33 (rows of 3-vectors)."""
    return np.stack([-a[:, 2] * b[:, 1] + a[:, 1] * b[:, 2],
                     a[:, 2] * b[:, 0] - a[:, 0] * b[:, 2],
                     -a[:, 1] * b[:, 0] + a[:, 0] * b[:, 1]], axis=1)


class OracleSD:
    def __init__(self, osim, pins=None, min_tau=1e-10, max_tau=1.0):
        self.o = osim
        self.N = osim.N
        self.gk = np.zeros((self.N, 3))
        self.tau = 0.0
        self.min_tau, self.max_tau = min_tau, max_tau
        self.steps = 0
        self.pins = np.zeros(self.N, dtype=bool) if pins is None else np.asarray(pins, dtype=bool)
        self.spin = osim.spin.reshape(-1, 3).copy()
        self.prespin = self.spin.copy()

    def _compute_tau(self, m, h):
        g = _cross(m, _cross(m, h))
        if self.steps == 0:
            self.gk = g
            self.tau = self.min_tau
            return
        s = m - self.prespin
        f = g - self.gk
        self.gk = g
        s1 = np.sum(s * s)
        s2 = np.sum(s * f)
        s3 = np.sum(f * f)
        tau1 = s1 / s2 if s2 != 0 else self.min_tau
        tau2 = s2 / s3 if s3 != 0 else self.min_tau
        self.tau = abs(tau2) if self.steps % 2 == 0 else abs(tau1)
        self.tau = min(self.tau, self.max_tau)

    def run_step(self):
        m = self.spin
        h = self.o.effective_field(m.reshape(-1)).reshape(-1, 3)
        self._compute_tau(m, h)
        self.prespin = m.copy()
        f = _cross(m, h)
        factor = 0.25 * (f * f).sum(axis=1) * self.tau ** 2
        new = ((1 - factor)[:, None] * m - self.tau * self.gk) / (1 + factor)[:, None]
        new[self.pins] = m[self.pins]
        self.spin = new
        self.steps += 1
        if self.steps % 10 == 0:
            m2 = (self.spin * self.spin).sum(axis=1)
            nz = m2 > 0
            self.spin[nz] *= (1 / np.sqrt(m2[nz]))[:, None]

    def relax(self, max_steps=10000, stopping_dmdt=0.01):
        dmdt_factor = (2 * np.pi / 360) * 1e9
        md = 0.0
        for i in range(1, max_steps + 1):
            self.run_step()
            step_size = self.tau / (2.21e5 / 2)
            md = np.linalg.norm(self.prespin - self.spin, axis=1).max() / step_size
            if md < stopping_dmdt * dmdt_factor:
                return i, md
        return max_steps, md
