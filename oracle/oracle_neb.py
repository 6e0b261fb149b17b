"""Nudged elastic band on the CPU oracle (test infrastructure only).

Restates src/neb/neb.jl (NEB constructor, init_images, effective_field,
effective_field_energy, compute_distance, compute_tangents, relax),
src/neb/neb_kernels.jl (compute_distance_kernel!, reduce_tangent_kernel!,
compute_tangents_kernel!), src/neb/math.jl (slerp_kernel!) and the BS23 stepper
of src/integrator/adaptive_rk.jl:93-165 + adaptive_rk_steps.jl:11-47 with the
NEB driver settings of neb.jl:146-152 (LLG, BS23, tol 1e-5, alpha 0.2,
precession off).  Field and energy of an image come from the C oracle
(OracleSim.effective_field / term_energies).

As in the reference: the field, energies and distances of every stage are
those of the accepted images (effective_field_energy and compute_distance read
neb.spin, not the stage vector), while tangents and the LLG right-hand side use
the stage vector; relax measures max|m - prespin| against the initial images
(an NEB's prespin is never updated).
"""
import numpy as np


def slerp(x1, x2, t):
    """slerp_kernel!: src/neb/math.jl:45-61 (cells with theta == 0 stay zero)."""
    a = x1.reshape(-1, 3)
    b = x2.reshape(-1, 3)
    inner = np.clip(a[:, 0] * b[:, 0] + a[:, 1] * b[:, 1] + a[:, 2] * b[:, 2], -1.0, 1.0)
    theta = np.arccos(inner)
    res = np.zeros_like(a)
    nz = theta != 0.0
    s = np.sin(theta[nz])
    ca = np.sin((1 - t) * theta[nz]) / s
    cb = np.sin(t * theta[nz]) / s
    res[nz] = ca[:, None] * a[nz] + cb[:, None] * b[nz]
    return res.reshape(-1)


def init_images(given, frames):
    """init_images: src/neb/neb.jl:155-175."""
    out = [given[0]]
    for i in range(len(given) - 1):
        if frames[i] > 0:
            for j in range(1, frames[i] + 1):
                out.append(slerp(given[i], given[i + 1], j / (frames[i] + 1)))
            out.append(given[i + 1])
    return out


def geodesic(m1, m2):
    """compute_distance_kernel!: src/neb/neb_kernels.jl:66-81, per cell."""
    a = m1.reshape(-1, 3)
    b = m2.reshape(-1, 3)
    cx = -b[:, 1] * a[:, 2] + a[:, 1] * b[:, 2]
    cy = b[:, 0] * a[:, 2] - a[:, 0] * b[:, 2]
    cz = -b[:, 0] * a[:, 1] + a[:, 0] * b[:, 1]
    mm = a[:, 0] * b[:, 0] + a[:, 1] * b[:, 1] + a[:, 2] * b[:, 2]
    d = np.sqrt(cx * cx + cy * cy + cz * cz)
    return np.arctan2(d, mm)


class OracleNEB:
    def __init__(self, osim, given_images, frames, spring_constant=1.0e5, climbing_image=-1, alpha=0.2,
                 gamma=2.21e5, tol=1e-5, pins=None):
        self.o = osim
        self.N = osim.N
        self.n = int(sum(frames) + len(given_images) - 2)
        aligned = []
        for g in given_images:
            osim.init_m0(g)
            aligned.append(osim.spin.copy())
        imgs = init_images(aligned, frames)
        assert len(imgs) == self.n + 2
        self.left, self.right = imgs[0], imgs[-1]
        self.spin = np.stack(imgs[1:-1])            # (n, 3N), the accepted images (neb.spin)
        self.prespin = self.spin.copy()
        self.k = spring_constant
        self.clib = climbing_image
        self.alpha, self.gamma, self.tol = alpha, gamma, tol
        self.pins = np.zeros(self.N, dtype=bool) if pins is None else np.asarray(pins, dtype=bool)
        self.energy = np.zeros(self.n + 2)
        self.energy[0] = osim.term_energies(self.left).sum()
        self.energy[-1] = osim.term_energies(self.right).sum()
        self.field = np.zeros_like(self.spin)
        self.distance = np.zeros(self.n + 1)
        self._field_energy()
        # AdaptiveRK state (adaptive_rk.jl:37-55)
        self.t = 0.0
        self.step = 0.0
        self.step_next = 0.0
        self.nsteps = 0
        self.nfevals = 0
        self.succeed = False
        self.k4 = None
        self.neb_steps = 0

    # effective_field_energy(neb, spin): reads the accepted images
    def _field_energy(self):
        for q in range(self.n):
            self.field[q] = self.o.effective_field(self.spin[q])
            self.energy[q + 1] = self.o.term_energies(self.spin[q]).sum()

    def _distance(self):
        for q in range(self.n + 1):
            m1 = self.left if q == 0 else self.spin[q - 1]
            m2 = self.right if q == self.n else self.spin[q]
            self.distance[q] = np.linalg.norm(geodesic(m1, m2))

    def _tangents(self, y):
        E = self.energy
        t = np.empty_like(y)
        for q in range(self.n):
            E1, E2, E3 = E[q], E[q + 1], E[q + 2]
            dEmax = max(abs(E3 - E2), abs(E2 - E1))
            dEmin = min(abs(E3 - E2), abs(E2 - E1))
            mkp = self.right if q == self.n - 1 else y[q + 1]
            mkm = self.left if q == 0 else y[q - 1]
            tip = mkp - y[q]
            tim = y[q] - mkm
            if E1 > E2 and E2 > E3:
                t[q] = tim
            elif E3 > E2 and E2 > E1:
                t[q] = tip
            elif E3 > E1:
                t[q] = dEmax * tip + dEmin * tim
            elif E3 < E1:
                t[q] = dEmin * tip + dEmax * tim
            else:
                t[q] = tim + tip
        # reduce_tangent_kernel!: t -= (t.m) m with the stage vector
        tv = t.reshape(self.n, -1, 3)
        mv = y.reshape(self.n, -1, 3)
        f = (tv * mv).sum(axis=2)
        tv -= f[:, :, None] * mv
        for q in range(self.n):
            t[q] = t[q] / np.linalg.norm(t[q])
        return t

    def force(self, y):
        """neb.field after effective_field(neb, y, t): src/neb/neb.jl:196-216."""
        self._field_energy()
        t = self._tangents(y)
        self._distance()
        f = self.field.copy()
        for q in range(self.n):
            if q + 1 == self.clib:
                ft = -2 * np.dot(f[q], t[q])
            else:
                ft = self.k * (self.distance[q + 1] - self.distance[q]) - np.dot(f[q], t[q])
            f[q] = f[q] + ft * t[q]
        return f

    def rhs(self, y):
        """llg_call_back with the NEB field, alpha 0.2, precession off (llg.jl:6-58)."""
        h = self.force(y)
        m = y.reshape(self.n, -1, 3)
        hv = h.reshape(self.n, -1, 3)
        a = -self.gamma / (1 + self.alpha * self.alpha)
        mm = (m * m).sum(axis=2)[:, :, None]
        mh = (m * hv).sum(axis=2)[:, :, None]
        hp = mm * hv - mh * m
        d = a * (-hp * self.alpha)
        d[:, self.pins, :] = 0.0
        self.nfevals += 1
        return d.reshape(self.n, -1)

    @staticmethod
    def _normalise(y):
        v = y.reshape(-1, 3)
        m2 = (v * v).sum(axis=1)
        nz = m2 > 0
        v[nz] *= (1.0 / np.sqrt(m2[nz]))[:, None]
        return y

    def _bs23_try(self, step, t):
        """rk_step_bs23!: src/integrator/adaptive_rk_steps.jl:11-47."""
        y = self.spin_current
        if self.succeed and t >= self.t and self.k4 is not None:
            k1 = self.k4
        else:
            k1 = self.rhs(y)
        k2 = self.rhs(y + step * 0.5 * k1)
        k3 = self.rhs(y + step * 0.75 * k2)
        y_next = y + step * (2 / 9 * k1 + 1 / 3 * k2 + 4 / 9 * k3)
        self._normalise(y_next)
        k4 = self.rhs(y_next)
        err = -(5 / 72) * k1 + (1 / 12) * k2 + (1 / 9) * k3 + (-1 / 8) * k4
        max_error = np.max(np.abs(err / (1 + np.abs(y_next)))) * step / self.tol
        return max_error, y_next, k4

    def _initial_step(self, dt):
        """compute_initial_step: src/integrator/adaptive_rk.jl:151-165 (order 3)."""
        f = self.rhs(self.spin_current)
        r = np.max(np.abs(f)) / (0.824 * self.tol ** (1 / 3))
        return min(dt, 0.001 / r) if dt * r > 0.001 else dt

    def advance_step(self):
        """advance_step(neb, ::AdaptiveRK): src/integrator/adaptive_rk.jl:93-148 (BS23)."""
        self.spin_current = self.spin.copy()
        t = self.t
        if self.step_next <= 0:
            self.step_next = self._initial_step(1e-12)
        self.step = self.step_next
        nstep = 1
        while True:
            max_error, y_next, k4 = self._bs23_try(self.step, t)
            if np.isnan(max_error):
                raise FloatingPointError("NaN error")
            nstep += 1
            if nstep > 50:
                raise RuntimeError("too many tries")
            self.succeed = max_error <= 1.0
            if self.succeed:
                self.spin = y_next
                self.k4 = k4
                self.nsteps += 1
                self.t += self.step
                factor = 0.824 * (1.0 / max_error) ** (1 / 3)
                self.step_next = self.step * min(5.0, max(0.2, factor))
                break
            factor = 0.824 * (1.0 / max_error) ** (1 / 2)
            self.step = self.step * min(5.0, max(0.2, factor))
        self.neb_steps += 1

    def relax(self, max_steps=10000, stopping_dmdt=0.01):
        """relax(::NEB): src/neb/neb.jl:231-296 with the LLG driver (no output)."""
        dmdt_factor = (2 * np.pi / 360) * 1e9
        max_dmdt = np.inf
        for i in range(max_steps + 1):
            self.advance_step()
            dm = np.linalg.norm((self.prespin - self.spin).reshape(-1, 3), axis=1)
            max_dmdt = dm.max() / self.step
            if max_dmdt < stopping_dmdt * dmdt_factor:
                return i, max_dmdt
        return max_steps, max_dmdt
