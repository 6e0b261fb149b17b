"""Twin monoclinic anisotropy of the oracle (add_twin_mono_anis,
twin_monoclinic_anisotropy_kernel! src/micro/kernels.jl:169-225) pinned by
what its energy fixes: E = Ka a^2 + Kb b^2 + Kaa a^4 + Kbb b^4 + Kab a^2 b^2
- Ku u^2 (a = m.axis_a, ...) times V, H = -(1/(mu0Ms V)) dE/dm (central
differences of the oracle's energy), the uniaxial special cases in closed
form, and the normalisation of the given axes.
"""
import numpy as np

from oracle import OracleSim

MU0 = 4e-7 * np.pi
MS = 5e5


def _cell(m, **kw):
    o = OracleSim(1, 1, 1, 2e-9, 2e-9, 3e-9)
    o.set_Ms(MS)
    o.init_m0(np.asarray(m, dtype=float), norm=False)
    return o, o.add_twin_mono_anis(**kw)


def test_uniaxial_special_cases():
    V = 12e-27
    m = np.array([0.6, -0.48, 0.64])
    o, k = _cell(m, Ka=3e4, axis_a=(2.0, 0, 0))  # axis normalised to x
    h, e = o.term_field(k)
    np.testing.assert_allclose(h, [-2 * 3e4 * 0.6 / (MU0 * MS), 0, 0], rtol=1e-13)
    np.testing.assert_allclose(e[0], 3e4 * 0.36 * V, rtol=1e-13)
    u = np.array([1.0, 1.0, 1.0]) / np.sqrt(3)
    o, k = _cell(m, Ku=2e4, axis_u111=(1, 1, 1))
    h, e = o.term_field(k)
    mu = m @ u
    np.testing.assert_allclose(h, 2 * 2e4 * mu * u / (MU0 * MS), rtol=1e-13)
    np.testing.assert_allclose(e[0], -2e4 * mu * mu * V, rtol=1e-13)


def test_field_is_minus_energy_gradient():
    rng = np.random.default_rng(21)
    V = 12e-27
    for _ in range(4):
        m = rng.standard_normal(3)
        kw = dict(Ka=rng.uniform(-5e4, 5e4), Kb=rng.uniform(-5e4, 5e4), Kaa=rng.uniform(-5e4, 5e4),
                  Kbb=rng.uniform(-5e4, 5e4), Kab=rng.uniform(-5e4, 5e4), Ku=rng.uniform(-5e4, 5e4),
                  axis_a=tuple(rng.standard_normal(3)), axis_b=tuple(rng.standard_normal(3)),
                  axis_u111=tuple(rng.standard_normal(3)))
        o, k = _cell(m, **kw)
        h, _ = o.term_field(k)
        g = np.empty(3)
        for a in range(3):
            d = np.zeros(3)
            d[a] = 1e-6
            g[a] = (o.term_field(k, spin=m + d)[1][0] - o.term_field(k, spin=m - d)[1][0]) / 2e-6
        np.testing.assert_allclose(h, -g / (MU0 * MS * V), rtol=1e-6, atol=1e-6 * np.abs(h).max())
