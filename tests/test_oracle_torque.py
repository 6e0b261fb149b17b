"""Spin-torque terms of the oracle (add_sot, add_stt) pinned by the reference's
own checks and by closed forms:

* SOT: test/test_sot.jl:9-48 (one cell, H = (aj m x p + bj p)/gamma) and the
  zero-polarisation case (:127-147);
* Slonczewski: test/test_slonczewski.jl:10-41 (analytic H with beta = hbar J /
  (mu0 e tf Ms));
* Zhang-Li: a helix m = (cos kx, sin kx, 0) along x, whose central difference
  is sin(k dx)/dx (-sin kx, cos kx, 0) inside and one-sided at the two ends,
  so H = (-xi A sin kx, xi A cos kx, A) with A = b J sin(k dx)/(dx gamma);
  the (P, Ms) form b = P mu_B / (e Ms) / (1 + xi^2) (add_field.jl:1164-1181).
"""
import numpy as np
import pytest

from oracle import OracleSim

MU0 = 4e-7 * np.pi
MU_B = 9.27400949e-24  # src/const.jl:2
C_E = 1.602176565e-19  # src/const.jl:4
H_BAR = 1.05457172647e-34  # src/const.jl:9


def _one_cell(m, norm):
    o = OracleSim(1, 1, 1, 1e-9, 1e-9, 1e-9)
    o.set_Ms(8e5)
    o.set_llg(0.05, 2.21e5, True)
    o.init_m0(np.asarray(m, dtype=float), norm=norm)
    o.add_zeeman((0, 0, 1e5))
    return o


def test_sot_one_cell_matches_reference_test():
    o = _one_cell((-0.3, 0.4, 0.8), norm=False)
    aj, bj, p = 1e9, 5e8, np.array([0.1, 0.2, 0.3])
    t = o.add_sot(aj, bj, p)
    h, e = o.term_field(t)
    expected = (aj * np.cross([-0.3, 0.4, 0.8], p) + bj * p) / 2.21e5
    assert np.allclose(h, expected, rtol=1e-13, atol=0)
    assert np.all(e == 0)
    # the torque adds into H_eff after the Zeeman field (insertion order)
    H = o.effective_field()
    assert np.allclose(H, expected + [0, 0, 1e5], rtol=1e-13, atol=0)


def test_sot_zero_polarisation():
    o = _one_cell((-0.3, 0.4, 0.8), norm=False)
    t = o.add_sot(1e9, 5e8, (0.0, 0.0, 0.0))
    assert np.all(o.term_field(t)[0] == 0)


def test_slonczewski_one_cell_matches_reference_test():
    o = _one_cell((0.6, 0.8, 0.0), norm=True)
    J, tf, Ms, P, xi, Lam = 1e12, 2e-9, 8e5, 0.7, 0.5, 3.7
    p = np.array([0.0, 0.0, 1.0])
    beta = H_BAR / (MU0 * Ms * C_E * tf)  # add_field.jl:1210
    t = o.add_stt_slonczewski(J, beta, Lam, xi, P, p)
    h, _ = o.term_field(t)
    m = o.spin
    beta_t = H_BAR * J / (MU0 * C_E * tf * Ms)
    eps = P * Lam ** 2 / (Lam ** 2 + 1 + (Lam ** 2 - 1) * m.dot(p))
    expected = beta_t * (eps * np.cross(m, p) + xi * p)
    assert np.allclose(h, expected, rtol=1e-13, atol=0)


def test_slonczewski_time_profile_scales_the_field():
    o = _one_cell((0.6, 0.8, 0.0), norm=True)
    beta = H_BAR / (MU0 * 8e5 * C_E * 2e-9)
    t = o.add_stt_slonczewski(np.full(1, 1e12), beta, 2.0, 0.1, 0.5, (1.0, 0.0, 0.0))
    h1 = o.term_field(t)[0]
    o.set_term_time_fn(t, lambda tt: 2.5 if tt > 1e-12 else 1.0)
    assert np.array_equal(o.term_field(t, t=0.0)[0], h1)
    assert np.allclose(o.term_field(t, t=2e-12)[0], 2.5 * h1, rtol=1e-15, atol=0)


def _helix(nx, k, dx):
    x = np.arange(nx) * dx
    m = np.stack([np.cos(k * x), np.sin(k * x), np.zeros(nx)], axis=1)
    return m.reshape(-1), k * x


@pytest.mark.parametrize("pbc", ["", "x"])
def test_zhang_li_helix_closed_form(pbc):
    nx, dx, gamma = 48, 2e-9, 2.21e5
    k = 2 * np.pi / (nx * dx) * (3 if pbc else 2.5)
    o = OracleSim(nx, 1, 1, dx, 3e-9, 1e-9, pbc=pbc)
    o.set_Ms(8e5)
    o.set_llg(0.1, gamma, True)
    m, th = _helix(nx, k, dx)
    o.init_m0(m, norm=False)
    b, Jx, xi = 7.3, 2.0, 0.15
    t = o.add_stt_zhangli((Jx, 0.5, -0.25), b, np.full(nx, xi), xi_factor=False)
    h = o.term_field(t)[0].reshape(nx, 3)
    A = b * Jx * np.sin(k * dx) / dx / gamma
    exp = np.stack([-xi * A * np.sin(th), xi * A * np.cos(th), np.full(nx, A)], axis=1)
    inner = slice(0, nx) if pbc else slice(1, nx - 1)
    assert np.allclose(h[inner], exp[inner], rtol=1e-9, atol=1e-9 * A)
    if not pbc:
        # one-sided ends: (m2 - m1)/dx
        for i, (i1, i2) in ((0, (0, 1)), (nx - 1, (nx - 2, nx - 1))):
            mr = m.reshape(nx, 3)
            f = b * Jx / dx / gamma * (mr[i2] - mr[i1])
            e = np.cross(mr[i], f) + xi * f
            assert np.allclose(h[i], e, rtol=1e-12, atol=1e-12 * A)


def test_zhang_li_P_Ms_form_scales_by_xi():
    nx, dx = 16, 2e-9
    o1 = OracleSim(nx, 1, 1, dx, dx, dx)
    o2 = OracleSim(nx, 1, 1, dx, dx, dx)
    m, _ = _helix(nx, 0.3e9, dx)
    xi = np.linspace(0.0, 0.4, nx)
    P, Ms = 0.5, 8e5
    b = P * MU_B / (C_E * Ms)
    for o in (o1, o2):
        o.set_Ms(Ms)
        o.init_m0(m)
    t1 = o1.add_stt_zhangli((1e12, 0, 0), b, xi, xi_factor=True)
    t2 = o2.add_stt_zhangli((1e12, 0, 0), b, xi, xi_factor=False)
    h1 = o1.term_field(t1)[0].reshape(nx, 3)
    h2 = o2.term_field(t2)[0].reshape(nx, 3)
    assert np.allclose(h1, h2 / (1 + xi ** 2)[:, None], rtol=1e-13, atol=1e-30)


def test_torques_carry_no_energy_and_enter_the_rhs():
    nx, ny = 6, 5
    o = OracleSim(nx, ny, 1, 2e-9, 2e-9, 2e-9)
    o.set_Ms(8e5)
    rng = np.random.default_rng(3)
    m = rng.standard_normal(3 * nx * ny)
    o.init_m0(m)
    o.add_exch(1.3e-11)
    t = o.add_sot(np.full(nx * ny, 3e9), 1e9, (0, 1, 0))
    E = o.term_energies()
    assert E[t] == 0
    H = o.effective_field()
    H_ex = o.term_field(0)[0]
    H_t = o.term_field(t)[0]
    assert np.allclose(H, H_ex + H_t, rtol=1e-15, atol=0)
