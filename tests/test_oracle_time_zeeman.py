"""Time-dependent Zeeman field of the oracle (add_zeeman(sim, H0, ft)) pinned by
the reference's test/test_zeeman.jl:8-63 (exact products H0 * f(t), energies
equal to the static term's when f = 1) and by a closed form: a macrospin with
alpha = 0 in H(t) = H0 f(t) z precesses at phi(t) = phi0 + gamma int_0^t H, so
phi(t) = gamma H0 (t + a (1 - cos wt) / w) for f = 1 + a sin wt -- a stepper
that evaluates the field at the wrong stage times misses it.
"""
import math

import numpy as np

from oracle import OracleSim


def time_fun(t):  # test/test_zeeman.jl:8-10
    return (math.cos(1e6 * t), math.cos(1e9 * t), 1.0)


def time_fun2(t):  # test/test_zeeman.jl:12-14
    return math.cos(1e9 * t)


def _sim():
    o = OracleSim(3, 2, 1, 2e-9, 2e-9, 2e-9, pbc="x")
    o.set_llg(0.01, 2.21e5, True)
    o.set_Ms(8.6e5)
    o.init_m0(np.tile([1.0, 1.0, 1.0], 6), norm=False)
    return o


def test_time_zeeman_matches_reference_test():
    o = _sim()
    z1 = o.add_zeeman((1, 2, 2e3))
    z2 = o.add_zeeman_time((1, 2, 2e3), time_fun)
    z3 = o.add_zeeman_time((1, 2, 2e3), time_fun2)
    t = 1.23e-9
    f1, e1 = o.term_field(z1, t=0.0)
    f2, e2 = o.term_field(z2, t=0.0)
    f3, _ = o.term_field(z3, t=t)
    assert f1[0] == 1.0 and f1[1] == 2.0 and f1[2] == 2e3
    assert np.array_equal(f1, f2)
    c = math.cos(1e9 * t)
    assert f3[0] == 1.0 * c and f3[1] == 2.0 * c and f3[2] == 2e3 * c
    assert abs(e1.sum()) > 0 and e1.sum() == e2.sum()
    z2 = o.add_zeeman_time((1e3, 1e4, 1e5), time_fun)
    f2, _ = o.term_field(z2, t=1.23e-11)
    assert f2[0] == 1e3 * math.cos(1.23e-11 * 1e6)
    assert f2[1] == 1e4 * math.cos(1.23e-2)
    assert f2[2] == 1e5


def test_time_zeeman_energy_is_minus_mu0_Ms_m_dot_h():
    o = OracleSim(4, 3, 2, 1e-9, 2e-9, 3e-9)
    rng = np.random.default_rng(4)
    Ms = rng.uniform(1e5, 9e5, 24)
    o.set_Ms(Ms)
    o.init_m0(rng.standard_normal(72))
    H = rng.standard_normal(72) * 1e4
    z = o.add_zeeman_time(H, lambda t: (2.0, -0.5, 3.0))
    h, e = o.term_field(z, t=7e-12)
    assert np.array_equal(h, H * np.tile([2.0, -0.5, 3.0], 24))
    m = o.spin.reshape(-1, 3)
    expect = -6e-27 * 4e-7 * np.pi * Ms * np.sum(m * h.reshape(-1, 3), axis=1)
    assert np.allclose(e, expect, rtol=1e-13, atol=0)


def test_macrospin_precession_in_modulated_field():
    gamma, H0, a, w = 2.21e5, 1e5, 0.5, 2 * math.pi / 4e-11
    o = OracleSim(1, 1, 1, 1e-9, 1e-9, 1e-9)
    o.set_Ms(8e5)
    o.set_llg(0.0, gamma, True)
    o.init_m0(np.array([1.0, 0.0, 0.0]))
    o.add_zeeman_time((0, 0, H0), lambda t: 1 + a * math.sin(w * t))
    o.set_integrator(0, 5e-14, 0)
    for _ in range(200):
        o.advance_step()
    t = 200 * 5e-14
    phi = gamma * H0 * (t + a * (1 - math.cos(w * t)) / w)
    m = o.spin
    got = math.atan2(m[1], m[0])
    assert abs(got - phi) < 1e-9 and abs(m[2]) < 1e-12
    # the field sampled only at step starts (t_n instead of the stage times) misses it
    phi_bad = gamma * H0 * sum(1 + a * math.sin(w * k * 5e-14) for k in range(200)) * 5e-14
    assert abs(phi_bad - phi) > 1e-4
