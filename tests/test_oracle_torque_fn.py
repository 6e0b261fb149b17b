"""add_torque (a user function of m and t, add_field.jl:998-1006) in the oracle:
the reference docstring's torque_fun gives H = (1/gamma) m x T (torque_kernel!,
kernels.jl:639-651); and with alpha = 0 and T = w z x m the LLG reduces to
dm/dt = T, a precession about z at angular rate w, which fixed-step RK4
follows to 1e-9 -- a wrong sign or a missing 1/gamma turns or speeds it
the wrong way.
"""
import math

import numpy as np

from oracle import OracleSim


def doc_torque(y, m, t):  # the example of add_field.jl:985-989
    y[:] = 0
    y[0::3] = m[0::3] * math.sin(t * 1e11)
    y[1::3] = 0.2 * math.cos(t * 1e11)


def test_field_of_docstring_torque():
    rng = np.random.default_rng(2)
    o = OracleSim(5, 4, 1, 1e-9, 1e-9, 1e-9)
    o.set_Ms(8e5)
    o.set_llg(0.1, 2.21e5, True)
    m = rng.standard_normal(60)
    o.init_m0(m)
    k = o.add_torque(doc_torque)
    t = 3.3e-12
    h, e = o.term_field(k, t=t)
    ms = o.spin.reshape(-1, 3)
    T = np.zeros_like(ms)
    T[:, 0] = ms[:, 0] * math.sin(t * 1e11)
    T[:, 1] = 0.2 * math.cos(t * 1e11)
    np.testing.assert_allclose(h.reshape(-1, 3), np.cross(ms, T) / 2.21e5, rtol=1e-13, atol=1e-16)
    assert np.all(e == 0)


def test_precession_closed_form():
    w = 2e10
    o = OracleSim(1, 1, 1, 1e-9, 1e-9, 1e-9)
    o.set_Ms(8e5)
    o.set_llg(0.0, 2.21e5, True)
    o.init_m0(np.array([1.0, 0.0, 0.0]))

    def rot(y, m, t):
        y[0::3] = -w * m[1::3]
        y[1::3] = w * m[0::3]
        y[2::3] = 0

    o.add_torque(rot)
    o.set_integrator(0, 1e-13, 0)
    for _ in range(100):
        o.advance_step()
    m = o.spin
    phi = w * 100 * 1e-13
    assert abs(m[0] - math.cos(phi)) < 1e-9 and abs(m[1] - math.sin(phi)) < 1e-9


def test_torque_buffer_persists_as_field():
    """field.jl:282-293 hands torque.field to the function and torque_kernel!
    (kernels.jl:639-651) then overwrites it in place with H = (1/gamma) m x T:
    the first call sees zeros, every later call the previous field."""
    seen = []

    def acc(y, m, t):
        seen.append(y.copy())
        y[:] = y + 1e9  # builds on what it was handed

    o = OracleSim(3, 2, 1, 1e-9, 1e-9, 1e-9)
    o.set_Ms(8e5)
    o.set_llg(0.1, 2.21e5, True)
    m = np.random.default_rng(4).standard_normal(18)
    o.init_m0(m)
    k = o.add_torque(acc)
    h1, _ = o.term_field(k)
    h2, _ = o.term_field(k)
    assert np.all(seen[0] == 0)
    np.testing.assert_array_equal(seen[1], h1)
    ms = o.spin.reshape(-1, 3)
    np.testing.assert_allclose(h2.reshape(-1, 3), np.cross(ms, h1.reshape(-1, 3) + 1e9) / 2.21e5, rtol=1e-13,
                               atol=1e-20)
