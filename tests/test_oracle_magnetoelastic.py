"""Magnetoelastic term of the oracle (add_magnetoelastic / add_mel,
kernels.jl:742-845) against the reference's test/test_magnetoelastic.jl (the
tensor model's energy and Hx for a uniaxial stress; the cubic model's B1 and
shear B2 cases) and H = -(1/(mu0Ms V)) dE/dm for random stress / strain.
"""
import numpy as np

from oracle import OracleSim

MU0 = 4e-7 * np.pi
MS = 8.6e5


def _line(m, model, **kw):
    o = OracleSim(10, 1, 1, 1e-9, 1e-9, 1e-9)
    o.set_Ms(MS)
    o.init_m0(np.tile(np.asarray(m, dtype=float), 10))
    return o, o.add_magnetoelastic(model=model, **kw)


def _voigt(comp, value, n=10):
    d = np.zeros((n, 6))
    d[:, comp] = value
    return d.reshape(-1)


def test_tensor_matches_reference_test():
    """test/test_magnetoelastic.jl:8-47."""
    ls = 30e-6
    o, k = _line((1, 0, 0), "tensor", lambda_s=ls, data=_voigt(0, 1e8))
    h, e = o.term_field(k)
    V = 1e-27
    np.testing.assert_allclose(e.sum(), -1.5 * ls * 1e8 * V * 10, rtol=1e-12)
    np.testing.assert_allclose(h[0], 3 * ls / (MU0 * MS) * 1e8, rtol=1e-12)


def test_cubic_matches_reference_test():
    """test/test_magnetoelastic.jl:49-113."""
    B1, B2, eps, V = 1e6, 0.5e6, 1e-3, 1e-27
    d = _voigt(0, eps) + _voigt(1, eps)
    o, k = _line((1, 0, 0), "cubic", B1=B1, B2=B2, data=d)
    h, e = o.term_field(k)
    np.testing.assert_allclose(e.sum(), B1 * eps * V * 10, rtol=1e-12)
    np.testing.assert_allclose(h[0], -2 / (MU0 * MS) * B1 * eps, rtol=1e-12)
    assert h[1] == 0 and h[2] == 0
    o, k = _line((1 / np.sqrt(2), 1 / np.sqrt(2), 0), "cubic", B1=0, B2=B2, data=_voigt(3, eps))
    h, e = o.term_field(k)
    np.testing.assert_allclose(e.sum(), B2 * eps * V * 10, rtol=1e-12)
    exp = -2 / (MU0 * MS) * B2 * eps / np.sqrt(2)
    np.testing.assert_allclose(h[:2], [exp, exp], rtol=1e-12)


def test_field_is_minus_energy_gradient():
    rng = np.random.default_rng(3)
    V = 1e-27
    for model, kw in (("tensor", dict(lambda_s=2e-5)), ("cubic", dict(B1=-7e6, B2=3e6))):
        for _ in range(3):
            m = rng.standard_normal(3)
            q = rng.uniform(-1e8, 1e8, 6) if model == "tensor" else rng.uniform(-1e-3, 1e-3, 6)
            o = OracleSim(1, 1, 1, 1e-9, 1e-9, 1e-9)
            o.set_Ms(MS)
            o.init_m0(m, norm=False)
            k = o.add_magnetoelastic(model=model, data=q, **kw)
            h, _ = o.term_field(k)
            g = np.empty(3)
            for a in range(3):
                d = np.zeros(3)
                d[a] = 1e-6
                g[a] = (o.term_field(k, spin=m + d)[1][0] - o.term_field(k, spin=m - d)[1][0]) / 2e-6
            np.testing.assert_allclose(h, -g / (MU0 * MS * V), rtol=1e-6, atol=1e-6 * np.abs(h).max())
