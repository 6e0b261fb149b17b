"""Spin anomalous Hall torque of the oracle (add_sahe_torque,
sahe_torque_kernel! src/micro/kernels.jl:589-613) pinned by the reference's
test/test_sahe_torque.jl:8-30 golden field (one cell, rtol 1e-6) and by two
properties of H = (beta sigma + m x sigma)/gamma: with beta = 0 the field is
perpendicular to m, and H is affine in beta with slope sigma/gamma.
"""
import numpy as np

from oracle import OracleSim


def _cell(m, beta, **kw):
    o = OracleSim(1, 1, 1, 1e-9, 1e-9, 1e-9)
    o.set_Ms(8e5)
    o.set_llg(0.05, 2.21e5, True)
    o.init_m0(np.asarray(m, dtype=float), norm=False)
    o.add_zeeman((0, 0, 1e5))
    return o, o.add_sahe_torque((0.1, 0.2, 0.3), sigma_sa1=(0.3, 0.4, 0.5), sigma_sa2=(0.6, 0.7, 0.8),
                                beta=beta, **kw)


def test_sahe_matches_reference_test():
    o, k = _cell((-0.3, 0.4, 0.8), 0.135)
    h, e = o.term_field(k)
    np.testing.assert_allclose(h * 1e7, [2.01628959, 3.74389140, 3.30678733], rtol=1e-6)
    assert np.all(e == 0)


def test_sahe_properties():
    rng = np.random.default_rng(5)
    for _ in range(4):
        m = rng.standard_normal(3)
        kw = dict(a1=tuple(rng.standard_normal(3)), a2=tuple(rng.standard_normal(3)),
                  a3=tuple(rng.standard_normal(3)))
        o0, k0 = _cell(m, 0.0, **kw)
        h0, _ = o0.term_field(k0)
        assert abs(h0 @ m) <= 1e-12 * np.linalg.norm(h0) * np.linalg.norm(m)
        o1, k1 = _cell(m, 0.7, **kw)
        h1, _ = o1.term_field(k1)
        o2, k2 = _cell(m, 1.4, **kw)
        h2, _ = o2.term_field(k2)
        np.testing.assert_allclose(h2 - h1, h1 - h0, rtol=1e-9, atol=1e-12 * np.abs(h1).max())
