"""Hexagonal anisotropy of the oracle (add_hex_anis, hexagonal_anisotropy_kernel!
src/micro/kernels.jl:142-167) pinned by what the energy fixes:

* in the plane, m = (cos p, sin p, 0): E = (K1 + K2 + K3 cos 6p) V and
  H = -(6 K3 / mu0Ms) (cos 5p, -sin 5p, 0) (the real and imaginary parts of
  (mx + i my)^5 and ^6);
* along z: E = 0 and H = (0, 0, 2 K1 / mu0Ms);
* for any m, H = -(1 / (mu0Ms V)) dE/dm, checked with central differences of
  the oracle's own energy (a dropped term, a wrong sign or coefficient in the
  field breaks this).
"""
import numpy as np

from oracle import OracleSim

K1, K2, K3 = 4e5, -1.2e5, 3e4
MS = 6e5
MU0 = 4e-7 * np.pi


def _cell(m):
    o = OracleSim(1, 1, 1, 2e-9, 3e-9, 4e-9)
    o.set_Ms(MS)
    o.init_m0(np.asarray(m, dtype=float), norm=False)
    k = o.add_hex_anis(K1, K2, K3)
    return o, k


def test_in_plane_closed_form():
    V = 24e-27
    for p in np.linspace(0, 2 * np.pi, 13):
        o, k = _cell((np.cos(p), np.sin(p), 0.0))
        h, e = o.term_field(k)
        c = -6 * K3 / (MU0 * MS)
        np.testing.assert_allclose(h, [c * np.cos(5 * p), -c * np.sin(5 * p), 0.0], rtol=1e-12, atol=1e-12 * abs(c))
        np.testing.assert_allclose(e[0], (K1 + K2 + K3 * np.cos(6 * p)) * V, rtol=1e-12, atol=1e-12 * K1 * V)


def test_along_z():
    o, k = _cell((0.0, 0.0, 1.0))
    h, e = o.term_field(k)
    np.testing.assert_allclose(h, [0, 0, 2 * K1 / (MU0 * MS)], rtol=1e-14)
    assert e[0] == 0


def test_field_is_minus_energy_gradient():
    rng = np.random.default_rng(12)
    V = 24e-27
    for _ in range(5):
        m = rng.standard_normal(3)
        o, k = _cell(m)
        h, _ = o.term_field(k)
        g = np.empty(3)
        for a in range(3):
            d = np.zeros(3)
            d[a] = 1e-6
            ep = o.term_field(k, spin=m + d)[1][0]
            em = o.term_field(k, spin=m - d)[1][0]
            g[a] = (ep - em) / 2e-6
        np.testing.assert_allclose(h, -g / (MU0 * MS * V), rtol=1e-6, atol=1e-6 * np.abs(h).max())
