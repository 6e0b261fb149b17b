"""CPU checks of the oracle's NEB restatement (oracle/neb.py) against facts the
geometry fixes: slerp end points and great-circle property, geodesic distance
of known rotations, tangents perpendicular to the images."""
import numpy as np

from oracle import OracleSim
from oracle_neb import OracleNEB, geodesic, init_images, slerp

MU0 = 4e-7 * np.pi


def _unit(rng, n):
    v = rng.standard_normal((n, 3))
    return (v / np.linalg.norm(v, axis=1, keepdims=True)).reshape(-1)


def test_slerp_endpoints_and_great_circle():
    rng = np.random.default_rng(1)
    a, b = _unit(rng, 50), _unit(rng, 50)
    np.testing.assert_allclose(slerp(a, b, 0.0), a, atol=1e-14)
    np.testing.assert_allclose(slerp(a, b, 1.0), b, atol=1e-14)
    m = slerp(a, b, 0.3).reshape(-1, 3)
    np.testing.assert_allclose(np.linalg.norm(m, axis=1), 1.0, atol=1e-14)
    # angle from a is 0.3 of the total angle
    ang = geodesic(a, b)
    np.testing.assert_allclose(geodesic(a, m.reshape(-1)), 0.3 * ang, atol=1e-12)
    # cells with theta == 0 are left at zero by slerp_kernel! (math.jl:52-60)
    e = np.tile([0.0, 0.0, 1.0], 4)
    assert np.all(slerp(e, e.copy(), 0.5) == 0)


def test_geodesic_of_rotation():
    th = np.linspace(0, np.pi, 7)
    a = np.tile([1.0, 0, 0], 7)
    b = np.stack([np.cos(th), np.sin(th), 0 * th], axis=1).reshape(-1)
    np.testing.assert_allclose(geodesic(a, b), th, atol=1e-15)


def test_init_images_count_and_order():
    rng = np.random.default_rng(2)
    g = [_unit(rng, 4) for _ in range(3)]
    imgs = init_images(g, [2, 3])
    assert len(imgs) == 3 + 5
    assert imgs[0] is g[0] and imgs[3] is g[1] and imgs[-1] is g[2]


def test_tangents_perpendicular_and_unit():
    nx, ny = 6, 5
    o = OracleSim(nx, ny, 1, 2e-9, 2e-9, 2e-9)
    o.set_Ms(np.full(nx * ny, 4e5))
    o.add_exch(5e-12)
    o.add_zeeman((0, 0, 0.1 / MU0))
    rng = np.random.default_rng(3)
    neb = OracleNEB(o, [_unit(rng, nx * ny), np.tile([0, 0, 1.0], nx * ny)], [4])
    t = neb._tangents(neb.spin)
    for q in range(neb.n):
        assert abs(np.linalg.norm(t[q]) - 1) < 1e-14
        dots = (t[q].reshape(-1, 3) * neb.spin[q].reshape(-1, 3)).sum(axis=1)
        assert np.abs(dots).max() < 1e-14
