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


# ---- the reference's own NEB checks (test/test_neb.jl) -------------------------
def test_slerp_reference_values():
    """test/test_neb.jl:29-39: a third of a quarter turn, and two thirds of a
    5/4 turn taken the short way round (-3/4 pi)."""
    t1 = np.pi / 2
    x1 = slerp(np.array([1, 0, 0.0]), np.array([np.cos(t1), np.sin(t1), 0.0]), 1 / 3)
    np.testing.assert_allclose(x1, [np.cos(t1 / 3), np.sin(t1 / 3), 0.0], atol=1e-7)
    t2 = 5 / 4 * np.pi
    t2c = t2 - 2 * np.pi
    x2 = slerp(np.array([1, 0, 0.0]), np.array([np.cos(t2), np.sin(t2), 0.0]), 2 / 3)
    np.testing.assert_allclose(x2, [np.cos(t2c * 2 / 3), np.sin(t2c * 2 / 3), 0.0], rtol=1e-8, atol=1e-15)


def test_init_images_reference_values():
    """test/test_neb.jl:41-53: (x1, x2, x3) with (2, 1) frames between them."""
    t1 = np.pi / 2
    x1, x2, x3 = np.array([1, 0, 0.0]), np.array([np.cos(t1), np.sin(t1), 0.0]), np.array([-1, 0.0, 0.0])
    imgs = init_images([x1, x2, x3], [2, 1])
    assert len(imgs) == 6
    np.testing.assert_allclose(imgs[0], x1)
    np.testing.assert_allclose(imgs[1], [np.cos(t1 / 3), np.sin(t1 / 3), 0.0], atol=1e-7)
    np.testing.assert_allclose(imgs[2], [np.cos(2 * t1 / 3), np.sin(2 * t1 / 3), 0.0], atol=1e-7)
    np.testing.assert_allclose(imgs[3], [np.cos(t1), np.sin(t1), 0.0], atol=1e-15)
    np.testing.assert_allclose(imgs[4], [np.cos(3 * np.pi / 4), np.sin(3 * np.pi / 4), 0.0], atol=1e-7)
    np.testing.assert_allclose(imgs[5], [-1, 0.0, 0.0])


def macrospin_oracle():
    """The macrospin of test/test_neb.jl:73-86: one 5 nm cube, Ms 8e5 A/m, Kx =
    5e4 J/m^3 along x and Kz = -2e4 J/m^3 along z."""
    o = OracleSim(1, 1, 1, 5e-9, 5e-9, 5e-9)
    o.set_Ms(np.array([8e5]))
    o.init_m0(np.array([0.6, 0, -0.8]))
    o.add_anis(np.array([5e4]), axis=(1, 0, 0))
    o.add_anis(np.array([-2e4]), axis=(0, 0, 1))
    return o


def test_neb_macrospin_barrier():
    """test/test_neb.jl:88-108 (LLG driver): 11 free images between (1,0,0),
    (0,1,1) and (-1,0,0); after relax(max_steps=500, stopping_dmdt=0.001) the
    band's barrier max(E) - E[1] equals Kx V = 5e4 (5e-9)^3 to 1e-3, and the
    initial band's barrier is above it."""
    o = macrospin_oracle()
    neb = OracleNEB(o, [np.array([1, 0, 0.0]), np.array([0, 1, 1.0]), np.array([-1, 0, 0.0])], [5, 5])
    assert neb.spin.shape == (11, 3)
    expected = 5e4 * 5e-9 ** 3
    init_diff = neb.energy.max() - neb.energy[0]
    neb.relax(max_steps=500, stopping_dmdt=0.001)
    final_diff = neb.energy.max() - neb.energy[0]
    print(f"init {init_diff:.6e} expected {expected:.6e} final {final_diff:.6e}")
    assert init_diff > expected
    assert abs(final_diff - expected) / expected < 1e-3
