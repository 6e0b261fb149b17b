"""Interlayer exchange and DMI of the oracle (add_exch_int / add_dmi_int,
kernels.jl:496-584) against the reference's own test/test_interlayer.jl:
the expected fields of test_interlayer (layers 1 and 3 of a 3x5x4 mesh with
m = (1,2,3) and (4,5,6)) and test_interlayer_exch's zero sum of the two
layers' fields for opposite magnetisations; the energies follow
E = -1/2 mu0Ms V m.H on the two layers.
"""
import numpy as np

from oracle import OracleSim

MU0 = 4e-7 * np.pi


def _layers(nx, ny, nz, fk):
    return np.array([fk(k) for k in range(1, nz + 1) for _ in range(ny) for _ in range(nx)], dtype=float)


def test_interlayer_matches_reference_test():
    """test/test_interlayer.jl:40-79."""
    J, Ms, dz = 1e-5, 8e5, 2e-9
    Dx, Dy, Dz = 2e-3, 1e-4, 3e-5
    nx, ny, nz = 3, 5, 4
    o = OracleSim(nx, ny, nz, 2e-9, 2e-9, 2e-9, pbc="xy")
    o.set_Ms(_layers(nx, ny, nz, lambda k: 8e5 if k in (1, 3) else 0.0))
    m0 = np.array([{1: (1, 2, 3), 3: (4, 5, 6)}.get(k, (1, 1, 1)) for k in range(1, nz + 1)
                   for _ in range(ny * nx)], dtype=float).reshape(-1)
    o.init_m0(m0, norm=False)
    ke = o.add_exch_int(J, k1=1, k2=3)
    kd = o.add_dmi_int((Dx, Dy, Dz), k1=1, k2=3)
    he, ee = o.term_field(ke)
    hd, _ = o.term_field(kd)
    c = J / (MU0 * Ms * dz)
    per_layer = [np.array([4, 5, 6]) * c, np.zeros(3), np.array([1, 2, 3]) * c, np.zeros(3)]
    exp_e = np.concatenate([np.tile(v, nx * ny) for v in per_layer])
    np.testing.assert_allclose(he, exp_e, rtol=1e-12, atol=0)
    g = 1 / (MU0 * Ms * dz)
    f1 = g * np.cross([Dx, Dy, Dz], [4.0, 5.0, 6.0])
    f2 = g * np.cross([1.0, 2.0, 3.0], [Dx, Dy, Dz])
    exp_d = np.concatenate([np.tile(v, nx * ny) for v in (f1, np.zeros(3), f2, np.zeros(3))])
    np.testing.assert_allclose(hd, exp_d, rtol=1e-12, atol=1e-12 * np.abs(exp_d).max())
    V = 8e-27
    m = m0.reshape(-1, 3)
    mu0Ms = MU0 * _layers(nx, ny, nz, lambda k: 8e5 if k in (1, 3) else 0.0)
    np.testing.assert_allclose(ee, -0.5 * mu0Ms * np.sum(m * he.reshape(-1, 3), axis=1) * V, rtol=1e-12)


def test_interlayer_exch_opposite_layers_cancel():
    """test/test_interlayer.jl:82-112: f(layer 1) + f(layer 3) == 0 for m(3) = -m(1)."""
    nx, ny, nz = 30, 80, 3
    o = OracleSim(nx, ny, nz, 2e-9, 2e-9, 2e-9, pbc="xy")
    o.set_Ms(_layers(nx, ny, nz, lambda k: 8e5 if k in (1, 3) else 0.0))
    m = np.zeros((nz, ny, nx, 3))
    for j in range(ny):
        for i in range(nx):
            inside = (i + 1 - 10) ** 2 + (j + 1 - 30) ** 2 < 400
            m[2, j, i] = (0, -0.01, -1) if inside else (0, 0.01, 1)
            m[0, j, i] = (0, 0.01, 1) if inside else (0, -0.01, -1)
    o.init_m0(m.reshape(-1))
    k = o.add_exch_int(np.full(nx * ny, 1e-5), k1=1, k2=3)
    f = o.term_field(k)[0].reshape(nz, ny, nx, 3)
    s = f[0] + f[2]
    assert s.max() == 0 and s.min() == 0
