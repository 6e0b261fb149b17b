"""The surface-charge field of rm_demag_charges (compute_demag_charge_field,
src/micro/demagc.jl:38-83), host code of the mirror, pinned by closed forms:
at the centre of a cube with charges +-Ms on the x faces (uniform m along x)
the field is +Ms/3 along x (it cancels the cube's demag field -Ms/3, N = 1/3
by symmetry) with no transverse part; the y and z faces give the same by
permutation; opposite charges on both faces (head-to-head) give no field at
the centre.
"""
import numpy as np

import mmx


def _centre(field, n):
    f = field.reshape(n, n, n, 3)
    return f[n // 2, n // 2, n // 2]


def test_cube_centre_is_a_third_of_Ms():
    n, Ms = 21, 8e5
    mesh = mmx.FDMesh(nx=n, ny=n, nz=n, dx=2e-9, dy=2e-9, dz=2e-9)
    for axis, kw in ((0, dict(x=(-1, 1))), (1, dict(y=(-1, 1))), (2, dict(z=(-1, 1)))):
        h = _centre(mmx.demag_charge_field(mesh, Ms, **kw), n)
        exp = np.zeros(3)
        exp[axis] = Ms / 3
        np.testing.assert_allclose(h, exp, rtol=1e-12, atol=1e-9 * Ms)


def test_head_to_head_centre_is_zero():
    n, Ms = 21, 8e5
    mesh = mmx.FDMesh(nx=n, ny=n, nz=n, dx=2e-9, dy=2e-9, dz=2e-9)
    h = _centre(mmx.demag_charge_field(mesh, Ms, x=(-1, -1)), n)
    np.testing.assert_allclose(h, 0.0, atol=1e-9 * Ms)
