"""Interlayer exchange and DMI (add_exch_int / add_dmi_int, kernels.jl:496-584)
on the HIP path, through the C ABI: the reference's test/test_interlayer.jl
checks, fields / energies / dm/dt against the oracle on a mesh with empty
cells and demag, and an RK4 trajectory.

FP64 gate: fields within 1e-12 (max-norm relative), trajectories within 1e-12
absolute; FP32 fields within 1e-5.
"""
import sys

import numpy as np
import pytest

import mmx

pytestmark = pytest.mark.gpu

MU0 = 4e-7 * np.pi


def note(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def relerr(a, b):
    den = np.max(np.abs(b))
    return np.max(np.abs(a - b)) / (den if den > 0 else 1.0)


def test_interlayer_reference_checks():
    """test/test_interlayer.jl:40-112 on the HIP path."""
    note("test_interlayer.jl on the HIP path")
    J, Ms, dz = 1e-5, 8e5, 2e-9
    Dx, Dy, Dz = 2e-3, 1e-4, 3e-5
    mesh = mmx.FDMesh(dx=2e-9, dy=2e-9, dz=2e-9, nx=3, ny=5, nz=4, pbc="xy")
    sim = mmx.Sim(mesh)
    mmx.set_Ms(sim, lambda i, j, k, dx, dy, dz: 8e5 if k in (1, 3) else 0.0)
    mmx.init_m0(sim, lambda i, j, k, dx, dy, dz: {1: (1, 2, 3), 3: (4, 5, 6)}.get(k, (1, 1, 1)), norm=False)
    exch = mmx.add_exch_int(sim, J, k1=1, k2=3)
    dmi = mmx.add_dmi_int(sim, (Dx, Dy, Dz), k1=1, k2=3)
    mmx.effective_field(sim)
    c = J / (MU0 * Ms * dz)
    exp_e = np.concatenate([np.tile(v, 15) for v in (np.array([4, 5, 6]) * c, np.zeros(3),
                                                      np.array([1, 2, 3]) * c, np.zeros(3))])
    np.testing.assert_allclose(exch.field, exp_e, rtol=1e-12, atol=0)
    g = 1 / (MU0 * Ms * dz)
    f1 = g * np.cross([Dx, Dy, Dz], [4.0, 5.0, 6.0])
    f2 = g * np.cross([1.0, 2.0, 3.0], [Dx, Dy, Dz])
    exp_d = np.concatenate([np.tile(v, 15) for v in (f1, np.zeros(3), f2, np.zeros(3))])
    np.testing.assert_allclose(dmi.field, exp_d, rtol=1e-12, atol=1e-12 * np.abs(exp_d).max())

    nx, ny, nz = 30, 80, 3
    mesh = mmx.FDMesh(dx=2e-9, dy=2e-9, dz=2e-9, nx=nx, ny=ny, nz=nz, pbc="xy")
    sim = mmx.Sim(mesh)
    mmx.set_Ms(sim, lambda i, j, k, dx, dy, dz: 8e5 if k in (1, 3) else 0.0)

    def skx(i, j, k, dx, dy, dz):
        inside = (i - 10) ** 2 + (j - 30) ** 2 < 400
        if k == 3:
            return (0, -0.01, -1) if inside else (0, 0.01, 1)
        if k == 1:
            return (0, 0.01, 1) if inside else (0, -0.01, -1)
        return (0, 0, 1)

    mmx.init_m0(sim, skx)
    exch = mmx.add_exch_int(sim, lambda i, j, k, dx, dy, dz: 1e-5, k1=1, k2=3)
    f = exch.field.reshape(nz, ny, nx, 3)
    s = f[0] + f[2]
    assert s.max() == 0 and s.min() == 0


def build(oracle, precision=64, demag=True, seed=17):
    nx, ny, nz = 10, 7, 5
    d = (2e-9, 2e-9, 1.5e-9)
    N = nx * ny * nz
    rng = np.random.default_rng(seed)
    Ms = np.where(rng.uniform(size=N) < 0.1, 0.0, 6e5)
    m0 = rng.standard_normal(3 * N)
    Js = rng.uniform(-2e-4, 2e-4, nx * ny)
    mesh = mmx.FDMesh(dx=d[0], dy=d[1], dz=d[2], nx=nx, ny=ny, nz=nz, pbc="x")
    g = mmx.Sim(mesh, precision=precision)
    o = oracle.OracleSim(nx, ny, nz, *d, pbc="x")
    mmx.set_Ms(g, Ms)
    o.set_Ms(Ms)
    mmx.init_m0(g, m0)
    o.init_m0(m0)
    g.driver.alpha, g.driver.gamma = 0.1, 2.21e5
    o.set_llg(0.1, 2.21e5, True)
    mmx.add_exch(g, 1.1e-11)
    o.add_exch(1.1e-11)
    mmx.add_exch_int(g, Js.reshape(ny, nx), k1=2, k2=4)
    o.add_exch_int(Js, k1=2, k2=4)
    mmx.add_dmi_int(g, (1e-3, -2e-3, 5e-4), k1=1)
    o.add_dmi_int((1e-3, -2e-3, 5e-4), k1=1)
    if demag:
        mmx.add_demag(g)
        o.add_demag()
    return g, o


@pytest.mark.parametrize("precision", [64, 32])
def test_interlayer_fields_match_oracle(oracle, precision):
    note(f"interlayer fields FP{precision}")
    g, o = build(oracle, precision=precision)
    tol = 1e-12 if precision == 64 else 1e-5
    for q in range(len(g.interactions)):
        fg, eg = g._term_field(q)
        fo, eo = o.term_field(q)
        assert relerr(fg, fo) <= tol, (q, relerr(fg, fo))
        assert relerr(eg, eo) <= tol, (q, relerr(eg, eo))
    assert relerr(mmx.llg_rhs(g), o.llg_rhs()) <= tol


def test_interlayer_rk4_trajectory(oracle):
    note("interlayer RK4 trajectory")
    g, o = build(oracle)
    mmx.set_driver(g, integrator="RungeKutta", step=2e-14)
    o.set_integrator(0, 2e-14, 0)
    for _ in range(30):
        o.advance_step()
    mmx.run_step(g, n=30)
    err = np.max(np.abs(g.spin - o.spin))
    print(f"max |dm| after 30 steps: {err:.2e}")
    assert err < 1e-12
