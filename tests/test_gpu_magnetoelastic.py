"""Magnetoelastic term (add_magnetoelastic / add_mel, kernels.jl:742-845) on
the HIP path, through the C ABI: the reference's test/test_magnetoelastic.jl
checks with (6, nx, ny, nz) input arrays, then both models with random
per-cell stress / strain against the oracle (FP64 1e-12, FP32 1e-5) and an
RK4 trajectory.
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


def test_magnetoelastic_reference_checks():
    note("test_magnetoelastic.jl on the HIP path")
    Ms, V = 8.6e5, 1e-27
    mesh = mmx.FDMesh(nx=10, ny=1, nz=1, dx=1e-9, dy=1e-9, dz=1e-9)
    sim = mmx.Sim(mesh)
    mmx.set_Ms(sim, Ms)
    sigma = np.zeros((6, 10, 1, 1))
    sigma[0] = 1e8
    mmx.init_m0(sim, (1, 0, 0))
    me = mmx.add_mel(sim, model="tensor", lambda_s=30e-6, sigma=sigma)
    mmx.effective_field(sim)
    np.testing.assert_allclose(me.energy.sum(), -1.5 * 30e-6 * 1e8 * V * 10, rtol=1e-12)
    np.testing.assert_allclose(me.field[0], 3 * 30e-6 / (MU0 * Ms) * 1e8, rtol=1e-12)

    B1, B2, eps = 1e6, 0.5e6, 1e-3
    sim = mmx.Sim(mesh)
    mmx.set_Ms(sim, Ms)
    strain = np.zeros((6, 10, 1, 1))
    strain[0] = eps
    strain[1] = eps
    mmx.init_m0(sim, (1, 0, 0))
    me = mmx.add_mel(sim, model="cubic", B1=B1, B2=B2, strain=strain)
    np.testing.assert_allclose(me.energy.sum(), B1 * eps * V * 10, rtol=1e-12)
    f = me.field
    np.testing.assert_allclose(f[0], -2 / (MU0 * Ms) * B1 * eps, rtol=1e-12)
    assert f[1] == 0 and f[2] == 0
    sim = mmx.Sim(mesh)
    mmx.set_Ms(sim, Ms)
    strain2 = np.zeros((6, 10, 1, 1))
    strain2[3] = eps
    mmx.init_m0(sim, (1 / np.sqrt(2), 1 / np.sqrt(2), 0))
    me = mmx.add_mel(sim, model="cubic", B1=0, B2=B2, strain=strain2)
    np.testing.assert_allclose(me.energy.sum(), B2 * eps * V * 10, rtol=1e-12)
    exp = -2 / (MU0 * Ms) * B2 * eps / np.sqrt(2)
    np.testing.assert_allclose(me.field[:2], [exp, exp], rtol=1e-12)
    with pytest.raises(ValueError):
        mmx.add_mel(sim, model="tensor", sigma=np.zeros((6, 9, 1, 1)))


@pytest.mark.parametrize("model", ["tensor", "cubic"])
@pytest.mark.parametrize("precision", [64, 32])
def test_magnetoelastic_matches_oracle(oracle, model, precision):
    note(f"magnetoelastic {model} FP{precision}")
    nx, ny, nz = 9, 7, 3
    N = nx * ny * nz
    rng = np.random.default_rng(41)
    Ms = np.where(rng.uniform(size=N) < 0.1, 0.0, 7e5)
    m0 = rng.standard_normal(3 * N)
    scale = 1e8 if model == "tensor" else 1e-3
    arr = rng.uniform(-scale, scale, (6, nx, ny, nz))
    data = arr.reshape(-1, order="F")  # cell-major Voigt, as Julia's (6, nx, ny, nz) array
    kw = dict(lambda_s=2e-5) if model == "tensor" else dict(B1=-7e6, B2=3e6)
    mesh = mmx.FDMesh(dx=2e-9, dy=2e-9, dz=2e-9, nx=nx, ny=ny, nz=nz, pbc="xy")
    g = mmx.Sim(mesh, precision=precision)
    o = oracle.OracleSim(nx, ny, nz, 2e-9, 2e-9, 2e-9, pbc="xy")
    mmx.set_Ms(g, Ms)
    o.set_Ms(Ms)
    mmx.init_m0(g, m0)
    o.init_m0(m0)
    g.driver.alpha, g.driver.gamma = 0.1, 2.21e5
    o.set_llg(0.1, 2.21e5, True)
    mmx.add_exch(g, 1e-11)
    o.add_exch(1e-11)
    if model == "tensor":
        mmx.add_mel(g, model=model, sigma=arr, **kw)
    else:
        mmx.add_mel(g, model=model, strain=arr, **kw)
    o.add_magnetoelastic(model=model, data=data, **kw)
    tol = 1e-12 if precision == 64 else 1e-5
    for q in range(2):
        fg, eg = g._term_field(q)
        fo, eo = o.term_field(q)
        assert relerr(fg, fo) <= tol, (q, relerr(fg, fo))
        assert relerr(eg, eo) <= tol, (q, relerr(eg, eo))
    assert relerr(mmx.llg_rhs(g), o.llg_rhs()) <= tol
    if precision == 64:
        mmx.set_driver(g, integrator="RungeKutta", step=1e-14)
        o.set_integrator(0, 1e-14, 0)
        for _ in range(20):
            o.advance_step()
        mmx.run_step(g, n=20)
        assert np.max(np.abs(g.spin - o.spin)) < 1e-12
