"""Time-dependent bulk DMI (add_dmi(sim, D, ft), add_field.jl:360-393) on the
HIP path, through the C ABI, against the oracle: fields and energies at
several times for a scalar, a 3-tuple and a per-cell D (TimeBulkDMI,
TimeSpatialBulkDMI), fixed-step and adaptive trajectories (stage times), the
time_D<name> table column, and the adjoint's refusal.

FP64 gate: fields within 1e-12 (max-norm relative), fixed-step trajectories
within 1e-12 absolute; FP32 fields within 1e-5.
"""
import math
import sys

import numpy as np
import pytest

import mmx

pytestmark = pytest.mark.gpu


def note(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def relerr(a, b):
    den = np.max(np.abs(b))
    return np.max(np.abs(a - b)) / (den if den > 0 else 1.0)


def ft(t):
    return 0.5 + math.cos(2 * math.pi * t / 1e-12)


def build(oracle, kind, precision=64, pbc="x", demag=True, seed=9):
    nx, ny, nz = 13, 9, 3
    d = (1e-9, 2e-9, 3e-9)
    N = nx * ny * nz
    rng = np.random.default_rng(seed)
    Ms = np.where(rng.uniform(size=N) < 0.1, 0.0, 7e5)
    m0 = rng.standard_normal(3 * N)
    mesh = mmx.FDMesh(dx=d[0], dy=d[1], dz=d[2], nx=nx, ny=ny, nz=nz, pbc=pbc or "open")
    g = mmx.Sim(mesh, precision=precision)
    o = oracle.OracleSim(nx, ny, nz, *d, pbc=pbc)
    mmx.set_Ms(g, Ms)
    o.set_Ms(Ms)
    mmx.init_m0(g, m0)
    o.init_m0(m0)
    g.driver.alpha, g.driver.gamma = 0.1, 2.21e5
    o.set_llg(0.1, 2.21e5, True)
    mmx.add_exch(g, 1.3e-11)
    o.add_exch(1.3e-11)
    D = {"scalar": 2.5e-3, "tuple": (2.5e-3, -1e-3, 4e-4),
         "spatial": rng.uniform(1e-3, 3e-3, N) * (rng.uniform(size=N) > 0.1)}[kind]
    mmx.add_dmi(g, D, ft)
    o.add_dmi_bulk_time(D, ft)
    if demag:
        mmx.add_demag(g)
        o.add_demag()
    return g, o


@pytest.mark.parametrize("kind", ["scalar", "tuple", "spatial"])
@pytest.mark.parametrize("precision", [64, 32])
def test_time_dmi_fields_match_oracle(oracle, kind, precision):
    note(f"time DMI fields {kind} FP{precision}")
    g, o = build(oracle, kind, precision=precision)
    tol = 1e-12 if precision == 64 else 1e-5
    for t in (0.0, 2.9e-13, 6.1e-13):
        for q in range(len(g.interactions)):
            fg, eg = g._term_field(q, t)
            fo, eo = o.term_field(q, t=t)
            assert relerr(fg, fo) <= tol, (q, t, relerr(fg, fo))
            assert relerr(eg, eo) <= tol, (q, t, relerr(eg, eo))
        assert relerr(mmx.llg_rhs(g, t=t), o.llg_rhs(t=t)) <= tol


@pytest.mark.parametrize("kind", ["tuple", "spatial"])
def test_time_dmi_rk4_trajectory(oracle, kind):
    note(f"time DMI RK4 trajectory {kind}")
    g, o = build(oracle, kind)
    mmx.set_driver(g, integrator="RungeKutta", step=2e-14)
    o.set_integrator(0, 2e-14, 0)
    for _ in range(30):
        o.advance_step()
    mmx.run_step(g, n=30)
    err = np.max(np.abs(g.spin - o.spin))
    print(f"max |dm| after 30 steps: {err:.2e}")
    assert err < 1e-12


def test_time_dmi_adaptive_trajectory(oracle):
    note("time DMI DormandPrince trajectory")
    g, o = build(oracle, "scalar", demag=False)
    mmx.set_driver(g, integrator="DormandPrince", tol=1e-6)
    o.set_integrator(2, 0, 1e-6)
    for i in range(1, 4):
        mmx.run_until(g, i * 3e-13)
        o.run_until(i * 3e-13)
    assert g._integrator.nsteps == o.state()["nsteps"]
    assert np.max(np.abs(g.spin - o.spin)) < 1e-10


def test_time_dmi_table_column_and_adjoint_refusal(tmp_path):
    note("time DMI table column, adjoint refusal")
    from mmx.io import read_table

    mesh = mmx.FDMesh(dx=2e-9, dy=2e-9, dz=2e-9, nx=6, ny=4, nz=2)
    sim = mmx.Sim(mesh, name=str(tmp_path / "td"))
    mmx.set_Ms(sim, 8e5)
    mmx.init_m0(sim, (1, 0.2, 0.3))
    mmx.add_exch(sim, 1e-11)
    mmx.add_dmi(sim, 2e-3, ft)
    for t in (1e-13, 2e-13):
        mmx.run_until(sim, t, save_data=True)
    data, units = read_table(sim.saver.name)
    assert list(data)[-2:] == ["E_time_dmi", "time_Dtime_dmi"]
    assert units["time_Dtime_dmi"] == "J/m^2"
    assert data["time_Dtime_dmi"][-1] == float("%.12e" % ft(2e-13))
    E = mmx.compute_system_energy(sim, t=2e-13)
    assert np.isclose(data["E_time_dmi"][-1], E[1], rtol=1e-11, atol=0)
    mmx.set_driver(sim, integrator="RungeKutta", step=1e-14)
    with pytest.raises(mmx.MmxError):
        mmx.rk4_trajectory_rrule(sim, 10, every=5)
