"""Time-dependent Zeeman field (add_zeeman(sim, H0, ft)) on the HIP path,
through the C ABI: the reference's test/test_zeeman.jl:8-63 checks, fields
against the oracle at several times, fixed-step and adaptive trajectories in a
modulated rotating field (stage times), and the macrospin closed form
phi(t) = gamma H0 (t + a (1 - cos wt) / w) for f = 1 + a sin wt, alpha = 0.

FP64 gate: fields exact for uniform H0 (one product), otherwise within 1e-12
(max-norm relative); fixed-step trajectories within 1e-12 absolute; FP32
fields within 1e-5.
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


def time_fun(t):  # test/test_zeeman.jl:8-10
    return (math.cos(1e6 * t), math.cos(1e9 * t), 1.0)


def time_fun2(t):  # test/test_zeeman.jl:12-14
    return math.cos(1e9 * t)


def rotating(t):
    w = 2 * math.pi / 2e-12
    return (math.cos(w * t), math.sin(w * t), 1 + 0.3 * math.sin(0.5 * w * t))


def test_time_zeeman_reference_checks():
    note("test_zeeman.jl on the HIP path")
    mesh = mmx.FDMesh(dx=2e-9, dy=2e-9, dz=2e-9, nx=3, ny=2, nz=1, pbc="x")
    sim = mmx.Sim(mesh)
    sim.driver.alpha, sim.driver.gamma = 0.01, 2.21e5
    mmx.set_Ms(sim, 8.6e5)
    mmx.init_m0(sim, (1, 1, 1), norm=False)
    z1 = mmx.add_zeeman(sim, (1, 2, 2e3))
    z2 = mmx.add_zeeman(sim, (1, 2, 2e3), time_fun)
    z3 = mmx.add_zeeman(sim, (1, 2, 2e3), time_fun2, name="z3")
    assert z2.name == "timezeeman"
    t = 1.23e-9
    f1, e1 = z1.effective_field(0.0)
    f2, e2 = z2.effective_field(0.0)
    f3, _ = z3.effective_field(t)
    assert f1[0] == 1.0 and f1[1] == 2.0 and f1[2] == 2e3
    assert np.array_equal(f1, f2)
    c = math.cos(1e9 * t)
    assert f3[0] == 1.0 * c and f3[1] == 2.0 * c and f3[2] == 2e3 * c
    assert abs(e1.sum()) > 0 and e1.sum() == e2.sum()
    mmx.update_zeeman(sim, (3, 4.9, -10.01))
    f1 = z1.field
    assert f1[0] == 3.0 and f1[1] == 4.9 and f1[2] == -10.01
    z2 = mmx.add_zeeman(sim, (1e3, 1e4, 1e5), time_fun)
    f2, _ = z2.effective_field(1.23e-11)
    assert f2[0] == 1e3 * math.cos(1.23e-11 * 1e6)
    assert f2[1] == 1e4 * math.cos(1.23e-2)
    assert f2[2] == 1e5
    with pytest.raises(ValueError):
        mmx.add_zeeman(sim, (1, 0, 0), lambda t: (1.0, 2.0))


def build(oracle, nx, ny, nz, pbc, precision=64, seed=5, demag=True, uniform=False):
    d = (2e-9, 2.5e-9, 3e-9)
    N = nx * ny * nz
    rng = np.random.default_rng(seed)
    Ms = np.where(rng.uniform(size=N) < 0.1, 0.0, 6e5)
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
    mmx.add_anis(g, 3e5, axis=(0, 0, 1))
    o.add_anis(3e5, (0, 0, 1))
    H = (2e4, -1e4, 5e4) if uniform else rng.uniform(-1, 1, 3 * N) * 5e4
    mmx.add_zeeman(g, H if uniform else H.reshape(N, 3), rotating)
    o.add_zeeman_time(H, rotating)
    mmx.add_zeeman(g, lambda x, y, z: (0.0, 0.0, 1e4), lambda t: math.exp(-t / 1e-12))
    o.add_zeeman_time((0.0, 0.0, 1e4), lambda t: math.exp(-t / 1e-12))
    if demag:
        mmx.add_demag(g)
        o.add_demag()
    return g, o


@pytest.mark.parametrize("uniform", [False, True])
@pytest.mark.parametrize("pbc", ["", "xy"])
@pytest.mark.parametrize("precision", [64, 32])
def test_time_zeeman_fields_match_oracle(oracle, pbc, precision, uniform):
    note(f"time Zeeman fields pbc={pbc!r} FP{precision} uniform={uniform}")
    g, o = build(oracle, 14, 10, 3, pbc, precision=precision, uniform=uniform)
    tol = 1e-12 if precision == 64 else 1e-5
    for t in (0.0, 3.7e-13, 1.1e-12):
        for q in range(len(g.interactions)):
            fg, eg = g._term_field(q, t)
            fo, eo = o.term_field(q, t=t)
            assert relerr(fg, fo) <= tol, (q, t, relerr(fg, fo))
            assert relerr(eg, eo) <= tol, (q, t, relerr(eg, eo))
        assert relerr(mmx.effective_field(g, t=t), o.effective_field(t=t)) <= tol
        assert relerr(mmx.llg_rhs(g, t=t), o.llg_rhs(t=t)) <= tol


@pytest.mark.parametrize("integrator", ["RungeKutta", "Heun"])
def test_time_zeeman_fixed_step_trajectory(oracle, integrator):
    note(f"time Zeeman trajectory {integrator}")
    g, o = build(oracle, 12, 10, 2, "x")
    kind = {"RungeKutta": 0, "Heun": 1}[integrator]
    mmx.set_driver(g, integrator=integrator, step=2e-14)
    o.set_integrator(kind, 2e-14, 0)
    for _ in range(40):
        o.advance_step()
    mmx.run_step(g, n=40)
    err = np.max(np.abs(g.spin - o.spin))
    print(f"max |dm| after 40 steps: {err:.2e}")
    assert err < 1e-12


@pytest.mark.parametrize("integrator", [("DormandPrince", 2), ("BS23", 3)])
def test_time_zeeman_adaptive_trajectory(oracle, integrator):
    name, kind = integrator
    note(f"time Zeeman trajectory {name}")
    g, o = build(oracle, 12, 10, 2, "", demag=False)
    mmx.set_driver(g, integrator=name, tol=1e-6)
    o.set_integrator(kind, 0, 1e-6)
    for i in range(1, 4):
        mmx.run_until(g, i * 4e-13)
        o.run_until(i * 4e-13)
    assert g._integrator.nsteps == o.state()["nsteps"]
    assert np.max(np.abs(g.spin - o.spin)) < 1e-10


def test_macrospin_precession_closed_form():
    note("macrospin precession in a modulated field")
    gamma, H0, a, w = 2.21e5, 1e5, 0.5, 2 * math.pi / 4e-11
    mesh = mmx.FDMesh(nx=1, ny=1, nz=1, dx=1e-9, dy=1e-9, dz=1e-9)
    sim = mmx.Sim(mesh, integrator="RungeKutta")
    mmx.set_Ms(sim, 8e5)
    sim.driver.alpha, sim.driver.gamma = 0.0, gamma
    mmx.init_m0(sim, (1, 0, 0))
    mmx.add_zeeman(sim, (0, 0, H0), lambda t: 1 + a * math.sin(w * t))
    mmx.set_driver(sim, integrator="RungeKutta", step=5e-14)
    mmx.run_step(sim, n=200)
    t = 200 * 5e-14
    phi = gamma * H0 * (t + a * (1 - math.cos(w * t)) / w)
    m = sim.spin
    assert abs(math.atan2(m[1], m[0]) - phi) < 1e-9 and abs(m[2]) < 1e-12


def test_batch_refuses_time_zeeman():
    note("batch refuses a time Zeeman member")
    mesh = mmx.FDMesh(nx=8, ny=8, nz=1, dx=2e-9, dy=2e-9, dz=2e-9)
    s = mmx.Sim(mesh, integrator="RungeKutta")
    mmx.set_Ms(s, 8e5)
    mmx.init_m0(s, (1, 0, 0))
    mmx.add_exch(s, 1e-11)
    mmx.add_zeeman(s, (0, 0, 1e5), lambda t: 1.0)
    with pytest.raises(mmx.MmxError):
        mmx.Batch([s])


def test_table_columns_of_zeeman_terms(tmp_path):
    """The columns add_zeeman registers (add_field.jl:28-41, 182-194): Hx, Hy, Hz
    of the term named "zeeman", <name>_Hx.. = H0 .* f(t) of a time Zeeman term,
    its E_<name> in "J"; no column for a torque; energies at the row's time."""
    note("table columns of Zeeman terms")
    import os

    from mmx.io import read_table

    mesh = mmx.FDMesh(dx=2e-9, dy=2e-9, dz=2e-9, nx=6, ny=4, nz=1)
    sim = mmx.Sim(mesh, name=str(tmp_path / "zt"))
    mmx.set_Ms(sim, 8e5)
    mmx.init_m0(sim, (1, 1, 1))
    mmx.add_exch(sim, 1e-11)
    mmx.add_zeeman(sim, (1e3, 0, 2e3))
    mmx.add_zeeman(sim, (0, 0, 5e4), time_fun, name="tz")
    mmx.add_sot(sim, 1e9, 0.0, (0, 1, 0))
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        for t in (1e-13, 2e-13):
            mmx.run_until(sim, t, save_data=True)
        E = mmx.compute_system_energy(sim, t=2e-13)
        data, units = read_table(sim.saver.name)
    finally:
        os.chdir(cwd)
    cols = list(data)
    assert cols == ["step", "time", "E_total", "m_x", "m_y", "m_z", "E_exch", "Hx", "Hy", "Hz", "E_zeeman",
                    "tz_Hx", "tz_Hy", "tz_Hz", "E_tz"]
    assert units["E_tz"] == "J" and units["Hx"] == "A/m"
    assert data["Hx"][-1] == 1e3 and data["Hz"][-1] == 2e3
    f = time_fun(2e-13)
    assert np.isclose(data["tz_Hz"][-1], 5e4 * f[2], rtol=1e-12, atol=0)
    assert np.isclose(data["E_tz"][-1], E[2], rtol=1e-11, atol=0)
    assert np.isclose(data["E_total"][-1], E.sum(), rtol=1e-11, atol=0)


@pytest.mark.parametrize("precision", [64, 32])
def test_uniform_time_zeeman_slot_form(oracle, precision):
    """One uniform time Zeeman term among the film's terms takes the specialised
    (slot-form) kernels with stage chaining: fields at several times and an RK4
    trajectory against the oracle."""
    note(f"uniform time Zeeman in the slot form FP{precision}")
    nx, ny, nz = 16, 12, 4
    N = nx * ny * nz
    rng = np.random.default_rng(8)
    m0 = rng.standard_normal(3 * N)
    mesh = mmx.FDMesh(dx=2e-9, dy=2e-9, dz=2e-9, nx=nx, ny=ny, nz=nz, pbc="xy")
    g = mmx.Sim(mesh, precision=precision)
    o = oracle.OracleSim(nx, ny, nz, 2e-9, 2e-9, 2e-9, pbc="xy")
    mmx.set_Ms(g, 5.8e5)
    o.set_Ms(5.8e5)
    mmx.init_m0(g, m0)
    o.init_m0(m0)
    g.driver.alpha, g.driver.gamma = 0.1, 2.21e5
    o.set_llg(0.1, 2.21e5, True)
    mmx.add_exch(g, 1.5e-11)
    o.add_exch(1.5e-11)
    mmx.add_dmi(g, 3e-3, type="interfacial")
    o.add_dmi(np.full(N, 3e-3), type="interfacial")
    mmx.add_anis(g, 8e5, axis=(0, 0, 1))
    o.add_anis(8e5, (0, 0, 1))
    mmx.add_zeeman(g, (1e4, -2e4, 8e4), rotating)
    o.add_zeeman_time((1e4, -2e4, 8e4), rotating)
    mmx.add_demag(g, Nx=1, Ny=1)
    o.add_demag(Nx=1, Ny=1)
    tol = 1e-12 if precision == 64 else 1e-5
    for t in (0.0, 4.4e-13):
        assert relerr(mmx.effective_field(g, t=t), o.effective_field(t=t)) <= tol
        assert relerr(mmx.llg_rhs(g, t=t), o.llg_rhs(t=t)) <= tol
    if precision == 64:
        mmx.set_driver(g, integrator="RungeKutta", step=2e-14)
        o.set_integrator(0, 2e-14, 0)
        for _ in range(30):
            o.advance_step()
        mmx.run_step(g, n=30)
        err = np.max(np.abs(g.spin - o.spin))
        print(f"max |dm| after 30 steps: {err:.2e}")
        assert err < 1e-12
