"""Spin-torque terms (add_sot, add_stt) on the HIP path, through the C ABI,
against the oracle: per-term fields, the aggregate field, trajectories with a
time-dependent current (stage times), the reference's one-cell checks, and the
Zhang-Li domain-wall velocity of test/test_stt.jl.

FP64 gate: fields and fixed-step trajectories within 1e-12 (max-norm relative
for fields, absolute for unit vectors); FP32 fields within 1e-5.
"""
import math
import sys

import numpy as np
import pytest

import mmx

pytestmark = pytest.mark.gpu

MU0 = 4e-7 * np.pi
C_E = 1.602176565e-19
H_BAR = 1.05457172647e-34


def note(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def relerr(a, b):
    den = np.max(np.abs(b))
    return np.max(np.abs(a - b)) / (den if den > 0 else 1.0)


def ft(t):
    return 1.0 + 0.5 * math.sin(2 * math.pi * t / 3e-13)


def build(oracle, nx, ny, nz, pbc, precision=64, seed=3, demag=True, time_fn=True):
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
    aj = rng.uniform(0.5e9, 2e9, N)
    J = rng.uniform(0.5e12, 1.5e12, N)
    Jv = rng.uniform(-1, 1, (N, 3)) * 1e12
    xi = rng.uniform(0, 0.3, N)
    mmx.add_exch(g, 1.3e-11)
    o.add_exch(1.3e-11)
    mmx.add_sot(g, aj, 4e8, (0.2, -0.3, 0.9))
    o.add_sot(aj, 4e8, (0.2, -0.3, 0.9))
    beta = mmx.h_bar / (mmx.mu_0 * 6e5 * mmx.c_e * 2e-9)
    tsl = mmx.add_stt(g, model="slonczewski", J=J, tf=2e-9, Ms=6e5, p=(0, 0, 1), P=0.6, xi=0.2, Lambda=1.7)
    isl = o.add_stt_slonczewski(J, beta, 1.7, 0.2, 0.6, (0, 0, 1))
    tzl = mmx.add_stt(g, model="zhang_li", P=0.5, Ms=6e5, J=Jv.reshape(-1), xi=xi)
    izl = o.add_stt_zhangli(Jv.reshape(-1), 0.5 * 9.27400949e-24 / (C_E * 6e5), xi, xi_factor=True)
    if time_fn:
        for tg, io in ((tsl, isl), (tzl, izl)):
            mmx._set_time_fn(g, tg, ft)
            o.set_term_time_fn(io, ft)
    if demag:
        mmx.add_demag(g)
        o.add_demag()
    return g, o


@pytest.mark.parametrize("pbc", ["", "xy"])
@pytest.mark.parametrize("precision", [64, 32])
def test_torque_fields_match_oracle(oracle, pbc, precision):
    note(f"torque fields pbc={pbc!r} FP{precision}")
    g, o = build(oracle, 14, 10, 3, pbc, precision=precision)
    tol = 1e-12 if precision == 64 else 1e-5
    for t in (0.0, 1.1e-13):
        for q in range(len(g.interactions)):
            fg, eg = g._term_field(q, t)
            fo, eo = o.term_field(q, t=t)
            assert relerr(fg, fo) <= tol, (q, t, relerr(fg, fo))
        assert relerr(mmx.effective_field(g, t=t), o.effective_field(t=t)) <= tol
        assert relerr(mmx.llg_rhs(g, t=t), o.llg_rhs(t=t)) <= tol


@pytest.mark.parametrize("integrator", ["RungeKutta", "Heun"])
def test_torque_fixed_step_trajectory(oracle, integrator):
    note(f"torque trajectory {integrator}")
    g, o = build(oracle, 12, 10, 2, "x")
    kind = {"RungeKutta": 0, "Heun": 1}[integrator]
    mmx.set_driver(g, integrator=integrator, step=2e-14)
    o.set_integrator(kind, 2e-14, 0)
    for _ in range(25):
        o.advance_step()
    mmx.run_step(g, n=25)
    err = np.max(np.abs(g.spin - o.spin))
    print(f"max |dm| after 25 steps: {err:.2e}")
    assert err < 1e-12


def test_torque_adaptive_trajectory(oracle):
    note("torque trajectory DormandPrince")
    g, o = build(oracle, 12, 10, 2, "", demag=False)
    mmx.set_driver(g, integrator="DormandPrince", tol=1e-6)
    o.set_integrator(2, 0, 1e-6)
    for i in range(1, 4):
        mmx.run_until(g, i * 2e-13)
        o.run_until(i * 2e-13)
    so = o.state()
    assert g._integrator.nsteps == so["nsteps"]
    assert np.max(np.abs(g.spin - o.spin)) < 1e-10


def test_sot_and_slonczewski_one_cell():
    """test/test_sot.jl:9-48 and test/test_slonczewski.jl:18-41 on the HIP path."""
    note("one-cell SOT / Slonczewski")
    mesh = mmx.FDMesh(nx=1, ny=1, nz=1, dx=1e-9, dy=1e-9, dz=1e-9)
    sim = mmx.Sim(mesh)
    mmx.set_Ms(sim, 8e5)
    sim.driver.alpha, sim.driver.gamma = 0.05, 2.21e5
    mmx.add_zeeman(sim, (0, 0, 1e5))
    t = mmx.add_sot(sim, 1e9, 5e8, (0.1, 0.2, 0.3), name="dft_torque")
    mmx.init_m0(sim, (-0.3, 0.4, 0.8), norm=False)
    mmx.effective_field(sim)
    expected = (1e9 * np.cross([-0.3, 0.4, 0.8], [0.1, 0.2, 0.3]) + 5e8 * np.array([0.1, 0.2, 0.3])) / 2.21e5
    assert np.allclose(t.field, expected, rtol=1e-12, atol=0)

    sim = mmx.Sim(mesh)
    mmx.set_Ms(sim, 8e5)
    sim.driver.alpha, sim.driver.gamma = 0.05, 2.21e5
    mmx.add_zeeman(sim, (0, 0, 1e5))
    t = mmx.add_stt(sim, model="slonczewski", J=1e12, tf=2e-9, Ms=8e5, p=(0, 0, 1), P=0.7, xi=0.5, Lambda=3.7)
    mmx.init_m0(sim, (0.6, 0.8, 0.0), norm=True)
    m = sim.spin
    beta = H_BAR * 1e12 / (MU0 * C_E * 2e-9 * 8e5)
    eps = 0.7 * 3.7 ** 2 / (3.7 ** 2 + 1 + (3.7 ** 2 - 1) * m[2])
    expected = beta * (eps * np.cross(m, [0, 0, 1.0]) + 0.5 * np.array([0, 0, 1.0]))
    assert np.allclose(t.field, expected, rtol=1e-12, atol=0)


def _init_dw(i, j, k, dx, dy, dz):  # test/test_stt.jl:9-17
    if i < 150:
        return (1, 0.1, 0)
    elif i < 160:
        return (0, 1, 0.1)
    return (-1, 0.1, 0)


def test_zhang_li_domain_wall_velocity():
    """test/test_stt.jl:19-69: a relaxed transverse wall in a 1 um strip moves
    at v = -(1 + alpha beta) / (1 + alpha^2) u under the Zhang-Li torque."""
    note("Zhang-Li domain-wall velocity")
    mesh = mmx.FDMesh(nx=500, ny=1, nz=11, dx=2e-9, dy=2e-9, dz=1e-9)
    sim = mmx.Sim(mesh, driver="SD")
    mmx.set_Ms(sim, 8.6e5)
    mmx.add_exch(sim, 1.3e-11)
    mmx.add_anis(sim, 1e5, axis=(1, 0, 0))
    mmx.init_m0(sim, _init_dw)
    mmx.relax(sim, max_steps=2000, stopping_dmdt=0.1)
    m_relaxed = sim.spin
    alpha = 0.1
    for beta, u in [(0, 10), (0.1, 3.2), (0.2, 4.7)]:
        sim = mmx.Sim(mesh, integrator="DormandPrince")
        mmx.set_Ms(sim, 8.6e5)
        sim.driver.alpha, sim.driver.gamma = alpha, 2.21e5
        mmx.add_stt(sim, model="zhang_li", b=u, J=(1.0, 0, 0), xi=beta)
        mmx.add_exch(sim, 1.3e-11)
        mmx.add_anis(sim, 1e5, axis=(1, 0, 0))
        mmx.init_m0(sim, m_relaxed)
        m_0 = sim.spin.reshape(-1, 3)[:, 0].mean()
        for i in range(1, 11):
            mmx.run_until(sim, 1e-11 * i)
        m_1 = sim.spin.reshape(-1, 3)[:, 0].mean()
        v = 0.5 * (m_1 - m_0) * 500 * 2e-9 / 1e-10
        v_expect = -(1 + alpha * beta) / (1 + alpha ** 2) * u
        print(f"beta={beta} u={u}: v={v:.4f} expected {v_expect:.4f}")
        assert abs((v - v_expect) / v_expect) < 0.006


def test_sahe_torque_reference_and_oracle(oracle):
    """test/test_sahe_torque.jl:8-30 on the HIP path, then per-cell vector fields
    against the oracle (FP64 1e-12, FP32 1e-5) and an RK4 trajectory."""
    note("spin anomalous Hall torque")
    mesh = mmx.FDMesh(nx=1, ny=1, nz=1, dx=1e-9, dy=1e-9, dz=1e-9)
    sim = mmx.Sim(mesh)
    mmx.set_Ms(sim, 8e5)
    sim.driver.alpha, sim.driver.gamma = 0.05, 2.21e5
    mmx.add_zeeman(sim, (0, 0, 1e5))
    t = mmx.add_sahe_torque(sim, (0.1, 0.2, 0.3), sigma_sa1=(0.3, 0.4, 0.5), sigma_sa2=(0.6, 0.7, 0.8), beta=0.135)
    mmx.init_m0(sim, (-0.3, 0.4, 0.8), norm=False)
    mmx.effective_field(sim)
    np.testing.assert_allclose(t.field * 1e7, [2.01628959, 3.74389140, 3.30678733], rtol=1e-6)

    nx, ny, nz = 9, 8, 2
    N = nx * ny * nz
    rng = np.random.default_rng(31)
    vs = [rng.standard_normal(3 * N) for _ in range(6)]
    Ms = np.where(rng.uniform(size=N) < 0.1, 0.0, 7e5)
    m0 = rng.standard_normal(3 * N)
    for precision in (64, 32):
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
        mmx.add_sahe_torque(g, *(v.reshape(N, 3) for v in vs[:1]), sigma_sa1=vs[1].reshape(N, 3),
                            sigma_sa2=vs[2].reshape(N, 3), a1=vs[3].reshape(N, 3), a2=vs[4].reshape(N, 3),
                            a3=vs[5].reshape(N, 3), beta=0.2)
        o.add_sahe_torque(vs[0], sigma_sa1=vs[1], sigma_sa2=vs[2], a1=vs[3], a2=vs[4], a3=vs[5], beta=0.2)
        tol = 1e-12 if precision == 64 else 1e-5
        for q in range(2):
            assert relerr(g._term_field(q)[0], o.term_field(q)[0]) <= tol, (precision, q)
        assert relerr(mmx.llg_rhs(g), o.llg_rhs()) <= tol
        if precision == 64:
            mmx.set_driver(g, integrator="RungeKutta", step=1e-14)
            o.set_integrator(0, 1e-14, 0)
            for _ in range(20):
                o.advance_step()
            mmx.run_step(g, n=20)
            assert np.max(np.abs(g.spin - o.spin)) < 1e-12
