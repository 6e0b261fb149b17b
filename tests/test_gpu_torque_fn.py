"""add_torque (user function of m and t, add_field.jl:998-1006) on the HIP path,
through the C ABI: fields at several times against the oracle (FP64 1e-12,
FP32 1e-5), an RK4 trajectory with a time-dependent torque (the function sees
each stage's state and time), the precession closed form, and the refusals of
the batch and the adjoint.
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


def torque(y, m, t):
    y[0::3] = m[1::3] * 3e10 * math.sin(t * 2e12)
    y[1::3] = 1e10 * math.cos(t * 1e12)
    y[2::3] = m[0::3] * m[2::3] * 2e10


def build(oracle, precision=64):
    nx, ny, nz = 10, 6, 2
    N = nx * ny * nz
    rng = np.random.default_rng(51)
    Ms = np.where(rng.uniform(size=N) < 0.1, 0.0, 7e5)
    m0 = rng.standard_normal(3 * N)
    mesh = mmx.FDMesh(dx=2e-9, dy=2e-9, dz=2e-9, nx=nx, ny=ny, nz=nz, pbc="x")
    g = mmx.Sim(mesh, precision=precision)
    o = oracle.OracleSim(nx, ny, nz, 2e-9, 2e-9, 2e-9, pbc="x")
    mmx.set_Ms(g, Ms)
    o.set_Ms(Ms)
    mmx.init_m0(g, m0)
    o.init_m0(m0)
    g.driver.alpha, g.driver.gamma = 0.1, 2.21e5
    o.set_llg(0.1, 2.21e5, True)
    mmx.add_exch(g, 1e-11)
    o.add_exch(1e-11)
    mmx.add_torque(g, torque)
    o.add_torque(torque)
    mmx.add_demag(g)
    o.add_demag()
    return g, o


@pytest.mark.parametrize("precision", [64, 32])
def test_torque_fn_fields_match_oracle(oracle, precision):
    note(f"add_torque fields FP{precision}")
    g, o = build(oracle, precision)
    tol = 1e-12 if precision == 64 else 1e-5
    for t in (0.0, 1.3e-12):
        for q in range(len(g.interactions)):
            assert relerr(g._term_field(q, t)[0], o.term_field(q, t=t)[0]) <= tol, (q, t)
        assert relerr(mmx.llg_rhs(g, t=t), o.llg_rhs(t=t)) <= tol


def test_torque_fn_rk4_trajectory(oracle):
    note("add_torque RK4 trajectory")
    g, o = build(oracle)
    mmx.set_driver(g, integrator="RungeKutta", step=1e-14)
    o.set_integrator(0, 1e-14, 0)
    for _ in range(20):
        o.advance_step()
    mmx.run_step(g, n=20)
    err = np.max(np.abs(g.spin - o.spin))
    print(f"max |dm| after 20 steps: {err:.2e}")
    assert err < 1e-12


def test_torque_fn_precession_and_refusals():
    note("add_torque precession closed form, refusals")
    w = 2e10
    mesh = mmx.FDMesh(nx=1, ny=1, nz=1, dx=1e-9, dy=1e-9, dz=1e-9)
    sim = mmx.Sim(mesh, integrator="RungeKutta")
    mmx.set_Ms(sim, 8e5)
    sim.driver.alpha, sim.driver.gamma = 0.0, 2.21e5
    mmx.init_m0(sim, (1, 0, 0))

    def rot(y, m, t):
        y[0::3] = -w * m[1::3]
        y[1::3] = w * m[0::3]
        y[2::3] = 0

    mmx.add_torque(sim, rot)
    mmx.set_driver(sim, integrator="RungeKutta", step=1e-13)
    mmx.run_step(sim, n=100)
    m = sim.spin
    phi = w * 100 * 1e-13
    assert abs(m[0] - math.cos(phi)) < 1e-9 and abs(m[1] - math.sin(phi)) < 1e-9
    with pytest.raises(mmx.MmxError):
        mmx.Batch([sim])
    with pytest.raises(mmx.MmxError):
        mmx.rk4_trajectory_rrule(sim, 4, every=2)
    with pytest.raises(mmx.MmxError):  # the band's fields are not evaluated per stage state
        mmx.NEB(sim, [(1, 0, 0), (0, 0, 1)], [3])


def test_torque_fn_buffer_and_errors(oracle):
    """The host buffer handed to the function holds the previous evaluation's
    field (as the reference's in-place torque_kernel! leaves it), so a function
    that builds on it gives the oracle's fields call after call; an exception in
    the function surfaces from the library call that ran it."""
    note("add_torque persistent buffer and exceptions")

    def acc(y, m, t):
        y[:] = 0.5 * y + 1e9 * m

    nx, ny, nz = 6, 5, 2
    N = nx * ny * nz
    m0 = np.random.default_rng(9).standard_normal(3 * N)
    mesh = mmx.FDMesh(dx=2e-9, dy=2e-9, dz=2e-9, nx=nx, ny=ny, nz=nz)
    g = mmx.Sim(mesh)
    o = oracle.OracleSim(nx, ny, nz, 2e-9, 2e-9, 2e-9)
    mmx.set_Ms(g, 7e5)
    o.set_Ms(np.full(N, 7e5))
    mmx.init_m0(g, m0)
    o.init_m0(m0)
    tg = mmx.add_torque(g, acc)
    to = o.add_torque(acc)
    for _ in range(3):
        fg, _ = g._term_field(tg.index)
        fo, _ = o.term_field(to)
        assert relerr(fg, fo) <= 1e-12
    boom = []

    def bad(y, m, t):
        boom.append(t)
        raise ValueError("torque function failed")

    mesh2 = mmx.FDMesh(dx=2e-9, dy=2e-9, dz=2e-9, nx=4, ny=3, nz=1)
    g2 = mmx.Sim(mesh2)
    mmx.set_Ms(g2, 7e5)
    mmx.init_m0(g2, (1, 0, 0))
    mmx.add_torque(g2, bad)
    with pytest.raises(ValueError, match="torque function failed"):
        mmx.effective_field(g2)
    assert boom
