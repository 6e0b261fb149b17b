"""Hexagonal anisotropy (add_hex_anis, kernels.jl:142-167) and twin monoclinic
anisotropy (add_twin_mono_anis, kernels.jl:169-225) on the HIP path, through
the C ABI, against the oracle: fields, energies and dm/dt with and without
demag, an RK4 trajectory, a batch member against the same member advanced
alone, and the adjoint's refusal.

FP64 gate: fields within 1e-12 (max-norm relative), trajectories within 1e-12
absolute; FP32 fields within 1e-5.
"""
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


def build(oracle, precision=64, demag=True, pbc="xy", seed=13):
    nx, ny, nz = 11, 12, 3
    d = (2e-9, 2e-9, 2.5e-9)
    N = nx * ny * nz
    rng = np.random.default_rng(seed)
    Ms = np.where(rng.uniform(size=N) < 0.1, 0.0, 6e5)
    m0 = rng.standard_normal(3 * N)
    mesh = mmx.FDMesh(dx=d[0], dy=d[1], dz=d[2], nx=nx, ny=ny, nz=nz, pbc=pbc)
    g = mmx.Sim(mesh, precision=precision)
    o = oracle.OracleSim(nx, ny, nz, *d, pbc=pbc)
    for s in (g,):
        mmx.set_Ms(s, Ms)
        mmx.init_m0(s, m0)
    o.set_Ms(Ms)
    o.init_m0(m0)
    g.driver.alpha, g.driver.gamma = 0.1, 2.21e5
    o.set_llg(0.1, 2.21e5, True)
    mmx.add_exch(g, 1.2e-11)
    o.add_exch(1.2e-11)
    mmx.add_hex_anis(g, K1=4e5, K2=-1.2e5, K3=3e4)
    o.add_hex_anis(4e5, -1.2e5, 3e4)
    axes = [rng.standard_normal(3 * N) for _ in range(3)]
    tw = dict(Ka=2e4, Kb=-1e4, Kaa=5e3, Kbb=-3e3, Kab=7e3, Ku=1.5e4)
    mmx.add_twin_mono_anis(g, **tw, axis_a=axes[0].reshape(N, 3), axis_b=axes[1].reshape(N, 3),
                           axis_u111=axes[2].reshape(N, 3))
    o.add_twin_mono_anis(**tw, axis_a=axes[0], axis_b=axes[1], axis_u111=axes[2])
    if demag:
        mmx.add_demag(g)
        o.add_demag()
    return g, o


@pytest.mark.parametrize("demag", [False, True])
@pytest.mark.parametrize("precision", [64, 32])
def test_hex_fields_match_oracle(oracle, precision, demag):
    note(f"hex anisotropy fields FP{precision} demag={demag}")
    g, o = build(oracle, precision=precision, demag=demag)
    tol = 1e-12 if precision == 64 else 1e-5
    for q in range(len(g.interactions)):
        fg, eg = g._term_field(q)
        fo, eo = o.term_field(q)
        assert relerr(fg, fo) <= tol, (q, relerr(fg, fo))
        assert relerr(eg, eo) <= tol, (q, relerr(eg, eo))
    assert relerr(mmx.effective_field(g), o.effective_field()) <= tol
    assert relerr(mmx.llg_rhs(g), o.llg_rhs()) <= tol


def test_hex_rk4_trajectory(oracle):
    note("hex anisotropy RK4 trajectory")
    g, o = build(oracle)
    mmx.set_driver(g, integrator="RungeKutta", step=2e-14)
    o.set_integrator(0, 2e-14, 0)
    for _ in range(30):
        o.advance_step()
    mmx.run_step(g, n=30)
    err = np.max(np.abs(g.spin - o.spin))
    print(f"max |dm| after 30 steps: {err:.2e}")
    assert err < 1e-12


def test_hex_batch_member_and_adjoint_refusal():
    note("hex anisotropy batch member, adjoint refusal")

    def member():
        mesh = mmx.FDMesh(dx=2e-9, dy=2e-9, dz=2e-9, nx=40, ny=24, nz=2, pbc="xy")
        s = mmx.Sim(mesh, integrator="RungeKutta")
        mmx.set_Ms(s, 6e5)
        mmx.init_m0(s, lambda i, j, k, dx, dy, dz: (np.cos(0.3 * i), np.sin(0.2 * j), 0.3))
        mmx.add_exch(s, 1e-11)
        mmx.add_hex_anis(s, K1=-2e5, K3=5e4)
        mmx.set_driver(s, integrator="RungeKutta", step=1e-14)
        return s

    a, b = member(), member()
    bt = mmx.Batch([a])
    bt.advance(7)
    mmx.run_step(b, n=7)
    assert np.array_equal(a.spin, b.spin)
    del bt
    with pytest.raises(mmx.MmxError):
        mmx.rk4_trajectory_rrule(b, 4, every=2)


def test_update_anis_equals_fresh_term():
    """update_anis(sim, Ku; name) (add_field.jl:514-525): the updated term's field is
    the field of a term built with the new Ku."""
    note("update_anis")
    mesh = mmx.FDMesh(dx=2e-9, dy=2e-9, dz=2e-9, nx=12, ny=6, nz=2)
    rng = np.random.default_rng(2)
    m0 = rng.standard_normal(3 * mesh.n_total)
    Ku2 = rng.uniform(1e5, 5e5, mesh.n_total)
    a = mmx.Sim(mesh)
    b = mmx.Sim(mesh)
    for s in (a, b):
        mmx.set_Ms(s, 6e5)
        mmx.init_m0(s, m0)
    ta = mmx.add_anis(a, 3e5, axis=(0, 1, 1))
    mmx.update_anis(a, Ku2)
    tb = mmx.add_anis(b, Ku2, axis=(0, 1, 1))
    assert np.array_equal(ta.field, tb.field) and np.array_equal(ta.energy, tb.energy)
