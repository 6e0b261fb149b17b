"""Steepest-descent driver (src/sd.jl) on the GPU against the oracle's
restatement (oracle/oracle_sd.py)."""
import numpy as np
import pytest

import mmx
from oracle import OracleSim
from oracle_sd import OracleSD

pytestmark = pytest.mark.gpu

MU0 = 4e-7 * np.pi


def _pair(demag):
    nx, ny, nz, d = 12, 10, 2, 2.5e-9
    N = nx * ny * nz
    rng = np.random.default_rng(21)
    m0 = np.array([1.0, 0.3, 0.2]) + 0.4 * rng.standard_normal((N, 3))
    pins = np.zeros(N, dtype=bool)
    pins[[0, 7]] = True
    o = OracleSim(nx, ny, nz, d, d, d, pbc="x")
    o.set_Ms(8e5)
    o.set_pins(pins)
    o.init_m0(m0.reshape(-1))
    o.add_exch(1.3e-11)
    o.add_dmi(np.full(N, 1e-3), type="interfacial")
    o.add_anis(np.full(N, 5e4), axis=(0, 0, 1))
    o.add_zeeman((1e4, 0, 0))
    mesh = mmx.FDMesh(dx=d, dy=d, dz=d, nx=nx, ny=ny, nz=nz, pbc="x")
    sim = mmx.Sim(mesh, driver="SD")
    mmx.set_Ms(sim, 8e5)
    mmx.set_pinning(sim, pins)
    mmx.init_m0(sim, m0.reshape(-1))
    mmx.add_exch(sim, 1.3e-11)
    mmx.add_dmi(sim, 1e-3, type="interfacial")
    mmx.add_anis(sim, 5e4, axis=(0, 0, 1))
    mmx.add_zeeman(sim, (1e4, 0, 0))
    if demag:
        o.add_demag()
        mmx.add_demag(sim)
    return OracleSD(o, pins=pins), sim


@pytest.mark.parametrize("demag", [False, True])
def test_sd_steps_match_oracle(demag):
    osd, sim = _pair(demag)
    for i in range(25):
        osd.run_step()
        mmx.run_step(sim)
        tau = sim._integrator.step
        assert abs(tau - osd.tau) <= 1e-9 * osd.tau, (i, tau, osd.tau)
    assert np.abs(sim.spin - osd.spin.reshape(-1)).max() < 1e-9


def test_sd_relax_matches_oracle():
    osd, sim = _pair(False)
    i1, md1 = osd.relax(max_steps=60, stopping_dmdt=1e-30)
    i2, md2 = mmx.relax(sim, max_steps=60, stopping_dmdt=1e-30)
    assert i1 == i2 == 60
    assert abs(md1 - md2) <= 1e-6 * md1
    assert np.abs(sim.spin - osd.spin.reshape(-1)).max() < 1e-8


@pytest.mark.parametrize("shape", [(30, 8, 3), (9, 30, 3), (3, 8, 30)])
def test_rm_demag_charges_relax(shape):
    """test/test_rm_charges.jl:8-44: with the charges on the long axis's faces
    removed, the SD-relaxed corners of a bar lie along that axis to 5e-3."""
    nx, ny, nz = shape
    axis = int(np.argmax(shape))
    print(f"[case] rm_demag_charges relax {shape}", flush=True)
    mesh = mmx.FDMesh(nx=nx, ny=ny, nz=nz, dx=2e-9, dy=2e-9, dz=2e-9)
    sim = mmx.Sim(mesh, driver="SD")
    mmx.set_Ms(sim, 8e5)
    mmx.add_exch(sim, 1.3e-12)
    mmx.add_demag(sim)
    kw = {0: dict(x=(-1, 1)), 1: dict(y=(-1, 1)), 2: dict(z=(-1, 1))}[axis]
    mmx.rm_demag_charges(sim, 8e5, **kw)
    mmx.init_m0(sim, (1, 1, 1))
    mmx.relax(sim, max_steps=20000, stopping_dmdt=0.01)
    m = sim.spin.reshape(nz, ny, nx, 3)
    corners = [m[0, 0, 0], m[0, -1, 0], m[0, 0, -1], m[0, -1, -1]]
    mp = max(np.sqrt(1 - c[axis] ** 2) for c in corners)
    print(f"max transverse part at the corners: {mp:.2e}")
    assert mp < 5e-3
