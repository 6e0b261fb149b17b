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
