"""CPU checks of the oracle's steepest-descent restatement (oracle/oracle_sd.py):
the update of src/sd.jl:39-55 keeps |m| = 1 exactly in exact arithmetic, the
first step uses min_tau, and a relaxation lowers the energy."""
import numpy as np

from oracle import OracleSim
from oracle_sd import OracleSD

MU0 = 4e-7 * np.pi


def _sys():
    nx, ny = 8, 6
    o = OracleSim(nx, ny, 1, 2e-9, 2e-9, 2e-9)
    o.set_Ms(np.full(nx * ny, 8e5))
    rng = np.random.default_rng(4)
    o.init_m0(rng.standard_normal(3 * nx * ny))
    o.add_exch(1.3e-11)
    o.add_anis(np.full(nx * ny, 5e4), axis=(0, 0, 1))
    o.add_zeeman((0, 0, 0.05 / MU0))
    return o


def test_sd_keeps_unit_length_and_lowers_energy():
    o = _sys()
    sd = OracleSD(o)
    e0 = o.term_energies(sd.spin.reshape(-1)).sum()
    sd.run_step()
    assert sd.tau == 1e-10
    for _ in range(8):  # fewer than 10 steps: no renormalisation yet
        sd.run_step()
    assert np.abs(np.linalg.norm(sd.spin, axis=1) - 1).max() < 1e-12
    for _ in range(200):
        sd.run_step()
    e1 = o.term_energies(sd.spin.reshape(-1)).sum()
    assert e1 < e0
