"""Jacobian products of one LLG right-hand side (mmx_llg_rhs_jvp / _vjp through
the C ABI) and the dynamic matrix built from them.

* JVP against central differences of the oracle's llg_rhs (the derivative the
  mathematics fixes), on a mesh with every covered term, pins and an empty cell.
* VJP against the JVP through <w, J v> = <J^T w, v>.
* test/eigen/test_demag.jl:21-62: the FMR frequency of a 5x1x1 nm macrospin
  from dynamic_matrix (ext/EnzymeExt.jl:32-59) equals the Kittel formula with
  Nx = 0.08831574004542228 to 100 eps, as the reference requires of Enzyme.
"""
import sys

import numpy as np
import pytest

import mmx
from oracle import OracleSim

pytestmark = pytest.mark.gpu

MU0 = 4e-7 * np.pi


def note(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def _system(precession=True):
    nx, ny, nz, d = 10, 8, 3, 2e-9
    N = nx * ny * nz
    rng = np.random.default_rng(21)
    Ms = rng.uniform(5e5, 8e5, N)
    Ms[13] = 0.0
    pins = np.zeros(N, dtype=bool)
    pins[[5, 70]] = True
    m0 = rng.standard_normal((N, 3)) + np.array([0.2, 0.1, 1.0])
    m0 = (m0 / np.linalg.norm(m0, axis=1, keepdims=True)).reshape(-1)
    Ku = rng.uniform(2e5, 6e5, N)
    mesh = mmx.FDMesh(dx=d, dy=d, dz=d, nx=nx, ny=ny, nz=nz, pbc="x")
    g = mmx.Sim(mesh)
    o = OracleSim(nx, ny, nz, d, d, d, pbc="x")
    mmx.set_Ms(g, Ms)
    o.set_Ms(Ms)
    mmx.set_pinning(g, pins)
    o.set_pins(pins.astype(np.uint8))
    mmx.init_m0(g, m0)
    o.init_m0(m0)
    g.driver.alpha, g.driver.gamma, g.driver.precession = 0.15, 2.21e5, precession
    o.set_llg(0.15, 2.21e5, precession)
    mmx.add_exch(g, 1.3e-11)
    o.add_exch(1.3e-11)
    mmx.add_dmi(g, 2e-3, type="interfacial")
    o.add_dmi(np.full(N, 2e-3), type="interfacial")
    mmx.add_dmi(g, 1e-3, type="bulk")
    o.add_dmi(1e-3, type="bulk")
    mmx.add_anis(g, Ku, axis=(0.3, 0.0, 1.0))
    o.add_anis(Ku, axis=(0.3, 0.0, 1.0))
    mmx.add_cubic_anis(g, -3e4)
    o.add_cubic_anis(np.full(N, -3e4))
    mmx.add_zeeman(g, (1e4, 0.0, 4e4))
    o.add_zeeman(np.tile([1e4, 0.0, 4e4], N).astype(float))
    mmx.add_demag(g)
    o.add_demag()
    return g, o, m0, N


@pytest.mark.parametrize("precession", [True, False])
def test_jvp_matches_oracle_differences(precession):
    note(f"llg_rhs JVP 10x8x3, precession {precession}")
    g, o, _, N = _system(precession)
    m0 = g.spin  # the library's state (init_m0 leaves empty cells at zero)
    rng = np.random.default_rng(5)
    for _ in range(3):
        v = rng.standard_normal(3 * N)
        jv = mmx.llg_rhs_jvp(g, v)
        eps = 1e-6
        fd = (o.llg_rhs(m0 + eps * v) - o.llg_rhs(m0 - eps * v)) / (2 * eps)
        err = np.abs(jv - fd).max() / np.abs(fd).max()
        print(f"JVP vs central difference of the oracle: {err:.2e}", flush=True)
        assert err < 1e-8


def test_vjp_is_the_transpose():
    note("llg_rhs VJP against JVP")
    g, o, m0, N = _system()
    rng = np.random.default_rng(6)
    for _ in range(3):
        v = rng.standard_normal(3 * N)
        w = rng.standard_normal(3 * N)
        a = w @ mmx.llg_rhs_jvp(g, v)
        b = mmx.llg_rhs_vjp(g, w) @ v
        print(f"<w, J v> {a:+.15e}  <J^T w, v> {b:+.15e}", flush=True)
        assert abs(a - b) <= 1e-12 * abs(a)
    # at an explicit state, too (zero at the empty cell, as init_m0 leaves it:
    # the pull-back carries no field cotangent into cells without Ms)
    m1 = np.roll(m0, 3).reshape(-1, 3)
    m1[g.mu0_Ms == 0] = 0.0
    m1 = m1.reshape(-1)
    v = rng.standard_normal(3 * N)
    w = rng.standard_normal(3 * N)
    a = w @ mmx.llg_rhs_jvp(g, v, spin=m1)
    b = mmx.llg_rhs_vjp(g, w, spin=m1) @ v
    assert abs(a - b) <= 1e-12 * abs(a)


def test_fmr_frequency_enzyme_reference():
    """test/eigen/test_demag.jl:21-62: f from dynamic_matrix vs the Kittel value."""
    note("FMR 5x1x1 nm macrospin (dynamic_matrix)")
    mesh = mmx.FDMesh(nx=1, ny=1, nz=1, dx=5e-9, dy=1e-9, dz=1e-9)
    sim = mmx.Sim(mesh)
    mmx.set_Ms(sim, 8e5)
    mmx.init_m0(sim, (1, 0, 0))
    H0 = 1.23e4
    mmx.add_zeeman(sim, (H0, 0, 0))
    mmx.add_demag(sim)
    B = mmx.dynamic_matrix(sim, gamma=2.21e5)
    f = np.max(np.abs(np.imag(np.linalg.eigvals(B)))) / 1e9 / (2 * np.pi)
    Nx = 0.08831574004542228
    Ny = (1 - Nx) / 2
    K = 2 * (0.5 * (Ny - Nx) * MU0 * 8e5 ** 2) / (MU0 * 8e5)
    fan = 2.21e5 * (H0 + K) / 1e9 / (2 * np.pi)
    print(f"f {f!r}  analytical {fan!r}  relative {abs(f - fan) / f:.2e}", flush=True)
    assert abs(f - fan) / f < 100 * np.finfo(float).eps
