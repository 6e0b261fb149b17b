"""Nudged elastic band on the GPU against the oracle's restatement of
src/neb/neb.jl (oracle/neb.py) on the same images."""
import numpy as np
import pytest

import mmx
from oracle import OracleSim
from oracle_neb import OracleNEB

pytestmark = pytest.mark.gpu

MU0 = 4e-7 * np.pi


def _skyrmion(nx, ny, nz, d, R=5e-9):
    x = (np.arange(nx) - (nx - 1) / 2) * d
    y = (np.arange(ny) - (ny - 1) / 2) * d
    X, Y = np.meshgrid(x, y)
    r = np.sqrt(X * X + Y * Y)
    th = np.pi * np.exp(-(r / R) ** 2)
    ph = np.arctan2(Y, X) + np.pi / 2
    m = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], axis=-1)
    return np.tile(m.reshape(-1, 3), (nz, 1)).reshape(-1)


def _setup(case, precision=64):
    if case == "film":
        nx, ny, nz, d, pbc = 12, 12, 1, 2e-9, "xy"
    elif case == "demag-ragged":  # rows and planes that fill no tile exactly
        nx, ny, nz, d, pbc = 12, 10, 3, 2e-9, "x"
    else:
        nx, ny, nz, d, pbc = 8, 8, 2, 2e-9, ""
    N = nx * ny * nz
    Ms = np.full(N, 3.84e5)
    pins = np.zeros(N, dtype=bool)
    if case == "demag":
        pins[0] = True
    o = OracleSim(nx, ny, nz, d, d, d, pbc=pbc)
    o.set_Ms(Ms)
    o.add_exch(3.25e-12)
    o.add_dmi(np.full(N, 5.83e-4), type="interfacial")
    o.add_zeeman((0, 0, 0.12 / MU0))
    mesh = mmx.FDMesh(dx=d, dy=d, dz=d, nx=nx, ny=ny, nz=nz, pbc=pbc)
    sim = mmx.Sim(mesh, precision=precision)
    mmx.set_Ms(sim, Ms)
    mmx.add_exch(sim, 3.25e-12)
    mmx.add_dmi(sim, 5.83e-4, type="interfacial")
    mmx.add_zeeman(sim, (0, 0, 0.12 / MU0))
    if case.startswith("demag"):
        o.add_demag()
        mmx.add_demag(sim)
        o.set_pins(pins)
        mmx.set_pinning(sim, pins)
    left = _skyrmion(nx, ny, nz, d)
    return o, sim, [left, (0.0, 0.0, 1.0)], pins


@pytest.mark.parametrize("case,clib", [("film", -1), ("demag", -1), ("film", 3), ("demag-ragged", 2)])
def test_neb_matches_oracle(case, clib):
    o, sim, given, pins = _setup(case)
    right = np.tile([0.0, 0.0, 1.0], o.N)
    on = OracleNEB(o, [given[0], right], [5], spring_constant=1e5, climbing_image=clib, pins=pins)
    gn = mmx.NEB(sim, given, [5], spring_constant=1e5, clib_image=clib)
    assert gn.n_images == on.n == 5
    np.testing.assert_allclose(gn.spin, on.spin, rtol=0, atol=1e-15)
    np.testing.assert_allclose(gn.energy, on.energy, rtol=1e-12, atol=0)
    for step in range(6):
        on.advance_step()
        gn.advance(1)
        st = gn.state
        print(f"[neb] {case} step {step}: step {st['step']:.6e} vs {on.step:.6e}", flush=True)
        assert st["nsteps"] == on.nsteps
        assert abs(st["step"] - on.step) <= 1e-9 * on.step
        assert np.abs(gn.spin - on.spin).max() < 1e-10
    on._field_energy()
    on._distance()
    np.testing.assert_allclose(gn.energy, on.energy, rtol=1e-10, atol=0)
    np.testing.assert_allclose(gn.distance, on.distance, rtol=1e-10, atol=0)


def test_neb_relax_matches_oracle():
    o, sim, given, pins = _setup("film")
    right = np.tile([0.0, 0.0, 1.0], o.N)
    on = OracleNEB(o, [given[0], right], [4])
    gn = mmx.NEB(sim, given, [4])
    i1, md1 = on.relax(max_steps=8, stopping_dmdt=1e-30)
    i2, md2 = gn.relax(max_steps=8, stopping_dmdt=1e-30)
    assert i1 == i2 == 8
    assert abs(md1 - md2) <= 1e-8 * md1
    assert np.abs(gn.spin - on.spin).max() < 1e-9


@pytest.mark.parametrize("case", ["film", "demag-ragged"])
def test_neb_fp32_against_fp64_oracle(case):
    """FP32 band (fields of all images in one launch per pass) against the FP64
    oracle: image energies within 1e-5 relative at the start, images within
    1e-5 after three steps (step control in FP64 on FP32 stage vectors)."""
    o, sim, given, pins = _setup(case, precision=32)
    right = np.tile([0.0, 0.0, 1.0], o.N)
    on = OracleNEB(o, [given[0], right], [5], spring_constant=1e5, pins=pins)
    gn = mmx.NEB(sim, given, [5], spring_constant=1e5)
    np.testing.assert_allclose(gn.energy, on.energy, rtol=1e-5, atol=0)
    for _ in range(3):
        on.advance_step()
        gn.advance(1)
    print(f"[neb fp32] {case}: max |dm| {np.abs(gn.spin - on.spin).max():.2e}", flush=True)
    assert gn.state["nsteps"] == on.nsteps
    assert np.abs(gn.spin - on.spin).max() < 1e-5


def test_neb_macrospin_barrier_reference():
    """test/test_neb.jl:56-108 on the GPU: NEB(sim, (m0, m1), (1,)) puts the
    one free image at (1,1,0)/sqrt(2); the macrospin band (Kx = 5e4 along x,
    Kz = -2e4 along z, 11 free images) relaxes to the barrier Kx V =
    5e4 (5e-9)^3 within 1e-3, and its saddle energy equals that of the
    oracle's band relaxed the same way to 1e-5 of the barrier (at stopping_dmdt
    = 0.001 the images still slide along the band, so the two runs stop a few
    steps apart with the intermediate images at slightly different places;
    step-for-step agreement is checked by test_neb_matches_oracle)."""
    from test_oracle_neb import macrospin_oracle

    print("[case] NEB macrospin (test_neb.jl)", flush=True)

    def sim(exch):
        mesh = mmx.FDMesh(dx=5e-9, dy=5e-9, dz=5e-9, nx=1, ny=1, nz=1)
        s = mmx.Sim(mesh)
        mmx.set_Ms(s, 8e5)
        mmx.init_m0(s, (0.6, 0, -0.8))
        if exch:
            mmx.add_exch(s, 1.3e-11)
        mmx.add_anis(s, 5e4, axis=(1, 0, 0), name="Kx")
        mmx.add_anis(s, -2e4, axis=(0, 0, 1), name="Kz")
        return s

    two = mmx.NEB(sim(True), [(1, 0, 0), (0, 1, 0)], [1])
    np.testing.assert_allclose(two.spin.reshape(-1), [0.7071067811865475, 0.7071067811865475, 0.0], rtol=1e-15)
    neb = mmx.NEB(sim(False), [(1, 0, 0), (0, 1, 1), (-1, 0, 0)], [5, 5])
    assert neb.spin.size == 3 * (5 + 5 + 1)
    expected = 5e4 * 5e-9 ** 3
    init_diff = neb.energy.max() - neb.energy[0]
    neb.relax(max_steps=500, stopping_dmdt=0.001)
    final_diff = neb.energy.max() - neb.energy[0]
    print(f"init {init_diff:.6e} expected {expected:.6e} final {final_diff:.6e}", flush=True)
    assert init_diff > expected
    assert abs(final_diff - expected) / expected < 1e-3
    o = macrospin_oracle()
    ref = OracleNEB(o, [np.array([1, 0, 0.0]), np.array([0, 1, 1.0]), np.array([-1, 0, 0.0])], [5, 5])
    ref.relax(max_steps=500, stopping_dmdt=0.001)
    assert abs(neb.energy.max() - ref.energy.max()) <= 1e-5 * expected
    assert neb.energy[0] == ref.energy[0] and neb.energy[-1] == ref.energy[-1]
