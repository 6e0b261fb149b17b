"""Parity at the sizes of the benchmarked workloads (HIP path through the C ABI
against the CPU oracle or fixtures the oracle wrote).

* 256x256x32 chiral film with 5x5 macro-PBC images: H_eff, dm/dt and the
  per-term energies to 1e-12 (max-norm relative), then RK4 steps through the
  compiled-signature stage kernels with chained spectra to 1e-12.
* Demag tensor where the Newell sums give way to the point-dipole form
  (R / max(d) > 60, src/micro/demag.jl:197-199,250-252): element-wise against
  the oracle, and the field of a mesh whose 5x5 images are almost all dipolar.
* 200x50x1 soft film, Dormand-Prince 5(4) tol 1e-6 from 0 to 1 ns, against
  the oracle's run (tests/golden/cfg1_soft_film.json): accepted steps and <m>.
* 512x512x64 precision form: the FP32 path's H_eff, dm/dt and energies against
  the FP64 oracle (tests/golden/cfg3_fields.npz), the FP64 path to 1e-12 at
  the same cells, and the FP32 trajectory error per Dormand-Prince step
  against the FP64 path.
* 128x128x8 film with Ku(x,y) and D(x,y) random fields and demag (the
  differentiation workload's forward path): FP64 to 1e-12, FP32 against the
  FP64 oracle to 1e-5 in H_eff and dm/dt.
Each case prints its name before it touches the GPU.
"""
import json
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
MU0 = 4e-7 * np.pi
TOL = 1e-12


def note(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def relerr(a, b):
    den = np.max(np.abs(b))
    return np.max(np.abs(a - b)) / (den if den > 0 else 1.0)


def film_m0(N, seed=2002):
    """Seeded random unit vectors (the benchmarked film's initial state)."""
    rng = np.random.default_rng(seed)
    z = rng.uniform(-1, 1, N)
    phi = rng.uniform(0, 2 * np.pi, N)
    r = np.sqrt(1 - z * z)
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1).reshape(-1)


def chiral_film(oracle, nx, ny, nz, images):
    import mmx

    N = nx * ny * nz
    m0 = film_m0(N)
    mesh = mmx.FDMesh(dx=2e-9, dy=2e-9, dz=2e-9, nx=nx, ny=ny, nz=nz, pbc="xy")
    g = mmx.Sim(mesh, integrator="RungeKutta")
    o = oracle.OracleSim(nx, ny, nz, 2e-9, 2e-9, 2e-9, pbc="xy")
    mmx.set_Ms(g, 5.8e5)
    o.set_Ms(np.full(N, 5.8e5))
    mmx.init_m0(g, m0)
    o.init_m0(m0)
    g.driver.alpha, g.driver.gamma, g.driver.precession = 0.1, 2.21e5, True
    o.set_llg(0.1, 2.21e5, True)
    H = (0, 0, 0.1 / MU0)
    mmx.add_exch(g, 1.5e-11)
    o.add_exch(1.5e-11)
    mmx.add_dmi(g, 3e-3, type="interfacial")
    o.add_dmi(np.full(N, 3e-3), type="interfacial")
    mmx.add_anis(g, 8e5, axis=(0, 0, 1))
    o.add_anis(np.full(N, 8e5), axis=(0, 0, 1))
    mmx.add_zeeman(g, H)
    o.add_zeeman(np.tile(H, N).astype(float))
    mmx.add_demag(g, Nx=images, Ny=images, Nz=0)
    o.add_demag(images, images, 0)
    return g, o


@pytest.mark.timeout(600)
def test_film_256x256x32_fields_energies_rk4(oracle):
    """The benchmarked workload itself: every kernel instantiation of the timed
    step (transform lengths 256/512/64, the compiled chiral-film signature,
    stage chaining) against the oracle."""
    import mmx

    note("256x256x32 chiral film, 5x5 images: fields, energies, 3 RK4 steps")
    g, o = chiral_film(oracle, 256, 256, 32, 5)
    eH = relerr(mmx.effective_field(g), o.effective_field())
    eR = relerr(mmx.llg_rhs(g), o.llg_rhs())
    Eg, Eo = mmx.compute_system_energy(g), o.term_energies()
    eE = np.max(np.abs(Eg - Eo) / np.abs(Eo))
    print(f"H_eff {eH:.2e}  dm/dt {eR:.2e}  energies {eE:.2e}", flush=True)
    assert eH <= TOL and eR <= TOL and eE <= TOL
    mmx.set_driver(g, integrator="RungeKutta", step=1e-14)
    o.set_integrator(oracle.RK4, 1e-14, 0)
    mmx.run_step(g, n=3)  # one call: the stages and the first two finals chain their spectra
    for _ in range(3):
        o.advance_step()
    e3 = np.max(np.abs(g.spin - o.spin))
    print(f"3 RK4 steps: max |dm| {e3:.2e}", flush=True)
    assert e3 <= TOL


def _grf(rng, nx, ny, dx, corr, mean, rel):
    """Seeded Gaussian random field: white noise filtered by a Gaussian of
    correlation length corr, unit variance, then mean * (1 + rel * g)."""
    g = rng.standard_normal((ny, nx))
    kx = np.fft.fftfreq(nx, d=dx) * 2 * np.pi
    ky = np.fft.fftfreq(ny, d=dx) * 2 * np.pi
    K2 = ky[:, None] ** 2 + kx[None, :] ** 2
    g = np.real(np.fft.ifft2(np.fft.fft2(g) * np.exp(-K2 * corr * corr / 4)))
    g = (g - g.mean()) / g.std()
    return mean * (1 + rel * g)


@pytest.mark.timeout(600)
def test_differentiation_film_128x128x8_forward(oracle):
    """The differentiation workload's forward path at its size: 128x128x8, 2 nm
    cells, Ku(x,y) and D(x,y) Gaussian random fields (means 6e5 J/m^3 and
    2.5e-3 J/m^2, 20 % std, 10 nm correlation), exchange + interfacial DMI +
    anisotropy + demag (transform lengths 128 / 256 / 16, the slot-form kernels
    with per-cell parameters); H_eff, dm/dt and 4 RK4 steps at dt 5e-14."""
    import mmx

    note("128x128x8 spatial Ku / D film with demag: fields and 4 RK4 steps")
    nx, ny, nz, d = 128, 128, 8, 2e-9
    N = nx * ny * nz
    rng = np.random.default_rng(5005)
    Ku = np.tile(_grf(rng, nx, ny, d, 10e-9, 6e5, 0.2).reshape(-1), nz)
    D = np.tile(_grf(np.random.default_rng(5006), nx, ny, d, 10e-9, 2.5e-3, 0.2).reshape(-1), nz)
    r = np.random.default_rng(5009)
    m0 = np.array([0.1, 0.1, 1.0]) + 0.3 * r.standard_normal((N, 3))
    m0 = (m0 / np.linalg.norm(m0, axis=1, keepdims=True)).reshape(-1)
    mesh = mmx.FDMesh(dx=d, dy=d, dz=d, nx=nx, ny=ny, nz=nz)
    g = mmx.Sim(mesh)
    o = oracle.OracleSim(nx, ny, nz, d, d, d)
    mmx.set_Ms(g, 6e5)
    o.set_Ms(np.full(N, 6e5))
    mmx.init_m0(g, m0)
    o.init_m0(m0)
    g.driver.alpha, g.driver.gamma, g.driver.precession = 0.1, 2.21e5, True
    o.set_llg(0.1, 2.21e5, True)
    mmx.add_exch(g, 1.3e-11)
    o.add_exch(1.3e-11)
    mmx.add_dmi(g, D, type="interfacial")
    o.add_dmi(D, type="interfacial")
    mmx.add_anis(g, Ku, axis=(0, 0, 1))
    o.add_anis(Ku, axis=(0, 0, 1))
    mmx.add_demag(g)
    o.add_demag(0, 0, 0)
    eH = relerr(mmx.effective_field(g), o.effective_field())
    eR = relerr(mmx.llg_rhs(g), o.llg_rhs())
    print(f"H_eff {eH:.2e}  dm/dt {eR:.2e}", flush=True)
    assert eH <= TOL and eR <= TOL
    mmx.set_driver(g, integrator="RungeKutta", step=5e-14)
    o.set_integrator(oracle.RK4, 5e-14, 0)
    mmx.run_step(g, n=4)
    for _ in range(4):
        o.advance_step()
    e4 = np.max(np.abs(g.spin - o.spin))
    print(f"4 RK4 steps: max |dm| {e4:.2e}", flush=True)
    assert e4 <= TOL


@pytest.mark.timeout(600)
def test_differentiation_film_128x128x8_forward_fp32(oracle):
    """The same 128x128x8 film through the FP32 kernels against the FP64
    oracle: H_eff and dm/dt within 1e-5 (max-norm relative, the bar of the
    512x512x64 FP32 case), and m after 4 RK4 steps within 1e-6."""
    import mmx

    note("128x128x8 spatial Ku / D film with demag, FP32: fields and 4 RK4 steps")
    nx, ny, nz, d = 128, 128, 8, 2e-9
    N = nx * ny * nz
    Ku = np.tile(_grf(np.random.default_rng(5005), nx, ny, d, 10e-9, 6e5, 0.2).reshape(-1), nz)
    D = np.tile(_grf(np.random.default_rng(5006), nx, ny, d, 10e-9, 2.5e-3, 0.2).reshape(-1), nz)
    r = np.random.default_rng(5009)
    m0 = np.array([0.1, 0.1, 1.0]) + 0.3 * r.standard_normal((N, 3))
    m0 = (m0 / np.linalg.norm(m0, axis=1, keepdims=True)).reshape(-1)
    mesh = mmx.FDMesh(dx=d, dy=d, dz=d, nx=nx, ny=ny, nz=nz)
    g = mmx.Sim(mesh, precision=mmx.F32)
    o = oracle.OracleSim(nx, ny, nz, d, d, d)
    mmx.set_Ms(g, 6e5)
    o.set_Ms(np.full(N, 6e5))
    mmx.init_m0(g, m0)
    o.init_m0(m0)
    g.driver.alpha, g.driver.gamma, g.driver.precession = 0.1, 2.21e5, True
    o.set_llg(0.1, 2.21e5, True)
    mmx.add_exch(g, 1.3e-11)
    o.add_exch(1.3e-11)
    mmx.add_dmi(g, D, type="interfacial")
    o.add_dmi(D, type="interfacial")
    mmx.add_anis(g, Ku, axis=(0, 0, 1))
    o.add_anis(Ku, axis=(0, 0, 1))
    mmx.add_demag(g)
    o.add_demag(0, 0, 0)
    eH = relerr(mmx.effective_field(g), o.effective_field())
    eR = relerr(mmx.llg_rhs(g), o.llg_rhs())
    print(f"FP32 H_eff {eH:.2e}  dm/dt {eR:.2e}", flush=True)
    assert eH <= 1e-5 and eR <= 1e-5
    mmx.set_driver(g, integrator="RungeKutta", step=5e-14)
    o.set_integrator(oracle.RK4, 5e-14, 0)
    mmx.run_step(g, n=4)
    for _ in range(4):
        o.advance_step()
    e4 = np.max(np.abs(g.spin - o.spin))
    print(f"FP32 4 RK4 steps: max |dm| {e4:.2e}", flush=True)
    assert e4 <= 1e-6


def dipole_xx(x, y, z, V):
    """Point-dipole Nxx of src/atomistic/demag.jl:106-113 times V/(4 pi)."""
    R = x * x + y * y + z * z
    return -(2 * x * x - y * y - z * z) / (R * R * np.sqrt(R)) * V / (4 * np.pi)


def test_demag_tensor_dipole_branch(oracle):
    """Tensor entries at R / max(d) > 60 take the dipole form; every entry of
    the library's tensor equals the oracle's to 1e-13 of itself, and the far
    entries equal the point-dipole closed form."""
    import mmx

    note("demag tensor 90x3x2 (dipole branch)")
    nx, ny, nz, d = 90, 3, 2, 2e-9
    N = nx * ny * nz
    mesh = mmx.FDMesh(dx=d, dy=d, dz=d, nx=nx, ny=ny, nz=nz)
    g = mmx.Sim(mesh)
    mmx.set_Ms(g, 8e5)
    mmx.init_m0(g, (1, 0.2, 0.1))
    t = mmx.add_demag(g)
    o = oracle.OracleSim(nx, ny, nz, d, d, d)
    o.set_Ms(np.full(N, 8e5))
    o.init_m0(np.tile([1, 0.2, 0.1], N).astype(float))
    o.add_demag(0, 0, 0)
    tg = mmx.demag_tensor(g, t)
    to = o.demag_tensor(0)
    nz_ = np.abs(to) > 0
    rel = np.max(np.abs(tg - to)[nz_] / np.abs(to)[nz_])
    print(f"element-wise relative difference {rel:.2e}", flush=True)
    assert rel <= 1e-13
    # entry (i, 0, 0) with i > 60 cells: the dipole form
    for i in (61, 75, 89):
        ref = dipole_xx(i * d, 0.0, 0.0, d ** 3)
        assert abs(tg[0, i] - ref) <= 1e-14 * abs(ref), (i, tg[0, i], ref)


def test_images_dipole_field(oracle):
    """A 16x16x2 film with 5x5 images: image displacements reach 80 cells, so
    most image sums use the dipole form; H_eff and dm/dt to 1e-12."""
    import mmx

    note("16x16x2 film, 5x5 images: dipolar image sums")
    nx, ny, nz = 16, 16, 2
    N = nx * ny * nz
    m0 = film_m0(N, seed=7)
    mesh = mmx.FDMesh(dx=2e-9, dy=2e-9, dz=2e-9, nx=nx, ny=ny, nz=nz, pbc="xy")
    g = mmx.Sim(mesh)
    o = oracle.OracleSim(nx, ny, nz, 2e-9, 2e-9, 2e-9, pbc="xy")
    mmx.set_Ms(g, 5.8e5)
    o.set_Ms(np.full(N, 5.8e5))
    mmx.init_m0(g, m0)
    o.init_m0(m0)
    mmx.add_exch(g, 1.5e-11)
    o.add_exch(1.5e-11)
    mmx.add_demag(g, Nx=5, Ny=5, Nz=0)
    o.add_demag(5, 5, 0)
    assert relerr(mmx.effective_field(g), o.effective_field()) <= TOL
    assert relerr(mmx.llg_rhs(g), o.llg_rhs()) <= TOL


def test_soft_film_dopri_1ns():
    """200x50x1 soft film, exchange + demag, DOPRI54 tol 1e-6, 0 -> 1 ns, against
    the oracle's run (tests/golden/cfg1_soft_film.json, make_cfg1_fixture.py):
    <m> at every 0.1 ns to 1e-7, accepted steps within 1 % and evaluations
    within 3 %.  The step sequences are not identical over a whole nanosecond:
    the embedded error estimate is a difference of nearly equal stage sums, so
    the ~1e-16 differences of the FFT demag become ~1e-7 relative differences
    of the error norm, the step sizes drift apart, and near the 0.1 ns marks
    (where the last step is clamped) a step or a rejected try more or less
    follows.  Identical step counts are checked on a short run of a smaller
    film in test_gpu_parity.test_adaptive_trajectory."""
    sys.path.insert(0, GOLDEN)
    import make_cfg1_fixture as F

    import mmx

    ref = json.load(open(os.path.join(GOLDEN, "cfg1_soft_film.json")))
    note("200x50x1 soft film DOPRI54 0 -> 1 ns")
    mesh = mmx.FDMesh(dx=F.DX, dy=F.DY, dz=F.DZ, nx=F.NX, ny=F.NY, nz=F.NZ)
    g = mmx.Sim(mesh, integrator="DormandPrince")
    mmx.set_Ms(g, F.MS)
    mmx.init_m0(g, F.m0())
    mmx.add_exch(g, F.A)
    mmx.add_demag(g)
    g.driver.alpha, g.driver.gamma = F.ALPHA, F.GAMMA
    mmx.set_driver(g, integrator="DormandPrince", tol=F.TOL)
    worst = 0.0
    for row in ref["rows"]:
        mmx.run_until(g, row["t"])
        ns, nf = g._integrator.nsteps, g._integrator.nfevals
        mavg = g.spin.reshape(-1, 3).mean(axis=0)
        worst = max(worst, float(np.max(np.abs(mavg - np.array(row["m"])))))
        print(f"t {row['t']:.1e}: steps {ns} (oracle {row['nsteps']}), evaluations {nf} (oracle {row['nfevals']}), "
              f"<m> {mavg}", flush=True)
        assert abs(ns - row["nsteps"]) <= 0.01 * row["nsteps"], (row["t"], ns, row["nsteps"])
        assert abs(nf - row["nfevals"]) <= 0.03 * row["nfevals"], (row["t"], nf, row["nfevals"])
    print(f"<m> max difference over 0.1 ns marks: {worst:.2e}", flush=True)
    assert worst <= 1e-7


def _cfg3_sim(F, precision):
    import mmx

    N = F.NX * F.NY * F.NZ
    mesh = mmx.FDMesh(dx=F.D, dy=F.D, dz=F.D, nx=F.NX, ny=F.NY, nz=F.NZ)
    g = mmx.Sim(mesh, integrator="DormandPrince", precision=precision)
    mmx.set_Ms(g, F.MS)
    mmx.init_m0(g, F.smooth_m0())
    mmx.add_exch(g, F.A)
    mmx.add_dmi(g, F.DB, type="bulk")
    mmx.add_cubic_anis(g, F.KC, (1, 0, 0), (0, 1, 0))
    mmx.add_demag(g)
    g.driver.alpha, g.driver.gamma, g.driver.precession = F.ALPHA, F.GAMMA, True
    mmx.set_driver(g, integrator="DormandPrince", tol=1e-6)
    assert g.n_total == N
    return g


@pytest.mark.timeout(900)
def test_precision_form_512x512x64():
    """FP32 stencils, LLG and FFT demag (FP64-built tensor, FP64 energies and
    error norm) on 16.8 M cells: H_eff, dm/dt and energies against the FP64
    oracle at 4096 cells; the FP64 path at the same cells to 1e-12; then both
    paths integrate with Dormand-Prince (their own step sequences) to four
    common times, where the FP32 state, field and energies are compared with
    the FP64 path's."""
    sys.path.insert(0, GOLDEN)
    import make_cfg3_fixture as F

    import mmx

    ref = np.load(os.path.join(GOLDEN, "cfg3_fields.npz"))
    idx = ref["cells"]
    Hr, Rr, Er = ref["H"], ref["dmdt"], ref["energies"]
    res = {}
    sims = {}
    for prec in (64, 32):
        note(f"512x512x64 precision form FP{prec}: set-up and fields")
        g = _cfg3_sim(F, mmx.F64 if prec == 64 else mmx.F32)
        H = mmx.effective_field(g).reshape(-1, 3)[idx]
        R = mmx.llg_rhs(g).reshape(-1, 3)[idx]
        E = mmx.compute_system_energy(g)
        res[prec] = (relerr(H, Hr), relerr(R, Rr), float(np.max(np.abs(E - Er) / np.abs(Er))))
        print(f"FP{prec} vs FP64 oracle at {len(idx)} cells: H_eff {res[prec][0]:.2e}  dm/dt {res[prec][1]:.2e}  "
              f"energies {res[prec][2]:.2e}", flush=True)
        sims[prec] = g
    assert max(res[64]) <= TOL
    assert res[32][0] < 1e-5 and res[32][1] < 1e-5 and res[32][2] < 1e-5
    g64, g32 = sims[64], sims[32]
    worst = 0.0
    for t in (1e-14, 5e-14, 1e-13, 2e-13):
        note(f"512x512x64 DOPRI54 to t = {t:.0e} s (FP64, FP32)")
        mmx.run_until(g64, t)
        mmx.run_until(g32, t)
        dm = float(np.max(np.abs(g64.spin - g32.spin)))
        eH = relerr(mmx.effective_field(g32), mmx.effective_field(g64))
        E64 = mmx.compute_system_energy(g64)
        dE = float(np.max(np.abs(E64 - mmx.compute_system_energy(g32)) / np.abs(E64)))
        print(f"t {t:.0e}: steps FP64 {g64._integrator.nsteps} FP32 {g32._integrator.nsteps}  max|m32 - m64| {dm:.2e}  "
              f"H_eff {eH:.2e}  energies {dE:.2e}", flush=True)
        worst = max(worst, dm)
    assert worst < 1e-4
