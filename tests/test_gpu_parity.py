"""Parity of the HIP path (through the C ABI) with the CPU oracle.

FP64 gate: max_i |gpu - oracle| / max_i |oracle| <= 1e-12 per field (H_eff,
each term's field, dm/dt) and <= 1e-12 relative for energies.  Trajectory gates
are stated per test.  Each case prints its name before touching the GPU.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-12


def note(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def relerr(a, b):
    den = np.max(np.abs(b))
    return np.max(np.abs(a - b)) / (den if den > 0 else 1.0)


def random_unit(rng, N):
    v = rng.standard_normal((N, 3))
    return (v / np.linalg.norm(v, axis=1, keepdims=True)).reshape(-1)


def build(oracle, nx, ny, nz, d=(2e-9, 2e-9, 2e-9), pbc="", Ms=None, m0=None, terms=(), seed=0,
          precision=64, llg=(0.1, 2.21e5, True), norm=True):
    """The same system in the HIP library and in the oracle."""
    import mmx

    rng = np.random.default_rng(seed)
    N = nx * ny * nz
    Ms = np.full(N, 8e5) if Ms is None else np.asarray(Ms, dtype=float)
    m0 = random_unit(rng, N) if m0 is None else np.asarray(m0, dtype=float)
    mesh = mmx.FDMesh(dx=d[0], dy=d[1], dz=d[2], nx=nx, ny=ny, nz=nz, pbc=pbc or "open")
    g = mmx.Sim(mesh, precision=precision)
    o = oracle.OracleSim(nx, ny, nz, *d, pbc=pbc)
    mmx.set_Ms(g, Ms)
    o.set_Ms(Ms)
    mmx.init_m0(g, m0, norm=norm)
    o.init_m0(m0, norm=norm)
    g.driver.alpha, g.driver.gamma, g.driver.precession = llg
    o.set_llg(*llg)
    for kind, kw in terms:
        kw = dict(kw)
        if kind == "exch":
            mmx.add_exch(g, kw["A"])
            o.add_exch(kw["A"])
        elif kind == "dmi":
            mmx.add_dmi(g, kw["D"], type=kw.get("type", "bulk"))
            o.add_dmi(kw["D"], type=kw.get("type", "bulk"))
        elif kind == "anis":
            mmx.add_anis(g, kw["Ku"], axis=kw.get("axis", (0, 0, 1)))
            o.add_anis(kw["Ku"], axis=kw.get("axis", (0, 0, 1)))
        elif kind == "cubic":
            mmx.add_cubic_anis(g, kw["Kc"], kw.get("axis1", (1, 0, 0)), kw.get("axis2", (0, 1, 0)))
            o.add_cubic_anis(kw["Kc"], kw.get("axis1", (1, 0, 0)), kw.get("axis2", (0, 1, 0)))
        elif kind == "zeeman":
            mmx.add_zeeman(g, kw["H"])
            o.add_zeeman(np.asarray(kw["H"], dtype=float))
        elif kind == "demag":
            mmx.add_demag(g, Nx=kw.get("Nx", 0), Ny=kw.get("Ny", 0), Nz=kw.get("Nz", 0))
            o.add_demag(kw.get("Nx", 0), kw.get("Ny", 0), kw.get("Nz", 0))
        elif kind == "thermal":
            mmx.add_thermal_noise(g, kw["T"], seed=kw["seed"])
            o.add_thermal(np.full(N, kw["T"]) if np.ndim(kw["T"]) == 0 else kw["T"], 0.0, kw["seed"])
        else:
            raise ValueError(kind)
    return g, o


def load_field_data():
    import os

    d = np.loadtxt(os.path.join(os.path.dirname(__file__), "golden", "test_fields.txt"), comments="#")
    n3 = d.shape[0] - 1
    return d[: n3 // 3, 0], d[:n3, 1], [d[:n3, c] for c in range(2, 7)], d[-1, 2:7]


def test_golden_fields_per_term(oracle):
    """test/test_fields.jl on the HIP path: every term vs Fidimag and vs the oracle."""
    import mmx

    note("golden 20x5x3 per-term fields")
    Ms, m0, (fd, fe, fdmi, fdmi2, fan), E = load_field_data()
    g, o = build(oracle, 20, 5, 3, d=(2.5e-9, 2.5e-9, 3e-9), Ms=Ms, m0=m0, norm=False,
                 terms=[("exch", dict(A=1.3e-11)), ("demag", {}), ("dmi", dict(D=4e-3)),
                        ("dmi", dict(D=2e-3, type="interfacial")), ("anis", dict(Ku=-3e4))])
    refs = [fe, fd, fdmi, fdmi2, fan]
    Eref = [E[1], E[0], E[2], E[3], E[4]]
    Eg = mmx.compute_system_energy(g)
    Eo = o.term_energies()
    for t, (ref, er) in enumerate(zip(refs, Eref)):
        fg, eg = g._term_field(t)
        fo, eo = o.term_field(t)
        assert np.allclose(fg, ref, rtol=0, atol=1e-5)
        assert relerr(fg, fo) <= TOL, (t, relerr(fg, fo))
        assert relerr(eg, eo) <= TOL, (t, relerr(eg, eo))
        assert abs(Eg[t] - Eo[t]) <= TOL * abs(Eo[t])
        assert abs(Eg[t] - er) <= (1e-11 if t == 1 else 1e-13) * abs(er)
    assert relerr(mmx.effective_field(g), o.effective_field()) <= TOL


def test_golden_demag_arrays(oracle):
    """test/test_demag.jl:8-57 hard-coded fields through the HIP demag."""
    import json
    import os

    import mmx

    cases = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "demag_fields.json")))["cases"]
    for c in cases:
        note(f"demag golden {c['mesh']}")
        nx, ny, nz = c["mesh"]
        N = nx * ny * nz
        g, o = build(oracle, nx, ny, nz, d=c["d"], Ms=np.full(N, c["Ms"]), m0=np.tile(c["m0"], N),
                     terms=[("demag", {})])
        f = mmx.effective_field(g)
        assert np.allclose(f, c["expected"], rtol=1e-8, atol=1e-6)
        assert relerr(f, o.effective_field()) <= TOL


CASES = [
    # (name, mesh, d, pbc, terms, images)
    ("all-terms open", (16, 12, 6), (2e-9, 2e-9, 2e-9), "",
     [("exch", dict(A=1.5e-11)), ("dmi", dict(D=3e-3, type="interfacial")), ("anis", dict(Ku=8e5)),
      ("zeeman", dict(H=(0, 0, 79577.47))), ("demag", {})]),
    ("all-terms pbc-xy images", (12, 10, 4), (2e-9, 2e-9, 2e-9), "xy",
     [("exch", dict(A=1.5e-11)), ("dmi", dict(D=3e-3, type="interfacial")), ("anis", dict(Ku=8e5)),
      ("zeeman", dict(H=(0, 0, 79577.47))), ("demag", dict(Nx=2, Ny=2))]),
    ("bulk+cubic pbc-xyz", (9, 7, 5), (3e-9, 2.5e-9, 2e-9), "xyz",
     [("exch", dict(A=1.3e-11)), ("dmi", dict(D=1e-3)), ("cubic", dict(Kc=-5e3)), ("demag", {})]),
    ("thin film nz=1", (24, 20, 1), (4e-9, 4e-9, 4e-9), "",
     [("exch", dict(A=1e-13)), ("zeeman", dict(H=(0, 100 / (4e-4 * np.pi), 0))), ("demag", {})]),
    ("1-d chain", (33, 1, 1), (2e-9, 1e-9, 1e-9), "x", [("exch", dict(A=1.3e-11)), ("dmi", dict(D=2e-3))]),
    ("single cell", (1, 1, 1), (5e-9, 1e-9, 1e-9), "", [("demag", {}), ("zeeman", dict(H=(1e4, 0, 2e4)))]),
    ("tiny pbc n=1,2", (1, 2, 3), (2e-9, 2e-9, 2e-9), "xyz",
     [("exch", dict(A=1.3e-11)), ("dmi", dict(D=2e-3)), ("dmi", dict(D=1e-3, type="interfacial")),
      ("demag", {})]),
    ("no demag local only", (40, 30, 3), (2e-9, 2e-9, 2e-9), "xy",
     [("exch", dict(A=1.5e-11)), ("dmi", dict(D=3e-3, type="interfacial")), ("anis", dict(Ku=8e5, axis=(1, 1, 1)))]),
    # long z transforms: Pz = 128 (two kx per wave, paired tensor loads) and 256 (one kx per wave)
    ("tall column nz=40", (6, 5, 40), (2e-9, 2e-9, 2e-9), "xy",
     [("exch", dict(A=1.5e-11)), ("dmi", dict(D=3e-3, type="interfacial")), ("demag", dict(Nx=1, Ny=1))]),
    ("tall column nz=80", (4, 3, 80), (2e-9, 2e-9, 1e-9), "",
     [("exch", dict(A=1.3e-11)), ("anis", dict(Ku=5e5)), ("demag", {})]),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_heff_and_rhs_parity(oracle, case):
    import mmx

    name, (nx, ny, nz), d, pbc, terms = case
    note(f"parity {name} {nx}x{ny}x{nz}")
    rng = np.random.default_rng(5)
    N = nx * ny * nz
    Ms = np.full(N, 5.8e5)
    if N > 8:
        Ms[rng.choice(N, N // 10, replace=False)] = 0.0
    g, o = build(oracle, nx, ny, nz, d=d, pbc=pbc, Ms=Ms, terms=terms, seed=9)
    pins = np.zeros(N, dtype=np.uint8)
    pins[:: max(1, N // 7)] = 1
    mmx.set_pinning(g, pins)
    o.set_pins(pins)
    Hg, Ho = mmx.effective_field(g), o.effective_field()
    assert relerr(Hg, Ho) <= TOL, relerr(Hg, Ho)
    for t in range(len(terms)):
        fg, eg = g._term_field(t)
        fo, eo = o.term_field(t)
        assert relerr(fg, fo) <= TOL, (t, relerr(fg, fo))
        assert relerr(eg, eo) <= TOL, (t, relerr(eg, eo))
    Eg, Eo = mmx.compute_system_energy(g), o.term_energies()
    assert np.all(np.abs(Eg - Eo) <= TOL * np.maximum(np.abs(Eo), 1e-300) + 1e-300)
    dg, do = mmx.llg_rhs(g), o.llg_rhs()
    assert relerr(dg, do) <= TOL, relerr(dg, do)
    assert np.all(dg.reshape(N, 3)[pins == 1] == 0)


def test_spatial_parameters(oracle):
    """Per-cell A, bulk D (vector), interfacial D, Ku with per-cell axis, Kc, Zeeman field."""
    note("spatial parameters 10x8x4")
    import mmx

    rng = np.random.default_rng(21)
    nx, ny, nz = 10, 8, 4
    N = nx * ny * nz
    A = rng.uniform(1e-11, 2e-11, N)
    A[rng.choice(N, 5, replace=False)] = 0
    D3 = tuple(rng.uniform(-3e-3, 3e-3, N) for _ in range(3))
    Di = rng.uniform(1e-3, 3e-3, N)
    Ku = rng.uniform(1e5, 6e5, N)
    ax = rng.standard_normal(3 * N)
    Kc = rng.uniform(-5e4, 5e4, N)
    H = rng.standard_normal(3 * N) * 1e5
    terms = [("exch", dict(A=A)), ("dmi", dict(D=D3)), ("dmi", dict(D=Di, type="interfacial")),
             ("anis", dict(Ku=Ku, axis=ax)), ("cubic", dict(Kc=Kc, axis1=(1, 1, 0), axis2=(1, -1, 0))),
             ("zeeman", dict(H=H))]
    g, o = build(oracle, nx, ny, nz, pbc="y", terms=terms, seed=4)
    for t in range(len(terms)):
        fg, eg = g._term_field(t)
        fo, eo = o.term_field(t)
        assert relerr(fg, fo) <= TOL, (t, relerr(fg, fo))
        assert relerr(eg, eo) <= TOL, (t, relerr(eg, eo))
    assert relerr(mmx.llg_rhs(g), o.llg_rhs()) <= TOL


def test_demag_tensor_matches_oracle(oracle):
    import mmx

    note("demag tensor 12x10x4 images")
    g, o = build(oracle, 12, 10, 4, pbc="xy", terms=[("demag", dict(Nx=3, Ny=2))])
    tg = mmx.demag_tensor(g, g.interactions[0])
    to = o.demag_tensor(0)
    assert np.max(np.abs(tg - to)) <= 1e-15 * np.max(np.abs(to))


@pytest.mark.parametrize("integrator", ["RungeKutta", "Heun"])
def test_fixed_step_trajectory(oracle, integrator):
    import mmx

    note(f"fixed-step trajectory {integrator}")
    terms = [("exch", dict(A=1.5e-11)), ("dmi", dict(D=3e-3, type="interfacial")), ("anis", dict(Ku=8e5)),
             ("zeeman", dict(H=(0, 0, 79577.47))), ("demag", dict(Nx=1, Ny=1))]
    g, o = build(oracle, 16, 16, 4, pbc="xy", Ms=np.full(16 * 16 * 4, 5.8e5), terms=terms, seed=2)
    kind = {"RungeKutta": 0, "Heun": 1}[integrator]
    mmx.set_driver(g, integrator=integrator, step=1e-14)
    o.set_integrator(kind, 1e-14, 0)
    for _ in range(50):
        o.advance_step()
    mmx.run_step(g, n=50)
    err = np.max(np.abs(g.spin - o.spin))
    print("max |dm| after 50 steps:", err)
    assert err < 1e-12
    assert g._integrator.nsteps == 50 and abs(g._integrator.t - o.state()["t"]) == 0


@pytest.mark.parametrize("integrator", ["RungeKutta", "Heun"])
@pytest.mark.parametrize("precision", [64, 32])
def test_chained_stage_spectra(oracle, integrator, precision):
    """Stages hand the next stage's x-spectrum over from the last pass (the
    first pass is skipped): an n-step call equals n one-step calls bit for bit,
    a state written between calls is picked up, and the trajectory matches the
    oracle (odd nx, vacuum cells)."""
    import mmx

    note(f"chained stage spectra {integrator} FP{precision}")
    nx, ny, nz = 15, 12, 3
    N = nx * ny * nz
    rng = np.random.default_rng(77)
    Ms = np.where(rng.uniform(size=N) < 0.1, 0.0, 6e5)
    terms = [("exch", dict(A=1.3e-11)), ("dmi", dict(D=2e-3, type="interfacial")), ("demag", {})]
    kind = {"RungeKutta": 0, "Heun": 1}[integrator]
    runs = []
    for split in (False, True):
        g, o = build(oracle, nx, ny, nz, Ms=Ms, terms=terms, seed=5, precision=precision)
        mmx.set_driver(g, integrator=integrator, step=1e-14)
        if split:
            for _ in range(6):
                mmx.run_step(g, n=1)
        else:
            mmx.run_step(g, n=6)
        runs.append((g, o))
    (g1, o1), (g2, _) = runs
    assert np.array_equal(g1.spin, g2.spin)
    o1.set_integrator(kind, 1e-14, 0)
    for _ in range(6):
        o1.advance_step()
    err = np.max(np.abs(g1.spin - o1.spin))
    print(f"max |dm| after 6 steps: {err:.2e}")
    assert err < (1e-12 if precision == 64 else 1e-5)
    # a state written between calls replaces the handed-over spectrum
    m_new = random_unit(np.random.default_rng(8), N)
    for g in (g1, g2):
        mmx.init_m0(g, m_new)
    o1.init_m0(m_new)
    mmx.run_step(g1, n=3)
    for _ in range(3):
        mmx.run_step(g2, n=1)
        o1.advance_step()
    assert np.array_equal(g1.spin, g2.spin)
    err = np.max(np.abs(g1.spin - o1.spin))
    assert err < (1e-12 if precision == 64 else 1e-5), err


@pytest.mark.parametrize("mesh", [(1, 1, 1), (3, 1, 1), (1, 3, 2), (2, 2, 1), (5, 3, 3)])
def test_tiny_mesh_chained_trajectory(oracle, mesh):
    """Chained stages on degenerate transform sizes (Px = 2, Py or Pz = 1)."""
    import mmx

    nx, ny, nz = mesh
    note(f"tiny mesh chained {mesh}")
    N = nx * ny * nz
    terms = [("exch", dict(A=1.3e-11)), ("anis", dict(Ku=2e5)), ("demag", {})]
    g, o = build(oracle, nx, ny, nz, Ms=np.full(N, 7e5), terms=terms, seed=12)
    mmx.set_driver(g, integrator="RungeKutta", step=5e-14)
    o.set_integrator(0, 5e-14, 0)
    mmx.run_step(g, n=8)
    for _ in range(8):
        o.advance_step()
    assert np.max(np.abs(g.spin - o.spin)) < 1e-12


@pytest.mark.parametrize("integrator", ["DormandPrince", "BS23"])
def test_adaptive_trajectory(oracle, integrator):
    import mmx

    note(f"adaptive trajectory {integrator}")
    terms = [("exch", dict(A=1.3e-11)), ("demag", {})]
    N = 40 * 10 * 1
    rng = np.random.default_rng(1001)
    m0 = np.tile([1.0, 0.25, 0.1], N) + 0.05 * rng.uniform(-1, 1, 3 * N)
    g, o = build(oracle, 40, 10, 1, d=(2.5e-9, 2.5e-9, 3e-9), Ms=np.full(N, 8e5), m0=m0, terms=terms,
                 llg=(0.5, 2.211e5, True))
    kind = {"DormandPrince": 2, "BS23": 3}[integrator]
    mmx.set_driver(g, integrator=integrator, tol=1e-6)
    o.set_integrator(kind, 0, 1e-6)
    for i in range(1, 6):
        mmx.run_until(g, i * 2e-12)
        o.run_until(i * 2e-12)
    so = o.state()
    print("steps gpu/oracle", g._integrator.nsteps, so["nsteps"], "max|dm|", np.max(np.abs(g.spin - o.spin)))
    assert g._integrator.nsteps == so["nsteps"]
    # the embedded error estimate is an O(h^5) difference of O(1) stages, so
    # rounding differences in k are amplified ~1e8 in it; the step size follows
    assert abs(g._integrator.step_next - so["step_next"]) <= 1e-6 * so["step_next"]
    assert np.max(np.abs(g.spin - o.spin)) < 1e-10


def test_macrospin_analytic():
    """test/test_llg.jl:23-60 on the HIP path."""
    import mmx

    note("macrospin analytic DOPRI54")
    mesh = mmx.FDMesh(nx=1, ny=1, dx=1e-9)
    sim = mmx.Sim(mesh)
    mmx.set_Ms(sim, 8e5)
    sim.driver.gamma = 2.21e5
    sim.driver.alpha = 0.05
    mmx.set_driver(sim, integrator="DormandPrince", tol=1e-8)
    mmx.add_zeeman(sim, (0, 0, 1e5))
    mmx.init_m0(sim, (1.0, 0, 0))
    for i in range(1, 11):
        mmx.run_until(sim, 1e-10 * i)
    b = 2.21e5 / (1 + 0.05 ** 2) * 1e5 * 1e-9
    exp = np.array([np.cos(b) / np.cosh(0.05 * b), np.sin(b) / np.cosh(0.05 * b), np.tanh(0.05 * b)])
    assert np.all(np.abs(sim.spin - exp) < 8e-7)


def test_relax_pins(oracle):
    """test/test_pins.jl (LLG, precession off) on the HIP path."""
    import mmx

    note("relax with pins")
    N = 100
    mesh = mmx.FDMesh(nx=N, ny=1, nz=1, dx=2e-9, dy=2e-9, dz=1e-9)
    sim = mmx.Sim(mesh)
    sim.driver.precession = False
    sim.driver.alpha = 0.5
    mmx.set_Ms(sim, 8.6e5)
    mmx.set_pinning(sim, lambda i, j, k, dx, dy, dz: i == 1 or i == 100)
    mmx.add_exch(sim, 1.3e-11)
    mmx.init_m0(sim, lambda i, j, k, dx, dy, dz: (0, 0, 1) if i == 1 else ((0, 1, 0) if i == 100 else (1, 1, 1)))
    steps, md = mmx.relax(sim, max_steps=10000, stopping_dmdt=0.1)
    m = sim.spin
    assert abs(m[0]) < 1e-15 and abs(m[1]) < 1e-15 and abs(m[2] - 1) < 1e-15
    assert md < 0.1 * (2 * np.pi / 360) * 1e9


def test_thermal_field_parity(oracle):
    import mmx

    note("thermal field 8x8x2")
    g, o = build(oracle, 8, 8, 2, terms=[("exch", dict(A=1e-11)), ("thermal", dict(T=300.0, seed=77))],
                 llg=(0.1, 2.21e5, True))
    mmx.set_driver(g, integrator="Heun", step=1e-14)
    o.set_integrator(1, 1e-14, 0)
    fg, _ = g._term_field(1)
    fo, _ = o.term_field(1)
    assert relerr(fg, fo) <= 1e-12
    mmx.run_step(g, n=3)
    for _ in range(3):
        o.advance_step()
    assert np.max(np.abs(g.spin - o.spin)) < 1e-11


@pytest.mark.parametrize("integrator", ["Heun", "RungeKutta"])
def test_thermal_demag_chained_trajectory(oracle, integrator):
    """sLLG with demag: noise regenerated per step while the stages hand their
    x-spectra over (the chained last pass)."""
    import mmx

    note(f"thermal + demag {integrator}")
    g, o = build(oracle, 10, 8, 3, pbc="x", terms=[("exch", dict(A=1e-11)), ("demag", {}),
                                                   ("thermal", dict(T=300.0, seed=91))],
                 llg=(0.1, 2.21e5, True), seed=6)
    kind = {"Heun": 1, "RungeKutta": 0}[integrator]
    mmx.set_driver(g, integrator=integrator, step=1e-14)
    o.set_integrator(kind, 1e-14, 0)
    mmx.run_step(g, n=6)
    for _ in range(6):
        o.advance_step()
    assert np.max(np.abs(g.spin - o.spin)) < 1e-11


def test_fp32_error_vs_fp64_oracle(oracle):
    """Precision form: FP32 stencils/LLG/FFT; error against the FP64 oracle is stated."""
    import mmx

    note("fp32 16x16x8")
    terms = [("exch", dict(A=1.3e-11)), ("dmi", dict(D=1e-3)), ("cubic", dict(Kc=-5e3)), ("demag", {})]
    g, o = build(oracle, 16, 16, 8, d=(3e-9, 3e-9, 3e-9), terms=terms, precision=32, seed=3)
    eH = relerr(mmx.effective_field(g), o.effective_field())
    eR = relerr(mmx.llg_rhs(g), o.llg_rhs())
    Eg, Eo = mmx.compute_system_energy(g), o.term_energies()
    eE = np.max(np.abs(Eg - Eo) / np.abs(Eo))
    print(f"FP32 vs FP64 oracle: H_eff {eH:.2e}  dm/dt {eR:.2e}  energy {eE:.2e}")
    assert eH < 1e-5 and eR < 1e-5 and eE < 1e-5


SIG_CASES = [
    # (name, mesh, pbc, terms, integrator) -- kernels compiled for term signatures
    # (bulk DMI + cubic + demag; the chiral film without demag; D and Ku maps
    # with demag), the order and per-cell parameters that take the
    # signature-free kernels
    ("bulk-cubic-demag", (12, 10, 4), "xy",
     [("exch", dict(A=1.3e-11)), ("dmi", dict(D=1e-3)), ("cubic", dict(Kc=-5e4)), ("demag", {})], "RungeKutta"),
    ("bulk-cubic-demag dopri", (12, 10, 4), "",
     [("exch", dict(A=1.3e-11)), ("dmi", dict(D=1e-3)), ("cubic", dict(Kc=-5e4)), ("demag", {})], "DormandPrince"),
    ("chiral-local", (20, 14, 3), "xy",
     [("exch", dict(A=1.5e-11)), ("dmi", dict(D=3e-3, type="interfacial")), ("anis", dict(Ku=8e5)),
      ("zeeman", dict(H=(0, 0, 79577.47)))], "Heun"),
    ("reordered terms", (12, 10, 4), "xy",
     [("zeeman", dict(H=(0, 0, 79577.47))), ("anis", dict(Ku=8e5)), ("exch", dict(A=1.5e-11)),
      ("demag", {}), ("dmi", dict(D=3e-3, type="interfacial"))], "RungeKutta"),
    ("per-cell Ku", (12, 10, 4), "xy",
     [("exch", dict(A=1.5e-11)), ("dmi", dict(D=3e-3, type="interfacial")),
      ("anis", dict(Ku=np.linspace(6e5, 9e5, 480))), ("zeeman", dict(H=(0, 0, 79577.47))), ("demag", {})],
     "RungeKutta"),
    # D(x,y) and Ku(x,y) maps with demag and no field: a compiled signature
    ("per-cell D and Ku", (12, 10, 4), "",
     [("exch", dict(A=1.3e-11)), ("dmi", dict(D=np.linspace(2e-3, 3e-3, 480), type="interfacial")),
      ("anis", dict(Ku=np.linspace(9e5, 6e5, 480))), ("demag", {})], "RungeKutta"),
    ("per-cell D and Ku dopri", (12, 10, 4), "xy",
     [("exch", dict(A=1.3e-11)), ("dmi", dict(D=np.linspace(2e-3, 3e-3, 480), type="interfacial")),
      ("anis", dict(Ku=np.linspace(9e5, 6e5, 480))), ("demag", {})], "DormandPrince"),
]


@pytest.mark.parametrize("case", SIG_CASES, ids=[c[0] for c in SIG_CASES])
def test_signature_kernels_trajectory(oracle, case):
    """Integrator stages through the signature-specialised and the generic slot
    kernels agree with the oracle's trajectory."""
    import mmx

    name, (nx, ny, nz), pbc, terms, integrator = case
    note(f"signature trajectory {name}")
    N = nx * ny * nz
    g, o = build(oracle, nx, ny, nz, pbc=pbc, Ms=np.full(N, 5.8e5), terms=terms, seed=4)
    kind = {"RungeKutta": 0, "Heun": 1, "DormandPrince": 2}[integrator]
    if kind < 2:
        mmx.set_driver(g, integrator=integrator, step=1e-14)
        o.set_integrator(kind, 1e-14, 0)
        for _ in range(20):
            o.advance_step()
        mmx.run_step(g, n=20)
    else:
        mmx.set_driver(g, integrator=integrator, tol=1e-6)
        o.set_integrator(kind, 0, 1e-6)
        # the oracle copies y_current into its spin array at the end of
        # run_until only (adaptive_rk.jl: advance_step leaves sim.spin alone)
        for i in range(1, 5):
            mmx.run_until(g, i * 1.5e-14)
            o.run_until(i * 1.5e-14)
        assert g._integrator.nsteps == o.state()["nsteps"]
    err = np.max(np.abs(g.spin - o.spin))
    print(f"{name}: max |dm| {err:.2e}")
    assert err < 1e-12


def test_spatial_llg_driver_set_alpha():
    """test/test_spatial_llg.jl:15-37 on the HIP path."""
    import mmx

    note("SpatialLLG set_alpha")
    mesh = mmx.FDMesh(nx=100, ny=1, nz=1, dx=2e-9, dy=2e-9, dz=1e-9)
    sim = mmx.Sim(mesh, name="test_alpha", driver="SpatialLLG")
    mmx.set_Ms(sim, 8.6e5)
    mmx.set_alpha(sim, lambda i, j, k, dx, dy, dz: 0.1)
    alpha = np.asarray(sim.driver.alpha)
    assert alpha.max() == alpha.min()
    mmx.set_alpha(sim, np.random.default_rng(1).uniform(size=100) * 0.2 + 0.01)
    alpha = np.asarray(sim.driver.alpha)
    assert alpha.max() <= 0.21 and alpha.min() >= 0.01
    mmx.set_alpha(sim, lambda i, j, k, dx, dy, dz: 0.01 if i < 10 else 0.99)
    alpha = np.asarray(sim.driver.alpha)
    assert alpha[8] == 0.01 and alpha[9] == 0.99
    with pytest.raises(mmx.MmxError):
        mmx.add_thermal_noise(sim, 300.0)


@pytest.mark.parametrize("form", ["slot", "list"])
@pytest.mark.parametrize("integrator", ["RungeKutta", "DormandPrince"])
def test_spatial_alpha_trajectory(oracle, integrator, form):
    """Per-cell damping (spatial_llg_rhs_kernel!, llg.jl:62-97) with demag
    against the oracle, through the slot-form kernels and (a kind of term
    added twice) through the term-list kernels."""
    import mmx

    note(f"spatial alpha {integrator} {form}")
    nx, ny, nz = 12, 9, 3
    N = nx * ny * nz
    terms = [("exch", dict(A=1.3e-11)), ("dmi", dict(D=2e-3, type="interfacial")), ("demag", {})]
    if form == "list":
        terms += [("zeeman", dict(H=(1e4, 0.0, 2e4))), ("zeeman", dict(H=(0.0, -3e4, 0.0)))]
    g, o = build(oracle, nx, ny, nz, pbc="y", Ms=np.full(N, 6e5), terms=terms, seed=21)
    alphas = np.random.default_rng(22).uniform(0.01, 0.8, N)
    if integrator == "RungeKutta":
        mmx.set_driver(g, driver="SpatialLLG", integrator=integrator, step=2e-14)
    else:
        mmx.set_driver(g, driver="SpatialLLG", integrator=integrator, tol=1e-6)
    mmx.set_alpha(g, alphas)
    o.set_alpha(alphas)
    assert relerr(mmx.llg_rhs(g), o.llg_rhs()) <= TOL
    if integrator == "RungeKutta":
        o.set_integrator(0, 2e-14, 0)
        mmx.run_step(g, n=15)
        for _ in range(15):
            o.advance_step()
        assert np.max(np.abs(g.spin - o.spin)) < 1e-12
    else:
        o.set_integrator(2, 0, 1e-6)
        for i in range(1, 4):
            mmx.run_until(g, i * 3e-13)
            o.run_until(i * 3e-13)
        assert g._integrator.nsteps == o.state()["nsteps"]
        assert np.max(np.abs(g.spin - o.spin)) < 1e-10
