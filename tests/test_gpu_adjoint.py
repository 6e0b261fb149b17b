"""Differentiation form: adjoint RK4 gradients on the GPU against central
finite differences of the CPU oracle's forward trajectory.

The reference has no trajectory gradient (its only derivative is Enzyme's
forward-mode Jacobian, ext/EnzymeExt.jl:6-59, which cannot run without Julia),
so the gradients are pinned by finite differences of the oracle, which follows
the reference's RK4 (src/integrator/rk.jl) and field kernels.
"""
import numpy as np
import pytest

import mmx
from oracle import OracleSim, RK4

pytestmark = pytest.mark.gpu

MU0 = 4e-7 * np.pi


def _smooth(rng, shape, mean, rel):
    g = rng.standard_normal(shape)
    # two passes of a 3-point box filter along x and y for some correlation
    for ax in (0, 1):
        g = (np.roll(g, 1, ax) + g + np.roll(g, -1, ax)) / 3
    g /= g.std()
    return mean * (1 + rel * g)


def _case_a(seed=11):
    """interfacial DMI(x,y) + Ku(x,y) + exchange + Zeeman + demag, pins, an empty cell, pbc x"""
    nx, ny, nz = 10, 8, 3
    rng = np.random.default_rng(seed)
    N = nx * ny * nz
    D = _smooth(rng, (nz, ny, nx), 3e-3, 0.2).reshape(-1)
    Ku = _smooth(rng, (nz, ny, nx), 6e5, 0.2).reshape(-1)
    Ms = np.full(N, 5.8e5)
    Ms[17] = 0.0
    pins = np.zeros(N, dtype=np.uint8)
    pins[[3, 40]] = 1
    m0 = (np.array([0.3, 0.2, 1.0]) + 0.5 * rng.standard_normal((N, 3)))
    m0 /= np.linalg.norm(m0, axis=1, keepdims=True)
    target = rng.uniform(-1, 1, N)
    window = np.zeros(N, dtype=bool)
    window.reshape(nz, ny, nx)[:, 2:6, 3:8] = True
    return dict(mesh=(nx, ny, nz), pbc="x", d=2e-9, Ms=Ms, pins=pins, m0=m0.reshape(-1), D=D, Ku=Ku,
                A=1.5e-11, H=(0.0, 0.0, 0.1 / MU0), demag=True, cubic=None, precession=True, alpha=0.3,
                target=target, window=window, dt=5e-14, nsteps=24, every=5)


def _case_b(seed=12):
    """spatial exchange A(x,y) + cubic Kc(x,y) + uniform interfacial DMI + Zeeman, no demag, precession off"""
    nx, ny, nz = 9, 7, 2
    rng = np.random.default_rng(seed)
    N = nx * ny * nz
    A = _smooth(rng, (nz, ny, nx), 1.3e-11, 0.2).reshape(-1)
    Kc = _smooth(rng, (nz, ny, nx), -5e4, 0.2).reshape(-1)
    Ms = np.full(N, 8e5)
    pins = np.zeros(N, dtype=np.uint8)
    m0 = (np.array([1.0, 0.2, 0.4]) + 0.6 * rng.standard_normal((N, 3)))
    m0 /= np.linalg.norm(m0, axis=1, keepdims=True)
    target = rng.uniform(-1, 1, N)
    window = np.ones(N, dtype=bool)
    return dict(mesh=(nx, ny, nz), pbc="", d=2.5e-9, Ms=Ms, pins=pins, m0=m0.reshape(-1), A=A, Kc=Kc,
                D=np.full(N, 2e-3), Ku=None, H=(2e4, 0.0, 5e4), demag=False, cubic=True, precession=False,
                alpha=0.5, target=target, window=window, dt=1e-13, nsteps=15, every=4)


def _case_c(seed=13):
    """spatial Ms(x,y) + vector bulk DMI D(x,y) + exchange + Ku + Zeeman + demag, an empty cell, pbc y"""
    nx, ny, nz = 9, 8, 3
    rng = np.random.default_rng(seed)
    N = nx * ny * nz
    Ms = _smooth(rng, (nz, ny, nx), 6e5, 0.2).reshape(-1)
    Ms[31] = 0.0
    Dv = tuple(_smooth(rng, (nz, ny, nx), 2e-3, 0.25).reshape(-1) for _ in range(3))
    m0 = (np.array([0.2, 0.1, 1.0]) + 0.5 * rng.standard_normal((N, 3)))
    m0 /= np.linalg.norm(m0, axis=1, keepdims=True)
    target = rng.uniform(-1, 1, N)
    window = np.zeros(N, dtype=bool)
    window.reshape(nz, ny, nx)[:, 2:7, 2:7] = True
    return dict(mesh=(nx, ny, nz), pbc="y", d=2e-9, Ms=Ms, pins=np.zeros(N, dtype=np.uint8), m0=m0.reshape(-1),
                Dbulk=Dv, Ku=np.full(N, 3e5), A=1.2e-11, H=(0.0, 2e4, 0.05 / MU0), demag=True, cubic=None,
                precession=True, alpha=0.2, target=target, window=window, dt=5e-14, nsteps=20, every=4)


def _case_d(seed=14):
    """uniform bulk DMI + exchange + cubic, no demag: the uniform D's gradient is the sum of the cells' shares"""
    nx, ny, nz = 8, 6, 2
    rng = np.random.default_rng(seed)
    N = nx * ny * nz
    m0 = (np.array([1.0, 0.3, 0.2]) + 0.6 * rng.standard_normal((N, 3)))
    m0 /= np.linalg.norm(m0, axis=1, keepdims=True)
    target = rng.uniform(-1, 1, N)
    return dict(mesh=(nx, ny, nz), pbc="", d=2e-9, Ms=np.full(N, 7e5), pins=np.zeros(N, dtype=np.uint8),
                m0=m0.reshape(-1), Dbulk=(2.5e-3, 2.5e-3, 2.5e-3), A=1.3e-11, Kc=np.full(N, -4e4), H=(0.0, 0.0, 3e4),
                demag=False, cubic=None, precession=True, alpha=0.3, target=target, window=np.ones(N, dtype=bool),
                dt=5e-14, nsteps=16, every=5)


def _loss(c, mT):
    mz = mT.reshape(-1, 3)[:, 2]
    w = c["window"]
    return float(np.sum((mz[w] - c["target"][w]) ** 2) / w.sum())


def _dloss(c, mT):
    mz = mT.reshape(-1, 3)[:, 2]
    w = c["window"]
    g = np.zeros_like(mT).reshape(-1, 3)
    g[w, 2] = 2 * (mz[w] - c["target"][w]) / w.sum()
    return g.reshape(-1)


def _oracle_run(c, m0=None, D=None, Ku=None, A=None, Kc=None, Ms=None, Dbulk=None):
    nx, ny, nz = c["mesh"]
    d = c["d"]
    o = OracleSim(nx, ny, nz, d, d, d, pbc=c["pbc"])
    o.set_Ms(c["Ms"] if Ms is None else Ms)
    o.set_pins(c["pins"])
    o.init_m0(c["m0"] if m0 is None else m0, norm=False)
    if "Dbulk" in c:
        o.add_exch(c["A"])
        o.add_dmi(c["Dbulk"] if Dbulk is None else Dbulk, type="bulk")
        if c.get("Ku") is not None:
            o.add_anis(c["Ku"], axis=(0, 0, 1))
        if c.get("Kc") is not None:
            o.add_cubic_anis(c["Kc"])
        o.add_zeeman(c["H"])
        if c["demag"]:
            o.add_demag()
    elif c["cubic"]:
        o.add_exch(c["A"] if A is None else A)
        o.add_cubic_anis(c["Kc"] if Kc is None else Kc)
        o.add_dmi(c["D"] if D is None else D, type="interfacial")
        o.add_zeeman(c["H"])
    else:
        o.add_exch(c["A"])
        o.add_dmi(c["D"] if D is None else D, type="interfacial")
        o.add_anis(c["Ku"] if Ku is None else Ku, axis=(0, 0, 1))
        o.add_zeeman(c["H"])
        if c["demag"]:
            o.add_demag()
    o.set_llg(alpha=c["alpha"], gamma=2.21e5, precession=c["precession"])
    o.set_integrator(RK4, step=c["dt"])
    for _ in range(c["nsteps"]):
        o.advance_step()
    return _loss(c, o.spin)


def _gpu_grads(c, precision=mmx.F64):
    nx, ny, nz = c["mesh"]
    d = c["d"]
    mesh = mmx.FDMesh(dx=d, dy=d, dz=d, nx=nx, ny=ny, nz=nz, pbc=c["pbc"])
    sim = mmx.Sim(mesh, integrator="RungeKutta", precision=precision)
    mmx.set_Ms(sim, c["Ms"])
    mmx.set_pinning(sim, c["pins"].astype(bool))
    mmx.init_m0(sim, c["m0"], norm=False)
    names = {}
    if "Dbulk" in c:
        mmx.add_exch(sim, c["A"])
        names["Dbulk"] = mmx.add_dmi(sim, c["Dbulk"], type="bulk").name
        if c.get("Ku") is not None:
            mmx.add_anis(sim, c["Ku"], axis=(0, 0, 1))
        if c.get("Kc") is not None:
            mmx.add_cubic_anis(sim, c["Kc"])
        mmx.add_zeeman(sim, c["H"])
        if c["demag"]:
            mmx.add_demag(sim)
    elif c["cubic"]:
        names["A"] = mmx.add_exch(sim, c["A"]).name
        names["Kc"] = mmx.add_cubic_anis(sim, c["Kc"]).name
        names["D"] = mmx.add_dmi(sim, 2e-3, type="interfacial").name
        mmx.add_zeeman(sim, c["H"])
    else:
        mmx.add_exch(sim, c["A"])
        names["D"] = mmx.add_dmi(sim, c["D"], type="interfacial").name
        names["Ku"] = mmx.add_anis(sim, c["Ku"], axis=(0, 0, 1)).name
        mmx.add_zeeman(sim, c["H"])
        if c["demag"]:
            mmx.add_demag(sim)
    sim.driver.alpha = c["alpha"]
    sim.driver.precession = c["precession"]
    mmx.set_driver(sim, integrator="RungeKutta", step=c["dt"])
    mT, pullback = mmx.rk4_trajectory_rrule(sim, c["nsteps"], c["every"])
    g = pullback(_dloss(c, mT))
    return sim, mT, {k: g[v] for k, v in names.items()} | {"m0": g["m0"], "Ms": g["Ms"]}


def _fd_check(c, grads, key, idx, rel_eps):
    base = {"D": c["D"], "Ku": c.get("Ku"), "A": c.get("A"), "Kc": c.get("Kc"), "m0": c["m0"]}[key]
    errs = []
    scale = np.abs(grads[key]).max()
    for i in idx:
        eps = rel_eps * (abs(base[i]) if key != "m0" else 1.0)
        p, q = base.copy(), base.copy()
        p[i] += eps
        q[i] -= eps
        fd = (_oracle_run(c, **{key: p}) - _oracle_run(c, **{key: q})) / (2 * eps)
        errs.append(abs(fd - grads[key][i]) / scale)
        print(f"[adj] {key}[{i}] adjoint {grads[key][i]:+.9e} fd {fd:+.9e}", flush=True)
    return max(errs)


def test_adjoint_ms_and_bulk_dmi_vs_oracle_fd():
    """dL/dMs per cell and dL/dD of a vector bulk DMI field (three components per
    cell) against central differences of the oracle's trajectory; the empty
    cell's gradients are zero."""
    c = _case_c()
    print("[case] adjoint Ms and vector bulk DMI 9x8x3", flush=True)
    sim, mT, grads = _gpu_grads(c)
    assert abs(_loss(c, mT) - _oracle_run(c)) < 1e-12
    rng = np.random.default_rng(98)
    live = np.flatnonzero(c["Ms"] > 0)
    scale = np.abs(grads["Ms"]).max()
    errs = []
    for i in rng.choice(live, 6, replace=False):
        eps = 1e-5 * c["Ms"][i]
        p, q = c["Ms"].copy(), c["Ms"].copy()
        p[i] += eps
        q[i] -= eps
        fd = (_oracle_run(c, Ms=p) - _oracle_run(c, Ms=q)) / (2 * eps)
        print(f"[adj] Ms[{i}] adjoint {grads['Ms'][i]:+.9e} fd {fd:+.9e}", flush=True)
        errs.append(abs(fd - grads["Ms"][i]) / scale)
    assert max(errs) < 1e-6, errs
    gD = grads["Dbulk"].reshape(-1, 3)
    scale = np.abs(gD).max()
    errs = []
    for i in rng.choice(live, 4, replace=False):
        for comp in range(3):
            base = c["Dbulk"][comp]
            eps = 1e-5 * abs(base[i])
            p = tuple(x.copy() for x in c["Dbulk"])
            q = tuple(x.copy() for x in c["Dbulk"])
            p[comp][i] += eps
            q[comp][i] -= eps
            fd = (_oracle_run(c, Dbulk=p) - _oracle_run(c, Dbulk=q)) / (2 * eps)
            print(f"[adj] D{'xyz'[comp]}[{i}] adjoint {gD[i, comp]:+.9e} fd {fd:+.9e}", flush=True)
            errs.append(abs(fd - gD[i, comp]) / scale)
    assert max(errs) < 1e-6, errs
    assert grads["Ms"][31] == 0 and np.all(gD[31] == 0)


def test_adjoint_uniform_bulk_dmi_share_sums_to_fd():
    """A uniform bulk D (one number for all three components): the sum of the
    cells' shares of the three components is dL/dD of the oracle."""
    c = _case_d()
    print("[case] adjoint uniform bulk DMI 8x6x2", flush=True)
    sim, mT, grads = _gpu_grads(c)
    g = grads["Dbulk"].sum()
    D0 = c["Dbulk"][0]
    eps = 1e-5 * D0
    fd = (_oracle_run(c, Dbulk=D0 + eps) - _oracle_run(c, Dbulk=D0 - eps)) / (2 * eps)
    print(f"[adj] uniform D adjoint {g:+.9e} fd {fd:+.9e}", flush=True)
    assert abs(g - fd) <= 1e-6 * abs(fd)
    # Ms gradient without demag
    i = 17
    eps = 1e-5 * c["Ms"][i]
    p, q = c["Ms"].copy(), c["Ms"].copy()
    p[i] += eps
    q[i] -= eps
    fd = (_oracle_run(c, Ms=p) - _oracle_run(c, Ms=q)) / (2 * eps)
    assert abs(grads["Ms"][i] - fd) <= 1e-6 * np.abs(grads["Ms"]).max()


@pytest.mark.parametrize("which", ["a", "b"])
def test_adjoint_vs_oracle_fd(which):
    c = _case_a() if which == "a" else _case_b()
    sim, mT, grads = _gpu_grads(c)
    rng = np.random.default_rng(99)
    N = c["Ms"].size
    keys = ["D", "Ku", "m0"] if which == "a" else ["A", "Kc", "m0"]
    for key in keys:
        if key == "m0":
            live = np.flatnonzero((c["Ms"] > 0).repeat(3))
            idx = rng.choice(live, 6, replace=False)
            err = _fd_check(c, grads, key, idx, 1e-6)
        else:
            idx = rng.choice(np.flatnonzero(c["Ms"] > 0), 6, replace=False)
            err = _fd_check(c, grads, key, idx, 1e-5)
        assert err < 1e-6, (key, err)
    # the forward trajectory itself agrees with the oracle
    assert abs(_loss(c, mT) - _oracle_run(c)) < 1e-12


def test_adjoint_pins_and_empty_cells():
    c = _case_a()
    _, _, g = _gpu_grads(c)
    g0 = g["m0"].reshape(-1, 3)
    # an empty cell has no influence and no parameter gradient
    assert np.all(g0[17] == 0) and g["Ku"][17] == 0 and g["D"][17] == 0
    # pinned cells still steer their neighbours: their m0 gradient is nonzero
    assert np.abs(g0[3]).max() > 0


def test_adjoint_backward_twice_is_identical():
    c = _case_a()
    sim, mT, _ = _gpu_grads(c)
    L = mmx._lib.load()
    lam = _dloss(c, mT)
    g1 = np.empty_like(lam)
    g2 = np.empty_like(lam)
    mmx._lib.check(L.mmx_adjoint_backward(sim._h, mmx._lib.dptr(lam), mmx._lib.dptr(g1)))
    mmx._lib.check(L.mmx_adjoint_backward(sim._h, mmx._lib.dptr(lam), mmx._lib.dptr(g2)))
    assert np.array_equal(g1, g2)


def test_adjoint_fp32_against_fp64():
    c = _case_a()
    _, _, g64 = _gpu_grads(c, mmx.F64)
    _, _, g32 = _gpu_grads(c, mmx.F32)
    for k in ("D", "Ku", "m0"):
        err = np.abs(g32[k] - g64[k]).max() / np.abs(g64[k]).max()
        print(f"[adj] FP32 {k} max-norm relative error {err:.2e}", flush=True)
        assert err < 2e-3, (k, err)


def test_adjoint_rejects_other_integrators():
    mesh = mmx.FDMesh(dx=1e-9, dy=1e-9, dz=1e-9, nx=4, ny=4, nz=1)
    sim = mmx.Sim(mesh, integrator="DormandPrince")
    mmx.set_Ms(sim, 8e5)
    mmx.init_m0(sim, (1, 0, 0))
    mmx.add_exch(sim, 1e-11)
    with pytest.raises(mmx.MmxError):
        mmx.rk4_trajectory_rrule(sim, 3, 1)
