"""Batching form: many independent systems of uneven size in one launch per
stage (members with demag: one launch per pass and stage for each geometry
group) must land exactly where each system advanced alone lands, and the
single-system path is itself pinned to the oracle (test_gpu_parity)."""
import numpy as np
import pytest

import mmx
from oracle import OracleSim, RK4

pytestmark = pytest.mark.gpu

MU0 = 4e-7 * np.pi


def _member(rng, integrator, precision=mmx.F64, thermal=False, params=None):
    nx, ny, nz = (int(v) for v in rng.integers([5, 4, 1], [41, 33, 6]))
    p = params or dict(Ms=rng.uniform(4e5, 1.2e6), A=rng.uniform(5e-12, 2e-11), D=rng.uniform(0, 4e-3),
                       Ku=rng.uniform(0, 1e6))
    pbc = "xy" if rng.uniform() < 0.5 else ""
    mesh = mmx.FDMesh(dx=2e-9, dy=2e-9, dz=2e-9, nx=nx, ny=ny, nz=nz, pbc=pbc)
    sim = mmx.Sim(mesh, integrator=integrator, precision=precision)
    N = nx * ny * nz
    Ms = np.full(N, p["Ms"])
    Ms[rng.integers(0, N)] = 0.0
    mmx.set_Ms(sim, Ms)
    m0 = rng.standard_normal((N, 3)) + np.array([0, 0, 2.0])
    mmx.init_m0(sim, m0.reshape(-1))
    mmx.add_exch(sim, p["A"])
    mmx.add_dmi(sim, p["D"], type="interfacial")
    mmx.add_anis(sim, p["Ku"], axis=(0, 0, 1))
    mmx.add_zeeman(sim, (0, 0, 0.05 / MU0))
    if thermal:
        mmx.add_thermal_noise(sim, 300.0, seed=int(rng.integers(1, 1 << 30)))
    mmx.set_driver(sim, integrator=integrator, step=1e-14 if integrator == "Heun" else 2e-14)
    return sim, dict(mesh=(nx, ny, nz), pbc=pbc, Ms=Ms, m0=m0.reshape(-1), params=p)


def _build(seed, n, integrator, precision=mmx.F64, thermal=False):
    rng = np.random.default_rng(seed)
    return [_member(rng, integrator, precision, thermal) for _ in range(n)]


@pytest.mark.parametrize("integrator,thermal", [("RungeKutta", False), ("Heun", True)])
def test_batch_matches_members_alone(integrator, thermal):
    batch_members = _build(7, 9, integrator, thermal=thermal)
    alone = _build(7, 9, integrator, thermal=thermal)
    b = mmx.Batch([s for s, _ in batch_members])
    cpt, grid, items, _ = b.info()
    shape = b.shape()
    print(f"[batch] tuned: {shape['block_threads']} threads/block, {cpt} cells/thread, grid {grid}, {items} work items",
          flush=True)
    assert shape["block_threads"] in (128, 256, 512) and shape["cells_per_thread"] == cpt
    assert shape["demag_groups"] == 0
    for n in (7, 4, 1):  # odd and even counts, several calls
        b.advance(n)
        for s, _ in alone:
            mmx.run_step(s, n=n)
        for (sb, _), (sa, _) in zip(batch_members, alone):
            assert np.array_equal(sb.spin, sa.spin)
            assert sb._integrator.nsteps == sa._integrator.nsteps
            assert sb._integrator.t == sa._integrator.t


def test_batch_member_steps_alone_in_between():
    mem = _build(8, 4, "RungeKutta")
    ref = _build(8, 4, "RungeKutta")
    b = mmx.Batch([s for s, _ in mem])
    b.advance(3)
    mmx.run_step(mem[1][0], n=1)  # one member moves on its own (odd number of steps)
    b.advance(2)
    for i, (s, _) in enumerate(ref):
        mmx.run_step(s, n=6 if i == 1 else 5)
    for (sb, _), (sa, _) in zip(mem, ref):
        assert np.array_equal(sb.spin, sa.spin)


def test_batch_member_against_oracle():
    mem = _build(9, 3, "RungeKutta")
    b = mmx.Batch([s for s, _ in mem])
    b.advance(10)
    s, c = mem[2]
    nx, ny, nz = c["mesh"]
    p = c["params"]
    o = OracleSim(nx, ny, nz, 2e-9, 2e-9, 2e-9, pbc=c["pbc"])
    o.set_Ms(c["Ms"])
    o.init_m0(c["m0"])
    o.add_exch(p["A"])
    o.add_dmi(np.full(nx * ny * nz, p["D"]), type="interfacial")
    o.add_anis(np.full(nx * ny * nz, p["Ku"]), axis=(0, 0, 1))
    o.add_zeeman((0, 0, 0.05 / MU0))
    o.set_llg(alpha=0.1, gamma=2.21e5, precession=True)
    o.set_integrator(RK4, step=2e-14)
    for _ in range(10):
        o.advance_step()
    assert np.abs(s.spin - o.spin).max() < 1e-12


def test_batch_fp32_matches_members_alone():
    mem = _build(10, 5, "RungeKutta", precision=mmx.F32)
    ref = _build(10, 5, "RungeKutta", precision=mmx.F32)
    b = mmx.Batch([s for s, _ in mem])
    b.advance(6)
    for s, _ in ref:
        mmx.run_step(s, n=6)
    for (sb, _), (sa, _) in zip(mem, ref):
        assert np.array_equal(sb.spin, sa.spin)


def test_batch_rejects_bad_members():
    rng = np.random.default_rng(3)
    s1, _ = _member(rng, "RungeKutta")
    s2, _ = _member(rng, "Heun")
    with pytest.raises(mmx.MmxError):
        mmx.Batch([s1, s2])
    with pytest.raises(mmx.MmxError):
        mmx.Batch([s1, s1])


def _demag_member(rng, integrator, mesh, pbc, images, thermal=False, twice=False):
    """A member with demag (grouped in the batch by mesh, cells, periodicity
    and images); twice: a second Zeeman term, which takes the term-list kernels."""
    nx, ny, nz = mesh
    p = dict(Ms=rng.uniform(4e5, 1.2e6), A=rng.uniform(5e-12, 2e-11), D=rng.uniform(0, 4e-3),
             Ku=rng.uniform(0, 1e6))
    m = mmx.FDMesh(dx=2e-9, dy=2e-9, dz=2e-9, nx=nx, ny=ny, nz=nz, pbc=pbc)
    sim = mmx.Sim(m, integrator=integrator)
    N = nx * ny * nz
    Ms = np.full(N, p["Ms"])
    Ms[rng.integers(0, N)] = 0.0
    mmx.set_Ms(sim, Ms)
    m0 = rng.standard_normal((N, 3)) + np.array([0, 0, 2.0])
    mmx.init_m0(sim, m0.reshape(-1))
    mmx.add_exch(sim, p["A"])
    mmx.add_dmi(sim, p["D"], type="interfacial")
    mmx.add_anis(sim, p["Ku"], axis=(0, 0, 1))
    mmx.add_zeeman(sim, (0, 0, 0.05 / MU0))
    mmx.add_demag(sim, Nx=images[0], Ny=images[1], Nz=images[2])
    if twice:
        mmx.add_zeeman(sim, (0.01 / MU0, 0, 0))
    if thermal:
        mmx.add_thermal_noise(sim, 300.0, seed=int(rng.integers(1, 1 << 30)))
    mmx.set_driver(sim, integrator=integrator, step=1e-14 if integrator == "Heun" else 2e-14)
    return sim, dict(mesh=mesh, pbc=pbc, images=images, Ms=Ms, m0=m0.reshape(-1), params=p, twice=twice)


def _mixed(seed, integrator, thermal):
    rng = np.random.default_rng(seed)
    out = [_demag_member(rng, integrator, (24, 20, 4), "xy", (1, 1, 0), thermal) for _ in range(3)]
    out += [_member(rng, integrator, thermal=thermal) for _ in range(2)]
    out += [_demag_member(rng, integrator, (16, 12, 2), "", (0, 0, 0), thermal, twice=(q == 1)) for q in range(2)]
    out += [_demag_member(rng, integrator, (24, 20, 4), "xy", (2, 2, 0), thermal)]
    return out


@pytest.mark.parametrize("integrator,thermal", [("RungeKutta", False), ("Heun", True)])
def test_batch_demag_groups_match_members_alone(integrator, thermal):
    """Members with demag in three geometry groups (one of them with a
    term-list member, one with other macro-PBC images) and two members
    without: every member lands exactly where it lands advanced alone."""
    mem = _mixed(11, integrator, thermal)
    ref = _mixed(11, integrator, thermal)
    b = mmx.Batch([s for s, _ in mem])
    assert b.shape()["demag_groups"] == 3
    for n in (5, 2, 1):
        b.advance(n)
        for s, _ in ref:
            mmx.run_step(s, n=n)
        for i, ((sb, _), (sa, _)) in enumerate(zip(mem, ref)):
            assert np.array_equal(sb.spin, sa.spin), f"member {i}"
            assert sb._integrator.nsteps == sa._integrator.nsteps


def test_batch_demag_member_against_oracle():
    mem = _mixed(12, "RungeKutta", False)
    b = mmx.Batch([s for s, _ in mem])
    b.advance(10)
    s, c = mem[1]  # second member of the first geometry group
    nx, ny, nz = c["mesh"]
    p = c["params"]
    o = OracleSim(nx, ny, nz, 2e-9, 2e-9, 2e-9, pbc=c["pbc"])
    o.set_Ms(c["Ms"])
    o.init_m0(c["m0"])
    o.add_exch(p["A"])
    o.add_dmi(np.full(nx * ny * nz, p["D"]), type="interfacial")
    o.add_anis(np.full(nx * ny * nz, p["Ku"]), axis=(0, 0, 1))
    o.add_zeeman((0, 0, 0.05 / MU0))
    o.add_demag(*c["images"])
    o.set_llg(alpha=0.1, gamma=2.21e5, precession=True)
    o.set_integrator(RK4, step=2e-14)
    for _ in range(10):
        o.advance_step()
    assert np.abs(s.spin - o.spin).max() < 1e-12


def test_batch_with_constant_spin_torques():
    """Members carrying SOT / Slonczewski / Zhang-Li terms (constant currents)
    take the term-list kernels in the batch and still land exactly where each
    member advanced alone lands; a time profile is refused."""

    def members(seed):
        rng = np.random.default_rng(seed)
        out = []
        for q in range(4):
            s, info = _member(rng, "RungeKutta")
            N = s.n_total
            if q % 3 == 0:
                mmx.add_sot(s, rng.uniform(1e9, 2e9, N), 3e8, (0, 1, 0))
            elif q % 3 == 1:
                mmx.add_stt(s, model="slonczewski", J=1e12, tf=2e-9, Ms=8e5, p=(0, 0, 1), P=0.6, xi=0.1)
            else:
                mmx.add_stt(s, model="zhang_li", b=5.0, J=(1.0, 0.5, 0), xi=0.05)
            out.append(s)
        return out

    bm, alone = members(11), members(11)
    b = mmx.Batch(bm)
    b.advance(5)
    for s in alone:
        mmx.run_step(s, n=5)
    for sb, sa in zip(bm, alone):
        assert np.array_equal(sb.spin, sa.spin)
    s = members(12)[2]
    mmx._set_time_fn(s, s.interactions[-1], lambda t: 2.0)
    with pytest.raises(mmx.MmxError):
        mmx.Batch([s])


def _cfg_member(rng, precision):
    """A stochastic-LLG member with the batching workload's ranges: mesh between
    32x32x4 and 192x192x16, Ms in [4e5, 1.2e6], A in [5e-12, 2e-11], D in
    [0, 4e-3], Ku in [0, 1e6], 300 K, alpha 0.1, Heun at 1e-14 s."""
    nx, ny, nz = (int(v) for v in rng.integers([32, 32, 4], [193, 193, 17]))
    p = dict(Ms=rng.uniform(4e5, 1.2e6), A=rng.uniform(5e-12, 2e-11), D=rng.uniform(0, 4e-3),
             Ku=rng.uniform(0, 1e6), seed=int(rng.integers(1, 1 << 30)))
    mesh = mmx.FDMesh(dx=2e-9, dy=2e-9, dz=2e-9, nx=nx, ny=ny, nz=nz, pbc="xy")
    s = mmx.Sim(mesh, integrator="Heun", precision=precision)
    N = nx * ny * nz
    mmx.set_Ms(s, p["Ms"])
    m0 = rng.standard_normal((N, 3)) + np.array([0, 0, 1.0])
    mmx.init_m0(s, m0.reshape(-1))
    mmx.add_exch(s, p["A"])
    mmx.add_dmi(s, p["D"], type="interfacial")
    mmx.add_anis(s, p["Ku"], axis=(0, 0, 1))
    mmx.add_thermal_noise(s, 300.0, seed=p["seed"])
    s.driver.alpha = 0.1
    mmx.set_driver(s, integrator="Heun", step=1e-14)
    return s, dict(mesh=(nx, ny, nz), m0=m0.reshape(-1), params=p)


@pytest.mark.parametrize("precision", [mmx.F64, mmx.F32])
def test_batch_workload_sizes(precision):
    """Members drawn from the batching workload's size range (up to 192x192x16)
    land exactly where each member advanced alone lands; the smallest FP64
    member follows the oracle's stochastic Heun steps."""
    print(f"[case] batch of workload-sized sLLG members FP{precision}", flush=True)
    mem = [_cfg_member(np.random.default_rng(40 + q), precision) for q in range(6)]
    ref = [_cfg_member(np.random.default_rng(40 + q), precision) for q in range(6)]
    b = mmx.Batch([s for s, _ in mem])
    b.advance(4)
    for s, _ in ref:
        mmx.run_step(s, n=4)
    for (sb, _), (sa, _) in zip(mem, ref):
        assert np.array_equal(sb.spin, sa.spin)
    if precision != mmx.F64:
        return
    s, c = min(mem, key=lambda x: x[0].n_total)
    nx, ny, nz = c["mesh"]
    p = c["params"]
    N = nx * ny * nz
    o = OracleSim(nx, ny, nz, 2e-9, 2e-9, 2e-9, pbc="xy")
    o.set_Ms(np.full(N, p["Ms"]))
    o.init_m0(c["m0"])
    o.add_exch(p["A"])
    o.add_dmi(np.full(N, p["D"]), type="interfacial")
    o.add_anis(np.full(N, p["Ku"]), axis=(0, 0, 1))
    o.add_thermal(np.full(N, 300.0), 0.0, p["seed"])
    o.set_llg(alpha=0.1, gamma=2.21e5, precession=True)
    o.set_integrator(1, step=1e-14)
    for _ in range(4):
        o.advance_step()
    err = np.abs(s.spin - o.spin).max()
    print(f"member {nx}x{ny}x{nz} vs oracle after 4 stochastic Heun steps: {err:.2e}", flush=True)
    assert err < 1e-12
