#!/usr/bin/env python3
"""Throughput of the fused field + LLG + RK4 path on one MI355X.

Workload (the metric's configuration): a 3-D chiral film on a 256x256x32 mesh of
2 nm cells, Ms 5.8e5 A/m, A 1.5e-11 J/m, interfacial D 3e-3 J/m^2, uniaxial Ku
8e5 J/m^3 along z, Zeeman (0, 0, 0.1 T), full FFT demag with 5x5 macro-PBC
images in x/y, periodic x/y stencils, seeded random unit vectors as m0, RK4 at
dt = 1e-14 s.  One step = one RK4 step = 4 field + LLG evaluations.
value = Mcell-evals/s (cells x evaluations / s) with m resident in HBM.

Also reported: per-pass kernel times from HIP events on the library's stream,
the roofline of the dominant kernel under the canonical byte model, the FP32
form, and the CPU oracle's rate on the host cores (a bounded sample).
"""
import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "micromagnetic.jl_amd"))

HBM_PEAK_GBS = 8000.0
MU0 = 4e-7 * np.pi
# The differentiation leg stays off until the adjoint pass of the term-list
# kernels at 128x128x8 is cleared (see ADJOINT_SKIP_REASON).
ADJOINT_ENABLED = False
# end time of the 200x50x1 soft film's CPU sample (the whole run is 1 ns)
SOFT_SAMPLE_T = 0.25e-9
ADJOINT_SKIP_REASON = ("the 128x128x8 adjoint leg faulted the GPU in the term-list last pass; it stays off "
                       "until that fault has a shown cause")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_system(mmx, nx, ny, nz, precision, images=5, seed=2002, dt=1e-14, demag=True):
    mesh = mmx.FDMesh(dx=2e-9, dy=2e-9, dz=2e-9, nx=nx, ny=ny, nz=nz, pbc="xy")
    sim = mmx.Sim(mesh, integrator="RungeKutta", precision=precision)
    rng = np.random.default_rng(seed)
    N = nx * ny * nz
    z = rng.uniform(-1, 1, N)
    phi = rng.uniform(0, 2 * np.pi, N)
    r = np.sqrt(1 - z * z)
    m0 = np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1).reshape(-1)
    mmx.set_Ms(sim, 5.8e5)
    mmx.init_m0(sim, m0)
    mmx.add_exch(sim, 1.5e-11)
    mmx.add_dmi(sim, 3e-3, type="interfacial")
    mmx.add_anis(sim, 8e5, axis=(0, 0, 1))
    mmx.add_zeeman(sim, (0, 0, 0.1 / MU0))
    if demag:
        mmx.add_demag(sim, Nx=images, Ny=images, Nz=0)
    mmx.set_driver(sim, integrator="RungeKutta", step=dt)
    return sim, m0


def byte_model(nx, ny, nz, Px, Py, Pz, w, p_arrays, chained=False):
    """Canonical algorithmic bytes of one RHS evaluation, per pass.

    S1 = Kx*ny*nz, S2 = Kx*Py*nz complex values per component; the k-space
    tensor is read folded (6 * Kx * (Py/2+1) * (Pz/2+1) reals); m is read by the
    first and the last pass; p_arrays per-cell parameter arrays of w bytes plus
    a 1-byte pin mask are read by the last pass, which writes 3N results.

    chained: an evaluation inside a chain of stages, whose first pass is
    skipped because the previous last pass already wrote its x-spectrum; the
    last pass then also writes the next spectrum (3 * 2w * S1).  The total
    leaves P1 out; its entry stays for the chain's first evaluation."""
    Kx = Px // 2 + 1
    N = nx * ny * nz
    S1 = Kx * ny * nz
    S2 = Kx * Py * nz
    c = 2 * w
    tensor = 6 * w * Kx * (Py // 2 + 1) * (Pz // 2 + 1)
    passes = {
        "P1": 3 * w * N + w * N + 3 * c * S1,
        "P2": 3 * c * S1 + 3 * c * S2,
        "P3": 3 * c * S2 + tensor + 3 * c * S2,
        "P4": 3 * c * S2 + 3 * c * S1,
        "P5": 3 * c * S1 + 3 * w * N + (p_arrays - 1) * w * N + N + 3 * w * N,
    }
    if chained:
        passes["P5"] += 3 * c * S1
        return passes, sum(v for k, v in passes.items() if k != "P1")
    return passes, sum(passes.values())


def rk4_stage_extra(N, w):
    """Stage traffic of one fused RK4 step beyond 4 evaluations (30 w N)."""
    return 30 * w * N


def time_steps(sim, mmx, steps, warmup):
    L = mmx._lib.load()
    h = sim._h
    if warmup:
        mmx._lib.check(L.mmx_advance_step(h, warmup))
    mmx._lib.check(L.mmx_synchronize(h))
    t0 = time.perf_counter()
    mmx._lib.check(L.mmx_advance_step(h, steps))  # ends with a stream synchronize
    return time.perf_counter() - t0


def profile_passes(sim, mmx, steps):
    import ctypes as C

    L = mmx._lib.load()
    mmx._lib.check(L.mmx_profile(sim._h, 1))
    mmx._lib.check(L.mmx_advance_step(sim._h, steps))
    ms = (C.c_double * 6)()
    n = (C.c_int64 * 6)()
    mmx._lib.check(L.mmx_profile_read(sim._h, ms, n))
    mmx._lib.check(L.mmx_profile(sim._h, 0))
    return {name: (ms[i] / n[i] if n[i] else None, n[i]) for i, name in enumerate(mmx._lib.PASSES)}


def _oracle_film(O, nx, ny, nz, m0, images):
    N = nx * ny * nz
    s = O.OracleSim(nx, ny, nz, 2e-9, 2e-9, 2e-9, pbc="xy")
    s.set_Ms(np.full(N, 5.8e5))
    s.init_m0(m0)
    s.add_exch(1.5e-11)
    s.add_dmi(np.full(N, 3e-3), type="interfacial")
    s.add_anis(np.full(N, 8e5), axis=(0, 0, 1))
    s.add_zeeman(np.tile([0, 0, 0.1 / MU0], N))
    s.add_demag(images, images, 0)
    s.set_llg(0.1, 2.21e5, True)
    return s


def cfg1_m0(nx=200, ny=50, seed=1001):
    """Initial state of the 200x50x1 soft film: normalise((1, 0.25, 0.1) + 0.05 U(-1,1)^3)."""
    rng = np.random.default_rng(seed)
    m = np.array([1.0, 0.25, 0.1]) + 0.05 * rng.uniform(-1, 1, (nx * ny, 3))
    m /= np.linalg.norm(m, axis=1, keepdims=True)
    return m.reshape(-1)


def _median_time(fn, reps, warmup=1):
    for _ in range(warmup):
        fn()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)), ts


def cpu_baseline(nx, ny, nz, m0, images, reps=5, soft_t_end=SOFT_SAMPLE_T):
    """The CPU oracle (a C restatement of the reference CPU path, FP64) on the
    host cores: the 256x256x32 film's field + LLG evaluation (median of `reps`
    after one warm-up) with OpenMP over all threads and with one thread, and
    the 200x50x1 soft film's Dormand-Prince run from 0 to soft_t_end (1 ns is
    the whole run; the default is a bounded sample of it)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    O.build()
    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    N = nx * ny * nz
    t0 = time.perf_counter()
    s = _oracle_film(O, nx, ny, nz, m0, images)
    t_setup = time.perf_counter() - t0
    cores = O.set_threads(threads)
    t_all, _ = _median_time(s.llg_rhs, reps)
    log(f"[bench] CPU eval ({cores} threads): {t_all:.3f} s")
    O.set_threads(1)
    t_one, _ = _median_time(s.llg_rhs, reps)
    log(f"[bench] CPU eval (1 thread): {t_one:.3f} s")
    del s
    # 200x50x1 soft film, exchange + demag, DOPRI54 tol 1e-6, 0 -> 1 ns
    O.set_threads(threads)
    n1 = 200 * 50
    o = O.OracleSim(200, 50, 1, 2.5e-9, 2.5e-9, 3e-9)
    o.set_Ms(np.full(n1, 8e5))
    o.init_m0(cfg1_m0())
    o.add_exch(1.3e-11)
    o.add_demag(0, 0, 0)
    o.set_llg(0.5, 2.211e5, True)
    o.set_integrator(O.DOPRI54, 0.0, 1e-6)
    t0 = time.perf_counter()
    o.run_until(soft_t_end)
    t_cfg1 = time.perf_counter() - t0
    st1 = o.state()
    log(f"[bench] CPU 200x50x1 DOPRI54 0 -> {soft_t_end:.3g} s: {t_cfg1:.2f} s, {st1['nsteps']} steps, "
        f"{st1['nfevals']} evaluations")
    return {"value": round(N / t_all / 1e6, 3), "unit": "Mcell-evals/s", "cores": cores, "kind": "port",
            "sample": f"{reps} field+LLG evaluations of the {nx}x{ny}x{nz} film after one warm-up (median "
                      f"{t_all:.3f} s each; demag set-up {t_setup:.1f} s not timed), FP64, OpenMP {cores} threads",
            "single_thread": {"value": round(N / t_one / 1e6, 3), "unit": "Mcell-evals/s", "cores": 1,
                              "median_s": round(t_one, 4)},
            "soft_film_200x50x1": {"t_end": soft_t_end, "seconds": round(t_cfg1, 3), "steps": st1["nsteps"],
                                       "evaluations": st1["nfevals"], "cores": cores,
                                       "Mcell_evals_per_s": round(n1 * st1["nfevals"] / t_cfg1 / 1e6, 3),
                                       "m_final_mean": [float(v) for v in o.spin.reshape(-1, 3).mean(axis=0)]}}


def grf(rng, nx, ny, dx, corr, mean, rel):
    """Seeded Gaussian random field on an nx x ny grid: white noise filtered with
    a Gaussian of correlation length `corr`, scaled to unit variance, then
    mean * (1 + rel * g)."""
    g = rng.standard_normal((ny, nx))
    kx = np.fft.fftfreq(nx, d=dx) * 2 * np.pi
    ky = np.fft.fftfreq(ny, d=dx) * 2 * np.pi
    K2 = ky[:, None] ** 2 + kx[None, :] ** 2
    g = np.real(np.fft.ifft2(np.fft.fft2(g) * np.exp(-K2 * corr * corr / 4)))
    g = (g - g.mean()) / g.std()
    return mean * (1 + rel * g)


def ensemble_specs(n_members=48, seed=4004):
    """Member meshes and parameters of the stochastic-LLG ensemble, drawn from
    a seed: meshes 32..192 x 32..192 x 4..16, Ms, A, D, Ku in the ranges of the
    batching workload."""
    rng = np.random.default_rng(seed)
    specs = []
    for q in range(n_members):
        nx, ny = (int(v) for v in rng.integers(32, 193, 2))
        nz = int(rng.integers(4, 17))
        specs.append(dict(nx=nx, ny=ny, nz=nz, Ms=rng.uniform(4e5, 1.2e6), A=rng.uniform(5e-12, 2e-11),
                          D=rng.uniform(0, 4e-3), Ku=rng.uniform(0, 1e6), seed=int(rng.integers(1, 1 << 30))))
    return specs


def ensemble_sims(mmx, precision, specs):
    """One Sim per member: Heun dt 1e-14, 300 K seeded thermal field, alpha 0.1."""
    sims = []
    for sp in specs:
        mesh = mmx.FDMesh(dx=2e-9, dy=2e-9, dz=2e-9, nx=sp["nx"], ny=sp["ny"], nz=sp["nz"], pbc="xy")
        s = mmx.Sim(mesh, integrator="Heun", precision=precision)
        mmx.set_Ms(s, sp["Ms"])
        mmx.init_m0(s, (0, 0, 1))
        mmx.add_exch(s, sp["A"])
        mmx.add_dmi(s, sp["D"], type="interfacial")
        mmx.add_anis(s, sp["Ku"], axis=(0, 0, 1))
        mmx.add_thermal_noise(s, 300.0, seed=sp["seed"])
        mmx.set_driver(s, integrator="Heun", step=1e-14)
        sims.append(s)
    return sims


def bench_batch(mmx, precision, steps, warmup, n_members=48, seed=4004):
    """Batching form: stochastic-LLG ensemble members of uneven size (Heun,
    300 K, alpha 0.1, seeded Philox noise) advanced by one launch per stage over
    all members, against the same members advanced one Sim at a time."""
    specs = ensemble_specs(n_members, seed)

    def make():
        return ensemble_sims(mmx, precision, specs)

    cells = sum(sp["nx"] * sp["ny"] * sp["nz"] for sp in specs)
    L = mmx._lib.load()
    sims = make()
    b = mmx.Batch(sims)
    b.advance(warmup)
    t0 = time.perf_counter()
    b.advance(steps)
    tb = time.perf_counter() - t0
    cpt, grid, items, tune = b.info()
    shape = b.shape()
    batched_spins = [s.spin for s in sims[:4]]
    del b, sims
    sims = make()
    for s in sims:
        mmx.run_step(s, n=warmup)
    for s in sims:
        mmx._lib.check(L.mmx_synchronize(s._h))
    t0 = time.perf_counter()
    for s in sims:
        mmx._lib.check(L.mmx_advance_step(s._h, steps))  # synchronous per member
    tp = time.perf_counter() - t0
    same = all(np.array_equal(a, s.spin) for a, s in zip(batched_spins, sims[:4]))
    del sims
    ev = cells * 2 * steps
    return {"workload": f"{n_members} sLLG members, meshes 32..192 x 32..192 x 4..16 (seed {seed}), "
                        f"{cells} cells, Heun dt 1e-14, 300 K",
            "value": round(ev / tb / 1e6, 1), "unit": "Mcell-evals/s",
            "plain_value": round(ev / tp / 1e6, 1), "speedup_vs_plain": round(tp / tb, 3),
            "identical_to_plain": same, "tuned": {"block_threads": shape["block_threads"], "cells_per_thread": cpt,
                                                  "grid": grid, "work_items": items}}


def bench_neb(mmx, precision, steps=20, warmup=3, chains=2, frames=22, seed=6006):
    """Batching form, NEB half: chains of 24 images (22 free) on a 128x128x4
    film between a seeded skyrmion and the uniform state, stepped by the NEB
    driver (BS23, tol 1e-5); the field of every free image is one launch per
    step, the band's tangent, force and stage updates one launch each per stage."""
    negs = neb_chains(mmx, precision, chains, frames, seed)
    N = 128 * 128 * 4
    for _, nb in negs:
        nb.advance(warmup)
    t0 = time.perf_counter()
    for _, nb in negs:
        nb.advance(steps)  # synchronous: one host sync per BS23 try
    t = time.perf_counter() - t0
    n_img = negs[0][1].n_images
    st = [nb.state for _, nb in negs]
    work = chains * n_img * N * steps
    return {"workload": f"{chains} NEB chains of {n_img + 2} images on a 128x128x4 film (seed {seed}), "
                        f"exch+iDMI+Ku+Zeeman, BS23 tol 1e-5",
            "value": round(work / t / 1e6, 1), "unit": "Mcell-image-steps/s",
            "ms_per_chain_step": round(t / (chains * steps) * 1e3, 4),
            "fevals_per_step": round(sum(x["nfevals"] for x in st) / sum(x["nsteps"] for x in st), 3)}


def bench_batch_demag(mmx, precision, steps=50, warmup=3, per_group=8, seed=4104):
    """Batching form with demag members: three geometry groups of per_group
    members each (128x128x4 periodic x/y with 1x1 images, 96x96x8 open,
    64x64x16 periodic x/y), every member with its own Ms, A, D, Ku, exchange +
    interfacial DMI + anisotropy + demag, Heun dt 1e-14 at 300 K; against the
    same members advanced one Sim at a time."""
    geoms = [((128, 128, 4), "xy", (1, 1, 0)), ((96, 96, 8), "", (0, 0, 0)), ((64, 64, 16), "xy", (0, 0, 0))]
    rng = np.random.default_rng(seed)
    specs = [dict(geom=g, Ms=rng.uniform(4e5, 1.2e6), A=rng.uniform(5e-12, 2e-11), D=rng.uniform(0, 4e-3),
                  Ku=rng.uniform(0, 1e6), seed=int(rng.integers(1, 1 << 30))) for g in geoms for _ in range(per_group)]

    def make():
        sims = []
        for sp in specs:
            (nx, ny, nz), pbc, img = sp["geom"]
            mesh = mmx.FDMesh(dx=2e-9, dy=2e-9, dz=2e-9, nx=nx, ny=ny, nz=nz, pbc=pbc)
            s = mmx.Sim(mesh, integrator="Heun", precision=precision)
            mmx.set_Ms(s, sp["Ms"])
            mmx.init_m0(s, (0, 0, 1))
            mmx.add_exch(s, sp["A"])
            mmx.add_dmi(s, sp["D"], type="interfacial")
            mmx.add_anis(s, sp["Ku"], axis=(0, 0, 1))
            mmx.add_demag(s, Nx=img[0], Ny=img[1], Nz=img[2])
            mmx.add_thermal_noise(s, 300.0, seed=sp["seed"])
            mmx.set_driver(s, integrator="Heun", step=1e-14)
            sims.append(s)
        return sims

    cells = sum(int(np.prod(sp["geom"][0])) for sp in specs)
    L = mmx._lib.load()
    sims = make()
    b = mmx.Batch(sims)
    b.advance(warmup)
    t0 = time.perf_counter()
    b.advance(steps)
    tb = time.perf_counter() - t0
    batched_spins = [s.spin for s in sims[::per_group]]
    del b, sims
    sims = make()
    for s in sims:
        mmx.run_step(s, n=warmup)
    for s in sims:
        mmx._lib.check(L.mmx_synchronize(s._h))
    t0 = time.perf_counter()
    for s in sims:
        mmx._lib.check(L.mmx_advance_step(s._h, steps))
    tp = time.perf_counter() - t0
    same = all(np.array_equal(a, s.spin) for a, s in zip(batched_spins, sims[::per_group]))
    del sims
    ev = cells * 2 * steps
    return {"workload": f"{len(specs)} members with demag in 3 geometry groups (128x128x4 pbc xy 1x1 images, "
                        f"96x96x8, 64x64x16 pbc xy), {cells} cells, Heun dt 1e-14, 300 K",
            "value": round(ev / tb / 1e6, 1), "unit": "Mcell-evals/s",
            "plain_value": round(ev / tp / 1e6, 1), "speedup_vs_plain": round(tp / tb, 3),
            "identical_to_plain": same}


def neb_chains(mmx, precision, chains=2, frames=22, seed=6006):
    """NEB chains of frames + 2 images on a 128x128x4 film (2 nm cells, periodic
    x/y): exchange, interfacial DMI, anisotropy along z and a 0.1 T field, each
    chain with its own seeded parameters, between a seeded skyrmion and the
    uniform state."""
    rng = np.random.default_rng(seed)
    nx, ny, nz, d = 128, 128, 4, 2e-9
    negs = []
    for _ in range(chains):
        mesh = mmx.FDMesh(dx=d, dy=d, dz=d, nx=nx, ny=ny, nz=nz, pbc="xy")
        s = mmx.Sim(mesh, precision=precision)
        mmx.set_Ms(s, rng.uniform(4e5, 1.2e6))
        mmx.add_exch(s, rng.uniform(5e-12, 2e-11))
        mmx.add_dmi(s, rng.uniform(1e-3, 4e-3), type="interfacial")
        mmx.add_anis(s, rng.uniform(2e5, 1e6), axis=(0, 0, 1))
        mmx.add_zeeman(s, (0, 0, 0.1 / mmx.mu_0))
        c = rng.uniform(-20e-9, 20e-9, 2)
        R = rng.uniform(6e-9, 12e-9)
        x = (np.arange(nx) - (nx - 1) / 2) * d - c[0]
        y = (np.arange(ny) - (ny - 1) / 2) * d - c[1]
        X, Y = np.meshgrid(x, y)
        th = np.pi * np.exp(-(X * X + Y * Y) / (R * R))
        ph = np.arctan2(Y, X) + np.pi / 2
        left = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], axis=-1)
        left = np.tile(left.reshape(-1, 3), (nz, 1)).reshape(-1)
        negs.append((s, mmx.NEB(s, [left, (0.0, 0.0, 1.0)], [frames])))
    return negs


def bench_all96(mmx, precision, ens_steps=100, neb_steps=40, warmup=3):
    """The whole batching workload: the 48 ensemble members and the 2 NEB
    chains of 24 images (44 free) at once.  The ensemble advances with one launch
    per stage over all members (a HIP graph on its own stream) while each
    chain steps its band on its own stream; the three are driven from three
    host threads (the library calls release the GIL), so the chains' small
    launches and host-side BS23 accept/reject fill the device around the
    ensemble's.  Reported: field+LLG cell evaluations of all 96 systems per
    second, against the same work run back to back."""
    import threading

    specs = ensemble_specs()
    ens_cells = sum(sp["nx"] * sp["ny"] * sp["nz"] for sp in specs)
    N = 128 * 128 * 4

    def setup():
        sims = ensemble_sims(mmx, precision, specs)
        b = mmx.Batch(sims)
        negs = neb_chains(mmx, precision)
        b.advance(warmup)
        for _, nb in negs:
            nb.advance(warmup)
        return sims, b, negs

    def fevals(negs):
        return [nb.state["nfevals"] for _, nb in negs]

    def work(negs, f0):
        n_img = negs[0][1].n_images
        neb = sum((f1 - a) * n_img * N for f1, a in zip(fevals(negs), f0))
        return ens_cells * 2 * ens_steps + neb

    # back to back
    sims, b, negs = setup()
    f0 = fevals(negs)
    t0 = time.perf_counter()
    b.advance(ens_steps)
    for _, nb in negs:
        nb.advance(neb_steps)
    t_seq = time.perf_counter() - t0
    w_seq = work(negs, f0)
    del b, negs, sims
    # concurrent
    sims, b, negs = setup()
    f0 = fevals(negs)
    errs = []

    def run(fn):
        try:
            fn()
        except Exception as e:  # re-raised below
            errs.append(e)

    th = [threading.Thread(target=run, args=(lambda: b.advance(ens_steps),))]
    th += [threading.Thread(target=run, args=(lambda nb=nb: nb.advance(neb_steps),)) for _, nb in negs]
    t0 = time.perf_counter()
    for x in th:
        x.start()
    for x in th:
        x.join()
    t_con = time.perf_counter() - t0
    if errs:
        raise errs[0]
    w_con = work(negs, f0)
    del b, negs, sims
    return {"workload": f"96 systems: 48 sLLG members ({ens_cells} cells, Heun dt 1e-14, 300 K, {ens_steps} steps) "
                        f"+ 2 NEB chains of 24 images on 128x128x4 ({neb_steps} BS23 steps each)",
            "value": round(w_con / t_con / 1e6, 1), "unit": "Mcell-evals/s",
            "back_to_back_value": round(w_seq / t_seq / 1e6, 1),
            "seconds": round(t_con, 4), "back_to_back_seconds": round(t_seq, 4)}


def bench_adjoint(mmx, precision, nsteps=2000, every=50, n_fd=8, seed=5005):
    """Differentiation form on the 128x128x8 film: Ku(x,y) and D(x,y) Gaussian
    random fields (means 6e5 J/m^3 and 2.5e-3 J/m^2, 20 % std, 10 nm
    correlation), exchange + interfacial DMI + anisotropy + demag, RK4 dt 5e-14
    for nsteps with a checkpoint every `every` steps; loss = mean over the
    centred 32x32 window (all layers) of (m_z(T) - target)^2.  Gradients with
    respect to Ku, D and m0 from the adjoint, checked against central finite
    differences of the same GPU forward path on n_fd cells per quantity."""
    nx, ny, nz, d = 128, 128, 8, 2e-9
    N = nx * ny * nz
    rng = np.random.default_rng(seed)
    Ku2 = grf(rng, nx, ny, d, 10e-9, 6e5, 0.2)
    D2 = grf(np.random.default_rng(seed + 1), nx, ny, d, 10e-9, 2.5e-3, 0.2)
    Ku = np.tile(Ku2.reshape(-1), nz)
    D = np.tile(D2.reshape(-1), nz)
    target = np.random.default_rng(seed + 2).uniform(-1, 1, N)
    win = np.zeros((nz, ny, nx), dtype=bool)
    win[:, ny // 2 - 16:ny // 2 + 16, nx // 2 - 16:nx // 2 + 16] = True
    win = win.reshape(-1)
    r = np.random.default_rng(seed + 4)
    m0 = np.array([0.1, 0.1, 1.0]) + 0.3 * r.standard_normal((N, 3))
    m0 = (m0 / np.linalg.norm(m0, axis=1, keepdims=True)).reshape(-1)

    mesh = mmx.FDMesh(dx=d, dy=d, dz=d, nx=nx, ny=ny, nz=nz)
    sim = mmx.Sim(mesh, integrator="RungeKutta", precision=precision)
    mmx.set_Ms(sim, 5.8e5)
    mmx.init_m0(sim, m0, norm=False)
    mmx.add_exch(sim, 1.5e-11)
    tD = mmx.add_dmi(sim, D, type="interfacial", name="D")
    tK = mmx.add_anis(sim, Ku, axis=(0, 0, 1), name="Ku")
    mmx.add_demag(sim)
    mmx.set_driver(sim, integrator="RungeKutta", step=5e-14)
    L = mmx._lib.load()

    def loss(mT):
        mz = mT.reshape(-1, 3)[:, 2]
        return float(np.sum((mz[win] - target[win]) ** 2) / win.sum())

    def reset(m=m0, Ku_=Ku, D_=D):
        mmx.init_m0(sim, m, norm=False)
        mmx.set_term_param(sim, tK, Ku_)
        mmx.set_term_param(sim, tD, D_)
        mmx.set_driver(sim, integrator="RungeKutta", step=5e-14)

    # plain forward for reference
    reset()
    mmx._lib.check(L.mmx_synchronize(sim._h))
    t0 = time.perf_counter()
    mmx._lib.check(L.mmx_advance_step(sim._h, nsteps))
    t_plain = time.perf_counter() - t0
    reset()
    t0 = time.perf_counter()
    mT, pullback = mmx.rk4_trajectory_rrule(sim, nsteps, every)
    t_fwd = time.perf_counter() - t0
    lam = np.zeros(3 * N)
    lam.reshape(-1, 3)[win, 2] = 2 * (mT.reshape(-1, 3)[win, 2] - target[win]) / win.sum()
    t0 = time.perf_counter()
    g = pullback(lam)
    t_bwd = time.perf_counter() - t0
    out = {"workload": f"128x128x8 film, 2 nm, exch+iDMI(x,y)+Ku(x,y)+demag, RK4 dt 5e-14, {nsteps} steps, "
                       f"checkpoint every {every}",
           "loss": loss(mT), "forward_record_s": round(t_fwd, 4), "backward_s": round(t_bwd, 4),
           "plain_forward_s": round(t_plain, 4),
           "value": round(N * 4 * nsteps / (t_fwd + t_bwd) / 1e6, 1), "unit": "Mcell-evals/s of the differentiated trajectory",
           "plain_value": round(N * 4 * nsteps / t_plain / 1e6, 1)}
    if n_fd:
        pr = np.random.default_rng(seed + 3)
        errs = {}
        for key in ("Ku", "D", "m0"):
            scale = np.abs(g[key]).max()
            worst = 0.0
            for i in pr.choice(3 * N if key == "m0" else N, n_fd, replace=False):
                if key == "m0":
                    eps = 1e-4
                    p_, q_ = m0.copy(), m0.copy()
                    p_[i] += eps
                    q_[i] -= eps
                    args = [dict(m=p_), dict(m=q_)]
                else:
                    base = Ku if key == "Ku" else D
                    eps = 1e-4 * abs(base[i])
                    p_, q_ = base.copy(), base.copy()
                    p_[i] += eps
                    q_[i] -= eps
                    kw = "Ku_" if key == "Ku" else "D_"
                    args = [{kw: p_}, {kw: q_}]
                vals = []
                for a in args:
                    reset(**a)
                    mmx._lib.check(L.mmx_advance_step(sim._h, nsteps))
                    vals.append(loss(sim.spin))
                fd = (vals[0] - vals[1]) / (2 * eps)
                worst = max(worst, abs(fd - g[key][i]) / scale)
            errs[key] = worst
        out["fd_cells_per_quantity"] = n_fd
        out["fd_max_rel_err"] = {k: float(f"{v:.3e}") for k, v in errs.items()}
    out["grads"] = g
    return out


def bench_diff_forward(mmx, precision, nsteps=2000, warmup=20, seed=5005):
    """Forward trajectory of the differentiation workload alone: the same
    128x128x8 film as bench_adjoint (spatial Ku and D, exchange + interfacial
    DMI + anisotropy + demag, RK4 dt 5e-14), nsteps plain RK4 steps with no
    recording and no backward pass."""
    nx, ny, nz, d = 128, 128, 8, 2e-9
    N = nx * ny * nz
    rng = np.random.default_rng(seed)
    Ku = np.tile(grf(rng, nx, ny, d, 10e-9, 6e5, 0.2).reshape(-1), nz)
    D = np.tile(grf(np.random.default_rng(seed + 1), nx, ny, d, 10e-9, 2.5e-3, 0.2).reshape(-1), nz)
    r = np.random.default_rng(seed + 4)
    m0 = np.array([0.1, 0.1, 1.0]) + 0.3 * r.standard_normal((N, 3))
    m0 = (m0 / np.linalg.norm(m0, axis=1, keepdims=True)).reshape(-1)
    mesh = mmx.FDMesh(dx=d, dy=d, dz=d, nx=nx, ny=ny, nz=nz)
    sim = mmx.Sim(mesh, integrator="RungeKutta", precision=precision)
    mmx.set_Ms(sim, 5.8e5)
    mmx.init_m0(sim, m0, norm=False)
    mmx.add_exch(sim, 1.5e-11)
    mmx.add_dmi(sim, D, type="interfacial", name="D")
    mmx.add_anis(sim, Ku, axis=(0, 0, 1), name="Ku")
    mmx.add_demag(sim)
    mmx.set_driver(sim, integrator="RungeKutta", step=5e-14)
    L = mmx._lib.load()
    mmx._lib.check(L.mmx_advance_step(sim._h, warmup))
    mmx._lib.check(L.mmx_synchronize(sim._h))
    t0 = time.perf_counter()
    mmx._lib.check(L.mmx_advance_step(sim._h, nsteps))
    mmx._lib.check(L.mmx_synchronize(sim._h))
    t = time.perf_counter() - t0
    mz = float(sim.spin.reshape(-1, 3)[:, 2].mean())
    return {"workload": f"128x128x8 film, 2 nm, exch+iDMI(x,y)+Ku(x,y)+demag, RK4 dt 5e-14, {nsteps} steps "
                        f"(forward only, no recording)",
            "value": round(N * 4 * nsteps / t / 1e6, 1), "unit": "Mcell-evals/s",
            "ms_per_step": round(t / nsteps * 1e3, 4), "mean_mz_end": round(mz, 6)}


def read_traffic(kernel_pass):
    """HBM bytes per launch of a pass from the committed PMC summary, if any."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.md")
    if not os.path.exists(path):
        return None
    # | pass | kernel(s) | launches | traffic (B/launch) | ...
    for line in open(path):
        m = re.match(r"\|\s*" + kernel_pass + r"\s*\|[^|]*\|\s*\d+\s*\|\s*([0-9.eE+]+)\s*\|", line)
        if m:
            return float(m.group(1))
    return None


def copy_ceiling(mmx, nbytes=2 << 30, reps=20):
    """Practical HBM ceiling: the library's streaming copy kernel (read + write
    of nbytes) at a few grid sizes; the best rate counts."""
    import ctypes as C

    L = mmx._lib.load()
    best = None
    for grid in (2048, 4096, 8192, 16384):
        ms = C.c_double()
        mmx._lib.check(L.mmx_diag_copy(0, nbytes, reps, grid, C.byref(ms)))
        gbs = 2 * nbytes / (ms.value * 1e-3) / 1e9
        if best is None or gbs > best["GBps"]:
            best = {"GBps": round(gbs, 1), "grid": grid, "ms": round(ms.value, 4), "bytes_copied": nbytes}
    return best


def rocfft_compare(nx, ny, nz, reps=20):
    """The reference's demag evaluation with rocFFT (torch.fft on ROCm), unpruned
    as src/micro/demag.jl:102-130 does it: fill the 2n-padded arrays (zeros +
    mu0 Ms m in the corner), 3 real-to-complex 3-D transforms, the 6-component
    real k-space tensor product, 3 complex-to-real transforms, and the collect
    of the corner; FP64, random data of the 256x256x32 shape."""
    import torch

    pad = lambda n: 2 * n if n > 3 else 2 * n - 1
    Px, Py, Pz = pad(nx), pad(ny), pad(nz)
    dev = "cuda"
    g = torch.Generator(device=dev)
    g.manual_seed(7)
    m = torch.randn(3, nz, ny, nx, dtype=torch.float64, device=dev, generator=g)
    ms = torch.rand(nz, ny, nx, dtype=torch.float64, device=dev, generator=g)
    Nk = torch.randn(6, Pz, Py, Px // 2 + 1, dtype=torch.float64, device=dev, generator=g)
    buf = torch.empty(3, Pz, Py, Px, dtype=torch.float64, device=dev)
    out = torch.empty(3, nz, ny, nx, dtype=torch.float64, device=dev)

    def once():
        buf.zero_()
        buf[:, :nz, :ny, :nx] = m * ms
        F = torch.fft.rfftn(buf, dim=(1, 2, 3))
        Hx = Nk[0] * F[0] + Nk[3] * F[1] + Nk[4] * F[2]
        Hy = Nk[3] * F[0] + Nk[1] * F[1] + Nk[5] * F[2]
        Hz = Nk[4] * F[0] + Nk[5] * F[1] + Nk[2] * F[2]
        h = torch.fft.irfftn(torch.stack([Hx, Hy, Hz]), s=(Pz, Py, Px), dim=(1, 2, 3))
        out.copy_(h[:, :nz, :ny, :nx])

    for _ in range(3):
        once()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        once()
    e1.record()
    torch.cuda.synchronize()
    ms_eval = e0.elapsed_time(e1) / reps
    del buf, Nk, m, ms, out
    torch.cuda.empty_cache()
    return {"ms_per_demag_eval": round(ms_eval, 4), "padded": [Px, Py, Pz], "library": "rocFFT via torch.fft",
            "what": "fill + 3 rfftn + 6-component tensor product + 3 irfftn + collect, FP64"}


def rocfft_child(mesh):
    """rocfft_compare in a child process: torch brings its own HIP runtime,
    which does not share the device with the library's in one process."""
    import subprocess
    import tempfile

    fd, path = tempfile.mkstemp(suffix=".json", prefix="mmx_rocfft_")
    os.close(fd)
    try:
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--rocfft-child", path, "--mesh", mesh],
                           timeout=300)
        if r.returncode != 0:
            raise RuntimeError(f"rocFFT child exited with status {r.returncode}")
        with open(path) as f:
            return json.load(f)
    finally:
        if os.path.exists(path):
            os.unlink(path)


def gpu_soft_film(mmx, t_end=1e-9):
    """The 200x50x1 soft film (exchange + demag, alpha 0.5, gamma 2.211e5) run
    0 -> t_end with Dormand-Prince, tol 1e-6, on the GPU."""
    mesh = mmx.FDMesh(dx=2.5e-9, dy=2.5e-9, dz=3e-9, nx=200, ny=50, nz=1)
    sim = mmx.Sim(mesh, integrator="DormandPrince")
    mmx.set_Ms(sim, 8e5)
    mmx.init_m0(sim, cfg1_m0())
    mmx.add_exch(sim, 1.3e-11)
    mmx.add_demag(sim)
    sim.driver.alpha = 0.5
    sim.driver.gamma = 2.211e5
    mmx.set_driver(sim, integrator="DormandPrince", tol=1e-6)
    t0 = time.perf_counter()
    mmx.run_until(sim, t_end)
    t = time.perf_counter() - t0
    import ctypes as C

    tt, step, sn, ns, nf = C.c_double(), C.c_double(), C.c_double(), C.c_int64(), C.c_int64()
    mmx._lib.check(mmx._lib.load().mmx_get_state(sim._h, C.byref(tt), C.byref(step), C.byref(sn), C.byref(ns),
                                                 C.byref(nf)))
    return {"t_end": t_end, "seconds": round(t, 4), "steps": ns.value, "evaluations": nf.value,
            "m_final_mean": [float(v) for v in sim.spin.reshape(-1, 3).mean(axis=0)]}


def run_forms(args, out_path):
    """The batching and differentiation forms (run in a child process so that a
    failure there cannot take the headline with it)."""
    import mmx

    res = {}

    def leg(name, fn):
        try:
            res[name] = fn()
        except Exception as e:  # recorded, and the child exits non-zero
            res[name] = {"error": f"{type(e).__name__}: {e}"}
            log(f"[bench] {name} failed: {e}")

    log("[bench] batching form (FP64, FP32)")
    leg("batch_fp64", lambda: bench_batch(mmx, mmx.F64, steps=100, warmup=5))
    leg("batch_fp32", lambda: bench_batch(mmx, mmx.F32, steps=100, warmup=5))
    for k in ("batch_fp64", "batch_fp32"):
        if "value" in res.get(k, {}):
            log(f"[bench] {k}: {res[k]['value']} Mcell-evals/s vs plain {res[k]['plain_value']}")
    leg("neb_fp64", lambda: bench_neb(mmx, mmx.F64))
    leg("neb_fp32", lambda: bench_neb(mmx, mmx.F32))
    for k in ("neb_fp64", "neb_fp32"):
        if "value" in res.get(k, {}):
            log(f"[bench] {k}: {res[k]['value']} {res[k]['unit']}, {res[k]['ms_per_chain_step']} ms per chain step")
    leg("batch_demag_fp64", lambda: bench_batch_demag(mmx, mmx.F64))
    leg("batch_demag_fp32", lambda: bench_batch_demag(mmx, mmx.F32))
    for k in ("batch_demag_fp64", "batch_demag_fp32"):
        if "value" in res.get(k, {}):
            log(f"[bench] {k}: {res[k]['value']} Mcell-evals/s vs plain {res[k]['plain_value']} "
                f"(identical: {res[k]['identical_to_plain']})")
    leg("all96_fp64", lambda: bench_all96(mmx, mmx.F64))
    leg("all96_fp32", lambda: bench_all96(mmx, mmx.F32))
    for k in ("all96_fp64", "all96_fp32"):
        if "value" in res.get(k, {}):
            log(f"[bench] {k}: {res[k]['value']} Mcell-evals/s over the 96 systems "
                f"(back to back {res[k]['back_to_back_value']})")
    leg("diff_forward_fp64", lambda: bench_diff_forward(mmx, mmx.F64))
    leg("diff_forward_fp32", lambda: bench_diff_forward(mmx, mmx.F32))
    for k in ("diff_forward_fp64", "diff_forward_fp32"):
        if "value" in res.get(k, {}):
            log(f"[bench] {k}: {res[k]['value']} Mcell-evals/s, {res[k]['ms_per_step']} ms per RK4 step "
                f"(128x128x8 differentiation film, forward only)")
    if ADJOINT_ENABLED:
        log("[bench] differentiation form (FP64 with finite-difference probes, FP32)")
        leg("adjoint_fp64", lambda: bench_adjoint(mmx, mmx.F64, n_fd=args.fd_cells))
        leg("adjoint_fp32", lambda: bench_adjoint(mmx, mmx.F32, n_fd=0))
        a64, a32 = res.get("adjoint_fp64", {}), res.get("adjoint_fp32", {})
        if "grads" in a64 and "grads" in a32:
            a32["grad_rel_err_vs_fp64"] = {
                k: float(f"{np.abs(a32['grads'][k] - a64['grads'][k]).max() / np.abs(a64['grads'][k]).max():.3e}")
                for k in a64["grads"]}
        for a in (a64, a32):
            a.pop("grads", None)
    else:
        res["adjoint"] = {"skipped": ADJOINT_SKIP_REASON}
    with open(out_path, "w") as f:
        json.dump(res, f)
    return 1 if any("error" in v for v in res.values() if isinstance(v, dict)) else 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--mesh", type=str, default="256x256x32")
    ap.add_argument("--images", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-forms", action="store_true", help="skip the batching and differentiation forms")
    ap.add_argument("--no-extras", action="store_true", help="skip the copy ceiling, rocFFT and soft-film legs")
    ap.add_argument("--local-only", action="store_true",
                    help="time the same film without its demag term (local-field kernel only) and exit")
    ap.add_argument("--term-list", action="store_true",
                    help="time the same film through the term-list kernels (a second, zero Zeeman term: a kind "
                         "that appears twice takes the term list) and exit")
    ap.add_argument("--diff-film", action="store_true",
                    help="time the forward RK4 steps of the 128x128x8 differentiation film (spatial Ku and D, "
                         "demag; FP64 and FP32) and exit")
    ap.add_argument("--fd-cells", type=int, default=64, help="finite-difference probes per gradient quantity")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run the library's streaming copy of 1 GiB three times (calibrates HBM counters)")
    ap.add_argument("--forms-child", type=str, default=None, help=argparse.SUPPRESS)
    ap.add_argument("--rocfft-child", type=str, default=None, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported by this build "
              "(the path is single-GPU; multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)

    if args.rocfft_child:  # torch's HIP runtime in a process of its own (libmmx not loaded)
        nx, ny, nz = (int(v) for v in args.mesh.split("x"))
        with open(args.rocfft_child, "w") as f:
            json.dump(rocfft_compare(nx, ny, nz), f)
        return

    import mmx

    if args.forms_child:
        sys.exit(run_forms(args, args.forms_child))
    if args.copy_probe:
        import ctypes as C

        ms = C.c_double()
        mmx._lib.check(mmx._lib.load().mmx_diag_copy(0, 1 << 30, 3, 8192, C.byref(ms)))
        log(f"[bench] copy probe: 1 GiB copied 4 times, {2 * (1 << 30) / (ms.value * 1e-3) / 1e9:.1f} GB/s")
        return

    nx, ny, nz = (int(v) for v in args.mesh.split("x"))
    N = nx * ny * nz
    if args.local_only:
        sim, _ = build_system(mmx, nx, ny, nz, mmx.F64, demag=False)
        el = time_steps(sim, mmx, args.steps, args.warmup)
        passes = profile_passes(sim, mmx, max(5, min(args.steps, 20)))
        ms = passes["LOCAL"][0]
        log(f"[bench] local-only FP64: {N * 4 * args.steps / el / 1e6:.1f} Mcell-evals/s, "
            f"local kernel {ms * 1e3:.1f} us per launch")
        return
    if args.term_list:
        sim, _ = build_system(mmx, nx, ny, nz, mmx.F64, images=args.images)
        mmx.add_zeeman(sim, (0.0, 0.0, 0.0))  # a second term of one kind: the term-list kernels
        el = time_steps(sim, mmx, args.steps, args.warmup)
        passes = profile_passes(sim, mmx, max(5, min(args.steps, 20)))
        log(f"[bench] term-list FP64: {N * 4 * args.steps / el / 1e6:.1f} Mcell-evals/s, "
            f"{el / args.steps * 1e3:.3f} ms/step; " +
            ", ".join(f"{k} {v[0] * 1e3:.1f}us" for k, v in passes.items() if v[0] is not None))
        return
    if args.diff_film:
        for prec, name in ((mmx.F64, "FP64"), (mmx.F32, "FP32")):
            r = bench_diff_forward(mmx, prec, nsteps=args.steps, warmup=args.warmup)
            log(f"[bench] differentiation film {name}: {r['value']} Mcell-evals/s, {r['ms_per_step']} ms per RK4 "
                f"step ({r['workload']})")
        return
    errors = {}
    log(f"[bench] FP64 set-up {nx}x{ny}x{nz}, images {args.images}")
    t0 = time.perf_counter()
    sim, m0 = build_system(mmx, nx, ny, nz, mmx.F64, images=args.images)
    log(f"[bench] set-up {time.perf_counter() - t0:.1f} s; warmup {args.warmup} + timed {args.steps} RK4 steps")
    elapsed = time_steps(sim, mmx, args.steps, args.warmup)
    evals = 4 * args.steps
    value = N * evals / elapsed / 1e6
    ms_step = elapsed / args.steps * 1e3
    log(f"[bench] FP64: {value:.1f} Mcell-evals/s, {ms_step:.3f} ms/step")

    # per-pass times from HIP events on the library stream (a second timed run)
    prof_steps = max(5, min(args.steps, 20))
    passes = profile_passes(sim, mmx, prof_steps)
    import ctypes as C

    Px, Py, Pz, KS = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
    mmx._lib.check(mmx._lib.load().mmx_demag_geometry(sim._h, C.byref(Px), C.byref(Py), C.byref(Pz), C.byref(KS)))
    # every evaluation of a timed call but the first is chained (no first pass)
    pb, b_eval = byte_model(nx, ny, nz, Px.value, Py.value, Pz.value, 8, p_arrays=2, chained=True)
    # the fused RK4 stage epilogue adds 30 w N per step (the canonical model's B_rk4)
    b_step = 4 * b_eval + rk4_stage_extra(N, 8)
    # ... all of it in the last pass, whose launches are what the pass timers
    # average over: a quarter per evaluation belongs to P5's algorithmic bytes
    pb["P5"] += rk4_stage_extra(N, 8) // 4
    ms_eval = ms_step / 4
    pass_report = {}
    for name, (avg_ms, n) in passes.items():
        if avg_ms is None:
            continue
        bytes_ = pb.get(name, 0)
        pass_report[name] = {"avg_ms": round(avg_ms, 5), "launches": n,
                             "algorithmic_bytes": bytes_, "GBps": round(bytes_ / (avg_ms * 1e-3) / 1e9, 1)}
    dom = max((k for k in pass_report if k != "P1"), key=lambda k: pass_report[k]["avg_ms"])
    dom_ms = pass_report[dom]["avg_ms"]
    achieved = pb[dom] / (dom_ms * 1e-3) / 1e9
    traffic = read_traffic(dom)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": dom,
                "algorithmic_bytes_per_launch": pb[dom], "avg_launch_ms": dom_ms}
    step_gbs = b_step / (ms_step * 1e-3) / 1e9
    log(f"[bench] passes: " + ", ".join(f"{k} {v['avg_ms']*1e3:.1f}us {v['GBps']:.0f}GB/s" for k, v in pass_report.items()))
    del sim

    fp32 = None
    if not args.no_fp32:
        try:
            log("[bench] FP32 set-up")
            sim32, _ = build_system(mmx, nx, ny, nz, mmx.F32, images=args.images)
            el32 = time_steps(sim32, mmx, args.steps, args.warmup)
            v32 = N * evals / el32 / 1e6
            _, b32 = byte_model(nx, ny, nz, Px.value, Py.value, Pz.value, 4, p_arrays=2, chained=True)
            b32_step = 4 * b32 + rk4_stage_extra(N, 4)
            fp32 = {"value": round(v32, 1), "unit": "Mcell-evals/s", "ms_per_step": round(el32 / args.steps * 1e3, 4),
                    "hbm_frac_step": round(b32_step / (el32 / args.steps) / 1e9 / HBM_PEAK_GBS, 4)}
            log(f"[bench] FP32: {v32:.1f} Mcell-evals/s")
            del sim32
        except Exception as e:
            errors["fp32"] = f"{type(e).__name__}: {e}"

    extras = {}
    if not args.no_extras:
        for name, fn in (("copy_ceiling", lambda: copy_ceiling(mmx)),
                         ("rocfft_demag", lambda: rocfft_child(args.mesh)),
                         ("soft_film_200x50x1_1ns", lambda: gpu_soft_film(mmx)),
                         ("soft_film_200x50x1_sample", lambda: gpu_soft_film(mmx, SOFT_SAMPLE_T))):
            try:
                extras[name] = fn()
                log(f"[bench] {name}: {extras[name]}")
            except Exception as e:
                errors[name] = f"{type(e).__name__}: {e}"
                log(f"[bench] {name} failed: {e}")
        if "rocfft_demag" in extras:
            extras["rocfft_demag"]["full_eval_ms_this_path"] = round(ms_eval, 4)
        if "copy_ceiling" in extras:
            roofline["copy_ceiling_GBps"] = extras["copy_ceiling"]["GBps"]
            roofline["frac_of_copy_ceiling"] = round(achieved / extras["copy_ceiling"]["GBps"], 4)

    forms = None
    if not args.no_forms:
        import subprocess
        import tempfile

        fd, path = tempfile.mkstemp(suffix=".json", prefix="mmx_forms_")
        os.close(fd)
        cmd = [sys.executable, os.path.abspath(__file__), "--forms-child", path, "--fd-cells", str(args.fd_cells)]
        try:
            r = subprocess.run(cmd, timeout=900)
            try:
                with open(path) as f:
                    forms = json.load(f)
            except (OSError, ValueError):
                forms = None
            if r.returncode != 0:
                errors["forms"] = f"forms child exited with status {r.returncode}"
        except subprocess.TimeoutExpired:
            errors["forms"] = "forms child timed out"
        finally:
            if os.path.exists(path):
                os.unlink(path)

    cpu = None
    if not args.no_cpu:
        try:
            log("[bench] CPU oracle baseline")
            cpu = cpu_baseline(nx, ny, nz, m0, args.images)
            log(f"[bench] CPU: {cpu['value']} Mcell-evals/s on {cpu['cores']} threads")
        except Exception as e:
            errors["cpu_baseline"] = f"{type(e).__name__}: {e}"

    out = {
        "metric": "Mcell field+LLG evals/s FP64/FP32 + % HBM peak; batched & adjoint vs plain",
        "value": round(value, 2),
        "unit": "Mcell-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"3-D chiral film {nx}x{ny}x{nz}, 2 nm cells, exch+iDMI+Ku+Zeeman+FFT demag "
                               f"({args.images}x{args.images} macro-PBC images), RK4 dt=1e-14",
                   "mesh": [nx, ny, nz], "integrator": "RK4", "evals_per_step": 4,
                   "padded": [Px.value, Py.value, Pz.value]},
        "hbm_frac_step": round(step_gbs / HBM_PEAK_GBS, 4),
        "bytes_per_eval": b_eval,
        "ms_per_eval": round(ms_eval, 4),
        "passes": pass_report,
        "roofline": roofline,
        "fp32": fp32,
        "extras": extras,
        "forms": forms,
        "cpu_baseline": cpu,
        "errors": errors or None,
    }
    print(json.dumps(out), flush=True)
    if errors:
        sys.exit(1)


if __name__ == "__main__":
    main()
