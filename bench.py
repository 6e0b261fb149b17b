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


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_system(mmx, nx, ny, nz, precision, images=5, seed=2002, dt=1e-14):
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
    mmx.add_demag(sim, Nx=images, Ny=images, Nz=0)
    mmx.set_driver(sim, integrator="RungeKutta", step=dt)
    return sim, m0


def byte_model(nx, ny, nz, Px, Py, Pz, w, p_arrays):
    """Canonical algorithmic bytes of one RHS evaluation, per pass.

    S1 = Kx*ny*nz, S2 = Kx*Py*nz complex values per component; the k-space
    tensor is read folded (6 * Kx * (Py/2+1) * (Pz/2+1) reals); m is read by the
    first and the last pass; p_arrays per-cell parameter arrays of w bytes plus
    a 1-byte pin mask are read by the last pass, which writes 3N results."""
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


def cpu_baseline(nx, ny, nz, m0, images, budget_s=30.0):
    """The CPU oracle (a C restatement of the reference CPU path) on the host:
    one field + LLG evaluation of the same workload, OpenMP over all cores."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    O.build()
    cores = O.set_threads(os.cpu_count() or 1)
    N = nx * ny * nz
    s = O.OracleSim(nx, ny, nz, 2e-9, 2e-9, 2e-9, pbc="xy")
    s.set_Ms(np.full(N, 5.8e5))
    s.init_m0(m0)
    s.add_exch(1.5e-11)
    s.add_dmi(np.full(N, 3e-3), type="interfacial")
    s.add_anis(np.full(N, 8e5), axis=(0, 0, 1))
    s.add_zeeman(np.tile([0, 0, 0.1 / MU0], N))
    t0 = time.perf_counter()
    s.add_demag(images, images, 0)
    t_setup = time.perf_counter() - t0
    s.set_llg(0.1, 2.21e5, True)
    times = []
    while True:
        t0 = time.perf_counter()
        s.llg_rhs()
        times.append(time.perf_counter() - t0)
        if sum(times) > budget_s / 3 or len(times) >= 3:
            break
    t = float(np.median(times))
    return {"value": round(N / t / 1e6, 3), "unit": "Mcell-evals/s", "cores": cores, "kind": "port",
            "sample": f"{len(times)} field+LLG evaluation(s) of the 256x256x32 workload (median {t:.2f} s each; "
                      f"demag set-up {t_setup:.1f} s not timed), FP64, OpenMP"}


def read_traffic(kernel_pass):
    """HBM bytes per launch of a pass from the committed PMC summary, if any."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        m = re.match(r"\|\s*" + kernel_pass + r"\s*\|\s*([0-9.eE+]+)\s*\|", line)
        if m:
            return float(m.group(1))
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--mesh", type=str, default="256x256x32")
    ap.add_argument("--images", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 1 GiB device copy (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch

        n = 1 << 27  # 128 Mi doubles = 1 GiB
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        log(f"[bench] copy probe: 3 copies of {n * 8} bytes")
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported by this build "
              "(the path is single-GPU; multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)

    import mmx

    nx, ny, nz = (int(v) for v in args.mesh.split("x"))
    N = nx * ny * nz
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
    pb, b_eval = byte_model(nx, ny, nz, Px.value, Py.value, Pz.value, 8, p_arrays=2)
    # the fused RK4 stage epilogue adds 30 w N per step (the canonical model's B_rk4)
    b_step = 4 * b_eval + rk4_stage_extra(N, 8)
    ms_eval = ms_step / 4
    pass_report = {}
    for name, (avg_ms, n) in passes.items():
        if avg_ms is None:
            continue
        bytes_ = pb.get(name, 0)
        pass_report[name] = {"avg_ms": round(avg_ms, 5), "launches": n,
                             "algorithmic_bytes": bytes_, "GBps": round(bytes_ / (avg_ms * 1e-3) / 1e9, 1)}
    dom = max(pass_report, key=lambda k: pass_report[k]["avg_ms"])
    dom_ms = pass_report[dom]["avg_ms"]
    achieved = pb[dom] / (dom_ms * 1e-3) / 1e9
    traffic = read_traffic(dom)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": dom}
    step_gbs = b_step / (ms_step * 1e-3) / 1e9
    log(f"[bench] passes: " + ", ".join(f"{k} {v['avg_ms']*1e3:.1f}us {v['GBps']:.0f}GB/s" for k, v in pass_report.items()))
    del sim

    fp32 = None
    if not args.no_fp32:
        log("[bench] FP32 set-up")
        sim32, _ = build_system(mmx, nx, ny, nz, mmx.F32, images=args.images)
        el32 = time_steps(sim32, mmx, args.steps, args.warmup)
        v32 = N * evals / el32 / 1e6
        _, b32 = byte_model(nx, ny, nz, Px.value, Py.value, Pz.value, 4, p_arrays=2)
        b32_step = 4 * b32 + rk4_stage_extra(N, 4)
        fp32 = {"value": round(v32, 1), "unit": "Mcell-evals/s", "ms_per_step": round(el32 / args.steps * 1e3, 4),
                "hbm_frac_step": round(b32_step / (el32 / args.steps) / 1e9 / HBM_PEAK_GBS, 4)}
        log(f"[bench] FP32: {v32:.1f} Mcell-evals/s")
        del sim32

    cpu = None
    if not args.no_cpu:
        log("[bench] CPU oracle baseline")
        cpu = cpu_baseline(nx, ny, nz, m0, args.images)
        log(f"[bench] CPU: {cpu['value']} Mcell-evals/s on {cpu['cores']} threads")

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
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
