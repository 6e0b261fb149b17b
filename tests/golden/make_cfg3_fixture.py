"""Writes cfg3_fields.npz: the oracle's FP64 effective field, dm/dt and
per-term energies of the 512x512x64 precision-form workload (3 nm cells, Ms
8e5 A/m, A 1.3e-11 J/m, bulk D 1e-3 J/m^2, cubic Kc -5e3 J/m^3 along x/y/z,
demag padded to 1024x1024x128, LLG alpha 0.5 gamma 2.211e5) at 4096 seeded
cells, for the GPU test that states the FP32 path's error against it.

The initial state is smooth_m0(): the normalised sum of 32 seeded random
Fourier modes.  Run: python tests/golden/make_cfg3_fixture.py (several minutes
on 8 cores; the oracle is FP64 C with OpenMP)."""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
NX, NY, NZ, D = 512, 512, 64, 3e-9
MS, A, DB, KC = 8e5, 1.3e-11, 1e-3, -5e3
ALPHA, GAMMA = 0.5, 2.211e5
NSAMPLE, SEED = 4096, 3003


def smooth_m0(nx=NX, ny=NY, nz=NZ, modes=32, seed=SEED):
    """Normalised sum of `modes` random Fourier modes: wave numbers up to 4
    periods along x and y and 2 along z, normal amplitudes per component,
    uniform phases (AoS, 3N values)."""
    rng = np.random.default_rng(seed)
    x = (np.arange(nx) + 0.5) / nx
    y = (np.arange(ny) + 0.5) / ny
    z = (np.arange(nz) + 0.5) / nz
    m = np.zeros((3, nz, ny, nx))
    for _ in range(modes):
        k = 2 * np.pi * np.array([rng.integers(-4, 5), rng.integers(-4, 5), rng.integers(0, 3)])
        amp = rng.standard_normal(3)
        ph = rng.uniform(0, 2 * np.pi)
        arg = (k[0] * x[None, None, :] + k[1] * y[None, :, None]) + k[2] * z[:, None, None] + ph
        c = np.cos(arg)
        for q in range(3):
            m[q] += amp[q] * c
    m = m.reshape(3, -1).T
    m /= np.linalg.norm(m, axis=1, keepdims=True)
    return np.ascontiguousarray(m.reshape(-1))


def sample_cells(n=NSAMPLE, seed=SEED + 1):
    return np.sort(np.random.default_rng(seed).choice(NX * NY * NZ, n, replace=False))


def main():
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "oracle"))
    import oracle as O

    N = NX * NY * NZ
    t0 = time.time()
    o = O.OracleSim(NX, NY, NZ, D, D, D)
    o.set_Ms(np.full(N, MS))
    o.init_m0(smooth_m0())
    o.add_exch(A)
    o.add_dmi(np.full(N, DB), type="bulk")
    o.add_cubic_anis(np.full(N, KC), (1, 0, 0), (0, 1, 0))
    o.add_demag(0, 0, 0)
    o.set_llg(ALPHA, GAMMA, True)
    print(f"set-up {time.time() - t0:.1f} s", flush=True)
    idx = sample_cells()
    H = o.effective_field().reshape(-1, 3)[idx]
    print(f"field {time.time() - t0:.1f} s", flush=True)
    R = o.llg_rhs().reshape(-1, 3)[idx]
    E = o.term_energies()
    print(f"done {time.time() - t0:.1f} s; energies {E}", flush=True)
    np.savez_compressed(os.path.join(HERE, "cfg3_fields.npz"), cells=idx, H=H, dmdt=R, energies=E,
                        m0_sum=np.sum(smooth_m0().reshape(-1, 3), axis=0))


if __name__ == "__main__":
    main()
