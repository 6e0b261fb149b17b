"""Writes cfg1_soft_film.json: the oracle's Dormand-Prince 5(4) run of the
200x50x1 soft film (2.5 x 2.5 x 3 nm cells, Ms 8e5 A/m, A 1.3e-11 J/m,
exchange + demag, LLG alpha 0.5 gamma 2.211e5, tol 1e-6) from 0 to 1 ns:
mu0Ms-weighted <m>, accepted steps and evaluations at every 0.1 ns, for the
GPU test of the same run.

The initial state is normalise((1, 0.25, 0.1) + 0.05 U(-1, 1)^3) per cell
with numpy's default_rng(1001).  Run: python tests/golden/make_cfg1_fixture.py
(a few minutes; the oracle is FP64 C)."""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
NX, NY, NZ = 200, 50, 1
DX, DY, DZ = 2.5e-9, 2.5e-9, 3e-9
MS, A, ALPHA, GAMMA, TOL = 8e5, 1.3e-11, 0.5, 2.211e5, 1e-6
TIMES = [round(0.1e-9 * k, 12) for k in range(1, 11)]


def m0(seed=1001):
    rng = np.random.default_rng(seed)
    m = np.array([1.0, 0.25, 0.1]) + 0.05 * rng.uniform(-1, 1, (NX * NY * NZ, 3))
    m /= np.linalg.norm(m, axis=1, keepdims=True)
    return m.reshape(-1)


def main():
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "oracle"))
    import oracle as O

    N = NX * NY * NZ
    o = O.OracleSim(NX, NY, NZ, DX, DY, DZ)
    o.set_Ms(np.full(N, MS))
    o.init_m0(m0())
    o.add_exch(A)
    o.add_demag(0, 0, 0)
    o.set_llg(ALPHA, GAMMA, True)
    o.set_integrator(O.DOPRI54, 0.0, TOL)
    rows = []
    t0 = time.time()
    for t in TIMES:
        o.run_until(t)
        st = o.state()
        mavg = o.spin.reshape(-1, 3).mean(axis=0)
        rows.append({"t": t, "m": [float(v) for v in mavg], "nsteps": int(st["nsteps"]),
                     "nfevals": int(st["nfevals"]), "step_next": float(st["step_next"])})
        print(f"t={t:.2e} steps {st['nsteps']} <m> {mavg} ({time.time() - t0:.0f} s)", flush=True)
    json.dump({"workload": "200x50x1 soft film, exchange + demag, DOPRI54 tol 1e-6, 0 -> 1 ns",
               "times": TIMES, "rows": rows}, open(os.path.join(HERE, "cfg1_soft_film.json"), "w"), indent=1)


if __name__ == "__main__":
    main()
