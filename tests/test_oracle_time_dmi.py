"""Time-dependent bulk DMI of the oracle (add_dmi(sim, D, ft), add_field.jl:360-393):
TimeBulkDMI (field.jl:145-163) scales (Dx, Dy, Dz) by time_D(t) before the
bulk kernel, TimeSpatialBulkDMI (:180-193) scales every per-cell D.  Pinned by
the definition through an independent path: the static bulk DMI terms (pinned
by the reference's golden fields) built with the scaled D give the same field
and energy bit for bit; a zero profile gives a zero field.
"""
import math

import numpy as np

from oracle import OracleSim


def _sim(seed=8):
    rng = np.random.default_rng(seed)
    o = OracleSim(7, 5, 3, 1e-9, 2e-9, 3e-9, pbc="x")
    N = o.N
    o.set_Ms(np.where(rng.uniform(size=N) < 0.15, 0.0, 7e5))
    o.init_m0(rng.standard_normal(3 * N))
    return o, rng


def ft(t):
    return 0.5 + math.cos(2 * math.pi * t / 1e-12)


def test_time_bulk_dmi_equals_static_at_scaled_D():
    for t in (0.0, 3.3e-13, 7.7e-13):
        o, _ = _sim()
        D = (2.5e-3, -1e-3, 4e-4)
        k = o.add_dmi_bulk_time(D, ft)
        f = ft(t)
        ref = o.add_dmi(tuple(f * d for d in D), type="bulk")
        h1, e1 = o.term_field(k, t=t)
        h2, e2 = o.term_field(ref, t=t)
        assert np.array_equal(h1, h2) and np.array_equal(e1, e2)


def test_time_spatial_bulk_dmi_equals_static_at_scaled_D():
    o, rng = _sim()
    D = rng.uniform(0, 3e-3, o.N) * (rng.uniform(size=o.N) > 0.1)
    k = o.add_dmi_bulk_time(D, ft)
    t = 4.1e-13
    f = ft(t)
    ref = o.add_dmi((f * D, f * D, f * D), type="bulk")
    h1, e1 = o.term_field(k, t=t)
    h2, e2 = o.term_field(ref, t=t)
    assert np.array_equal(h1, h2) and np.array_equal(e1, e2)


def test_time_dmi_zero_profile():
    o, rng = _sim()
    k = o.add_dmi_bulk_time((3e-3, 3e-3, 3e-3), lambda t: 0.0)
    h, e = o.term_field(k, t=1e-13)
    assert np.all(h == 0) and np.all(e == 0)
    k = o.add_dmi_bulk_time(rng.uniform(1e-3, 2e-3, o.N), lambda t: 0.0)
    h, e = o.term_field(k, t=1e-13)
    assert np.all(h == 0) and np.all(e == 0)
