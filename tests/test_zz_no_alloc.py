"""The allocation-free energy entry point (mmx_compute_energies_noalloc, through
the C ABI) against mmx_compute_energies on two calls in a row, the second on a
smaller mesh with fewer terms; device memory for its output comes from the HIP
runtime the library itself uses."""
import ctypes as C
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def note(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def hip():
    lib = C.CDLL(os.environ.get("MMX_HIP_LIBRARY", "libamdhip64.so"))
    lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    lib.hipFree.argtypes = [C.c_void_p]
    lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    lib.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    return lib


def energies_noalloc(sim, n, H):
    import mmx

    L = mmx._lib.load()
    d = C.c_void_p()
    assert H.hipMalloc(C.byref(d), 8 * n) == 0
    try:
        assert H.hipMemset(d, 0xFF, 8 * n) == 0  # garbage the entry point must overwrite
        mmx._lib.check(L.mmx_compute_energies_noalloc(sim._h, 0.0, d))
        mmx._lib.check(L.mmx_synchronize(sim._h))
        out = np.empty(n)
        assert H.hipMemcpy(out.ctypes.data, d, 8 * n, 2) == 0  # device to host
    finally:
        H.hipFree(d)
    return out


def film(nx, ny, nz, terms, seed):
    import mmx

    rng = np.random.default_rng(seed)
    N = nx * ny * nz
    mesh = mmx.FDMesh(dx=2e-9, dy=2e-9, dz=2e-9, nx=nx, ny=ny, nz=nz, pbc="xy")
    s = mmx.Sim(mesh)
    Ms = np.full(N, 5.8e5)
    Ms[rng.integers(0, N)] = 0.0
    mmx.set_Ms(s, Ms)
    mmx.init_m0(s, rng.standard_normal(3 * N))
    mmx.add_exch(s, 1.5e-11)
    if "dmi" in terms:
        mmx.add_dmi(s, 3e-3, type="interfacial")
    if "anis" in terms:
        mmx.add_anis(s, rng.uniform(6e5, 9e5, N), axis=(0, 0, 1))
    if "zeeman" in terms:
        mmx.add_zeeman(s, (0, 0, 8e4))
    if "demag" in terms:
        mmx.add_demag(s, Nx=1, Ny=1)
    return s


def test_energies_noalloc_matches_compute_energies():
    import mmx

    H = hip()
    for (nx, ny, nz), terms, seed in (((40, 32, 6), ("dmi", "anis", "zeeman", "demag"), 1),
                                      ((12, 10, 2), ("anis",), 2)):
        note(f"energies without allocation {nx}x{ny}x{nz} {terms}")
        s = film(nx, ny, nz, terms, seed)
        ref = mmx.compute_system_energy(s)
        got = energies_noalloc(s, len(ref), H)
        print("noalloc", got, "reference", ref, flush=True)
        assert np.array_equal(got, ref)
