"""ctypes view of the CPU oracle (test infrastructure only).

The oracle restates MicroMagnetic.jl's CPU path in C (``mm_oracle.c``).  Only the
tests, the smoke check and the CPU-baseline leg of ``bench.py`` use this module,
always as the checker.  ``build()`` compiles ``liboracle.so`` with the Makefile
beside this file.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

MU0 = 4e-7 * np.pi
RK4, HEUN, DOPRI54, BS23 = 0, 1, 2, 3

_lib = None
_dp = C.POINTER(C.c_double)


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.c_int] * 3 + [C.c_double] * 3 + [C.c_int]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_num_threads.argtypes = [C.c_int]
        L.orc_num_threads.restype = C.c_int
        for name, args in {
            "orc_neighbours": [C.c_void_p, C.c_void_p],
            "orc_set_Ms": [C.c_void_p, _dp],
            "orc_set_mu0Ms": [C.c_void_p, _dp],
            "orc_set_pins": [C.c_void_p, C.c_void_p],
            "orc_set_m": [C.c_void_p, _dp],
            "orc_get_m": [C.c_void_p, _dp],
            "orc_effective_field": [C.c_void_p, _dp, C.c_double, _dp],
            "orc_term_field": [C.c_void_p, C.c_int, _dp, C.c_double, _dp, _dp],
            "orc_term_energies": [C.c_void_p, _dp, C.c_double, _dp],
            "orc_set_llg": [C.c_void_p, C.c_double, C.c_double, C.c_int],
            "orc_llg_rhs": [C.c_void_p, _dp, C.c_double, _dp],
            "orc_set_integrator": [C.c_void_p, C.c_int, C.c_double, C.c_double],
            "orc_integrator_state": [C.c_void_p, _dp, _dp, _dp, C.c_void_p, C.c_void_p],
            "orc_fft": [_dp, C.c_int, C.c_int],
            "orc_philox_normal": [C.c_uint64, C.c_uint64, C.c_uint64, _dp],
            "orc_philox4x32_10": [C.c_void_p, C.c_void_p, C.c_void_p],
            "orc_set_alpha": [C.c_void_p, _dp],
        }.items():
            getattr(L, name).argtypes = args
            getattr(L, name).restype = None
        for name, args in {
            "orc_init_m0": [C.c_void_p, _dp, C.c_int],
            "orc_add_exch": [C.c_void_p, C.c_double, C.c_double, C.c_double],
            "orc_add_exch_spatial": [C.c_void_p, _dp],
            "orc_add_dmi_bulk": [C.c_void_p, C.c_double, C.c_double, C.c_double],
            "orc_add_dmi_bulk_spatial": [C.c_void_p, _dp, _dp, _dp],
            "orc_add_dmi_interfacial": [C.c_void_p, _dp],
            "orc_add_anis": [C.c_void_p, _dp, C.c_double, C.c_double, C.c_double],
            "orc_add_anis_spatial": [C.c_void_p, _dp, _dp],
            "orc_add_cubic_anis": [C.c_void_p, _dp, _dp, _dp],
            "orc_add_zeeman": [C.c_void_p, _dp],
            "orc_add_demag": [C.c_void_p, C.c_int, C.c_int, C.c_int],
            "orc_add_demag_direct": [C.c_void_p, C.c_int, C.c_int, C.c_int],
            "orc_add_thermal": [C.c_void_p, _dp, C.c_double, C.c_uint64],
            "orc_add_sot": [C.c_void_p, _dp] + [C.c_double] * 4,
            "orc_add_stt_slonczewski": [C.c_void_p, _dp] + [C.c_double] * 7,
            "orc_add_stt_zhangli": [C.c_void_p, _dp, C.c_double, _dp, C.c_int],
            "orc_set_term_time_fn": [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p],
            "orc_add_zeeman_time": [C.c_void_p, _dp, C.c_void_p, C.c_void_p],
            "orc_add_hex_anis": [C.c_void_p, C.c_double, C.c_double, C.c_double],
            "orc_add_exch_int": [C.c_void_p, _dp, C.c_int, C.c_int],
            "orc_add_twin_mono_anis": [C.c_void_p, _dp, _dp, _dp, _dp],
            "orc_add_sahe_torque": [C.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp, C.c_double],
            "orc_add_magnetoelastic": [C.c_void_p, C.c_int, _dp, C.c_double, C.c_double, C.c_double],
            "orc_add_torque": [C.c_void_p, C.c_void_p, C.c_void_p],
            "orc_add_dmi_int": [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int],
            "orc_num_terms": [C.c_void_p],
            "orc_demag_tensor": [C.c_void_p, C.c_int, _dp],
            "orc_advance_step": [C.c_void_p],
            "orc_run_until": [C.c_void_p, C.c_double],
            "orc_relax": [C.c_void_p, C.c_int, C.c_double, C.c_void_p, C.c_void_p],
        }.items():
            getattr(L, name).argtypes = args
            getattr(L, name).restype = C.c_int
        _lib = L
    return _lib


_TIMEFN = C.CFUNCTYPE(C.c_double, C.c_double, C.c_void_p)
_TIMEFN3 = C.CFUNCTYPE(None, C.c_double, _dp, C.c_void_p)
_TORQUEFN = C.CFUNCTYPE(None, C.c_double, _dp, _dp, C.c_void_p)


def _factor3(ft):
    """A time_fun returning a number or a 3-tuple, as (fx, fy, fz): src/micro/field.jl:16-26."""

    def cb(t, f, user):
        v = ft(t)
        if np.ndim(v) == 0:
            f[0] = f[1] = f[2] = float(v)
        else:
            f[0], f[1], f[2] = map(float, v)

    return _TIMEFN3(cb)


def _p(a):
    return a.ctypes.data_as(_dp)


def _f64(a, n):
    a = np.ascontiguousarray(np.broadcast_to(np.asarray(a, dtype=np.float64), (n,)))
    return a


def set_threads(n):
    return lib().orc_num_threads(int(n))


def fft(x, sign=-1):
    """Unnormalised complex transform of a 1-D array (oracle's own FFT)."""
    x = np.asarray(x, dtype=np.complex128)
    buf = np.empty(2 * x.size)
    buf[0::2] = x.real
    buf[1::2] = x.imag
    lib().orc_fft(_p(buf), x.size, sign)
    return buf[0::2] + 1j * buf[1::2]


def philox4x32_10(ctr, key):
    """The raw Philox4x32-10 block of a 4-word counter and a 2-word key."""
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    out = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, out)
    return list(out)


def philox_normal(seed, step, n):
    out = np.empty(n)
    lib().orc_philox_normal(seed, step, n, _p(out))
    return out


class OracleSim:
    """One simulated system; mirrors the reference's ``Sim``/``MicroSim`` state."""

    def __init__(self, nx, ny, nz, dx=1e-9, dy=1e-9, dz=1e-9, pbc=""):
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.N = self.nx * self.ny * self.nz
        mask = (1 if "x" in pbc else 0) | (2 if "y" in pbc else 0) | (4 if "z" in pbc else 0)
        self.h = lib().orc_create(self.nx, self.ny, self.nz, dx, dy, dz, mask)
        self.dx, self.dy, self.dz = dx, dy, dz

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    # -- state ---------------------------------------------------------------
    def neighbours(self):
        out = np.empty((self.N, 6), dtype=np.int32)
        lib().orc_neighbours(self.h, out.ctypes.data)
        return out

    def set_Ms(self, Ms):
        lib().orc_set_Ms(self.h, _p(_f64(Ms, self.N)))

    def set_mu0Ms(self, v):
        lib().orc_set_mu0Ms(self.h, _p(_f64(v, self.N)))

    def set_pins(self, pins):
        p = np.ascontiguousarray(np.broadcast_to(np.asarray(pins, dtype=np.uint8), (self.N,)))
        lib().orc_set_pins(self.h, p.ctypes.data)

    def init_m0(self, m, norm=True):
        m = _f64(np.asarray(m, dtype=np.float64).reshape(-1), 3 * self.N)
        rc = lib().orc_init_m0(self.h, _p(m), int(bool(norm)))
        if rc != 0:
            raise ValueError("NaN is given by the input m0!")

    def set_m(self, m):
        lib().orc_set_m(self.h, _p(_f64(np.asarray(m).reshape(-1), 3 * self.N)))

    @property
    def spin(self):
        out = np.empty(3 * self.N)
        lib().orc_get_m(self.h, _p(out))
        return out

    # -- terms ---------------------------------------------------------------
    def add_exch(self, A):
        if np.ndim(A) == 0:
            return lib().orc_add_exch(self.h, A, A, A)
        if isinstance(A, tuple) and len(A) == 3:
            return lib().orc_add_exch(self.h, *map(float, A))
        return lib().orc_add_exch_spatial(self.h, _p(_f64(A, self.N)))

    def add_dmi(self, D, type="bulk"):
        if type == "interfacial":
            return lib().orc_add_dmi_interfacial(self.h, _p(_f64(D, self.N)))
        if np.ndim(D) == 0:
            return lib().orc_add_dmi_bulk(self.h, D, D, D)
        if isinstance(D, tuple) and len(D) == 3 and all(np.ndim(d) == 0 for d in D):
            return lib().orc_add_dmi_bulk(self.h, *map(float, D))
        if isinstance(D, tuple):
            Dx, Dy, Dz = (_f64(d, self.N) for d in D)
        else:
            Dx = Dy = Dz = _f64(D, self.N)
        return lib().orc_add_dmi_bulk_spatial(self.h, _p(Dx), _p(Dy), _p(Dz))

    def add_torque(self, fun):
        """add_torque(sim, fun): src/micro/add_field.jl:998-1006; fun(y, m, t) fills y (3N AoS)."""
        n = 3 * self.N

        def cb(t, m, y, user):
            fun(np.ctypeslib.as_array(y, shape=(n,)), np.ctypeslib.as_array(m, shape=(n,)), t)

        f = _TORQUEFN(cb)
        self._keep = getattr(self, "_keep", []) + [f]
        return lib().orc_add_torque(self.h, C.cast(f, C.c_void_p), None)

    def add_magnetoelastic(self, model="tensor", lambda_s=0.0, B1=0.0, B2=0.0, data=None):
        """add_magnetoelastic: src/micro/add_field.jl:1367-1405; data (6N, Voigt per cell)."""
        d = _f64(np.asarray(data, dtype=np.float64).reshape(-1), 6 * self.N)
        return lib().orc_add_magnetoelastic(self.h, int(model == "cubic"), _p(d), float(lambda_s), float(B1),
                                            float(B2))

    def add_sahe_torque(self, sigma_s, sigma_sa1=(0, 0, 0), sigma_sa2=(0, 0, 0), a1=(1, 0, 0), a2=(0, 0, 1),
                        a3=(0, 0, 1), beta=0.0):
        """add_sahe_torque: src/micro/add_field.jl:892-936 (3-tuples or 3N AoS)."""
        vs = []
        for v in (sigma_s, sigma_sa1, sigma_sa2, a1, a2, a3):
            v = np.asarray(v, dtype=np.float64)
            vs.append(np.ascontiguousarray(np.tile(v, self.N) if v.size == 3 else v.reshape(-1)))
        return lib().orc_add_sahe_torque(self.h, *(_p(v) for v in vs), float(beta))

    def add_twin_mono_anis(self, Ka=0, Kb=0, Kaa=0, Kbb=0, Kab=0, Ku=0, axis_a=(1, 0, 0), axis_b=(0, 1, 0),
                           axis_u111=(0, 0, 1)):
        """add_twin_mono_anis: src/micro/add_field.jl:629-660 (axes: 3-tuples or 3N AoS)."""
        K = np.array([Ka, Kb, Kaa, Kbb, Kab, Ku], dtype=np.float64)
        ax = []
        for v in (axis_a, axis_b, axis_u111):
            v = np.asarray(v, dtype=np.float64)
            ax.append(np.ascontiguousarray(np.tile(v, self.N) if v.size == 3 else v.reshape(-1)))
        return lib().orc_add_twin_mono_anis(self.h, _p(K), *(_p(a) for a in ax))

    def add_exch_int(self, J, k1=1, k2=-1):
        """add_exch_int(sim, J; k1, k2): src/micro/add_field.jl:789-816 (J: one value per column)."""
        return lib().orc_add_exch_int(self.h, _p(_f64(J, self.nx * self.ny)), int(k1), int(k2))

    def add_dmi_int(self, D, k1=1, k2=-1):
        """add_dmi_int(sim, D; k1, k2): src/micro/add_field.jl:835-860."""
        return lib().orc_add_dmi_int(self.h, *map(float, D), int(k1), int(k2))

    def add_hex_anis(self, K1=0.0, K2=0.0, K3=0.0):
        """add_hex_anis(sim; K1, K2, K3): src/micro/add_field.jl:590-606."""
        return lib().orc_add_hex_anis(self.h, float(K1), float(K2), float(K3))

    def add_dmi_bulk_time(self, D, ft):
        """add_dmi(sim, D, ft): src/micro/add_field.jl:360-393 (a number or 3-tuple:
        TimeBulkDMI; per cell: TimeSpatialBulkDMI with D for all three components)."""
        if np.ndim(D) == 0 or (isinstance(D, tuple) and len(D) == 3):
            Dv = (float(D),) * 3 if np.ndim(D) == 0 else tuple(map(float, D))
            k = lib().orc_add_dmi_bulk(self.h, *Dv)
        else:
            a = _f64(D, self.N)
            k = lib().orc_add_dmi_bulk_spatial(self.h, _p(a), _p(a), _p(a))
        self.set_term_time_fn(k, ft)
        return k

    def add_anis(self, Ku, axis=(0, 0, 1)):
        Ku = _f64(Ku, self.N)
        if isinstance(axis, tuple):
            return lib().orc_add_anis(self.h, _p(Ku), *map(float, axis))
        ax = _f64(np.asarray(axis).reshape(-1), 3 * self.N).copy()
        return lib().orc_add_anis_spatial(self.h, _p(Ku), _p(ax))

    def add_cubic_anis(self, Kc, axis1=(1, 0, 0), axis2=(0, 1, 0)):
        a1 = np.array(axis1, dtype=np.float64)
        a2 = np.array(axis2, dtype=np.float64)
        r = lib().orc_add_cubic_anis(self.h, _p(_f64(Kc, self.N)), _p(a1), _p(a2))
        if r < 0:
            raise ValueError("The axis1 and axis2 are not perpendicular to each other!")
        return r

    def add_zeeman(self, H):
        H = np.asarray(H, dtype=np.float64)
        if H.size == 3:
            H = np.tile(H, self.N)
        return lib().orc_add_zeeman(self.h, _p(_f64(H.reshape(-1), 3 * self.N)))

    def add_zeeman_time(self, H, ft):
        """add_zeeman(sim, H0, ft): src/micro/add_field.jl:161-200."""
        H = np.asarray(H, dtype=np.float64)
        if H.size == 3:
            H = np.tile(H, self.N)
        cb = _factor3(ft)
        self._keep = getattr(self, "_keep", []) + [cb]
        return lib().orc_add_zeeman_time(self.h, _p(_f64(H.reshape(-1), 3 * self.N)), C.cast(cb, C.c_void_p), None)

    def add_demag(self, Nx=0, Ny=0, Nz=0, fft=True):
        f = lib().orc_add_demag if fft else lib().orc_add_demag_direct
        return f(self.h, int(Nx), int(Ny), int(Nz))

    def add_thermal(self, T, T0=0.0, seed=0):
        return lib().orc_add_thermal(self.h, _p(_f64(T, self.N)), float(T0), int(seed))

    def add_sot(self, aj, bj, p):
        """add_sot(sim, aj, bj, p): src/micro/add_field.jl:951-966."""
        return lib().orc_add_sot(self.h, _p(_f64(aj, self.N)), float(bj), *map(float, p))

    def add_stt_slonczewski(self, J, beta, Lambda, xi, P, p):
        """add_stt(model=:slonczewski): src/micro/add_field.jl:1195-1227 (beta precomputed)."""
        return lib().orc_add_stt_slonczewski(self.h, _p(_f64(J, self.N)), float(beta), float(Lambda),
                                             float(xi), float(P), *map(float, p))

    def add_stt_zhangli(self, J, b, xi, xi_factor):
        """add_stt(model=:zhang_li): src/micro/add_field.jl:1150-1192; J 3-vector or 3N AoS."""
        J = np.asarray(J, dtype=np.float64)
        if J.size == 3:
            J = np.tile(J, self.N)
        return lib().orc_add_stt_zhangli(self.h, _p(_f64(J.reshape(-1), 3 * self.N)), float(b),
                                         _p(_f64(xi, self.N)), int(bool(xi_factor)))

    def set_term_time_fn(self, term, fn):
        """ufun(t) of a torque term or time_D(t) of a bulk DMI term; fn(t) -> float, or None for 1."""
        if fn is None:
            lib().orc_set_term_time_fn(self.h, term, None, None)
            return
        cb = _TIMEFN(lambda t, user: float(fn(t)))
        self._keep = getattr(self, "_keep", []) + [cb]
        lib().orc_set_term_time_fn(self.h, term, C.cast(cb, C.c_void_p), None)

    def demag_tensor(self, term):
        out = np.empty(6 * self.N)
        if lib().orc_demag_tensor(self.h, term, _p(out)) != 0:
            raise ValueError("not a demag term")
        return out.reshape(6, self.N)

    # -- evaluation ----------------------------------------------------------
    def effective_field(self, spin=None, t=0.0):
        spin = self.spin if spin is None else _f64(spin, 3 * self.N)
        out = np.empty(3 * self.N)
        lib().orc_effective_field(self.h, _p(spin), t, _p(out))
        return out

    def term_field(self, term, spin=None, t=0.0):
        spin = self.spin if spin is None else _f64(spin, 3 * self.N)
        f = np.empty(3 * self.N)
        e = np.empty(self.N)
        lib().orc_term_field(self.h, term, _p(spin), t, _p(f), _p(e))
        return f, e

    def term_energies(self, spin=None, t=0.0):
        spin = self.spin if spin is None else _f64(spin, 3 * self.N)
        out = np.empty(lib().orc_num_terms(self.h))
        lib().orc_term_energies(self.h, _p(spin), t, _p(out))
        return out

    def set_llg(self, alpha=0.1, gamma=2.21e5, precession=True):
        lib().orc_set_llg(self.h, alpha, gamma, int(bool(precession)))

    def set_alpha(self, alpha):
        """Per-cell damping of the SpatialLLG driver (None: uniform again)."""
        self._alpha = None if alpha is None else _f64(alpha, self.N).copy()
        lib().orc_set_alpha(self.h, None if alpha is None else _p(self._alpha))

    def llg_rhs(self, spin=None, t=0.0):
        spin = self.spin if spin is None else _f64(spin, 3 * self.N)
        out = np.empty(3 * self.N)
        lib().orc_llg_rhs(self.h, _p(spin), t, _p(out))
        return out

    def set_integrator(self, kind, step=0.0, tol=1e-6):
        lib().orc_set_integrator(self.h, int(kind), float(step), float(tol))

    def advance_step(self):
        return lib().orc_advance_step(self.h)

    def run_until(self, t_end):
        return lib().orc_run_until(self.h, float(t_end))

    def relax(self, max_steps=10000, stopping_dmdt=0.01):
        steps = C.c_int(0)
        md = C.c_double(0)
        rc = lib().orc_relax(self.h, int(max_steps), float(stopping_dmdt), C.byref(steps), C.byref(md))
        return rc, steps.value, md.value

    def state(self):
        t, step, step_next = C.c_double(), C.c_double(), C.c_double()
        ns, nf = C.c_int64(), C.c_int64()
        lib().orc_integrator_state(self.h, C.byref(t), C.byref(step), C.byref(step_next),
                                   C.byref(ns), C.byref(nf))
        return dict(t=t.value, step=step.value, step_next=step_next.value,
                    nsteps=ns.value, nfevals=nf.value)
