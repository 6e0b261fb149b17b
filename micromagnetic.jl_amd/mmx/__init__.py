"""mmx -- MicroMagnetic.jl's finite-difference API over the MI355X library.

The names, argument meanings and error behaviour mirror the reference's Julia
API for this path (src/micro/mesh.jl FDMesh, src/sim.jl Sim / set_Ms / init_m0
/ set_pinning / relax / run_until / run_sim, src/micro/add_field.jl add_*,
src/micro/field.jl effective_field / compute_system_energy), so that code
written against the reference reads the same here.  All arithmetic happens in
libmmx (HIP kernels for gfx950); this module only expands parameters the way
the reference's init_scalar! / init_vector! do (src/helper.jl:5-102) and moves
arrays across the C ABI.
"""
import ctypes as C
import os
import math

import numpy as np

from . import _lib
from ._lib import BS23, DOPRI54, F32, F64, HEUN, RK4, MmxError, check, dptr

mu_0 = 4 * math.pi * 1e-7       # src/const.jl:1
mu_B = 9.27400949e-24           # src/const.jl:2
k_B = 1.3806505e-23             # src/const.jl:3
c_e = 1.602176565e-19           # src/const.jl:4
h_bar = 1.05457172647e-34       # src/const.jl:9
mT = 0.001 / mu_0               # src/const.jl:15

_INTEGRATORS = {"RungeKutta": RK4, "Heun": HEUN, "DormandPrince": DOPRI54, "DOPRI54": DOPRI54,
                "BS23": BS23}
# create_driver defaults: src/driver.jl:126-145
_DEFAULT_STEP = {RK4: 5e-14, HEUN: 1e-14}


class FDMesh:
    """FDMesh(; dx, dy, dz, nx, ny, nz, pbc, x0, y0, z0): src/micro/mesh.jl:69-98."""

    def __init__(self, dx=1e-9, dy=1e-9, dz=1e-9, nx=1, ny=1, nz=1, pbc="open", x0=None, y0=None, z0=None):
        self.dx, self.dy, self.dz = float(dx), float(dy), float(dz)
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.xperiodic, self.yperiodic, self.zperiodic = "x" in pbc, "y" in pbc, "z" in pbc
        self.x0 = -self.nx * self.dx / 2 if x0 is None else float(x0)
        self.y0 = -self.ny * self.dy / 2 if y0 is None else float(y0)
        self.z0 = -self.nz * self.dz / 2 if z0 is None else float(z0)
        self.n_total = self.nx * self.ny * self.nz
        self.volume = self.dx * self.dy * self.dz
        self.regions = np.zeros(self.n_total, dtype=np.int32)

    @property
    def pbc_mask(self):
        return int(self.xperiodic) | (int(self.yperiodic) << 1) | (int(self.zperiodic) << 2)


def _ijk(mesh):
    k, j, i = np.meshgrid(np.arange(1, mesh.nz + 1), np.arange(1, mesh.ny + 1), np.arange(1, mesh.nx + 1),
                          indexing="ij")
    return i.reshape(-1), j.reshape(-1), k.reshape(-1)


def init_scalar(mesh, v, dtype=np.float64):
    """init_scalar!: src/helper.jl:5-44 (number, array, f(i,j,k,dx,dy,dz), f(x,y,z), f(region))."""
    N = mesh.n_total
    if callable(v):
        import inspect
        nargs = len(inspect.signature(v).parameters)
        i, j, k = _ijk(mesh)
        if nargs == 6:
            out = [v(a, b, c, mesh.dx, mesh.dy, mesh.dz) for a, b, c in zip(i, j, k)]
        elif nargs == 3:
            out = [v(mesh.x0 + (a - 0.5) * mesh.dx, mesh.y0 + (b - 0.5) * mesh.dy, mesh.z0 + (c - 0.5) * mesh.dz)
                   for a, b, c in zip(i, j, k)]
        elif nargs == 1:
            out = [v(r) for r in mesh.regions]
        else:
            raise ValueError("The input function should have 6, 3 or 1 arguments.")
        return np.asarray(out, dtype=dtype)
    a = np.asarray(v, dtype=dtype)
    if a.ndim == 0:
        return np.full(N, a, dtype=dtype)
    a = np.ascontiguousarray(a.reshape(-1))
    if a.size != N:
        raise ValueError(f"expected {N} values, got {a.size}")
    return a


def init_vector(mesh, v):
    """init_vector!: src/helper.jl:64-102, :239-253 (tuple, array, f(i,j,k,...), f(x,y,z))."""
    N = mesh.n_total
    if callable(v):
        import inspect
        nargs = len(inspect.signature(v).parameters)
        i, j, k = _ijk(mesh)
        out = np.zeros((N, 3))
        for q, (a, b, c) in enumerate(zip(i, j, k)):
            if nargs == 6:
                val = v(a, b, c, mesh.dx, mesh.dy, mesh.dz)
            elif nargs == 3:  # centred coordinates, helper.jl:84-86
                val = v((a - 0.5 - mesh.nx / 2) * mesh.dx, (b - 0.5 - mesh.ny / 2) * mesh.dy,
                        (c - 0.5 - mesh.nz / 2) * mesh.dz)
            else:
                raise ValueError("The input function should have either 6 or 3 arguments.")
            if val is not None:
                out[q] = val
        out = out.reshape(-1)
        if np.isnan(out).any():
            raise ValueError("NaN is given by the input function.")
        return out
    a = np.asarray(v, dtype=np.float64)
    if a.size == 3:
        return np.tile(a.reshape(3), N)
    a = np.ascontiguousarray(a.reshape(-1))
    if a.size != 3 * N:
        raise ValueError(f"expected {3 * N} values, got {a.size}")
    return a


class Term:
    """One interaction term (the reference's Zeeman / UniformExchange / ... structs)."""

    def __init__(self, sim, index, name):
        self.sim, self.index, self.name = sim, index, name
        # table columns the reference's add_* call registers for this term:
        # (names, units, fn(sim, saver)) before and after its E_<name> column
        self.items_before, self.items_after = [], []
        self.energy_unit = "<J>"  # None: the term registers no energy column

    @property
    def field(self):
        """term.field after effective_field(term, sim, spin, t) (3N, AoS)."""
        return self.sim._term_field(self.index)[0]

    @property
    def energy(self):
        return self.sim._term_field(self.index)[1]

    def effective_field(self, t=0.0):
        """effective_field(term, sim, sim.spin, t): the term's (field, energy) at time t."""
        return self.sim._term_field(self.index, t)


class LLGDriver:
    """The LLG driver fields (src/driver.jl:15-21); writes go to the library."""

    def __init__(self, sim):
        object.__setattr__(self, "_sim", sim)
        object.__setattr__(self, "alpha", 0.1)
        object.__setattr__(self, "gamma", 2.21e5)
        object.__setattr__(self, "precession", True)

    def __setattr__(self, k, v):
        object.__setattr__(self, k, v)
        if k in ("alpha", "gamma", "precession"):
            self._sim._push_llg()

    @property
    def integrator(self):
        return self._sim._integrator


class Integrator:
    def __init__(self, sim, kind, step, tol):
        object.__setattr__(self, "_sim", sim)
        object.__setattr__(self, "kind", kind)
        object.__setattr__(self, "_step", step)
        object.__setattr__(self, "_tol", tol)

    def _state(self):
        t, st, sn = C.c_double(), C.c_double(), C.c_double()
        ns, nf = C.c_int64(), C.c_int64()
        check(_lib.load().mmx_get_state(self._sim._h, C.byref(t), C.byref(st), C.byref(sn), C.byref(ns),
                                        C.byref(nf)))
        return t.value, st.value, sn.value, ns.value, nf.value

    t = property(lambda self: self._state()[0])
    step_next = property(lambda self: self._state()[2])
    nsteps = property(lambda self: self._state()[3])
    nfevals = property(lambda self: self._state()[4])

    @property
    def step(self):
        return self._state()[1]

    def __setattr__(self, k, v):
        if k == "step":
            object.__setattr__(self, "_step", float(v))
            check(_lib.load().mmx_set_integrator(self._sim._h, self.kind, float(v), self._tol))
        elif k == "tol":
            object.__setattr__(self, "_tol", float(v))
            check(_lib.load().mmx_set_integrator(self._sim._h, self.kind, self._step, float(v)))
        else:
            object.__setattr__(self, k, v)


class Sim:
    """Sim(mesh; driver="LLG", integrator="DormandPrince"): src/sim.jl:48-107."""

    def __init__(self, mesh, driver="LLG", name="dyn", integrator="DormandPrince", precision=F64, device=0):
        if driver not in ("LLG", "SD", "SpatialLLG"):
            raise ValueError(f"driver {driver!r} is not available on this path (LLG, SpatialLLG and SD)")
        if integrator not in _INTEGRATORS:
            raise ValueError("Supported integrators: " + " ".join(_INTEGRATORS))
        self.mesh, self.name = mesh, name
        self.n_total = mesh.n_total
        self.precision = precision
        L = _lib.load()
        d = _lib.MeshDesc(mesh.nx, mesh.ny, mesh.nz, mesh.pbc_mask, mesh.dx, mesh.dy, mesh.dz, int(precision),
                          int(device))
        h = C.c_void_p()
        check(L.mmx_create(C.byref(h), C.byref(d)))
        self._h = h
        self.interactions = []
        self.mu0_Ms = np.zeros(self.n_total)
        self.driver = LLGDriver(self)
        kind = _lib.SD if driver == "SD" else _INTEGRATORS[integrator]
        self.driver_name = driver
        self._integrator = Integrator(self, kind, _DEFAULT_STEP.get(kind, 0.0), 1e-6)
        check(L.mmx_set_integrator(self._h, kind, _DEFAULT_STEP.get(kind, 0.0), 1e-6))
        self._push_llg()
        if driver == "SpatialLLG":  # alpha .= 0.1 per cell, src/driver.jl:149-152
            set_alpha(self, 0.1)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            _lib.load().mmx_destroy(h)
            self._h = None

    def _push_llg(self):
        d = self.driver
        if np.ndim(d.alpha) == 0:
            check(_lib.load().mmx_set_llg(self._h, float(d.alpha), float(d.gamma), int(bool(d.precession))))
        else:  # SpatialLLG: the per-cell array travels through mmx_set_alpha
            check(_lib.load().mmx_set_llg(self._h, 0.1, float(d.gamma), int(bool(d.precession))))
            a = np.ascontiguousarray(d.alpha, dtype=np.float64)
            check(_lib.load().mmx_set_alpha(self._h, dptr(a)))

    def _term_field(self, index, t=0.0):
        N = self.n_total
        f = np.empty(3 * N)
        e = np.empty(N)
        check(_lib.load().mmx_term_field(self._h, index, float(t), dptr(f), dptr(e)))
        return f, e

    def _add(self, index, name):
        t = Term(self, check(index), name)
        self.interactions.append(t)
        return t

    @property
    def spin(self):
        out = np.empty(3 * self.n_total)
        check(_lib.load().mmx_get_m(self._h, dptr(out)))
        return out

    @property
    def field(self):
        return effective_field(self)


# ---- the reference's free functions ----------------------------------------
def set_Ms(sim, Ms):
    a = init_scalar(sim.mesh, Ms)
    if np.isnan(a).any():
        raise ValueError("NaN is given by the input Ms!")
    check(_lib.load().mmx_set_Ms(sim._h, dptr(a)))
    sim.mu0_Ms = a * mu_0
    return True


def set_pinning(sim, ids):
    p = np.ascontiguousarray(init_scalar(sim.mesh, ids, dtype=np.float64) != 0, dtype=np.uint8)
    check(_lib.load().mmx_set_pinning(sim._h, p.ctypes.data))
    return True


def init_m0(sim, m0, norm=True):
    m = np.ascontiguousarray(init_vector(sim.mesh, m0), dtype=np.float64)
    check(_lib.load().mmx_init_m0(sim._h, dptr(m), int(bool(norm))))
    return True


def add_exch(sim, A, name="exch"):
    L = _lib.load()
    if np.ndim(A) == 0 and not callable(A):
        return sim._add(L.mmx_add_exch(sim._h, None, float(A), float(A), float(A)), name)
    if isinstance(A, tuple) and len(A) == 3:
        return sim._add(L.mmx_add_exch(sim._h, None, *map(float, A)), name)
    a = init_scalar(sim.mesh, A)
    return sim._add(L.mmx_add_exch(sim._h, dptr(a), 0.0, 0.0, 0.0), name)


def add_dmi(sim, D, ft=None, name=None, type="bulk"):
    """add_dmi(sim, D; name="dmi", type): src/micro/add_field.jl:293-335; with a
    time function add_dmi(sim, D, ft; name="time_dmi") (add_field.jl:360-393):
    always bulk DMI, D (number, 3-tuple, or per cell for all three
    components) scaled by ft(t) at every evaluation time."""
    L = _lib.load()
    if ft is not None:
        name = name or "time_dmi"
        if np.ndim(D) == 0 and not callable(D):
            idx = L.mmx_add_dmi(sim._h, _lib.DMI_BULK, None, None, None, float(D), float(D), float(D))
        elif isinstance(D, tuple) and len(D) == 3:
            idx = L.mmx_add_dmi(sim._h, _lib.DMI_BULK, None, None, None, *map(float, D))
        else:
            a = init_scalar(sim.mesh, D)
            idx = L.mmx_add_dmi(sim._h, _lib.DMI_BULK, dptr(a), dptr(a), dptr(a), 0.0, 0.0, 0.0)
        t = sim._add(idx, name)
        _set_time_fn(sim, t, ft)
        # add_field.jl:384-388: time_D<name> = ft(integrator.t)
        t.items_after.append((f"time_D{name}", "<J/m^2>", lambda o, sv: float(ft(sv.t))))
        return t
    name = name or "dmi"
    if type == "interfacial":
        if np.ndim(D) == 0 and not callable(D):
            return sim._add(L.mmx_add_dmi(sim._h, _lib.DMI_INTERFACIAL, None, None, None, float(D), 0.0, 0.0), name)
        a = init_scalar(sim.mesh, D)
        return sim._add(L.mmx_add_dmi(sim._h, _lib.DMI_INTERFACIAL, dptr(a), None, None, 0.0, 0.0, 0.0), name)
    if type not in ("bulk", "D2d"):
        raise ValueError("Supported DMI type: interfacial bulk D2d")
    if type == "D2d":
        if np.ndim(D) != 0:
            raise ValueError("D2d only support uniform DMI!")
        idx = L.mmx_add_dmi(sim._h, _lib.DMI_BULK, None, None, None, -float(D), float(D), 0.0)
    elif np.ndim(D) == 0 and not callable(D):
        idx = L.mmx_add_dmi(sim._h, _lib.DMI_BULK, None, None, None, float(D), float(D), float(D))
    elif isinstance(D, tuple) and len(D) == 3 and all(np.ndim(d) == 0 and not callable(d) for d in D):
        idx = L.mmx_add_dmi(sim._h, _lib.DMI_BULK, None, None, None, *map(float, D))
    else:
        if isinstance(D, tuple) and len(D) == 3:
            Dx, Dy, Dz = (init_scalar(sim.mesh, d) for d in D)
        else:
            Dx = Dy = Dz = init_scalar(sim.mesh, D)
        idx = L.mmx_add_dmi(sim._h, _lib.DMI_BULK, dptr(Dx), dptr(Dy), dptr(Dz), 0.0, 0.0, 0.0)
    t = sim._add(idx, name)
    t.kind = "bulk_dmi"  # its parameter gradient has three components per cell
    return t


def add_anis(sim, Ku, axis=(0, 0, 1), name="anis"):
    """add_anis: src/micro/add_field.jl:430-465.  A number Ku stays a uniform
    parameter (the same value in every cell, without a per-cell array)."""
    L = _lib.load()
    uniform = np.ndim(Ku) == 0 and not callable(Ku)
    Kua = None if uniform else init_scalar(sim.mesh, Ku)
    Ku0 = float(Ku) if uniform else 0.0
    if isinstance(axis, tuple):
        return sim._add(L.mmx_add_anis(sim._h, dptr(Kua), Ku0, None, *map(float, axis)), name)
    ax = np.ascontiguousarray(init_vector(sim.mesh, axis))
    return sim._add(L.mmx_add_anis(sim._h, dptr(Kua), Ku0, dptr(ax), 0.0, 0.0, 0.0), name)


def update_anis(sim, Ku, name="anis"):
    """update_anis(sim, Ku; name): src/micro/add_field.jl:514-525 -- new per-cell Ku for the
    anisotropy term of that name (silently nothing when there is none, as the reference)."""
    for t in sim.interactions:
        if t.name == name:
            set_term_param(sim, t, init_scalar(sim.mesh, Ku))
            return None
    return None


def add_cubic_anis(sim, Kc, axis1=(1, 0, 0), axis2=(0, 1, 0), name="cubic"):
    """add_cubic_anis: src/micro/add_field.jl:541-574 (a number Kc stays uniform)."""
    L = _lib.load()
    uniform = np.ndim(Kc) == 0 and not callable(Kc)
    a = None if uniform else init_scalar(sim.mesh, Kc)
    a1 = np.array(axis1, dtype=np.float64)
    a2 = np.array(axis2, dtype=np.float64)
    return sim._add(L.mmx_add_cubic_anis(sim._h, dptr(a), float(Kc) if uniform else 0.0, dptr(a1), dptr(a2)), name)


_TORQUEFN = None


def add_torque(sim, fun, name="torque"):
    """add_torque(sim, fun; name): src/micro/add_field.jl:998-1006 -- fun(y, m, t) fills the
    torque y for the state m (numpy views, 3N AoS) at every evaluation time; the term is
    H = (1/gamma) m x y (torque_kernel!, kernels.jl:639-651).  No table column."""
    global _TORQUEFN
    if _TORQUEFN is None:
        _TORQUEFN = C.CFUNCTYPE(None, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_void_p)
    n = 3 * sim.n_total

    def cb(t, m, y, user):
        # an exception cannot cross the C frames: it is kept and raised by the
        # check() of the library call that ran this evaluation
        try:
            fun(np.ctypeslib.as_array(y, shape=(n,)), np.ctypeslib.as_array(m, shape=(n,)), t)
        except BaseException as e:  # noqa: BLE001
            _lib.pending_errors.append(e)

    f = _TORQUEFN(cb)
    t = sim._add(_lib.load().mmx_add_torque(sim._h, C.cast(f, C.c_void_p), None), name)
    t._ft = f  # the library keeps only the pointer
    t.energy_unit = None
    return t


def _six_component(mesh, data):
    """_init_six_component_field: src/micro/add_field.jl:1413-1435 -- a (6, nx, ny, nz)
    array (Fortran order, as Julia holds it) or f(i,j,k,dx,dy,dz) / f(x,y,z) with
    centred coordinates (init_vector6!, src/helper.jl:104-138); 6N values out."""
    N = mesh.n_total
    if callable(data):
        import inspect

        nargs = len(inspect.signature(data).parameters)
        i, j, k = _ijk(mesh)
        out = np.zeros((N, 6))
        for q, (a, b, c) in enumerate(zip(i, j, k)):
            if nargs == 6:
                v = data(a, b, c, mesh.dx, mesh.dy, mesh.dz)
            elif nargs == 3:
                v = data((a - 0.5 - mesh.nx / 2) * mesh.dx, (b - 0.5 - mesh.ny / 2) * mesh.dy,
                         (c - 0.5 - mesh.nz / 2) * mesh.dz)
            else:
                raise ValueError("The input function should have either 6 or 3 arguments.")
            if v is not None:
                out[q] = v
        return np.ascontiguousarray(out.reshape(-1))
    a = np.asarray(data, dtype=np.float64)
    if a.shape != (6, mesh.nx, mesh.ny, mesh.nz):
        raise ValueError(f"Array size must be (6, nx, ny, nz) = (6, {mesh.nx}, {mesh.ny}, {mesh.nz})")
    return np.ascontiguousarray(a.reshape(-1, order="F"))


def add_magnetoelastic(sim, model="tensor", lambda_s=0.0, B1=0.0, B2=0.0, sigma=None, strain=None,
                       name="magnetoelastic"):
    """add_magnetoelastic(sim; model, lambda_s, B1, B2, sigma, strain, name) (add_mel):
    src/micro/add_field.jl:1367-1405."""
    model = str(model).lstrip(":")
    if model not in ("tensor", "cubic"):
        raise ValueError(f"Unknown model: {model}. Use :tensor or :cubic")
    data = _six_component(sim.mesh, sigma if model == "tensor" else strain)
    kind = _lib.MEL_TENSOR if model == "tensor" else _lib.MEL_CUBIC
    return sim._add(_lib.load().mmx_add_magnetoelastic(sim._h, kind, dptr(data), float(lambda_s), float(B1),
                                                       float(B2)), name)


add_mel = add_magnetoelastic


def add_sahe_torque(sim, sigma_s, sigma_sa1=(0, 0, 0), sigma_sa2=(0, 0, 0), a1=(1, 0, 0), a2=(0, 0, 1),
                    a3=(0, 0, 1), beta=0, name="she"):
    """add_sahe_torque(sim, sigma_s; sigma_sa1, sigma_sa2, a1, a2, a3, beta, name):
    src/micro/add_field.jl:892-936 (tuples, arrays or functions; no table column)."""
    vs = [np.ascontiguousarray(init_vector(sim.mesh, v)) for v in (sigma_s, sigma_sa1, sigma_sa2, a1, a2, a3)]
    t = sim._add(_lib.load().mmx_add_sahe_torque(sim._h, *(dptr(v) for v in vs), float(beta)), name)
    t.energy_unit = None
    return t


def add_twin_mono_anis(sim, Ka=0, Kb=0, Kaa=0, Kbb=0, Kab=0, Ku=0, axis_a=(1, 0, 0), axis_b=(0, 1, 0),
                       axis_u111=(0, 0, 1), name="twin_mono_anis"):
    """add_twin_mono_anis(sim; Ka, Kb, Kaa, Kbb, Kab, Ku, axis_a, axis_b, axis_u111, name):
    src/micro/add_field.jl:629-660; axes are tuples, arrays or functions (init_vector)."""
    K = np.array([Ka, Kb, Kaa, Kbb, Kab, Ku], dtype=np.float64)
    ax = [np.ascontiguousarray(init_vector(sim.mesh, v)) for v in (axis_a, axis_b, axis_u111)]
    return sim._add(_lib.load().mmx_add_twin_mono_anis(sim._h, dptr(K), *(dptr(a) for a in ax)), name)


def add_exch_int(sim, J, k1=1, k2=-1, name="exch_int"):
    """add_exch_int(sim, J; k1, k2, name): src/micro/add_field.jl:789-816.  J is
    expanded on the one-layer mesh (nx, ny, 1) as the reference does."""
    m = sim.mesh
    flat = FDMesh(dx=m.dx, dy=m.dy, dz=m.dz, nx=m.nx, ny=m.ny, nz=1)
    Js = np.ascontiguousarray(init_scalar(flat, J))
    return sim._add(_lib.load().mmx_add_exch_int(sim._h, dptr(Js), int(k1), int(k2)), name)


def add_dmi_int(sim, D, k1=1, k2=-1, name="dmi_int"):
    """add_dmi_int(sim, D; k1, k2, name): src/micro/add_field.jl:835-860 (D a 3-tuple)."""
    if not (isinstance(D, tuple) and len(D) == 3):
        raise ValueError("add_dmi_int takes D as a 3-tuple")
    return sim._add(_lib.load().mmx_add_dmi_int(sim._h, *map(float, D), int(k1), int(k2)), name)


def add_hex_anis(sim, K1=0.0, K2=0.0, K3=0.0, name="hex"):
    """add_hex_anis(sim; K1, K2, K3, name): src/micro/add_field.jl:590-606."""
    return sim._add(_lib.load().mmx_add_hex_anis(sim._h, float(K1), float(K2), float(K3)), name)


def add_zeeman(sim, H0, ft=None, name=None):
    """add_zeeman(sim, H0; name="zeeman"): src/micro/add_field.jl:11-45, and with a
    time function add_zeeman(sim, H0, ft; name="timezeeman"): add_field.jl:161-200;
    ft(t) returns a number or a 3-tuple scaling (Hx, Hy, Hz)."""
    L = _lib.load()
    if ft is not None:
        return _add_time_zeeman(sim, H0, ft, name or "timezeeman")
    name = name or "zeeman"
    if isinstance(H0, tuple) and len(H0) == 3:
        t = sim._add(L.mmx_add_zeeman(sim._h, None, *map(float, H0)), name)
        t.H0 = tuple(map(float, H0))
    else:
        H = np.ascontiguousarray(init_vector(sim.mesh, H0))
        t = sim._add(L.mmx_add_zeeman(sim._h, dptr(H), 0.0, 0.0, 0.0), name)
        t.H0 = (0.0, 0.0, 0.0)
    # add_field.jl:28-41: Hx, Hy, Hz (prefixed unless the name is "zeeman") = H0
    cols = ("Hx", "Hy", "Hz") if name == "zeeman" else tuple(f"{name}_H{c}" for c in "xyz")
    t.items_before.append((cols, ("<A/m>",) * 3, lambda o, sv, t=t: t.H0))
    return t


def _surface_x(x, y, z, b, c, sigma):
    """Field of a charged rectangle (half sides b, c) in the plane x = 0 at (x, y, z):
    field_x_surface, src/micro/demagc.jl:1-26 (F = atan(yz/(xr)), G = -log|r + z|)."""

    def F(x, y, z):
        r = np.sqrt(x * x + y * y + z * z)
        return np.arctan(y * z / (x * r))

    def G(x, y, z):
        r = np.sqrt(x * x + y * y + z * z)
        return -np.log(np.abs(r + z))

    def comb(f, x, y, z, b, c):
        return (f(x, y + b, z + c) - f(x, y - b, z + c) - f(x, y + b, z - c) + f(x, y - b, z - c)) / (4 * math.pi)

    return (-comb(F, x, y, z, b, c) * sigma, -comb(G, x, y, z, b, c) * sigma, -comb(G, x, z, y, c, b) * sigma)


def demag_charge_field(m, Ms, x=(0, 0), y=(0, 0), z=(0, 0)):
    """compute_demag_charge_field, src/micro/demagc.jl:38-83: the field that removes the
    surface charges (x-, x+), (y-, y+), (z-, z+) (in units of Ms) of mesh m, at the cell
    centres (lengths in nm, as the reference); 3N values, AoS."""
    dx, dy, dz = m.dx * 1e9, m.dy * 1e9, m.dz * 1e9
    Lx, Ly, Lz = m.nx * dx, m.ny * dy, m.nz * dz
    a, b, c = Lx / 2, Ly / 2, Lz / 2
    K, J, I = np.meshgrid(np.arange(1, m.nz + 1), np.arange(1, m.ny + 1), np.arange(1, m.nx + 1), indexing="ij")
    H = [np.zeros(I.shape) for _ in range(3)]
    (xm, xp), (ym, yp), (zm, zp) = ((v[0] * Ms, v[1] * Ms) for v in (x, y, z))
    if xm != 0 or xp != 0:
        X, Y, Z = (I - 0.5) * dx, (J - 0.5) * dy - b, (K - 0.5) * dz - c
        f1, f2 = _surface_x(X, Y, Z, b, c, xm), _surface_x(X - Lx, Y, Z, b, c, xp)
        for q in range(3):
            H[q] += f1[q] + f2[q]
    if ym != 0 or yp != 0:
        X, Y, Z = (I - 0.5) * dx - a, (J - 0.5) * dy, (K - 0.5) * dz - c
        g1, g2 = _surface_x(Y, Z, X, c, a, ym), _surface_x(Y - Ly, Z, X, c, a, yp)  # (Hy, Hz, Hx)
        for q, src in ((1, 0), (2, 1), (0, 2)):
            H[q] += g1[src] + g2[src]
    if zm != 0 or zp != 0:
        X, Y, Z = (I - 0.5) * dx - a, (J - 0.5) * dy - b, (K - 0.5) * dz
        h1, h2 = _surface_x(Z, X, Y, a, b, zm), _surface_x(Z - Lz, X, Y, a, b, zp)  # (Hz, Hx, Hy)
        for q, src in ((2, 0), (0, 1), (1, 2)):
            H[q] += h1[src] + h2[src]
    return np.ascontiguousarray(np.stack([h.reshape(-1) for h in H], axis=1).reshape(-1))


def rm_demag_charges(sim, Ms, x=(0, 0), y=(0, 0), z=(0, 0), name="demagc"):
    """rm_demag_charges(sim, Ms; x, y, z, name): src/micro/add_field.jl:1277-1300 -- the
    field of demag_charge_field as a static Zeeman term (energy column only)."""
    for v, n in ((x, "x"), (y, "y"), (z, "z")):
        if len(v) != 2:
            raise ValueError(f"{n} must be a 2-element tuple")
    field = demag_charge_field(sim.mesh, Ms, x, y, z)
    t = sim._add(_lib.load().mmx_add_zeeman(sim._h, dptr(field), 0.0, 0.0, 0.0), name)
    t.H0 = (0.0, 0.0, 0.0)
    return t


def update_zeeman(sim, H0, name="zeeman"):
    L = _lib.load()
    for t in sim.interactions:
        if t.name == name:
            if isinstance(H0, tuple) and len(H0) == 3:
                check(L.mmx_update_zeeman(sim._h, t.index, None, *map(float, H0)))
                t.H0 = tuple(map(float, H0))  # add_field.jl:72-76
            else:
                H = np.ascontiguousarray(init_vector(sim.mesh, H0))
                check(L.mmx_update_zeeman(sim._h, t.index, dptr(H), 0.0, 0.0, 0.0))
            return True
    raise ValueError(f"no Zeeman term named {name!r}")


def add_demag(sim, name="demag", Nx=0, Ny=0, Nz=0, fft=True):
    if not fft:
        raise ValueError("the direct O(N^2) demag is not part of this path; use fft=True")
    return sim._add(_lib.load().mmx_add_demag(sim._h, int(Nx), int(Ny), int(Nz)), name)


def add_thermal_noise(sim, Temp, name="thermal", T0=0.0, seed=0):
    L = _lib.load()
    if np.ndim(Temp) == 0 and not callable(Temp):
        t = sim._add(L.mmx_add_thermal_noise(sim._h, None, float(Temp), float(T0), int(seed)), name)
        Tm = float(Temp)
    else:
        a = init_scalar(sim.mesh, Temp)
        t = sim._add(L.mmx_add_thermal_noise(sim._h, dptr(a), 0.0, float(T0), int(seed)), name)
        Tm = float(np.mean(a))
    # add_field.jl:763-766: T_<name> = mean temperature (scaling factor 1)
    t.items_after.append((f"T_{name}", "<K>", lambda o, sv: Tm))
    return t


_TIMEFN = None
_TIMEFN3 = None


def _add_time_zeeman(sim, H0, ft, name):
    global _TIMEFN3
    import ctypes as C

    f0 = ft(0.0)
    n = 1 if np.ndim(f0) == 0 else len(f0)
    if n not in (1, 3):
        raise ValueError("The time_fun function should return either 1 value (scalar) or 3 values (as a tuple).")
    if _TIMEFN3 is None:
        _TIMEFN3 = C.CFUNCTYPE(None, C.c_double, C.POINTER(C.c_double), C.c_void_p)

    def factors(t, f, user):
        v = ft(t)
        if n == 1:
            f[0] = f[1] = f[2] = float(v)
        else:
            f[0], f[1], f[2] = map(float, v)

    cb = _TIMEFN3(factors)
    L = _lib.load()
    if isinstance(H0, tuple) and len(H0) == 3:
        idx = L.mmx_add_zeeman_time(sim._h, None, *map(float, H0), C.cast(cb, C.c_void_p), None)
    else:
        H = np.ascontiguousarray(init_vector(sim.mesh, H0))
        idx = L.mmx_add_zeeman_time(sim._h, dptr(H), 0.0, 0.0, 0.0, C.cast(cb, C.c_void_p), None)
    t = sim._add(idx, name)
    t._ft = cb  # the library keeps only the pointer
    t.energy_unit = "J"  # add_field.jl:192-194
    if isinstance(H0, tuple) and len(H0) == 3:  # add_field.jl:182-190: H0 .* f(t)
        H0f = tuple(map(float, H0))

        def cols(o, sv):
            v = ft(sv.t)
            f = (v, v, v) if n == 1 else v
            return tuple(h * float(x) for h, x in zip(H0f, f))

        t.items_before.append((tuple(f"{name}_H{c}" for c in "xyz"), ("<A/m>",) * 3, cols))
    return t


def _set_time_fn(sim, term, ft):
    """ft(t) -> float as the term's ufun (called at every evaluation time)."""
    global _TIMEFN
    import ctypes as C

    if _TIMEFN is None:
        _TIMEFN = C.CFUNCTYPE(C.c_double, C.c_double, C.c_void_p)
    cb = _TIMEFN(lambda t, user: float(ft(t)))
    term._ft = cb  # the library keeps only the pointer
    check(_lib.load().mmx_set_term_time_fn(sim._h, term.index, C.cast(cb, C.c_void_p), None))


def set_alpha(sim, alpha):
    """set_alpha(sim, alpha): src/sim.jl:213-222 (SpatialLLG driver only; a
    number, an array or a function of (i,j,k,dx,dy,dz) / (x,y,z))."""
    if sim.driver_name != "SpatialLLG":
        import warnings

        warnings.warn("set_alpha only works for the SpatialLLG driver")
        return None
    a = init_scalar(sim.mesh, alpha)
    sim.driver.alpha = a
    return True


def add_sot(sim, aj, bj, p, name="sot"):
    """add_sot(sim, aj, bj, p; name): src/micro/add_field.jl:951-966.  The torque
    enters as the field H = (1/gamma)(aj m x p + bj p) (kernels.jl:619-633)."""
    L = _lib.load()
    uniform = np.ndim(aj) == 0 and not callable(aj)
    a = None if uniform else init_scalar(sim.mesh, aj)
    t = sim._add(L.mmx_add_sot(sim._h, dptr(a), float(aj) if uniform else 0.0, float(bj), *map(float, p)), name)
    t.energy_unit = None  # the torques register no table column
    return t


def add_stt(sim, model="slonczewski", **params):
    """add_stt(sim; model, kwargs...): src/micro/add_field.jl:1134-1147 with the
    Zhang-Li (:1150-1192) and Slonczewski (:1195-1227) models."""
    model = str(model).lstrip(":")
    if "J" not in params:
        raise ValueError("Current density parameter `J` is required.")
    if model == "zhang_li":
        return _add_zhang_li(sim, params.pop("name", "zhangli"), params)
    if model == "slonczewski":
        return _add_slonczewski(sim, params.pop("name", "slonczewski"), params)
    raise ValueError(f"Unsupported STT model: {model}. Supported models are: [:zhang_li, :slonczewski]")


def _add_zhang_li(sim, name, params):
    has_b = "b" in params
    has_P = "P" in params and "Ms" in params
    if has_b:
        b = params["b"]
    elif has_P:
        b = params["P"] * mu_B / (c_e * params["Ms"])
    else:
        raise ValueError("Zhang-Li model requires either `b` or `(P, Ms)` parameters")
    L = _lib.load()
    xi = params.get("xi", 0)
    xi_uniform = np.ndim(xi) == 0 and not callable(xi)
    xa = None if xi_uniform else init_scalar(sim.mesh, xi)
    J = params["J"]
    if isinstance(J, tuple) and len(J) == 3:
        Ja, J3 = None, tuple(map(float, J))
    else:
        Ja, J3 = np.ascontiguousarray(init_vector(sim.mesh, J)), (0.0, 0.0, 0.0)
    t = sim._add(L.mmx_add_stt_zhangli(sim._h, dptr(Ja), *J3, float(b), dptr(xa),
                                       float(xi) if xi_uniform else 0.0, int(has_P and not has_b)), name)
    t.energy_unit = None
    if "ft" in params:
        _set_time_fn(sim, t, params["ft"])
    return t


def _add_slonczewski(sim, name, params):
    for k in ("tf", "Ms", "p", "P"):
        if k not in params:
            raise ValueError(f"Parameter {k} is not provided")
    xi = params.get("xi", 0)
    Lambda = params.get("Lambda", 2.0)
    beta = h_bar / (mu_0 * params["Ms"] * c_e * params["tf"])
    L = _lib.load()
    J = params["J"]
    uniform = np.ndim(J) == 0 and not callable(J)
    Ja = None if uniform else init_scalar(sim.mesh, J)
    t = sim._add(L.mmx_add_stt_slonczewski(sim._h, dptr(Ja), float(J) if uniform else 0.0, beta, float(Lambda),
                                           float(xi), float(params["P"]), *map(float, params["p"])), name)
    t.energy_unit = None
    if "ft" in params:
        _set_time_fn(sim, t, params["ft"])
    return t


def effective_field(sim, spin=None, t=0.0):
    """effective_field(sim, spin, t): src/micro/field.jl:496-505; returns sim.field."""
    if spin is not None:
        raise ValueError("pass the state through init_m0; the library evaluates the current spin")
    out = np.empty(3 * sim.n_total)
    check(_lib.load().mmx_effective_field(sim._h, float(t), dptr(out)))
    return out


def compute_system_energy(sim, t=0.0):
    """compute_system_energy: src/micro/field.jl:520-536; returns sum(term.energy) per term."""
    n = len(sim.interactions)
    out = np.empty(max(n, 1))
    check(_lib.load().mmx_compute_energies(sim._h, float(t), dptr(out)))
    return out[:n]


def llg_rhs(sim, spin=None, t=0.0):
    out = np.empty(3 * sim.n_total)
    m = None if spin is None else np.ascontiguousarray(spin, dtype=np.float64)
    check(_lib.load().mmx_llg_rhs(sim._h, float(t), dptr(m), dptr(out)))
    return out


def set_driver(sim, driver="LLG", integrator="DormandPrince", step=None, tol=1e-6):
    """set_driver: src/sim.jl:589-620 ("LLG" or "SpatialLLG" with an integrator,
    or "SD").  A change of driver creates a new one (create_driver): the
    SpatialLLG damping starts at 0.1 per cell, the LLG one is uniform again."""
    if driver not in ("LLG", "SD", "SpatialLLG"):
        raise ValueError("Supported drivers: LLG SpatialLLG SD")
    kind = _lib.SD if driver == "SD" else _INTEGRATORS[integrator]
    changed = sim.driver_name != driver
    sim.driver_name = driver
    if changed and driver == "SpatialLLG":
        set_alpha(sim, 0.1)
    elif changed and np.ndim(sim.driver.alpha) != 0:
        check(_lib.load().mmx_set_alpha(sim._h, None))
        sim.driver.alpha = 0.1
    st = _DEFAULT_STEP.get(kind, 0.0) if step is None else float(step)
    sim._integrator = Integrator(sim, kind, st, tol)
    check(_lib.load().mmx_set_integrator(sim._h, kind, st, float(tol)))


def run_step(sim, driver=None, n=1):
    check(_lib.load().mmx_advance_step(sim._h, int(n)))
    return True


advance_step = run_step


def _saver(sim):
    if getattr(sim, "saver", None) is None:
        from .io import DataSaver

        sim.saver = DataSaver.for_sim(sim)
    return sim.saver


def run_until(sim, t_end, save_data=False):
    """run_until(sim, t_end; save_data): src/sim.jl:547-579 (one table row when save_data)."""
    check(_lib.load().mmx_run_until(sim._h, float(t_end)))
    sv = _saver(sim)
    sv.t = float(t_end)
    if save_data:
        sv.write(sim)
    sv.nsteps += 1


def run_sim(sim, steps=10, dt=1e-10, save_data=True, save_m_every=-1, saver_item=None, call_back=None,
            output="ovf"):
    """run_sim: src/sim.jl:813-857 (tables and OVF snapshots; no VTK or live view)."""
    from .io import SaverItem, save_ovf

    sv = _saver(sim)
    if isinstance(saver_item, SaverItem):
        sv.items.append(saver_item)
    elif isinstance(saver_item, (list, tuple)):
        sv.items.extend(saver_item)
    folder = f"{sim.name}_LLG"
    if save_m_every >= 0:
        os.makedirs(folder, exist_ok=True)
    for i in range(0, steps + 1):
        run_until(sim, i * dt, save_data=save_data)
        if call_back is not None:
            call_back(sim, i * dt)
        if save_m_every > 0 and (i - 1) % save_m_every == 0 and output == "ovf":
            save_ovf(sim, os.path.join(folder, "m_%08d.ovf" % i))


def relax(sim, max_steps=10000, stopping_dmdt=0.01):
    st = C.c_int64()
    md = C.c_double()
    check(_lib.load().mmx_relax(sim._h, int(max_steps), float(stopping_dmdt), C.byref(st), C.byref(md)))
    return st.value, md.value


def average_m(sim):
    """average_m: src/sim.jl:251-255 (weighted by mu0 Ms)."""
    b = sim.spin.reshape(-1, 3)
    ms = sim.mu0_Ms
    return tuple((b * ms[:, None]).sum(axis=0) / ms.sum())


def demag_tensor(sim, term):
    out = np.empty(6 * sim.n_total)
    check(_lib.load().mmx_demag_tensor(sim._h, term.index, dptr(out)))
    return out.reshape(6, sim.n_total)


def set_term_param(sim, term, values):
    """Replace the per-cell scalar parameter of a term (A, D, Ku, Kc or T) in place."""
    v = init_scalar(sim.mesh, values)
    check(_lib.load().mmx_set_term_param(sim._h, term.index, dptr(v)))


# ---- differentiation form ----------------------------------------------------
def rk4_trajectory_rrule(sim, nsteps, every=50):
    """Run `nsteps` RK4 steps from the current state and return (m_T, pullback).

    The analogue of a ChainRules `rrule` for the map m0, Ms, {Ku, D, Kc, A} ->
    m(T): `pullback(dL_dmT)` takes the cotangent of the final state (3N, AoS)
    and returns {"m0": dL/dm0 (3N, AoS), "Ms": dL/dMs (N), term name: dL/dp
    per cell} for every term with a parameter field (anisotropy Ku, cubic Kc,
    interfacial DMI D, spatial exchange A: N values; bulk DMI: 3N values AoS,
    the Dx, Dy, Dz components, per-cell shares of the gradient when D is
    uniform).  The backward pass recomputes the trajectory from checkpoints
    kept every `every` steps; it may be called more than once.
    """
    L = _lib.load()
    check(L.mmx_adjoint_record(sim._h, int(nsteps), int(every)))
    mT = sim.spin
    N = sim.n_total

    def pullback(dL_dmT):
        lam = np.ascontiguousarray(np.asarray(dL_dmT, dtype=np.float64).reshape(-1))
        if lam.size != 3 * N:
            raise ValueError(f"expected {3 * N} cotangent values, got {lam.size}")
        g0 = np.empty(3 * N)
        check(L.mmx_adjoint_backward(sim._h, dptr(lam), dptr(g0)))
        out = {"m0": g0}
        gms = np.empty(N)
        check(L.mmx_adjoint_ms_grad(sim._h, dptr(gms)))
        out["Ms"] = gms
        for t in sim.interactions:
            g = np.empty(3 * N)  # bulk DMI: (dL/dDx, dL/dDy, dL/dDz) per cell; N values otherwise
            if L.mmx_adjoint_param_grad(sim._h, t.index, dptr(g)) >= 0:
                out[t.name] = g if getattr(t, "kind", None) == "bulk_dmi" else g[:N].copy()
        return out

    return mT, pullback


def llg_rhs_jvp(sim, v, spin=None, t=0.0):
    """(d dm/dt / dm) v for one llg_call_back evaluation (src/llg/llg.jl:45-58)
    at `spin` (or the current state): 3N AoS in, 3N AoS out."""
    N = sim.n_total
    v = np.ascontiguousarray(np.asarray(v, dtype=np.float64).reshape(-1))
    if v.size != 3 * N:
        raise ValueError(f"expected {3 * N} values, got {v.size}")
    m = None if spin is None else np.ascontiguousarray(np.asarray(spin, dtype=np.float64).reshape(-1))
    out = np.empty(3 * N)
    check(_lib.load().mmx_llg_rhs_jvp(sim._h, float(t), dptr(m), dptr(v), dptr(out)))
    return out


def llg_rhs_vjp(sim, a, spin=None, t=0.0):
    """(d dm/dt / dm)^T a for one llg_call_back evaluation at `spin` (or the
    current state): 3N AoS in, 3N AoS out."""
    N = sim.n_total
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1))
    if a.size != 3 * N:
        raise ValueError(f"expected {3 * N} values, got {a.size}")
    m = None if spin is None else np.ascontiguousarray(np.asarray(spin, dtype=np.float64).reshape(-1))
    out = np.empty(3 * N)
    check(_lib.load().mmx_llg_rhs_vjp(sim._h, float(t), dptr(m), dptr(a), dptr(out)))
    return out


def rotation_matrix(mx, my, mz):
    """rotation_matrix: src/eigen/util.jl:117-139 (third column along m)."""
    norm = np.sqrt(mx * mx + my * my + mz * mz)
    if norm < np.finfo(float).eps:
        return np.eye(3)
    theta = np.arccos(mz / norm)
    phi = np.arctan2(my, mx)
    return np.array([[np.cos(phi) * np.cos(theta), -np.sin(phi), np.sin(theta) * np.cos(phi)],
                     [np.sin(phi) * np.cos(theta), np.cos(phi), np.sin(theta) * np.sin(phi)],
                     [-np.sin(theta), 0.0, np.cos(theta)]])


def dynamic_matrix(sim, gamma=2.21e5, alpha=0.01):
    """dynamic_matrix(sim; gamma, alpha): ext/EnzymeExt.jl:32-59.  The Jacobian A
    of y = gamma m x H_eff(m) at the current state (H_eff without the constant
    Zeeman field's derivative) built column by column, then for every pair of
    cells the transverse 2x2 block of R_i^T A_ij R_j (2N x 2N).  The columns
    come from the library's JVP of the LLG right-hand side with alpha = 0
    (-gamma m x h there, so A = -J on the transverse directions); alpha is
    unused, as in the reference."""
    N = sim.n_total
    m0 = sim.spin.copy()
    d = sim.driver
    saved = (d.alpha, d.gamma, d.precession)
    try:
        d.alpha, d.gamma, d.precession = 0.0, float(gamma), True
        A = np.zeros((3 * N, 3 * N))
        e = np.zeros(3 * N)
        for i in range(3 * N):
            e[:] = 0.0
            e[i] = 1.0
            A[:, i] = -llg_rhs_jvp(sim, e, spin=m0)
    finally:
        d.alpha, d.gamma, d.precession = saved
    Rs = [rotation_matrix(*m0[3 * i:3 * i + 3]) for i in range(N)]
    Cm = np.zeros((2 * N, 2 * N))
    for i in range(N):
        for j in range(N):
            B = Rs[i].T @ A[3 * i:3 * i + 3, 3 * j:3 * j + 3] @ Rs[j]
            Cm[2 * i:2 * i + 2, 2 * j:2 * j + 2] = B[:2, :2]
    return Cm


# ---- batching form -------------------------------------------------------------
class Batch:
    """Many independent Sims advanced together (one launch per RK stage over all
    members without demag; members with demag grouped by geometry, one launch
    per demag pass and stage for each group; replayed from a HIP graph).  Each
    member must use the RungeKutta or Heun integrator with a step set and share
    precision and device with the others; after `advance(n)` each member is
    exactly where `run_step(sim, n=n)` would have left it."""

    def __init__(self, sims):
        self.sims = list(sims)
        arr = (C.c_void_p * len(self.sims))(*[s._h.value for s in self.sims])
        h = C.c_void_p()
        check(_lib.load().mmx_batch_create(C.byref(h), arr, len(self.sims)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            _lib.load().mmx_batch_destroy(h)
            self._h = None

    def advance(self, nsteps=1):
        check(_lib.load().mmx_batch_advance(self._h, int(nsteps)))

    def info(self):
        """Launch shape chosen by the set-up's timing: (cells per thread, grid,
        work items, candidate times in ms)."""
        c, g, w = C.c_int32(), C.c_int32(), C.c_int32()
        t = (C.c_float * 12)()
        check(_lib.load().mmx_batch_info(self._h, C.byref(c), C.byref(g), C.byref(w), t, 12))
        return c.value, g.value, w.value, list(t)

    def shape(self):
        """The whole tuned launch shape: dict(block_threads, cells_per_thread,
        grid, demag_groups)."""
        b, c, g, n = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
        check(_lib.load().mmx_batch_shape(self._h, C.byref(b), C.byref(c), C.byref(g), C.byref(n)))
        return dict(block_threads=b.value, cells_per_thread=c.value, grid=g.value, demag_groups=n.value)


# ---- nudged elastic band -----------------------------------------------------
def _slerp(x1, x2, t):
    """slerp_kernel!: src/neb/math.jl:45-61 (cells with theta == 0 stay zero)."""
    a = x1.reshape(-1, 3)
    b = x2.reshape(-1, 3)
    inner = np.clip(a[:, 0] * b[:, 0] + a[:, 1] * b[:, 1] + a[:, 2] * b[:, 2], -1.0, 1.0)
    theta = np.arccos(inner)
    res = np.zeros_like(a)
    nz = theta != 0.0
    s = np.sin(theta[nz])
    res[nz] = (np.sin((1 - t) * theta[nz]) / s)[:, None] * a[nz] + (np.sin(t * theta[nz]) / s)[:, None] * b[nz]
    return res.reshape(-1)


class NEB:
    """NEB(sim, init_images, frames_between_images; spring_constant, clib_image):
    src/neb/neb.jl:84-153 with the LLG driver (BS23, tol 1e-5, alpha 0.2,
    precession off).  The images are aligned as init_m0 does (normalised, zero
    where Ms = 0) and interpolated with slerp (init_images, neb.jl:155-175) on
    the host; stepping runs on the GPU."""

    def __init__(self, sim, init_images, frames_between_images, spring_constant=1.0e5, clib_image=-1):
        N = sim.n_total
        ms = sim.mu0_Ms
        aligned = []
        for img in init_images:
            v = init_vector(sim.mesh, img).reshape(-1, 3).copy()
            m2 = (v * v).sum(axis=1)
            nz = m2 > 0
            v[nz] *= (1 / np.sqrt(m2[nz]))[:, None]
            v[ms == 0] = 0.0
            aligned.append(v.reshape(-1))
        images = [aligned[0]]
        for i in range(len(aligned) - 1):
            if frames_between_images[i] > 0:
                for j in range(1, frames_between_images[i] + 1):
                    images.append(_slerp(aligned[i], aligned[i + 1], j / (frames_between_images[i] + 1)))
                images.append(aligned[i + 1])
        self.n_images = len(images) - 2
        if self.n_images <= 0:
            raise ValueError("The number of free images should larger than 1 (typically 10~20).")
        self.sim, self.N = sim, N
        arr = np.ascontiguousarray(np.concatenate(images))
        h = C.c_void_p()
        check(_lib.load().mmx_neb_create(C.byref(h), sim._h, dptr(arr), self.n_images, float(spring_constant),
                                         int(clib_image)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            _lib.load().mmx_neb_destroy(h)
            self._h = None

    def set_driver(self, alpha=0.2, gamma=2.21e5, precession=False, tol=1e-5):
        check(_lib.load().mmx_neb_set_driver(self._h, float(alpha), float(gamma), int(bool(precession)),
                                             float(tol)))

    def advance(self, nsteps=1):
        check(_lib.load().mmx_neb_advance(self._h, int(nsteps)))

    def relax(self, max_steps=10000, stopping_dmdt=0.01):
        st, md = C.c_int64(), C.c_double()
        check(_lib.load().mmx_neb_relax(self._h, int(max_steps), float(stopping_dmdt), C.byref(st), C.byref(md)))
        return st.value, md.value

    def _get(self):
        n, N = self.n_images, self.N
        imgs = np.empty(n * 3 * N)
        E = np.empty(n + 2)
        d = np.empty(n + 1)
        s = np.empty(5)
        check(_lib.load().mmx_neb_get(self._h, dptr(imgs), dptr(E), dptr(d), dptr(s)))
        return imgs.reshape(n, 3 * N), E, d, s

    @property
    def spin(self):
        return self._get()[0]

    @property
    def energy(self):
        return self._get()[1]

    @property
    def distance(self):
        return self._get()[2]

    @property
    def state(self):
        s = self._get()[3]
        return dict(t=s[0], step=s[1], step_next=s[2], nsteps=int(s[3]), nfevals=int(s[4]))
