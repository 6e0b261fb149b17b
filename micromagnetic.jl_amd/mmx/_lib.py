"""ctypes binding of libmmx (the C ABI declared in include/mmx.h).

The library is the product: there is no fallback.  Loading fails loudly when
libmmx.so has not been built, and every call that returns a negative status
raises MmxError with the library's own message.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(HERE)
LIB_PATH = os.environ.get("MMX_LIBRARY") or os.path.join(PKG_ROOT, "libmmx.so")
REPO_ROOT = os.path.dirname(PKG_ROOT)
HEADER = os.path.join(REPO_ROOT, "include", "mmx.h")

OK, EINVAL, EHIP, ENAN, ETRIES, ESTATE, ENOMEM = 0, -1, -2, -3, -4, -5, -6
F64, F32 = 64, 32
DMI_BULK, DMI_INTERFACIAL = 0, 1
MEL_TENSOR, MEL_CUBIC = 0, 1
RK4, HEUN, DOPRI54, BS23, SD = 0, 1, 2, 3, 4
PASSES = ("P1", "P2", "P3", "P4", "P5", "LOCAL")


class MmxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libmmx error {code}: {msg}")
        self.code = code


class MeshDesc(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("pbc", C.c_int32),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
                ("precision", C.c_int32), ("device", C.c_int32)]


_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)
_vp = C.c_void_p

# name -> argument types (all return int)
SIGNATURES = {
    "mmx_create": [C.POINTER(_vp), C.POINTER(MeshDesc)],
    "mmx_destroy": [_vp],
    "mmx_set_Ms": [_vp, _dp],
    "mmx_set_pinning": [_vp, C.c_void_p],
    "mmx_init_m0": [_vp, _dp, C.c_int],
    "mmx_get_m": [_vp, _dp],
    "mmx_add_exch": [_vp, _dp, C.c_double, C.c_double, C.c_double],
    "mmx_add_dmi": [_vp, C.c_int, _dp, _dp, _dp, C.c_double, C.c_double, C.c_double],
    "mmx_add_anis": [_vp, _dp, C.c_double, _dp, C.c_double, C.c_double, C.c_double],
    "mmx_add_cubic_anis": [_vp, _dp, C.c_double, _dp, _dp],
    "mmx_add_zeeman": [_vp, _dp, C.c_double, C.c_double, C.c_double],
    "mmx_update_zeeman": [_vp, C.c_int, _dp, C.c_double, C.c_double, C.c_double],
    "mmx_add_demag": [_vp, C.c_int, C.c_int, C.c_int],
    "mmx_add_thermal_noise": [_vp, _dp, C.c_double, C.c_double, C.c_uint64],
    "mmx_add_sot": [_vp, _dp] + [C.c_double] * 5,
    "mmx_add_stt_slonczewski": [_vp, _dp] + [C.c_double] * 8,
    "mmx_add_stt_zhangli": [_vp, _dp] + [C.c_double] * 4 + [_dp, C.c_double, C.c_int],
    "mmx_set_term_time_fn": [_vp, C.c_int, C.c_void_p, C.c_void_p],
    "mmx_add_zeeman_time": [_vp, _dp, C.c_double, C.c_double, C.c_double, C.c_void_p, C.c_void_p],
    "mmx_add_hex_anis": [_vp, C.c_double, C.c_double, C.c_double],
    "mmx_add_exch_int": [_vp, _dp, C.c_int, C.c_int],
    "mmx_add_twin_mono_anis": [_vp, _dp, _dp, _dp, _dp],
    "mmx_add_sahe_torque": [_vp, _dp, _dp, _dp, _dp, _dp, _dp, C.c_double],
    "mmx_add_magnetoelastic": [_vp, C.c_int, _dp, C.c_double, C.c_double, C.c_double],
    "mmx_add_torque": [_vp, C.c_void_p, C.c_void_p],
    "mmx_compute_energies_noalloc": [_vp, C.c_double, C.c_void_p],
    "mmx_adjoint_ms_grad": [_vp, _dp],
    "mmx_llg_rhs_jvp": [_vp, C.c_double, _dp, _dp, _dp],
    "mmx_llg_rhs_vjp": [_vp, C.c_double, _dp, _dp, _dp],
    "mmx_add_dmi_int": [_vp, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int],
    "mmx_set_alpha": [_vp, _dp],
    "mmx_num_terms": [_vp],
    "mmx_effective_field": [_vp, C.c_double, _dp],
    "mmx_term_field": [_vp, C.c_int, C.c_double, _dp, _dp],
    "mmx_compute_energies": [_vp, C.c_double, _dp],
    "mmx_set_llg": [_vp, C.c_double, C.c_double, C.c_int],
    "mmx_llg_rhs": [_vp, C.c_double, _dp, _dp],
    "mmx_set_integrator": [_vp, C.c_int, C.c_double, C.c_double],
    "mmx_advance_step": [_vp, C.c_int64],
    "mmx_run_until": [_vp, C.c_double],
    "mmx_relax": [_vp, C.c_int64, C.c_double, _i64p, _dp],
    "mmx_get_state": [_vp, _dp, _dp, _dp, _i64p, _i64p],
    "mmx_demag_tensor": [_vp, C.c_int, _dp],
    "mmx_demag_geometry": [_vp, _i32p, _i32p, _i32p, _i32p],
    "mmx_synchronize": [_vp],
    "mmx_profile": [_vp, C.c_int],
    "mmx_profile_read": [_vp, _dp, _i64p],
    "mmx_diag_copy": [C.c_int, C.c_int64, C.c_int, C.c_int, _dp],
    "mmx_set_term_param": [_vp, C.c_int, _dp],
    "mmx_adjoint_record": [_vp, C.c_int64, C.c_int64],
    "mmx_adjoint_backward": [_vp, _dp, _dp],
    "mmx_adjoint_param_grad": [_vp, C.c_int, _dp],
    "mmx_batch_create": [C.POINTER(_vp), C.POINTER(_vp), C.c_int],
    "mmx_batch_destroy": [_vp],
    "mmx_batch_advance": [_vp, C.c_int64],
    "mmx_batch_info": [_vp, _i32p, _i32p, _i32p, C.POINTER(C.c_float), C.c_int32],
    "mmx_batch_shape": [_vp, _i32p, _i32p, _i32p, _i32p],
    "mmx_neb_create": [C.POINTER(_vp), _vp, _dp, C.c_int, C.c_double, C.c_int],
    "mmx_neb_destroy": [_vp],
    "mmx_neb_set_driver": [_vp, C.c_double, C.c_double, C.c_int, C.c_double],
    "mmx_neb_advance": [_vp, C.c_int64],
    "mmx_neb_relax": [_vp, C.c_int64, C.c_double, _i64p, _dp],
    "mmx_neb_get": [_vp, _dp, _dp, _dp, _dp],
}

_lib = None


def load(path=LIB_PATH):
    """Load libmmx.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise OSError(f"libmmx.so not found at {path}: build it first "
                      "(python -c 'import __graft_entry__ as g; g.build()')")
    L = C.CDLL(path)
    for name, args in SIGNATURES.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    L.mmx_last_error.argtypes = []
    L.mmx_last_error.restype = C.c_char_p
    _lib = L
    return L


# exceptions raised inside host callbacks (add_torque functions) during a
# library call; the next check() raises the first of them
pending_errors = []


def check(rc):
    if pending_errors:
        e = pending_errors[0]
        pending_errors.clear()
        raise e
    if rc < 0:
        raise MmxError(rc, load().mmx_last_error().decode(errors="replace"))
    return rc


def dptr(a):
    """Pointer to a C-contiguous float64 array (the ABI reads plain dense arrays)."""
    if a is None:
        return None
    if a.dtype != "float64" or not a.flags["C_CONTIGUOUS"]:
        raise TypeError("arrays passed to libmmx must be C-contiguous float64")
    return a.ctypes.data_as(_dp)
