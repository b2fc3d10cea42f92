"""ctypes binding of liblf2mi.so (the C ABI declared in include/lf2mi.h).

The library is loaded from the package directory only; a missing library raises
immediately -- there is no fallback implementation.
"""
from __future__ import annotations

import ctypes
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "liblf2mi.so")

F32 = 0
F64 = 1
FFT_FORWARD = 1
FFT_BACKWARD = -1
PROJ_SCALAR, PROJ_VECTOR, PROJ_VECVEC, PROJ_VECTOR_NGP, PROJ_SYMTENSOR_NGP = 0, 1, 2, 3, 4
DEP_UNFOLDED, DEP_PERIODIC, DEP_PERIODIC_SET = 0, 1, 2
ACC_NATIVE, ACC_FIXED64 = 0, 1
PARTS_ID, PARTS_MASS = 1, 2
MASS_INFO, MASS_PART, MASS_PART_F32 = 0, 1, 2
# updateVel / moveParticles statistics (particles/LATfield2_Particles.hpp:81-86)
PQ_V2, PQ_MVX, PQ_MVY, PQ_MVZ, PQ_MASS, PQ_EKIN = 0, 1, 2, 3, 4, 5
SUM, MIN, MAX, SUM_LOCAL, MIN_LOCAL, MAX_LOCAL = 1, 2, 4, 8, 16, 32

_vp = ctypes.c_void_p
_i = ctypes.c_int
_i64 = ctypes.c_int64
_d = ctypes.c_double
_u64 = ctypes.c_uint64
_dp = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)
_i32p = ctypes.POINTER(ctypes.c_int32)

# name -> argtypes (restype int for all)
SIGNATURES = {
    "lf2mi_version": [],
    "lf2mi_last_error": [],
    "lf2mi_device_init": [_i],
    "lf2mi_synchronize": [],
    "lf2mi_device_alloc": [ctypes.POINTER(_vp), _i64],
    "lf2mi_device_free": [_vp],
    "lf2mi_copy": [_vp, _vp, _i64, _i, _vp],
    "lf2mi_profile_enable": [_i],
    "lf2mi_profile_reset": [],
    "lf2mi_profile_get": [_i, ctypes.c_char_p, _i, _dp, _i64p],
    "lf2mi_zero": [_vp, _i64, _i, _vp],
    "lf2mi_ghost_fill": [_vp, _i, _i, _i, _i, _vp],
    "lf2mi_ghost_fill_box": [_vp, _i, _i, _i, _i, _i, _i, _vp],
    "lf2mi_projection_comm": [_vp, _i, _i, _i, _i, _vp],
    "lf2mi_field_sum": [_vp, _i, _i, _i, _i, _i, _dp, _vp],
    "lf2mi_field_affine": [_vp, _i, _i, _i, _i, _d, _d, _i, _vp],
    "lf2mi_fft_plan_create": [_i, _i, _i, _i, _i, ctypes.POINTER(_vp)],
    "lf2mi_fft_plan_create_c2c": [_i, _i, _i, _i, _i, ctypes.POINTER(_vp)],
    "lf2mi_fft_plan_destroy": [_vp],
    "lf2mi_fft_plan_bytes": [_vp, _i64p],
    "lf2mi_fft_execute": [_vp, _i, _vp, _vp, _vp],
    "lf2mi_poisson_green": [_vp, _i, _i, _i, _d, _d, _i, _vp],
    "lf2mi_poisson_green_noalloc": [_vp, _i, _i, _i, _d, _d, _i, _vp],
    "lf2mi_poisson_solve": [_vp, _vp, _vp, _d, _d, _vp],
    "lf2mi_parts_create": [_i, _i64, _i, _i, ctypes.POINTER(_vp)],
    "lf2mi_parts_destroy": [_vp],
    "lf2mi_parts_set_mass": [_vp, _d],
    "lf2mi_parts_count": [_vp, _i64p],
    "lf2mi_parts_bytes": [_vp, _i64p],
    "lf2mi_parts_upload": [_vp, _i64, _dp, _dp, _dp, _i64p, _i64p, _vp],
    "lf2mi_parts_download": [_vp, _dp, _dp, _dp, _i64p, _i32p, _vp],
    "lf2mi_parts_cell_counts": [_vp, _vp, _vp],
    "lf2mi_parts_download_box": [_vp, _i32p, _i32p, _i64, _dp, _dp, _dp, _i64p, _vp],
    "lf2mi_parts_set_payload": [_vp, _i],
    "lf2mi_parts_upload_aos": [_vp, _i64, _vp, _i64, _i, _i64, _i64, _i64, _i64, _i64p, _vp],
    "lf2mi_parts_download_aos": [_vp, _vp, _i64, _i64, _i, _i64, _i64, _i64, _i64p, _vp],
    "lf2mi_parts_mass_from": [_vp, _i, _i64, _vp],
    "lf2mi_parts_vector_from": [_vp, _i64, _i],
    "lf2mi_bin_positions": [_vp, _i64, _i, _vp, _vp, _vp],
    "lf2mi_parts_ic_lattice": [_vp, _i, _d, _d, _u64, _vp],
    "lf2mi_parts_ic_uniform": [_vp, _i64, _d, _u64, _vp],
    "lf2mi_parts_ic_blobs": [_vp, _i64, _i, _d, _d, _d, _d, _u64, _vp],
    "lf2mi_deposit": [_vp, _i, _vp, _i, _i, _i, _vp],
    "lf2mi_gather_kick_drift": [_vp, _vp, _i, _d, _d, _dp, _vp],
    "lf2mi_parts_reduce": [_vp, _i, _i32p, _i32p, _dp, _vp],
    "lf2mi_pm_step": [_vp, _vp, _vp, _vp, _i, _d, _d, _i, _vp],
    "lf2mi_diffpm_create": [_i, _i64, _i, _d, _d, _i, ctypes.POINTER(_vp)],
    "lf2mi_diffpm_destroy": [_vp],
    "lf2mi_diffpm_bytes": [_vp, _i64p],
    "lf2mi_diffpm_forward": [_vp, _vp, _vp, _vp, _i, _vp, _dp, _vp],
    "lf2mi_diffpm_backward": [_vp, _vp, _vp, _vp],
    "lf2mi_diffpm_state": [_vp, _i, _vp, _vp, _i, _vp],
}

_lib = None


class Lf2miError(RuntimeError):
    pass


def load():
    """Load liblf2mi.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise Lf2miError(f"{LIB_PATH} is missing: build it with `python -m latfield2_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = ctypes.c_char_p if name == "lf2mi_last_error" else ctypes.c_int
        _lib = L
    return _lib


def call(name: str, *args) -> int:
    rc = getattr(load(), name)(*args)
    if rc != 0:
        msg = load().lf2mi_last_error().decode(errors="replace")
        raise Lf2miError(f"{name} failed ({rc}): {msg}")
    return rc


def profile_enable(on: bool = True) -> None:
    call("lf2mi_profile_enable", 1 if on else 0)


def profile_reset() -> None:
    call("lf2mi_profile_reset")


def profile_read() -> dict:
    """{kernel name: (total_ms, launches)} for every timed kernel"""
    out = {}
    i = 0
    L = load()
    while True:
        name = ctypes.create_string_buffer(128)
        ms = ctypes.c_double(0.0)
        cnt = ctypes.c_int64(0)
        rc = L.lf2mi_profile_get(i, name, 128, ctypes.byref(ms), ctypes.byref(cnt))
        if rc != 0:
            break
        out[name.value.decode()] = (ms.value, cnt.value)
        i += 1
    return out
