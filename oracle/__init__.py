"""FP64 CPU oracle for the LATfield2 particle-mesh path -- TEST INFRASTRUCTURE ONLY.

This package is the checker: only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it.  The product path
(``latfield2_amd``) never calls into it.

Thin numpy/ctypes wrapper over ``liblf2_oracle.so`` (built from ``lf2_oracle.c`` by
``oracle/Makefile``).  Each function restates a reference routine; the citation is
in ``lf2_oracle.h``/``lf2_oracle.c``.

Array conventions
-----------------
* real fields: numpy float64 arrays of shape ``(G, G, G, C)`` indexed
  ``[z, y, x, c]`` with ``G = N + 2h`` (storage order of ``Field<Real>``);
* k-fields: complex128 arrays of shape ``(N+2kh, N+2kh, N/2+1+2kh, C)`` indexed
  ``[ky, kz, kx, c]`` (storage order of the R2C ``Field<Imag>``);
* complex fields of the C2C transform: complex128 arrays ``(G, G, G, C)``, position space
  ``[z, y, x, c]`` and k space ``[ky, kx, kz, c]`` (storage order read by ``cKSite``);
* particles: ``pos``/``vel`` float64 arrays of shape ``(np, 3)`` in box units.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liblf2_oracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_i64 = ctypes.c_longlong


def build(force: bool = False) -> str:
    """Compile liblf2_oracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        sig = {
            "orc_set_threads": [ctypes.c_int],
            "orc_get_threads": [],
            "orc_zero": [_dp, _i64],
            "orc_ghost_fill": [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int],
            "orc_bin": [_dp, _i64, ctypes.c_int, _ip],
            "orc_cell_counts": [_dp, _i64, ctypes.c_int, _ip],
            "orc_scalar_project": [_dp, _dp, ctypes.c_double, _i64, _dp, ctypes.c_int, ctypes.c_int],
            "orc_vector_project": [_dp, _dp, _dp, ctypes.c_double, _i64, _dp, ctypes.c_int, ctypes.c_int],
            "orc_vecvec_project": [_dp, _dp, _dp, ctypes.c_double, _i64, _dp, ctypes.c_int, ctypes.c_int],
            "orc_vector_ngp_project": [_dp, _dp, _dp, ctypes.c_double, _i64, _dp, ctypes.c_int, ctypes.c_int],
            "orc_symtensor_ngp_project": [_dp, _dp, _dp, ctypes.c_double, _i64, _dp, ctypes.c_int,
                                          ctypes.c_int],
            "orc_vector_ngp_comm": [_dp, ctypes.c_int, ctypes.c_int],
            "orc_symtensor_ngp_comm": [_dp, ctypes.c_int, ctypes.c_int],
            "orc_scalar_comm": [_dp, ctypes.c_int, ctypes.c_int],
            "orc_vector_comm": [_dp, ctypes.c_int, ctypes.c_int],
            "orc_vecvec_comm": [_dp, ctypes.c_int, ctypes.c_int],
            "orc_brute_deposit": [_dp, _dp, _dp, ctypes.c_double, _i64, ctypes.c_int, _dp,
                                  ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp],
            "orc_fft_r2c": [_dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int],
            "orc_fft_c2r": [_dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int],
            "orc_fft_c2c_forward": [_dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int],
            "orc_fft_c2c_backward": [_dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int],
            "orc_fft_c2c_backward_reference": [_dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int],
            "orc_poisson_green": [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                  ctypes.c_double],
            "orc_force_field": [_dp, _dp, ctypes.c_int, ctypes.c_int],
            "orc_gather_kick": [_dp, _dp, _i64, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                _dp, _dp],
            "orc_drift": [_dp, _dp, _i64, ctypes.c_double],
            "orc_pm_step": [_dp, _dp, _dp, ctypes.c_double, _i64, ctypes.c_int, ctypes.c_double,
                            ctypes.c_double, _dp, _dp, _dp, _dp],
            "orc_cic_density": [_dp, _i64, ctypes.c_int, ctypes.c_double, _dp],
            "orc_diffpm_forward": [_dp, _dp, _i64, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_int, _dp],
            "orc_diffpm_loss": [_dp, _dp, _i64, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                ctypes.c_int, _dp, _dp],
            "orc_diffpm_loss_grad": [_dp, _dp, _i64, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_int, _dp, _dp, _dp, _dp],
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "oracle arrays must be C-contiguous"
    if a.dtype == np.float64 or a.dtype == np.complex128:
        return a.ctypes.data_as(_dp)
    if a.dtype == np.int32:
        return a.ctypes.data_as(_ip)
    raise TypeError(f"unsupported dtype {a.dtype}")


def _check(rc, name):
    if rc != 0:
        raise RuntimeError(f"{name} failed with code {rc}")


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def get_threads() -> int:
    return int(lib().orc_get_threads())


def real_field(N: int, h: int = 1, C: int = 1) -> np.ndarray:
    G = N + 2 * h
    return np.zeros((G, G, G, C), dtype=np.float64)


def k_field(N: int, kh: int = 0, C: int = 1) -> np.ndarray:
    return np.zeros((N + 2 * kh, N + 2 * kh, N // 2 + 1 + 2 * kh, C), dtype=np.complex128)


def interior(f: np.ndarray, h: int) -> np.ndarray:
    """view of the interior [z, y, x, c] of a real field with ghost width h"""
    G = f.shape[0]
    return f[h:G - h, h:G - h, h:G - h, :]


def ghost_fill(f: np.ndarray, N: int, h: int) -> np.ndarray:
    _check(lib().orc_ghost_fill(_p(f), N, h, f.shape[-1]), "orc_ghost_fill")
    return f


def bin_keys(pos: np.ndarray, N: int) -> np.ndarray:
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    key = np.empty(pos.shape[0], dtype=np.int32)
    _check(lib().orc_bin(_p(pos), pos.shape[0], N, _p(key)), "orc_bin")
    return key


def cell_counts(pos: np.ndarray, N: int) -> np.ndarray:
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    cnt = np.empty(N * N * N, dtype=np.int32)
    _check(lib().orc_cell_counts(_p(pos), pos.shape[0], N, _p(cnt)), "orc_cell_counts")
    return cnt


def _mass_arg(mass):
    if mass is None:
        return None
    return np.ascontiguousarray(mass, dtype=np.float64)


def scalar_project(pos, rho, N, h=1, mass=None, mass_global=1.0):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    m = _mass_arg(mass)
    _check(lib().orc_scalar_project(_p(pos), _p(m), mass_global, pos.shape[0], _p(rho), N, h),
           "orc_scalar_project")
    return rho


def vector_project(pos, vel, B, N, h=1, mass=None, mass_global=1.0):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    vel = np.ascontiguousarray(vel, dtype=np.float64)
    m = _mass_arg(mass)
    _check(lib().orc_vector_project(_p(pos), _p(vel), _p(m), mass_global, pos.shape[0], _p(B), N, h),
           "orc_vector_project")
    return B


def vecvec_project(pos, vel, T, N, h=1, mass=None, mass_global=1.0):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    vel = np.ascontiguousarray(vel, dtype=np.float64)
    m = _mass_arg(mass)
    _check(lib().orc_vecvec_project(_p(pos), _p(vel), _p(m), mass_global, pos.shape[0], _p(T), N, h),
           "orc_vecvec_project")
    return T


def vector_ngp_project(pos, vel, B, N, h=1, mass=None, mass_global=1.0):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    vel = np.ascontiguousarray(vel, dtype=np.float64)
    m = _mass_arg(mass)
    _check(lib().orc_vector_ngp_project(_p(pos), _p(vel), _p(m), mass_global, pos.shape[0], _p(B), N, h),
           "orc_vector_ngp_project")
    return B


def symtensor_ngp_project(pos, vel, T, N, h=1, mass=None, mass_global=1.0):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    vel = np.ascontiguousarray(vel, dtype=np.float64)
    m = _mass_arg(mass)
    _check(lib().orc_symtensor_ngp_project(_p(pos), _p(vel), _p(m), mass_global, pos.shape[0], _p(T), N, h),
           "orc_symtensor_ngp_project")
    return T


def vector_ngp_comm(B, N, h=1):
    _check(lib().orc_vector_ngp_comm(_p(B), N, h), "orc_vector_ngp_comm")
    return B


def symtensor_ngp_comm(T, N, h=1):
    _check(lib().orc_symtensor_ngp_comm(_p(T), N, h), "orc_symtensor_ngp_comm")
    return T


def scalar_comm(rho, N, h=1):
    _check(lib().orc_scalar_comm(_p(rho), N, h), "orc_scalar_comm")
    return rho


def vector_comm(B, N, h=1):
    _check(lib().orc_vector_comm(_p(B), N, h), "orc_vector_comm")
    return B


def vecvec_comm(T, N, h=1):
    _check(lib().orc_vecvec_comm(_p(T), N, h), "orc_vecvec_comm")
    return T


def brute_deposit(pos, N, stagger=(0.0, 0.0, 0.0), vel=None, va=-1, vb=-1, mass=None, mass_global=1.0,
                  ngp=0):
    """periodic hat-function deposit of one component (nearest-grid-point along the axes
    in the bit mask ``ngp``); returns [z, y, x] interior array"""
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    v = None if vel is None else np.ascontiguousarray(vel, dtype=np.float64)
    s = np.ascontiguousarray(stagger, dtype=np.float64)
    out = np.zeros((N, N, N), dtype=np.float64)
    m = _mass_arg(mass)
    _check(lib().orc_brute_deposit(_p(pos), _p(v), _p(m), mass_global, pos.shape[0], N, _p(s),
                                   va, vb, ngp, _p(out)), "orc_brute_deposit")
    return out


def fft_r2c(r, N, rh=1, kh=0):
    C = r.shape[-1]
    k = k_field(N, kh, C)
    _check(lib().orc_fft_r2c(_p(r), _p(k), N, C, rh, kh), "orc_fft_r2c")
    return k


def fft_c2r(k, r, N, rh=1, kh=0):
    C = k.shape[-1]
    assert r.shape[-1] == C
    _check(lib().orc_fft_c2r(_p(k), _p(r), N, C, rh, kh), "orc_fft_c2r")
    return r


def complex_field(N: int, h: int = 1, C: int = 1) -> np.ndarray:
    """a Field<Imag> on an N^3 lattice (position space, or the C2C k-lattice)"""
    G = N + 2 * h
    return np.zeros((G, G, G, C), dtype=np.complex128)


def fft_c2c_forward(f, N, rh=1, kh=0):
    """C2C forward; returns the k-field [ky, kx, kz, c] (storage order (kz, kx, ky))"""
    f = np.ascontiguousarray(f, dtype=np.complex128)
    C = f.shape[-1]
    k = complex_field(N, kh, C)
    _check(lib().orc_fft_c2c_forward(_p(f.view(np.float64)), _p(k.view(np.float64)), N, C, rh, kh),
           "orc_fft_c2c_forward")
    return k


def fft_c2c_backward(k, f, N, rh=1, kh=0):
    k = np.ascontiguousarray(k, dtype=np.complex128)
    C = k.shape[-1]
    _check(lib().orc_fft_c2c_backward(_p(k.view(np.float64)), _p(f.view(np.float64)), N, C, rh, kh),
           "orc_fft_c2c_backward")
    return f


def fft_c2c_backward_reference(k, f, N, h=0):
    k = np.ascontiguousarray(k, dtype=np.complex128)
    C = k.shape[-1]
    _check(lib().orc_fft_c2c_backward_reference(_p(k.view(np.float64)), _p(f.view(np.float64)), N, C, h),
           "orc_fft_c2c_backward_reference")
    return f


def poisson_green(k, N, kh=0, res=None, scale=1.0):
    res = 1.0 / N if res is None else res
    _check(lib().orc_poisson_green(_p(k), N, k.shape[-1], kh, res, scale), "orc_poisson_green")
    return k


def force_field(phi, N, h=1):
    F = real_field(N, h, 3)
    _check(lib().orc_force_field(_p(phi), _p(F), N, h), "orc_force_field")
    return F


def gather_kick(pos, vel, F, N, h=1, dt_kick=0.0):
    """vel is updated in place; returns (acc, vmax2)"""
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    assert vel.flags["C_CONTIGUOUS"] and vel.dtype == np.float64
    acc = np.zeros_like(pos)
    vm = ctypes.c_double(0.0)
    _check(lib().orc_gather_kick(_p(pos), _p(vel), pos.shape[0], _p(F), N, h, dt_kick, _p(acc),
                                 ctypes.byref(vm)), "orc_gather_kick")
    return acc, vm.value


def drift(pos, vel, dt):
    assert pos.flags["C_CONTIGUOUS"] and pos.dtype == np.float64
    vel = np.ascontiguousarray(vel, dtype=np.float64)
    _check(lib().orc_drift(_p(pos), _p(vel), pos.shape[0], dt), "orc_drift")
    return pos


def pm_step(pos, vel, N, dt_kick, dt_drift, mass=None, mass_global=1.0, want_acc=True):
    """One FP64 PM step in place on pos/vel.  Returns dict with rho, phi, acc, stage_ms."""
    assert pos.flags["C_CONTIGUOUS"] and vel.flags["C_CONTIGUOUS"]
    rho = real_field(N, 1, 1)
    phi = real_field(N, 1, 1)
    acc = np.zeros_like(pos) if want_acc else None
    st = np.zeros(8, dtype=np.float64)
    m = _mass_arg(mass)
    _check(lib().orc_pm_step(_p(pos), _p(vel), _p(m), mass_global, pos.shape[0], N, dt_kick,
                             dt_drift, _p(rho), _p(phi), _p(acc), _p(st)), "orc_pm_step")
    names = ["deposit", "fold", "mean", "fft_fwd", "green", "fft_bwd", "halo_force", "gather_kick_drift"]
    return {"rho": rho, "phi": phi, "acc": acc, "stage_ms": dict(zip(names, st.tolist()))}


# ----------------------------------------------------------------------------- differentiable PM
def cic_density(pos, N, mass):
    """periodic CIC density (mass * N^3 per particle) on [z, y, x]"""
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    out = np.zeros((N, N, N))
    _check(lib().orc_cic_density(_p(pos), pos.shape[0], N, mass, _p(out)), "orc_cic_density")
    return out


def diffpm_forward(pos, vel, N, mass, dt, nsteps):
    pos = np.ascontiguousarray(pos, dtype=np.float64).copy()
    vel = np.ascontiguousarray(vel, dtype=np.float64).copy()
    _check(lib().orc_diffpm_forward(_p(pos), _p(vel), pos.shape[0], N, mass, dt, nsteps, None),
           "orc_diffpm_forward")
    return pos, vel


def diffpm_loss(pos0, vel0, N, mass, dt, nsteps, target):
    pos0 = np.ascontiguousarray(pos0, dtype=np.float64)
    vel0 = np.ascontiguousarray(vel0, dtype=np.float64)
    target = np.ascontiguousarray(target, dtype=np.float64)
    L = ctypes.c_double(0.0)
    _check(lib().orc_diffpm_loss(_p(pos0), _p(vel0), pos0.shape[0], N, mass, dt, nsteps, _p(target),
                                 ctypes.byref(L)), "orc_diffpm_loss")
    return L.value


def diffpm_loss_grad(pos0, vel0, N, mass, dt, nsteps, target):
    """(loss, dL/dpos0, dL/dvel0) in box units by the FP64 adjoint"""
    pos0 = np.ascontiguousarray(pos0, dtype=np.float64)
    vel0 = np.ascontiguousarray(vel0, dtype=np.float64)
    target = np.ascontiguousarray(target, dtype=np.float64)
    L = ctypes.c_double(0.0)
    gp = np.zeros_like(pos0)
    gv = np.zeros_like(vel0)
    _check(lib().orc_diffpm_loss_grad(_p(pos0), _p(vel0), pos0.shape[0], N, mass, dt, nsteps, _p(target),
                                      ctypes.byref(L), _p(gp), _p(gv)), "orc_diffpm_loss_grad")
    return L.value, gp, gv
