"""latfield2_amd -- MI355X-native particle-mesh path of LATfield2.

Python mirror of the reference's host interface for the PM path, driving the
HIP kernels of ``liblf2mi.so`` through the C ABI of ``include/lf2mi.h``.  torch
supplies device memory and the current stream; it does no arithmetic here.

Names follow the reference (paths relative to the LATfield2 tree):
  Lattice              LATfield2_Lattice_decl.hpp:13-218
  Field                LATfield2_Field_decl.hpp:61-430 (AoS storage, updateHalo)
  KField               Field<Imag> on Lattice::initializeRealFFT (dims kx, kz, ky)
  PlanFFT              LATfield2_PlanFFT_decl.hpp:83-1577 (R2C), FFT_FORWARD/BACKWARD
  Particles            particles/LATfield2_Particles.hpp:131-344
  projection_init, scalarProjectionCIC_project/_comm, vectorProjectionCIC_project/_comm,
  VecVecProjectionCIC_project/_comm   particles/projections.hpp
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import (ACC_FIXED64, ACC_NATIVE, DEP_PERIODIC, DEP_PERIODIC_SET, DEP_UNFOLDED, F32, F64, MASS_INFO,
                   MASS_PART, MASS_PART_F32, FFT_BACKWARD, FFT_FORWARD, MAX, MAX_LOCAL, MIN, MIN_LOCAL, PARTS_ID, PARTS_MASS, PQ_EKIN,
                   PQ_MASS, PQ_MVX, PQ_MVY, PQ_MVZ, PQ_V2, PROJ_SCALAR, PROJ_SYMTENSOR_NGP, PROJ_VECTOR,
                   PROJ_VECTOR_NGP, PROJ_VECVEC, SUM, SUM_LOCAL, Lf2miError, call, load)

__all__ = [
    "Lattice", "Field", "KField", "CField", "PlanFFT", "Particles", "projection_init", "bin_positions",
    "scalarProjectionCIC_project", "scalarProjectionCIC_comm", "vectorProjectionCIC_project",
    "vectorProjectionCIC_comm", "VecVecProjectionCIC_project", "VecVecProjectionCIC_comm",
    "vectorProjectionCICNGP_project", "vectorProjectionCICNGP_comm", "symtensorProjectionCICNGP_project",
    "symtensorProjectionCICNGP_comm", "poisson_green", "poisson_solve", "pm_step", "FFT_FORWARD", "FFT_BACKWARD", "F32", "F64",
    "Lf2miError", "load", "updateVel_simple", "updateVel_cicKick", "move_particles_simple",
    "SUM", "MIN", "MAX", "SUM_LOCAL", "MIN_LOCAL", "MAX_LOCAL",
]

FROM_INFO = 0  # particles/projections.hpp:39-40 (global mass)
FROM_PART = 1  # per-particle mass


def _torch():
    import torch
    return torch


def _stream():
    torch = _torch()
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr())


def _dtype_code(dtype) -> int:
    torch = _torch()
    if dtype in (torch.float32, np.float32, "f32", F32):
        return F32
    if dtype in (torch.float64, np.float64, "f64", F64):
        return F64
    raise ValueError(f"unsupported precision {dtype}")


def _torch_dtype(code: int):
    torch = _torch()
    return torch.float32 if code == F32 else torch.float64


class Lattice:
    """Cubic 3-D lattice of N sites per axis with ghost width `halo` (box length 1)."""

    def __init__(self, dim: int, N: int, halo: int):
        if dim != 3:
            raise ValueError("the PM path is three-dimensional")
        self.dim, self.N, self.halo = dim, int(N), int(halo)

    def size(self, i: int = 0) -> int:
        return self.N

    def sizeLocal(self, i: int = 0) -> int:
        return self.N

    def jump(self, i: int) -> int:
        return (self.N + 2 * self.halo) ** i

    def sitesLocalGross(self) -> int:
        return (self.N + 2 * self.halo) ** 3

    def sites(self) -> int:
        return self.N ** 3


class Field:
    """Field<Real> with C components on the device: storage [z, y, x, c] with ghosts."""

    def __init__(self, lat: Lattice, components: int = 1, dtype=F32, device: str = "cuda"):
        torch = _torch()
        self.lat, self.C, self.code = lat, int(components), _dtype_code(dtype)
        G = lat.N + 2 * lat.halo
        self.t = torch.zeros((G, G, G, self.C), dtype=_torch_dtype(self.code), device=device)

    @property
    def N(self):
        return self.lat.N

    @property
    def halo(self):
        return self.lat.halo

    def components(self) -> int:
        return self.C

    def data(self):
        return self.t

    def interior(self):
        h = self.halo
        G = self.N + 2 * h
        return self.t[h:G - h, h:G - h, h:G - h, :]

    def updateHalo(self):
        call("lf2mi_ghost_fill", _ptr(self.t), self.N, self.halo, self.C, self.code, _stream())

    def zero(self):
        call("lf2mi_zero", _ptr(self.t), self.t.numel(), self.code, _stream())

    def sum(self, comp: int = 0) -> float:
        out = ctypes.c_double(0.0)
        call("lf2mi_field_sum", _ptr(self.t), self.N, self.halo, self.C, comp, self.code,
             ctypes.byref(out), _stream())
        return out.value

    def affine(self, mul: float, add: float, comp: int = 0):
        call("lf2mi_field_affine", _ptr(self.t), self.N, self.halo, self.C, comp, float(mul),
             float(add), self.code, _stream())


class KField:
    """Field<Imag> on the R2C k-lattice: storage [ky, kz, kx, c, re/im] with khalo ghosts."""

    def __init__(self, lat: Lattice, khalo: int = 0, components: int = 1, dtype=F32, device: str = "cuda"):
        torch = _torch()
        self.N, self.kh, self.C, self.code = lat.N, int(khalo), int(components), _dtype_code(dtype)
        KG = self.N + 2 * self.kh
        R = self.N // 2 + 1 + 2 * self.kh
        self.t = torch.zeros((KG, KG, R, self.C, 2), dtype=_torch_dtype(self.code), device=device)

    def data(self):
        return self.t

    def interior(self):
        kh, N = self.kh, self.N
        return self.t[kh:kh + N, kh:kh + N, kh:kh + N // 2 + 1]

    def as_complex(self):
        return _torch().view_as_complex(self.t)


class CField:
    """Field<Imag> on an N^3 lattice with C components: storage [d2, d1, d0, c, re/im] with
    ghosts.  Position space is [z, y, x]; the C2C k-lattice (Lattice::initializeComplexFFT)
    is [ky, kx, kz] (cKSite, LATfield2_Site.hpp:219-263)."""

    def __init__(self, lat: Lattice, components: int = 1, dtype=F32, device: str = "cuda"):
        torch = _torch()
        self.lat, self.C, self.code = lat, int(components), _dtype_code(dtype)
        G = lat.N + 2 * lat.halo
        self.t = torch.zeros((G, G, G, self.C, 2), dtype=_torch_dtype(self.code), device=device)

    @property
    def N(self):
        return self.lat.N

    @property
    def halo(self):
        return self.lat.halo

    @property
    def kh(self):
        return self.lat.halo

    def data(self):
        return self.t

    def interior(self):
        h, G = self.halo, self.N + 2 * self.halo
        return self.t[h:G - h, h:G - h, h:G - h]

    def as_complex(self):
        return _torch().view_as_complex(self.t)


class PlanFFT:
    """PlanFFT<Imag>(Field<Real>*, Field<Imag>*) -- R2C/C2R on the device; with two
    complex fields (CField, CField) the C2C transform (LATfield2_PlanFFT_decl.hpp:600-719),
    whose backward is the inverse of its forward (times N^3)."""

    def __init__(self, rfield, kfield=None):
        if kfield is not None and (kfield.N != rfield.N or kfield.C != rfield.C or kfield.code != rfield.code):
            raise ValueError("PlanFFT: real and k fields disagree on size, components or precision")
        self.r, self.k = rfield, kfield
        self.N, self.C, self.code = rfield.N, rfield.C, rfield.code
        kh = kfield.kh if kfield is not None else 0
        self.c2c = isinstance(rfield, CField)
        if self.c2c and not isinstance(kfield, CField):
            raise ValueError("PlanFFT: a complex position-space field needs a C2C k-field (CField)")
        h = ctypes.c_void_p()
        call("lf2mi_fft_plan_create_c2c" if self.c2c else "lf2mi_fft_plan_create", self.N, self.C, rfield.halo, kh,
             self.code, ctypes.byref(h))
        self.h = h

    def execute(self, fft_type: int):
        call("lf2mi_fft_execute", self.h, int(fft_type), _ptr(self.r.t), _ptr(self.k.t), _stream())

    def bytes(self) -> int:
        b = ctypes.c_int64(0)
        call("lf2mi_fft_plan_bytes", self.h, ctypes.byref(b))
        return b.value

    def __del__(self):
        try:
            if getattr(self, "h", None) is not None and _lib._lib is not None:
                _lib._lib.lf2mi_fft_plan_destroy(self.h)
                self.h = None
        except Exception:
            pass


def poisson_green(k: KField, res: float | None = None, scale: float = 1.0):
    """examples/poissonSolver.cpp:99-111 on the device (lambda and scale/lambda in FP64)."""
    res = 1.0 / k.N if res is None else res
    call("lf2mi_poisson_green", _ptr(k.t), k.N, k.C, k.kh, float(res), float(scale), k.code, _stream())


def poisson_solve(plan: PlanFFT, rho: Field, phi: Field, shift: float, scale: float):
    """Fused phi = C2R(scale * (-1/K^2) * R2C(rho - shift)) with the zero mode removed."""
    call("lf2mi_poisson_solve", plan.h, _ptr(rho.t), _ptr(phi.t), float(shift), float(scale), _stream())


def _dp(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class _DeviceFunctor:
    """A particle functor of the reference's updateVel / moveParticles interface
    (particles/LATfield2_Particles.hpp:217-244) that runs on the device.  Its per-particle
    outputs are, by slot: |v|^2, m v_x, m v_y, m v_z, m, m |v|^2 / 2 after the update."""

    def __init__(self, name: str, kind: str):
        self.name, self.kind = name, kind

    def __call__(self, *args, **kwargs):
        raise Lf2miError(f"{self.name} runs on the device: pass it to Particles.{self.kind}")

    def __repr__(self):
        return f"<device functor {self.name}>"


# updateVel_function.hpp:13-38 (the identity on the velocities)
updateVel_simple = _DeviceFunctor("updateVel_simple", "updateVel")
# the PM gather + kick: v += dtau * CIC(-grad phi) of the potential fields[0]
updateVel_cicKick = _DeviceFunctor("updateVel_cicKick", "updateVel")
# move_function.hpp:15-32: pos += dtau * vel, wrapped and re-binned
move_particles_simple = _DeviceFunctor("move_particles_simple", "moveParticles")
_OUTPUT_SLOTS = (PQ_V2, PQ_MVX, PQ_MVY, PQ_MVZ, PQ_MASS, PQ_EKIN)


def _reduce_outputs(parts, output, reduce_type, noutput):
    """max |v|^2 and the functor outputs of slots 0..noutput-1, reduced by reduce_type
    (particles/LATfield2_Particles.hpp:788-807, :846-885)"""
    if noutput > len(_OUTPUT_SLOTS):
        raise Lf2miError(f"the device functors have {len(_OUTPUT_SLOTS)} output slots")
    if noutput and (output is None or reduce_type is None):
        raise Lf2miError("output and reduce_type are needed for noutput > 0")
    q = [PQ_V2] + list(_OUTPUT_SLOTS[:noutput])
    rt = [MAX] + [int(reduce_type[i]) for i in range(noutput)]
    init = [0.0] + [float(output[i]) for i in range(noutput)]
    res = parts.reduce(q, rt, init)
    for i in range(noutput):
        output[i] = res[i + 1]
    return res[0] if parts.count() > 0 else 0.0



class Particles:
    """Particles<part_simple, part_simple_info, ...> on the device (box length 1)."""

    def __init__(self, N: int, capacity: int, dtype=F32, ids: bool = False, per_particle_mass: bool = False,
                 mass: float = 1.0):
        self.N, self.code = int(N), _dtype_code(dtype)
        flags = (PARTS_ID if ids else 0) | (PARTS_MASS if per_particle_mass else 0)
        self.flags = flags
        h = ctypes.c_void_p()
        call("lf2mi_parts_create", self.N, int(capacity), self.code, flags, ctypes.byref(h))
        self.h = h
        call("lf2mi_parts_set_mass", self.h, float(mass))

    def __del__(self):
        try:
            if getattr(self, "h", None) is not None and _lib._lib is not None:
                _lib._lib.lf2mi_parts_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def count(self) -> int:
        n = ctypes.c_int64(0)
        call("lf2mi_parts_count", self.h, ctypes.byref(n))
        return n.value

    def bytes(self) -> int:
        n = ctypes.c_int64(0)
        call("lf2mi_parts_bytes", self.h, ctypes.byref(n))
        return n.value

    def addParticles(self, pos, vel, mass=None, ids=None) -> int:
        """addParticle_global for arrays of particles; returns the number kept."""
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        vel = np.ascontiguousarray(vel, dtype=np.float64)
        m = None if mass is None else np.ascontiguousarray(mass, dtype=np.float64)
        i = None if ids is None else np.ascontiguousarray(ids, dtype=np.int64)
        kept = ctypes.c_int64(0)
        call("lf2mi_parts_upload", self.h, pos.shape[0], _dp(pos), _dp(vel), _dp(m),
             None if i is None else i.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), ctypes.byref(kept),
             _stream())
        return kept.value

    def download(self):
        """dict of numpy arrays in container order: pos (box units), vel, mass, cell, id (if kept)"""
        n = self.count()
        pos = np.zeros((n, 3))
        vel = np.zeros((n, 3))
        mass = np.zeros(n)
        cell = np.zeros(n, dtype=np.int32)
        ids = np.zeros(n, dtype=np.int64) if self.flags & PARTS_ID else None
        call("lf2mi_parts_download", self.h, _dp(pos), _dp(vel), _dp(mass),
             None if ids is None else ids.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
             cell.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), _stream())
        out = {"pos": pos, "vel": vel, "mass": mass, "cell": cell}
        if ids is not None:
            out["id"] = ids
        return out

    def download_box(self, lo, hi, max_out: int):
        """particles whose cell lies in [lo, hi) per axis: dict of pos, vel, mass (numpy)"""
        lo_a = (ctypes.c_int32 * 3)(*[int(v) for v in lo])
        hi_a = (ctypes.c_int32 * 3)(*[int(v) for v in hi])
        pos = np.zeros((max_out, 3))
        vel = np.zeros((max_out, 3))
        mass = np.zeros(max_out)
        n = ctypes.c_int64(0)
        call("lf2mi_parts_download_box", self.h, lo_a, hi_a, int(max_out), _dp(pos), _dp(vel), _dp(mass),
             ctypes.byref(n), _stream())
        k = n.value
        return {"pos": pos[:k], "vel": vel[:k], "mass": mass[:k]}

    # ---- user structs (Particles<part, ...> with any part type; the mass by byte offset of
    # projections.hpp:108-126)
    def set_payload(self, nbytes: int):
        """keep each particle's whole struct (nbytes, a multiple of 8) on the device"""
        call("lf2mi_parts_set_payload", self.h, int(nbytes))
        self._payload = int(nbytes)

    @staticmethod
    def _fields(dtype):
        names = dtype.names or ()
        def off(name):
            return dtype.fields[name][1] if name in names else -1
        real = dtype.fields["pos"][0].base
        code = F32 if real == np.float32 else F64
        return code, off("pos"), off("vel"), off("ID"), off("mass")

    def upload_structs(self, parts) -> int:
        """addParticle_global from a numpy structured array with fields pos (3 Real), vel
        (3 Real), optional ID (int64) and mass (float64, or float32 read with its own type);
        returns the particles kept"""
        a = np.ascontiguousarray(parts)
        code, op, ov, oi, om = self._fields(a.dtype)
        f4 = om >= 0 and a.dtype.fields["mass"][0] == np.float32
        kept = ctypes.c_int64(0)
        call("lf2mi_parts_upload_aos", self.h, a.shape[0], a.ctypes.data_as(ctypes.c_void_p), a.dtype.itemsize,
             code, op, ov, oi, -1 if f4 else om, ctypes.byref(kept), _stream())
        if f4:
            # gathered from the kept structs on the device; the mass source stays as it was
            if not getattr(self, "_payload", 0):
                raise ValueError("a float32 mass field needs the structs kept on the device (set_payload)")
            self.mass_from(MASS_PART_F32, om)
            if not self.flags & PARTS_MASS:
                self.mass_from(MASS_INFO)
        return kept.value

    def download_structs(self, dtype):
        """the particles as a numpy structured array of `dtype` (container order): the kept
        structs with pos, vel and ID brought up to date"""
        dtype = np.dtype(dtype)
        code, op, ov, oi, _ = self._fields(dtype)
        out = np.zeros(self.count(), dtype=dtype)
        n = ctypes.c_int64(0)
        call("lf2mi_parts_download_aos", self.h, out.ctypes.data_as(ctypes.c_void_p), out.shape[0], dtype.itemsize,
             code, op, ov, oi, ctypes.byref(n), _stream())
        return out[:n.value]

    def mass_from(self, where: int, offset: int = -1):
        """mass of later deposits: MASS_INFO (global), MASS_PART (the double at byte offset of
        each kept struct, gathered into the per-particle masses; -1: those masses as they
        stand) or MASS_PART_F32 (the float at byte offset)"""
        call("lf2mi_parts_mass_from", self.h, int(where), int(offset), _stream())

    def vector_from(self, offset: int = -1, real: int = F32):
        """vector of later vector/tensor deposits (a projection's oset[1]): the velocity
        (offset -1 or the uploaded vel offset), else the 3 Reals of type `real` at byte
        offset of each kept struct"""
        call("lf2mi_parts_vector_from", self.h, int(offset), int(real))

    def cell_counts(self):
        torch = _torch()
        c = torch.empty(self.N ** 3, dtype=torch.int32, device="cuda")
        call("lf2mi_parts_cell_counts", self.h, _ptr(c), _stream())
        return c

    def ic_lattice(self, ng: int, sigma_cells: float, sigma_v: float, seed: int):
        call("lf2mi_parts_ic_lattice", self.h, int(ng), float(sigma_cells), float(sigma_v), int(seed), _stream())

    def ic_uniform(self, np_: int, sigma_v: float, seed: int):
        call("lf2mi_parts_ic_uniform", self.h, int(np_), float(sigma_v), int(seed), _stream())

    def ic_blobs(self, np_: int, nblobs: int = 4096, alpha: float = 1.5, sigma_cells: float = 2.0,
                 frac_uniform: float = 0.2, sigma_v: float = 0.0, seed: int = 1):
        call("lf2mi_parts_ic_blobs", self.h, int(np_), int(nblobs), float(alpha), float(sigma_cells),
             float(frac_uniform), float(sigma_v), int(seed), _stream())

    def deposit(self, kind: int, field: Field, target: int = DEP_UNFOLDED, acc: int = ACC_NATIVE):
        call("lf2mi_deposit", self.h, int(kind), _ptr(field.t), field.halo, int(target), int(acc), _stream())

    def gather_kick_drift(self, phi: Field, dt_kick: float, dt_drift: float, want_vmax: bool = False):
        vm = ctypes.c_double(0.0)
        call("lf2mi_gather_kick_drift", self.h, None if phi is None else _ptr(phi.t), 1 if phi is None else phi.halo,
             float(dt_kick), float(dt_drift),
             ctypes.byref(vm) if want_vmax else None, _stream())
        return vm.value if want_vmax else None

    def updateVel(self, funct, dtau, fields=None, nfields=0, params=None, output=None, reduce_type=None,
                             noutput=0):
        """Particles::updateVel (particles/LATfield2_Particles.hpp:753-893) with a device functor;
        returns sqrt(max |v|^2) after the update and fills output[:noutput]"""
        if funct is updateVel_cicKick:
            if not fields or fields[0].C != 1:
                raise Lf2miError("updateVel_cicKick needs the potential as fields[0]")
            self.gather_kick_drift(fields[0], float(dtau), 0.0)
        elif funct is not updateVel_simple:
            raise Lf2miError("this functor has no device implementation (updateVel_simple and updateVel_cicKick have)")
        return float(np.sqrt(_reduce_outputs(self, output, reduce_type, noutput)))


    def moveParticles(self, funct, dtau, fields=None, nfields=0, params=None, output=None, reduce_type=None,
                                 noutput=0):
        """Particles::moveParticles (particles/LATfield2_Particles.hpp:1811-2803) with a device functor"""
        if funct is not move_particles_simple:
            raise Lf2miError("this functor has no device implementation (move_particles_simple has)")
        self.gather_kick_drift(None, 0.0, float(dtau))
        if noutput:
            _reduce_outputs(self, output, reduce_type, noutput)


    def reduce(self, quantities, reduce_types, init=None):
        """lf2mi_parts_reduce: FP64 statistics of the particles (PQ_* quantities, SUM/MIN/MAX
        reduction bits); a slot whose type has none of the bits keeps its `init` value"""
        n = len(quantities)
        q = (ctypes.c_int32 * n)(*[int(v) for v in quantities])
        rt = (ctypes.c_int32 * n)(*[int(v) for v in reduce_types])
        out = (ctypes.c_double * n)(*([0.0] * n if init is None else [float(v) for v in init]))
        call("lf2mi_parts_reduce", self.h, n, q, rt, out, _stream())
        return list(out)


def bin_positions(pos, N: int, want_keys: bool = True, want_counts: bool = True):
    """Particles::getPartCoord keys of FP64 positions (a (np, 3) device tensor) and the
    per-cell histogram (particles/LATfield2_Particles.hpp:464-467); key -1 for a negative
    coordinate.  Returns (keys, counts), either None when not asked for."""
    torch = _torch()
    pos = pos.contiguous()
    if pos.dtype != torch.float64 or pos.dim() != 2 or pos.shape[1] != 3:
        raise ValueError("bin_positions takes an (np, 3) float64 tensor")
    keys = torch.empty(pos.shape[0], dtype=torch.int32, device=pos.device) if want_keys else None
    counts = torch.empty(N ** 3, dtype=torch.int32, device=pos.device) if want_counts else None
    call("lf2mi_bin_positions", _ptr(pos), pos.shape[0], int(N), None if keys is None else _ptr(keys),
         None if counts is None else _ptr(counts), _stream())
    return keys, counts


def projection_init(f: Field):
    """particles/projections.hpp:31-34"""
    f.zero()


def scalarProjectionCIC_project(parts: Particles, rho: Field):
    parts.deposit(PROJ_SCALAR, rho, DEP_UNFOLDED)


def scalarProjectionCIC_comm(rho: Field):
    call("lf2mi_projection_comm", _ptr(rho.t), rho.N, rho.halo, PROJ_SCALAR, rho.code, _stream())


def vectorProjectionCIC_project(parts: Particles, B: Field):
    parts.deposit(PROJ_VECTOR, B, DEP_UNFOLDED)


def vectorProjectionCIC_comm(B: Field):
    call("lf2mi_projection_comm", _ptr(B.t), B.N, B.halo, PROJ_VECTOR, B.code, _stream())


def VecVecProjectionCIC_project(parts: Particles, T: Field):
    parts.deposit(PROJ_VECVEC, T, DEP_UNFOLDED)


def VecVecProjectionCIC_comm(T: Field):
    call("lf2mi_projection_comm", _ptr(T.t), T.N, T.halo, PROJ_VECVEC, T.code, _stream())


def vectorProjectionCICNGP_project(parts: Particles, B: Field):
    """particles/projections.hpp:422-549: component i nearest-grid-point along axis i"""
    parts.deposit(PROJ_VECTOR_NGP, B, DEP_UNFOLDED)


def vectorProjectionCICNGP_comm(B: Field):
    """particles/projections.hpp:558-681"""
    call("lf2mi_projection_comm", _ptr(B.t), B.N, B.halo, PROJ_VECTOR_NGP, B.code, _stream())


def symtensorProjectionCICNGP_project(parts: Particles, T: Field):
    """particles/projections.hpp:709-847: T_ii CIC, T_ij NGP along i and j"""
    parts.deposit(PROJ_SYMTENSOR_NGP, T, DEP_UNFOLDED)


def symtensorProjectionCICNGP_comm(T: Field):
    """particles/projections.hpp:856-976"""
    call("lf2mi_projection_comm", _ptr(T.t), T.N, T.halo, PROJ_SYMTENSOR_NGP, T.code, _stream())


def pm_step(parts: Particles, plan: PlanFFT, rho: Field, phi: Field, dt_kick: float, dt_drift: float,
            acc: int = ACC_NATIVE):
    call("lf2mi_pm_step", parts.h, plan.h, _ptr(rho.t), _ptr(phi.t), rho.halo, float(dt_kick), float(dt_drift),
         int(acc), _stream())
