"""ctypes binding of libjja (include/jja.h).

The library is built in-tree (pyjjasim_amd/libjja.so) by ``_build.build_library``.
There is no fallback: if the library or a HIP device is missing, the GPU entry
points raise.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libjja.so")

JJA_OK = 0
JJA_EINVAL = -1
JJA_ENOMEM = -2
JJA_EFACTOR = -3
JJA_EHIP = -4
JJA_EUNSUPPORTED = -5
JJA_EBUSY = -6

i32p = C.POINTER(C.c_int32)
i64p = C.POINTER(C.c_int64)
f64p = C.POINTER(C.c_double)


class Slice(C.Structure):
    _fields_ = [("base", C.c_int64), ("ibase", C.c_int64), ("rowbase", C.c_int32),
                ("r0", C.c_int32), ("n_iter", C.c_int32), ("flags", C.c_int32)]


SLICE_BLOCK = 0x100
SLICE_FINISH = 0x200
SLICE_BARRIER = 0x400
SLICE_WG_BARRIER = 0x800
SLICE_NARROW = 0x1000


class ProblemDesc(C.Structure):
    _fields_ = [
        ("n_junctions", C.c_int32), ("n_faces", C.c_int32),
        ("stencil_width", C.c_int32), ("block_size", C.c_int32),
        ("Ic", f64p), ("Rv", f64p), ("coef", f64p),
        ("junc_code", i32p), ("face_of_pos", i32p),
        ("n_face_slices", C.c_int32), ("n_solve_slices", C.c_int32),
        ("face_slices", C.c_void_p), ("solve_slices", C.c_void_p),
        ("n_face_entries", C.c_int64), ("face_code", i32p),
        ("n_slice_rows", C.c_int64), ("slice_rows", i32p),
        ("n_entries", C.c_int64), ("ent_val", f64p), ("n_index", C.c_int64),
        ("ent_off", C.c_void_p),
        ("max_stage_rows", C.c_int32), ("face_copy_rows", C.c_int32),
        ("solve_wave_ptr", i32p), ("index_bytes", C.c_int32), ("parts", C.c_int32),
        ("column_pairs", C.c_int32), ("max_columns", C.c_int32),
        ("cpr_terms", C.c_int32), ("cpr_pad_", C.c_int32), ("cpr_coef", C.c_double * 8),
        ("lattice_rows", C.c_int32), ("lattice_cols", C.c_int32),
        ("lattice_sine", f64p), ("lattice_inv", f64p), ("lattice_g", C.c_double),
    ]


class Source(C.Structure):
    _fields_ = [
        ("base", C.c_void_p), ("s_i", C.c_int64), ("s_w", C.c_int64), ("s_t", C.c_int64),
        ("amp", C.c_void_p), ("a_w", C.c_int64), ("a_t", C.c_int64), ("t0", C.c_int64),
        ("zero", C.c_int32), ("pad_", C.c_int32),
    ]


class ProblemRun(C.Structure):
    _fields_ = [
        ("W", C.c_int32), ("column_offset", C.c_int32), ("dt", C.c_double),
        ("theta", C.c_void_p), ("theta_s", C.c_void_p), ("u", C.c_void_p),
        ("Is", Source), ("Vs", Source), ("f", Source), ("T", Source), ("sigma", Source),
        ("noise", C.c_void_p), ("noise_t0", C.c_int64), ("seed", C.c_uint64),
        ("th_slot", C.c_void_p), ("I_slot", C.c_void_p), ("slot_t0", C.c_int64),
        ("th_out", C.c_void_p), ("I_out", C.c_void_p), ("Ic", C.c_void_p), ("zbuf", C.c_void_p),
    ]


class AdjointRun(C.Structure):
    _fields_ = [
        ("W", C.c_int32), ("prime", C.c_int32), ("Ic", C.c_void_p), ("hist", C.c_void_p),
        ("hist_m0", C.c_int64), ("lam", C.c_void_p), ("u", C.c_void_p), ("g", C.c_void_p),
        ("g_slot", C.c_void_p), ("slot_t0", C.c_int64), ("g_lo", C.c_int64), ("g_hi", C.c_int64),
        ("grad_Ic", C.c_void_p), ("grad_Is", C.c_void_p), ("grad_f", C.c_void_p),
        ("zbuf", C.c_void_p),
    ]


class NativeError(RuntimeError):
    pass


_lib = None

EXPORTS = ["jja_cholesky", "jja_factor_get", "jja_factor_free", "jja_factor_last_error",
           "jja_plan_validate", "jja_plan_create", "jja_plan_destroy", "jja_plan_lds_bytes",
           "jja_plan_zbuf_doubles", "jja_plan_status", "jja_plan_set_wait_limit",
           "jja_plan_one_xcd_parts", "jja_evolve",
           "jja_vortex", "jja_adjoint", "jja_solve", "jja_solve_noalloc", "jja_last_error"]


def lib():
    """Load libjja.so once and declare the C signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise NativeError(f"{LIB_PATH} is not built; run pyjjasim_amd._build.build_library()")
    L = C.CDLL(LIB_PATH)
    L.jja_cholesky.argtypes = [C.c_int32, i64p, i32p, f64p, C.POINTER(C.c_void_p), i64p]
    L.jja_factor_get.argtypes = [C.c_void_p, i64p, i32p, f64p]
    L.jja_factor_free.argtypes = [C.c_void_p]
    L.jja_factor_free.restype = None
    L.jja_factor_last_error.restype = C.c_char_p
    L.jja_plan_validate.argtypes = [C.c_int32, C.POINTER(ProblemDesc)]
    L.jja_plan_create.argtypes = [C.c_int32, C.POINTER(ProblemDesc), C.POINTER(C.c_void_p)]
    L.jja_plan_destroy.argtypes = [C.c_void_p]
    L.jja_plan_lds_bytes.argtypes = [C.c_void_p, C.c_int32]
    L.jja_plan_lds_bytes.restype = C.c_int64
    L.jja_plan_zbuf_doubles.argtypes = [C.c_void_p, C.c_int32]
    L.jja_plan_zbuf_doubles.restype = C.c_int64
    L.jja_plan_status.argtypes = [C.c_void_p]
    L.jja_plan_status.restype = C.c_int
    L.jja_plan_set_wait_limit.argtypes = [C.c_void_p, C.c_uint32]
    L.jja_plan_set_wait_limit.restype = C.c_int
    L.jja_plan_one_xcd_parts.argtypes = [C.c_void_p]
    L.jja_plan_one_xcd_parts.restype = C.c_int64
    L.jja_evolve.argtypes = [C.c_void_p, C.POINTER(ProblemRun), C.c_int32, C.c_int64, C.c_void_p]
    L.jja_vortex.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]
    L.jja_adjoint.argtypes = [C.c_void_p, C.POINTER(AdjointRun), C.c_int64, C.c_int64, C.c_void_p]
    L.jja_solve.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_int32,
                            C.c_void_p, C.c_void_p]
    L.jja_solve_noalloc.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64,
                                    C.c_int32, C.c_void_p]
    L.jja_last_error.restype = C.c_char_p
    _lib = L
    return L


def check(rc, factor=False):
    """Map a libjja status code to the exception the reference raises for the
    same class of problem (ValueError for bad input, RuntimeError otherwise)."""
    if rc == JJA_OK:
        return
    msg = (lib().jja_factor_last_error() if factor else lib().jja_last_error()).decode()
    if rc == JJA_EINVAL:
        raise ValueError(msg)
    if rc == JJA_EFACTOR:
        raise ValueError(msg)
    raise NativeError(f"libjja error {rc}: {msg}")


def ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))
