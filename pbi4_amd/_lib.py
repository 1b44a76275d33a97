"""ctypes binding of libpbi4hip.so (declarations in include/pbi4_amd.h).

torch is imported first so that the library binds to the HIP runtime torch already loaded; the
library itself is mandatory -- there is no fallback implementation.
"""
import ctypes
import os

import torch  # noqa: F401  (loads the HIP runtime the library must share)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpbi4hip.so")

PBI4_OK = 0
PREC_FP32 = 0
PREC_BF16X3 = 1
RSD_NONE, RSD_HZ, RSD_VEL = 0, 1, 2
POS_NORMALISED = 0
POS_BOX = 1
NOALLOC_SCRATCH_BYTES = 64 << 20      # PBI4_NOALLOC_SCRATCH_BYTES

c_int, c_float, c_double, c_int64, c_size_t, c_void_p = (ctypes.c_int, ctypes.c_float,
                                                         ctypes.c_double, ctypes.c_int64,
                                                         ctypes.c_size_t, ctypes.c_void_p)


class Pbi4Error(RuntimeError):
    pass


class AssignDesc(ctypes.Structure):
    _fields_ = [("grid", c_int), ("order", c_int), ("interlacing", c_int), ("pos_mode", c_int),
                ("box", c_float), ("rsd", c_float * 3)]


class DensityDesc(ctypes.Structure):
    _fields_ = [("grid", c_int), ("order", c_int), ("interlacing", c_int), ("box", c_float),
                ("npart", c_double)]


class PkDesc(ctypes.Structure):
    _fields_ = [("grid", c_int), ("box", c_float), ("kbin", c_float), ("kcenter", c_float),
                ("nbins", c_int), ("multipoles", c_int), ("los", c_double * 3)]


class BkDesc(ctypes.Structure):
    _fields_ = [("nbins", c_int), ("grid", c_int), ("kbin", c_float), ("kcenter", c_float),
                ("iopen", c_int), ("kf", c_float), ("itot", c_int64), ("npart", c_int64),
                ("multipoles", c_int), ("los", c_double * 3), ("precision", c_int),
                ("maps_grid", c_int)]


# every symbol include/pbi4_amd.h declares, with (restype, argtypes)
P = ctypes.POINTER
SIGNATURES = {
    "pbi4_version": (c_int, []),
    "pbi4_ctx_create": (c_int, [c_int, P(c_void_p)]),
    "pbi4_ctx_destroy": (c_int, [c_void_p]),
    "pbi4_last_error": (ctypes.c_char_p, [c_void_p]),
    "pbi4_assign_workspace": (c_size_t, [P(AssignDesc), c_int64]),
    "pbi4_assign": (c_int, [c_void_p, P(AssignDesc), c_void_p, c_void_p, c_int64, c_void_p,
                            c_void_p, c_size_t, c_void_p]),
    "pbi4_density_workspace": (c_size_t, [c_void_p, P(DensityDesc)]),
    "pbi4_compute_density": (c_int, [c_void_p, P(DensityDesc), c_void_p, c_void_p, c_void_p,
                                     c_size_t, c_void_p]),
    "pbi4_assign_density": (c_int, [c_void_p, P(AssignDesc), P(DensityDesc), c_void_p, c_void_p,
                                    c_int64, c_void_p, c_void_p, c_void_p, c_size_t, c_void_p]),
    "pbi4_fcomb": (c_int, [c_void_p, c_int, c_int, c_int, c_float, c_void_p, c_void_p, c_void_p]),
    "pbi4_compute_qk": (c_int, [c_void_p, c_int, c_int, P(c_int), c_void_p, c_void_p, c_void_p]),
    "pbi4_catalogue_positions": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_int, c_int64, c_int,
                                         P(ctypes.c_double), ctypes.c_double, ctypes.c_double,
                                         ctypes.c_double, c_void_p, c_int, c_void_p]),
    "pbi4_reduce_array": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_int, c_int, c_void_p,
                                  c_void_p, c_void_p]),
    "pbi4_pk_workspace": (c_size_t, [c_void_p, P(PkDesc)]),
    "pbi4_measure_pk": (c_int, [c_void_p, P(PkDesc), c_void_p, c_void_p, c_void_p, c_void_p,
                                c_void_p, c_void_p, c_size_t, c_void_p]),
    "pbi4_count_triangles": (c_int, [c_void_p, c_int, c_int, c_float, c_float, c_int, c_int,
                                     c_int64, P(c_int64)]),
    "pbi4_bk_workspace": (c_size_t, [c_void_p, P(BkDesc)]),
    "pbi4_bk_maps_grid": (c_int, [c_void_p, P(BkDesc)]),
    "pbi4_measure_bk": (c_int, [c_void_p, P(BkDesc), c_void_p, c_void_p, c_void_p, c_void_p,
                                c_void_p, c_void_p, c_size_t, c_void_p]),
    "pbi4_bk_cross_workspace": (c_size_t, [c_void_p, P(BkDesc)]),
    "pbi4_measure_bk_cross": (c_int, [c_void_p, P(BkDesc), c_void_p, c_void_p, c_void_p, c_void_p,
                                      c_void_p, c_size_t, c_void_p]),
    "pbi4_pk_bk_workspace": (c_size_t, [c_void_p, P(PkDesc), P(BkDesc)]),
    "pbi4_measure_pk_bk": (c_int, [c_void_p, P(PkDesc), P(BkDesc), c_void_p, c_void_p, c_void_p,
                                   c_void_p, c_void_p, c_size_t, c_void_p]),
    "pbi4_triangle_sums_workspace": (c_size_t, [c_int, c_int, c_int64, c_int64, c_int]),
    "pbi4_triangle_sums": (c_int, [c_void_p, c_int, c_int, c_int64, c_int64, c_int, c_void_p,
                                   c_void_p, c_void_p, c_size_t, c_void_p]),
    "pbi4_triangle_sums_noalloc": (c_int, [c_void_p, c_int, c_int, c_int64, c_int64, c_int,
                                           c_void_p, c_void_p, c_void_p]),
    "pbi4_timing_enable": (c_int, [c_void_p, ctypes.c_char_p]),
    "pbi4_timing_read": (c_int, [c_void_p, P(c_double), P(c_int64)]),
}

_LIB = None


def lib():
    """The loaded library; raises if it has not been built (python -m pbi4_amd.build)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise Pbi4Error(f"{LIB_PATH} is missing: build it with `python -m pbi4_amd.build` "
                            "(the HIP library is required; there is no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB
