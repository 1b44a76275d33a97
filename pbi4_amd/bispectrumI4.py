"""The f2py module bispectrumI4 of the reference (pbi4/F90_modules/bispectrumI4.f90), on the GPU.

Same function names and argument order as the f2py wrappers imported at pbi4/pbi4.py:14 (npool
is accepted and ignored).  counts are the reference's counts = N_tr N^3 (float64, exact here);
they are handed to the kernels as REAL(4), as f2py does.  Triangle sums use FP32 MFMA
(``PRECISION = PREC_FP32``) unless the module attribute is set to ``PREC_BF16X3``; the shell maps
are built on the band-limited grid unless ``MAPS_GRID`` is set.
"""
import numpy as np
import torch

from ._lib import PREC_BF16X3, PREC_FP32  # noqa: F401
from .core import Pbi4Error, default_context
from .fields import half_from_array

PRECISION = PREC_FP32
MAPS_GRID = 0


def count_triangles(nbins, grid, kbin, kcenter, iopen, cross, npool, itot):
    """bispectrumI4.f90:763-915 -> counts(itot) = N_tr N^3 (exact integers as float64)."""
    ntr = default_context().count_triangles(int(nbins), int(grid), float(kbin), float(kcenter),
                                            bool(iopen), bool(cross), int(itot))
    return ntr.astype(np.float64) * float(grid) ** 3


def _bk(nbins, grid, kbin, kcenter, iopen, kf, itot, npart, counts, dcl, dcl2, dcl4):
    ctx = default_context()
    d = ctx.device
    mult = dcl2 is not None
    desc = ctx.bk_desc(int(nbins), int(grid), float(kbin), float(kcenter), bool(iopen),
                       np.float32(kf), int(itot), int(npart), mult, (0.0, 0.0, 1.0), PRECISION,
                       MAPS_GRID)
    c = torch.from_numpy(np.ascontiguousarray(np.asarray(counts).astype(np.float32))).to(d)
    if c.numel() != int(itot):
        raise Pbi4Error("counts must have itot entries")
    f2 = half_from_array(dcl2, d) if mult else None
    f4 = half_from_array(dcl4, d) if mult else None
    return np.asfortranarray(ctx.bk(half_from_array(dcl, d), c, desc, delta2=f2,
                                    delta4=f4).cpu().numpy().T)


def measure_bk(nbins, grid, kbin, kcenter, iopen, kf, npool, itot, npart, counts, dcl):
    """bispectrumI4.f90:1-198 -> bout (9, itot)."""
    return _bk(nbins, grid, kbin, kcenter, iopen, kf, itot, npart, counts, dcl, None, None)


def measure_bk_multipoles(nbins, grid, kbin, kcenter, iopen, kf, npool, itot, npart, counts, dcl,
                          dcl2, dcl4):
    """bispectrumI4.f90:200-516 -> bout (11, itot)."""
    return _bk(nbins, grid, kbin, kcenter, iopen, kf, itot, npart, counts, dcl, dcl2, dcl4)


def measure_bk_cross(nbins, grid, kbin, kcenter, iopen, kf, npool, itot, npart, counts, dcl1, dcl2):
    """bispectrumI4.f90:518-761 -> bout (5, itot): k1, k2, k3, B (dcl2 on the third bin), N_tr."""
    ctx = default_context()
    d = ctx.device
    desc = ctx.bk_desc(int(nbins), int(grid), float(kbin), float(kcenter), bool(iopen),
                       np.float32(kf), int(itot), int(npart), False, (0.0, 0.0, 1.0), PRECISION,
                       MAPS_GRID)
    c = torch.from_numpy(np.ascontiguousarray(np.asarray(counts).astype(np.float32))).to(d)
    if c.numel() != int(itot):
        raise Pbi4Error("counts must have itot entries")
    return np.asfortranarray(ctx.bk_cross(half_from_array(dcl1, d), half_from_array(dcl2, d), c,
                                          desc).cpu().numpy().T)
