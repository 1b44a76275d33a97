"""The f2py module powerI4 of the reference (pbi4/F90_modules/powerI4.f90), on the GPU.

Same function names, argument order and return values as the f2py wrappers imported at
pbi4/pbi4.py:12-13; inputs are numpy arrays in the reference's layouts (Fortran-order
(N, N, N) grids, (3, npart) positions) and outputs are new numpy arrays.  Arithmetic is FP32
on the device with FP64 reductions; mode counts are exact.
"""
import numpy as np
import torch

from .core import Pbi4Error, default_context
from .fields import full_from_array, full_to_array, half_from_array


def _pos(npart, pos):
    p = np.asarray(pos)
    if p.shape != (3, npart):
        raise Pbi4Error("pos must have shape (3, npart)")
    return torch.from_numpy(np.ascontiguousarray(p, dtype=np.float32)).to(default_context().device)


def _weights(weight):
    if weight is not None and not np.all(np.asarray(weight) == 1.0):
        raise Pbi4Error("per-particle weights other than 1 are not supported")


def assign_single_grid(grid, nintp, npart, pos, weight=None):
    """powerI4.f90:68-140 -> real (grid, grid, grid) Fortran-order float64."""
    _weights(weight)
    ctx = default_context()
    g = ctx.assign(_pos(npart, pos), int(grid), int(nintp), False)
    return np.asfortranarray(g[0].double().cpu().numpy().transpose(2, 1, 0))


def assign_double_grid(grid, nintp, npart, pos, weight=None):
    """powerI4.f90:142-249 -> complex (grid, grid, grid): grid 1 + i (half-cell shifted grid)."""
    _weights(weight)
    ctx = default_context()
    g = ctx.assign(_pos(npart, pos), int(grid), int(nintp), True).double().cpu().numpy()
    return np.asfortranarray((g[0] + 1j * g[1]).transpose(2, 1, 0))


def fcomb(grid, nintp, interlacing, box, fdcl):
    """powerI4.f90:251-378 on a full complex array (the FFT of the assigned grid / npart)."""
    ctx = default_context()
    out = ctx.fcomb_full(full_from_array(fdcl, ctx.device), int(grid), int(nintp),
                         bool(interlacing), float(np.float32(box)))
    return full_to_array(out)


def compute_q2k(grid, box, deltak, i, j):
    """powerI4.f90:409-458."""
    ctx = default_context()
    return full_to_array(ctx.compute_qk(full_from_array(deltak, ctx.device), int(grid), (i, j)))


def compute_q4k(grid, box, deltak, i, j, l, k):
    """powerI4.f90:460-508."""
    ctx = default_context()
    return full_to_array(ctx.compute_qk(full_from_array(deltak, ctx.device), int(grid),
                                        (i, j, l, k)))


def measure_pk(grid, box, kbin, kcenter, nbins, dcl1, dcl2):
    """powerI4.f90:510-594 -> pout (5, nbins): avgk, P, Nmodes, 0, 0."""
    ctx = default_context()
    a = half_from_array(dcl1, ctx.device)
    b = None if dcl2 is dcl1 else half_from_array(dcl2, ctx.device)
    return np.asfortranarray(ctx.pk(a, int(grid), float(box), float(kbin), float(kcenter),
                                    int(nbins), False, delta_b=b).cpu().numpy().T)


def measure_pk_multipoles(grid, box, kbin, kcenter, nbins, dcl, dcl2, dcl4):
    """powerI4.f90:596-693 -> pout (5, nbins): avgk, P0, P2, P4, Nmodes."""
    ctx = default_context()
    d = ctx.device
    return np.asfortranarray(ctx.pk(half_from_array(dcl, d), int(grid), float(box), float(kbin),
                                    float(kcenter), int(nbins), True,
                                    delta2=half_from_array(dcl2, d),
                                    delta4=half_from_array(dcl4, d)).cpu().numpy().T)


def reduce_array(L, din, grid):
    """powerI4.f90:380-407: the grid^3 sub-array of an L^3 density keeping the lowest
    frequencies, as a complex128 (grid, grid, grid) Fortran-order array whose x > grid/2 half is
    zero (the reference's zero-filled intent(out) array).  The copy runs on the device in FP64
    (bit-exact); only the x <= grid/2 slab of din is sent."""
    ctx = default_context()
    L, grid = int(L), int(grid)
    a = np.asarray(din)
    if a.shape != (L, L, L):
        raise Pbi4Error(f"din must have shape ({L}, {L}, {L})")
    slab = np.ascontiguousarray(a[:grid // 2 + 1].transpose(2, 1, 0), dtype=np.complex128)
    out = ctx.reduce_array(torch.from_numpy(slab).to(ctx.device), L, grid, nx=grid)
    return np.asfortranarray(out.cpu().numpy().transpose(2, 1, 0))
