"""One catalogue through the whole measurement path on one HIP stream.

A Measurement describes what a request asks for (the parameters of IO.derived_params,
pbi4/pbi4.py:86-119, plus the fields to measure); CataloguePlan owns everything that is shared
between catalogues of that shape (triangle list and exact counts, bispectrum descriptors) and
enqueues, for every field, assignment -> density -> P_l -> B_l with no host synchronisation.
"""
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from .catalogue import rsd_factor
from .core import Context, kf3_f32


def itot_count(nbins, kbin, kcenter, iopen, cross=False):
    """Number of triangles of the reference loop (pbi4/pbi4.py:105-117)."""
    ishift = int(kcenter / kbin) + (1 if iopen else 0)
    n = 0
    for l in range(1, nbins + 1):
        for j in range(1, l + 1):
            imax = min(nbins, j + l - 1 + ishift) if cross else j
            n += max(0, imax - max(1, l - j + 1 - ishift) + 1)
    return n


@dataclass(frozen=True)
class Field:
    """A density field to measure: real space, or redshift space along `los` (already divided
    by |los|^2 as pbi4/pbi4.py:103 does)."""
    rsd: bool = False
    los: tuple = (0.0, 0.0, 1.0)


@dataclass(frozen=True)
class Measurement:
    grid: int = 256
    order: int = 4
    interlacing: bool = True
    box: float = 1000.0
    kbin: float = 2.0
    kcenter: float = 2.0
    nbins_ps: int = 0            # 0 = floor((N/2 + kbin - kcenter)/kbin)
    nbins_bisp: int = 16         # 0 = floor((N/3 + kbin - kcenter)/kbin)
    iopen: bool = True
    fields: tuple = field(default_factory=lambda: (Field(False), Field(True)))
    precision: int = _lib.PREC_BF16X3
    maps_grid: int = 0
    z: float = 0.5
    om: float = 0.3
    ol: float = 0.7

    def nb_ps(self):
        return self.nbins_ps or int((self.grid / 2 + self.kbin - self.kcenter) // self.kbin)

    def nb_bisp(self):
        return self.nbins_bisp or int((self.grid / 3 + self.kbin - self.kcenter) // self.kbin)

    def itot(self):
        return itot_count(self.nb_bisp(), self.kbin, self.kcenter, self.iopen)

    def kf(self):
        return np.float32(2.0 * np.pi / self.box)


class CataloguePlan:
    """Shared state for catalogues of one Measurement shape on one context."""

    def __init__(self, ctx: Context, m: Measurement):
        self.ctx, self.m = ctx, m
        nb = m.nb_bisp()
        self.itot = m.itot()
        ntr = ctx.count_triangles(nb, m.grid, m.kbin, m.kcenter, m.iopen, False, self.itot)
        self.ntr = ntr
        # the reference's counts(m) = N_tr N^3 as REAL(4) (bispectrumI4.f90:8, f2py coercion)
        counts = (ntr.astype(np.float64) * float(m.grid) ** 3).astype(np.float32)
        self.counts = torch.from_numpy(counts).to(ctx.device)
        self.desc = {}
        self.pkdesc = {}
        for f in m.fields:
            self.desc[f] = ctx.bk_desc(nb, m.grid, m.kbin, m.kcenter, m.iopen, m.kf(), self.itot,
                                       1, f.rsd, f.los, m.precision, m.maps_grid)
            self.pkdesc[f] = ctx.pk_desc(m.grid, m.box, m.kbin, m.kcenter, m.nb_ps(), f.rsd, f.los)
        self.maps_grid = ctx.maps_grid(next(iter(self.desc.values())))
        self.bk_ws_bytes = max(ctx.pk_bk_workspace(self.pkdesc[f], self.desc[f]) for f in m.fields)
        self.rsd_coef = rsd_factor(m.z, m.om, m.ol)
        self._layouts = {}            # npart -> arena_layout

    def arena_layout(self, npart):
        """Byte offsets of one catalogue's device buffers inside one workspace: the assignment
        grids, the half-spectrum delta and one scratch area that the assignment, the density and
        the fused P_l + B pass use in turn (they run in stream order).  Returns
        (grids_off, delta_off, scratch_off, total)."""
        hit = self._layouts.get(int(npart))
        if hit is not None:
            return hit
        m, ctx = self.m, self.ctx
        N = m.grid
        ng = 2 if m.interlacing else 1
        al = lambda x: (int(x) + 255) // 256 * 256
        scratch = max(ctx.assign_workspace(N, m.order, m.interlacing, npart),
                      ctx.density_workspace(N, m.order, m.interlacing, m.box, max(int(npart), 1)),
                      self.bk_ws_bytes)
        g_off = 0
        d_off = al(g_off + 4 * ng * N ** 3)
        s_off = al(d_off + 8 * N * N * (N // 2 + 1))
        out = self._layouts[int(npart)] = (g_off, d_off, s_off, al(s_off + scratch))
        return out

    def bytes_per_catalogue(self, npart):
        """Device memory one catalogue of npart particles needs while in flight (its workspace
        plus the small outputs)."""
        return self.arena_layout(npart)[3] + (16 << 20)

    def run(self, pos, vel, npart, stream=None, precision=None, arena=None):
        """Enqueue every field of one catalogue; returns {field: (pout, bout)} device tensors.
        pos/vel: (3, npart) float32 box-unit positions and velocities (vel may be None when no
        field is in redshift space).  arena: optional uint8 device tensor of at least
        arena_layout(npart)[3] bytes holding every intermediate buffer (else they are allocated);
        it must not be touched by other streams until this catalogue's work has finished."""
        m, ctx = self.m, self.ctx
        N = m.grid
        ng = 2 if m.interlacing else 1
        out = {}
        s = stream if stream is not None else torch.cuda.current_stream()
        grids = delta = scratch = None
        if arena is not None:
            g_off, d_off, s_off, total = self.arena_layout(npart)
            if arena.numel() < total:
                raise ValueError(f"arena holds {arena.numel()} bytes, {total} needed")
            grids = arena[g_off:g_off + 4 * ng * N ** 3].view(torch.float32).view(ng, N, N, N)
            delta = arena[d_off:d_off + 8 * N * N * (N // 2 + 1)].view(torch.complex64)
            delta = delta.view(N, N, N // 2 + 1)
            scratch = arena[s_off:total]
        with torch.cuda.stream(s):
            for f in m.fields:
                rsd = tuple(float(np.float32(self.rsd_coef * l)) for l in f.los) if f.rsd else None
                d = ctx.assign_density(pos, N, m.order, m.interlacing, m.box, npart,
                                       vel=vel if f.rsd else None, rsd=rsd, normalised=False,
                                       grids=grids, out=delta, ws=scratch, stream=s)
                desc = self.desc[f]
                desc.npart = int(npart)
                if precision is not None:
                    desc.precision = int(precision)
                pout, bout = ctx.pk_bk(d, self.counts, self.pkdesc[f], desc, ws=scratch, stream=s)
                desc.precision = int(m.precision)
                out[f] = (pout, bout)
        return out
