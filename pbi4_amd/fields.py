"""Device-resident stand-ins for the arrays the reference passes between its methods.

The reference hands complex128 (N, N, N) Fortran-order numpy arrays from one call to the next
(DENSITIES.assign_grid -> compute_density -> compute_density_multipoles -> MEASUREMENTS.*,
pbi4/pbi4.py:290-553).  Here the data stay on the GPU in the layout the kernels use; each object
converts to exactly the reference's array when numpy asks for it (np.asarray(obj)), so code that
inspects the arrays keeps working.
"""
import numpy as np
import torch

from .core import Pbi4Error


def _signed(N):
    i = np.arange(N)
    return np.where(i <= N // 2, i, i - N).astype(np.float64)


def window_1d(N, order):
    """tWk of fcomb (powerI4.f90:278-292), FP64."""
    tp = np.float64(np.float32(6.28319))
    i = np.arange(N // 2 + 1, dtype=np.float64)
    rk = tp / N * i
    with np.errstate(divide="ignore", invalid="ignore"):
        x = np.sin(rk / 2.0) / (rk / 2.0)
    w = x ** order
    w[0] = 1.0
    return w


class DeviceCatalogue:
    """Output of IO.read_catalogues: the normalised positions (3, npart) as a device tensor
    (float32, or float64 where the reference computes in float64).  np.asarray gives the
    reference's array, which is float64 (its final np.minimum with the float64 scalar
    1 - nextafter(0, 1) promotes); shape, dtype, T, astype, len and indexing behave as on it."""

    def __init__(self, pos):
        self.pos = pos

    @property
    def shape(self):
        return tuple(self.pos.shape)

    @property
    def dtype(self):
        return np.dtype(np.float64)

    @property
    def T(self):
        return np.asarray(self).T

    def astype(self, dtype):
        return np.asarray(self).astype(dtype)

    def __len__(self):
        return self.pos.shape[0]

    def __getitem__(self, idx):
        return np.asarray(self)[idx]

    def __array__(self, dtype=None, copy=None):
        a = self.pos.cpu().numpy().astype(np.float64)
        return a.astype(dtype) if dtype is not None else a


class DeviceGrid:
    """Output of assign_grid: grid 1 (and the interlacing grid 2) as float32 [z][y][x]."""

    def __init__(self, grids, grid, order, interlacing, npart):
        self.grids, self.grid, self.order = grids, int(grid), int(order)
        self.interlacing, self.npart = bool(interlacing), int(npart)

    def __array__(self, dtype=None, copy=None):
        g = self.grids.double().cpu().numpy().transpose(0, 3, 2, 1)     # [g][x][y][z]
        a = g[0] + 1j * g[1] if self.interlacing else g[0]
        a = np.asfortranarray(a)
        return a.astype(dtype) if dtype is not None else a


class DeviceDensity:
    """Output of compute_density: the ix <= N/2+1 half of the reference's dcl as a complex64
    [z][y][x'] half spectrum (the only half any later routine reads)."""

    def __init__(self, delta, grid, order, interlacing, box, npart, upper_zero=False):
        self.delta, self.grid, self.order = delta, int(grid), int(order)
        self.interlacing, self.box, self.npart = bool(interlacing), float(box), npart
        # a density reduced from a larger grid (read_density) has a zero x > N/2 half, as the
        # reference's reduce_array output does
        self.upper_zero = bool(upper_zero)

    def full(self):
        """The reference's full dcl (N, N, N) complex128, indexed [ix, iy, iz]."""
        N = self.grid
        h = N // 2 + 1
        low = self.delta.cpu().numpy().astype(np.complex128).transpose(2, 1, 0)   # [x][y][z]
        D = np.empty((N, N, N), dtype=np.complex128, order="F")
        D[:h] = low
        if self.upper_zero:
            D[h:] = 0.0
        elif N > 2:
            xs = np.arange(h, N)
            iy = (N - np.arange(N)) % N
            mirror = low[N - xs][:, iy][:, :, iy]
            if self.interlacing:
                # stores 5-8 of powerI4.f90:339-342: conj of the mirrored cell
                D[h:] = np.conj(mirror)
            else:
                # upper half left unscaled by the non-interlaced fcomb (powerI4.f90:351-374)
                kf3 = np.float64(np.float32(np.float32(np.float32(np.float32(
                    np.float32(6.28319) * np.float32(6.28319)) * np.float32(6.28319)) /
                    np.float32(self.box)) / np.float32(self.box)) / np.float32(self.box))
                w = window_1d(N, self.order)
                a = np.minimum(np.arange(N), N - np.arange(N))
                W = w[N - xs][:, None, None] * w[a][None, :, None] * w[a][None, None, :]
                D[h:] = np.conj(mirror * W * kf3)
        return D

    def __array__(self, dtype=None, copy=None):
        a = self.full()
        return a.astype(dtype) if dtype is not None else a


class DeviceMultipole:
    """delta_2 or delta_4 of compute_density_multipoles for a fixed line of sight: L_l(mu) delta
    (pbi4/pbi4.py:343-388), kept symbolic; the kernels apply L_l per mode."""

    def __init__(self, density, ell, los):
        self.density, self.ell, self.los = density, int(ell), np.asarray(los, dtype=np.float64)

    def __array__(self, dtype=None, copy=None):
        d = self.density
        N = d.grid
        h = N // 2 + 1
        D = d.full()
        f = _signed(N)
        kx, ky, kz = f[:h, None, None], f[None, :, None], f[None, None, :]
        k2 = kx * kx + ky * ky + kz * kz
        l = self.los
        with np.errstate(invalid="ignore", divide="ignore"):
            mu2 = (l[0] * kx + l[1] * ky + l[2] * kz) ** 2 / k2
        mu2[0, 0, 0] = 0.0
        L2 = -0.5 + 1.5 * mu2
        s2 = float(np.sum(l)) ** 2
        c2 = -0.5 + 1.5 * s2           # the unfiltered ix > N/2+1 half of pbi4.py:362-369
        if self.ell == 2:
            out = D * c2
            out[:h] = D[:h] * L2
        else:
            L4 = 0.375 - 3.75 * mu2 + 4.375 * mu2 * mu2
            out = D * (-7.0 / 8.0 - 2.5 * c2 + 35.0 / 8.0 * s2 * s2)
            out[:h] = D[:h] * L4
        out = np.asfortranarray(out)
        return out.astype(dtype) if dtype is not None else out


def half_from_array(a, device):
    """Device complex64 [z][y][x'] half spectrum from a reference-layout (N, N, N) array."""
    a = np.asarray(a)
    if a.ndim != 3 or a.shape[0] != a.shape[1] or a.shape[1] != a.shape[2]:
        raise Pbi4Error("expected an (N, N, N) array")
    N = a.shape[0]
    low = np.ascontiguousarray(a[:N // 2 + 1].transpose(2, 1, 0), dtype=np.complex64)
    return torch.from_numpy(low).to(device)


def full_from_array(a, device):
    """Device complex64 [z][y][x] copy of a reference-layout (N, N, N) array."""
    a = np.asarray(a)
    return torch.from_numpy(np.ascontiguousarray(a.transpose(2, 1, 0), dtype=np.complex64)).to(device)


def full_to_array(t):
    """Reference-layout complex128 (N, N, N) Fortran-order array from a [z][y][x] tensor."""
    return np.asfortranarray(t.cpu().numpy().astype(np.complex128).transpose(2, 1, 0))
