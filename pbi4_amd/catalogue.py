"""Seeded synthetic catalogues: Zel'dovich-displaced particle lattices with velocities.

Gaussian linear field on an n^3 particle lattice in a periodic box L with
P(k) = amp * (k / kpivot)^slope  [(Mpc/h)^3], cut at the lattice Nyquist |k| <= pi n / L;
displacement Psi(k) = i k / k^2 delta(k), positions x = q + Psi (wrapped into [0, L)),
velocities v = f a H Psi + N(0, sigma_v^2) per component [km/s], f = Omega_m(z)^0.55.
The redshift-space shift of the reference, pos += v * los / H(z) * (1 + z)
(pbi4/pbi4.py:154), is then a shift of f Psi_los plus the dispersion term.

Runs on any torch device; positions and velocities are float32 (3, n^3) arrays in box units.
The noise comes from torch's generator on that device, so a catalogue is reproducible for a
given (seed, device) pair.
"""
import math

import torch

OM, OL, Z = 0.3, 0.7, 0.5      # Tutorials/Notebook/config.json


def hubble(z=Z, om=OM, ol=OL):
    """H(z) in (km/s)/(Mpc/h), as IO.derived_params (pbi4/pbi4.py:102)."""
    return 100.0 * math.sqrt(om * (1.0 + z) ** 3 + ol)


def growth_rate(z=Z, om=OM, ol=OL):
    omz = om * (1.0 + z) ** 3 / (om * (1.0 + z) ** 3 + ol)
    return omz ** 0.55


def rsd_factor(z=Z, om=OM, ol=OL):
    """(1+z)/H(z): velocity [km/s] -> comoving displacement [Mpc/h] along the line of sight."""
    return (1.0 + z) / hubble(z, om, ol)


def zeldovich(n, box=1000.0, amp=2.0e4, slope=-1.5, kpivot=0.1, seed=0, sigma_v=300.0,
              z=Z, om=OM, ol=OL, device="cpu", velocities=True):
    dev = torch.device(device)
    g = torch.Generator(device=dev)
    g.manual_seed(int(seed))
    noise = torch.randn((n, n, n), generator=g, device=dev, dtype=torch.float32)
    wk = torch.fft.rfftn(noise)
    del noise
    kf = 2.0 * math.pi / box
    kx = torch.fft.fftfreq(n, d=1.0 / n, device=dev) * kf
    kz = torch.fft.rfftfreq(n, d=1.0 / n, device=dev) * kf
    k2 = kx[:, None, None] ** 2 + kx[None, :, None] ** 2 + kz[None, None, :] ** 2
    k = torch.sqrt(k2)
    knyq = math.pi * n / box
    pk = torch.where((k > 0) & (k <= knyq), amp * (k.clamp_min(1e-30) / kpivot) ** slope,
                     torch.zeros_like(k))
    # delta_lattice = irfftn(w_k sqrt(P n^3 / V)) has variance sum_k P / V
    amp_k = torch.sqrt(pk * (n ** 3) / box ** 3)
    dk = wk * amp_k
    del wk, pk, amp_k
    inv_k2 = torch.where(k2 > 0, 1.0 / k2.clamp_min(1e-30), torch.zeros_like(k2))
    q = (torch.arange(n, device=dev, dtype=torch.float32) + 0.5) * (box / n)
    pos = torch.empty((3, n ** 3), dtype=torch.float32, device=dev)
    vel = torch.empty((3, n ** 3), dtype=torch.float32, device=dev) if velocities else None
    f = growth_rate(z, om, ol)
    fah = f * hubble(z, om, ol) / (1.0 + z)
    axes = [kx[:, None, None], kx[None, :, None], kz[None, None, :]]
    grids = torch.meshgrid(q, q, q, indexing="ij")
    for d in range(3):
        psi = torch.fft.irfftn(1j * axes[d] * inv_k2 * dk, s=(n, n, n)).float()
        # axis d of the lattice array is coordinate d
        x = grids[d] + psi
        pos[d] = torch.remainder(x, box).reshape(-1)
        if velocities:
            v = fah * psi + sigma_v * torch.randn((n, n, n), generator=g, device=dev)
            vel[d] = v.reshape(-1)
        del psi, x
    pos = torch.where(pos >= box, pos - box, pos)
    return pos, vel


def normalise(pos, box):
    """Positions in [0, 1) as IO.read_catalogues returns them (pbi4/pbi4.py:242-243)."""
    p = torch.remainder(pos + box, box) / box
    return torch.clamp(p, max=1.0)
