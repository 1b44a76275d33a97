"""Interlaced density with the FFT's z axis fused into the combine, against rocFFT's 3D transform.

On the grids 128, 192, 256, 384 and 512 pbi4_compute_density has rocFFT transform every z plane of
grid1 + i grid2 over (y, x) and one kernel does the forward DFT along z in LDS and the
interlacing combine (k_zfft_combine); PBI4_DENSITY_FFT=rocfft3d keeps rocFFT's 3D c2c and the
separate combine.  Both are unnormalised FP32 transforms of the same grids followed by the same
combine arithmetic, so the half spectra agree to FP32 rounding of the transforms: measured at
1.2e-7 to 2.3e-7 of the largest mode on Zel'dovich catalogues, held here to 1e-6 of it.  P(k)
and the mode counts of both, binned by the same kernel, agree to 1e-5 (counts bitwise).
The half spectra against the FP64 oracle are covered by the configuration tests (cfg2 at 256^3,
cfg4 at 128^3-384^3) that run on this path.
"""
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from pbi4_amd.core import Context
    return Context(0)


def _density(ctx, g, N, n, mode):
    old = os.environ.get("PBI4_DENSITY_FFT")
    try:
        if mode is None:
            os.environ.pop("PBI4_DENSITY_FFT", None)
        else:
            os.environ["PBI4_DENSITY_FFT"] = mode
        return ctx.density(g, N, 4, True, 1000.0, n)
    finally:
        if old is None:
            os.environ.pop("PBI4_DENSITY_FFT", None)
        else:
            os.environ["PBI4_DENSITY_FFT"] = old


@pytest.mark.parametrize("N", [128, 192, 256, 384, 512])
def test_fused_z_transform_matches_rocfft_3d(ctx, N):
    from pbi4_amd.catalogue import zeldovich
    print(f"fused z transform + combine vs rocFFT 3D: N={N}", flush=True)
    p, _ = zeldovich(min(N, 256), box=1000.0, seed=40 + N // 64, device="cuda", velocities=False)
    n = p.shape[1]
    g = ctx.assign(p, N, 4, True, box=1000.0, normalised=False)
    fused = _density(ctx, g, N, n, None)
    ref = _density(ctx, g, N, n, "rocfft3d")
    a, b = fused.cpu().numpy(), ref.cpu().numpy()
    assert np.all(np.isfinite(a))
    scale = float(np.abs(b).max())
    err = float(np.abs(a - b).max()) / scale
    print(f"  max |fused - rocFFT 3D| / max |delta| = {err:.3g}", flush=True)
    assert err <= 1e-6
    # P_l with multipoles (columns k, P0, P2, P4, N_modes) binned by the same kernel from both
    pa = ctx.pk(fused, N, 1000.0, 2.0, 2.0, N // 4, multipoles=True).cpu().numpy()
    pb = ctx.pk(ref, N, 1000.0, 2.0, 2.0, N // 4, multipoles=True).cpu().numpy()
    assert np.array_equal(pa[:, 4], pb[:, 4]) and np.array_equal(pa[:, 0], pb[:, 0])
    p0 = np.abs(pb[:, 1])
    for c in (1, 2, 3):
        e = float((np.abs(pa[:, c] - pb[:, c]) / p0).max())
        print(f"  P_{2 * (c - 1)}: max |fused - rocFFT 3D| / |P_0| = {e:.3g}", flush=True)
        assert e <= 1e-5


def _two_step(ctx, p, N, order, interlacing, n, vel=None, rsd=None):
    g = ctx.assign(p, N, order, interlacing, vel=vel, box=1000.0, rsd=rsd, normalised=False)
    return g, ctx.density(g, N, order, interlacing, 1000.0, n)


@pytest.mark.parametrize("N,npart_side,order,rsd", [
    (128, 128, 4, False), (128, 192, 3, True), (192, 256, 4, True), (256, 256, 4, True),
    (384, 192, 2, False), (512, 256, 4, True)])
def test_assign_density_fused_gather_matches_two_steps(ctx, N, npart_side, order, rsd):
    """pbi4_assign_density on the fused route (gather + x transform, y transform, z transform +
    combine) against pbi4_assign + pbi4_compute_density (gather, rocFFT over (y, x), z transform +
    combine).  The cells the fused gather transforms are summed in k_gather's order, so the two
    differ by the FP32 rounding of the transforms only: the (y, x) transform left in the grids
    buffer is held to 1e-6 of its largest mode against torch.fft.fft2 of the gathered grids, the
    half spectra to 1e-6 of the largest |delta|, P_l to 1e-5 of |P_0| (mode counts bitwise).
    The particle counts give tile edges 16, 8 and 12 (128^3 and 192^3 particles on 128^3, 256^3
    on 192^3)."""
    from pbi4_amd.catalogue import zeldovich
    print(f"assign_density fused: N={N} npart={npart_side}^3 order={order} rsd={rsd}", flush=True)
    p, v = zeldovich(npart_side, box=1000.0, seed=7 + N, device="cuda", velocities=rsd)
    n = p.shape[1]
    los = (0.0, 0.0, 0.01) if rsd else None
    g, ref = _two_step(ctx, p, N, order, True, n, vel=v if rsd else None, rsd=los)
    grids = torch.empty((2, N, N, N), dtype=torch.float32, device="cuda")
    fused = ctx.assign_density(p, N, order, True, 1000.0, n, vel=v if rsd else None, rsd=los,
                               normalised=False, grids=grids)
    torch.cuda.synchronize()
    # grids holds the (y, x) transform of grid1 + i grid2 as interleaved complex [z][y][x]
    G = torch.view_as_complex(grids.view(-1).view(N, N, N, 2))
    want = torch.fft.fft2(torch.complex(g[0], g[1]), dim=(1, 2))
    gs = float(want.abs().max())
    ge = float((G - want).abs().max()) / gs
    print(f"  (y, x) transform: max |fused - torch fft2| / max = {ge:.3g}", flush=True)
    assert ge <= 1e-6
    a, b = fused.cpu().numpy(), ref.cpu().numpy()
    assert np.all(np.isfinite(a))
    err = float(np.abs(a - b).max()) / float(np.abs(b).max())
    print(f"  max |fused - two steps| / max |delta| = {err:.3g}", flush=True)
    assert err <= 1e-6
    pa = ctx.pk(fused, N, 1000.0, 2.0, 2.0, N // 4, multipoles=True).cpu().numpy()
    pb = ctx.pk(ref, N, 1000.0, 2.0, 2.0, N // 4, multipoles=True).cpu().numpy()
    assert np.array_equal(pa[:, 4], pb[:, 4]) and np.array_equal(pa[:, 0], pb[:, 0])
    p0 = np.abs(pb[:, 1])
    for c in (1, 2, 3):
        assert float((np.abs(pa[:, c] - pb[:, c]) / p0).max()) <= 1e-5


@pytest.mark.parametrize("N,interlacing", [(64, True), (128, False), (96, True)])
def test_assign_density_other_shapes_are_the_two_steps(ctx, N, interlacing):
    """Off the fused route (single grid, or a grid the fused kernels are not built for)
    pbi4_assign_density runs pbi4_assign then pbi4_compute_density: bitwise the same."""
    from pbi4_amd.catalogue import zeldovich
    print(f"assign_density fallback: N={N} interlacing={interlacing}", flush=True)
    p, _ = zeldovich(64, box=1000.0, seed=3, device="cuda", velocities=False)
    n = p.shape[1]
    _, ref = _two_step(ctx, p, N, 3, interlacing, n)
    got = ctx.assign_density(p, N, 3, interlacing, 1000.0, n, normalised=False)
    assert torch.equal(torch.view_as_real(got), torch.view_as_real(ref))


def test_assign_density_rejects_mismatched_descriptors(ctx):
    """The assignment and density descriptors must name the same grid, order and interlacing:
    a mismatch is refused before any launch, with the reason in pbi4_last_error."""
    import ctypes
    from pbi4_amd import _lib
    from pbi4_amd._lib import AssignDesc, DensityDesc, lib
    print("assign_density: mismatched descriptors", flush=True)
    n = 1000
    pos = torch.rand((3, n), dtype=torch.float32, device="cuda")
    ad = AssignDesc(128, 4, 1, _lib.POS_NORMALISED, 1.0, (ctypes.c_float * 3)(0.0, 0.0, 0.0))
    dd = DensityDesc(256, 4, 1, 1000.0, float(n))
    grids = torch.empty((2, 256, 256, 256), dtype=torch.float32, device="cuda")
    delta = torch.empty((256, 256, 129), dtype=torch.complex64, device="cuda")
    ws = torch.empty(1 << 20, dtype=torch.uint8, device="cuda")
    rc = lib().pbi4_assign_density(ctx.h, ctypes.byref(ad), ctypes.byref(dd),
                                   ctypes.c_void_p(pos.data_ptr()), None, n,
                                   ctypes.c_void_p(grids.data_ptr()),
                                   ctypes.c_void_p(delta.data_ptr()),
                                   ctypes.c_void_p(ws.data_ptr()), ws.numel(), None)
    assert rc == -1                                     # PBI4_EINVAL
    assert b"differ" in lib().pbi4_last_error(ctx.h)
