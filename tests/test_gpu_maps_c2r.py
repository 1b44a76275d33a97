"""Shell maps: the library's own inverse transforms against rocFFT's.

Band-limited maps grid (3R < N' < N): the shell filter does the inverse transform over z and a 2D
c2r over (y, x) finishes each map plane; for the maps grids 100, 150 and 200 that c2r is k_c2r2d
(four-step transforms in registers and LDS over the x' <= R columns that can be nonzero, then a
half-length complex DFT per row).  Full grid (N' = N, the aliasing requests of the reference's
default shell counts) with 2R < N on 128, 192, 256, 384: the same fused filter, then a column pass
(inverse DFTs along y, in place) and a row pass (each row's c2r) through HBM, instead of rocFFT's
3D c2r.  PBI4_MAPS_C2R=rocfft, and every other size, keep rocFFT.  Both are unnormalised FP32
transforms of the same spectra, so the bispectrum must agree to FP32 rounding: rows that do not
involve the maps (k's, shell powers, shot noise, N_tr) bitwise, B, B_2, B_4 within 3e-5
(band-limited) or 1e-4 (full grid) of max(|B_l|, |B_0|, sigma_B) (sigma_B =
sqrt(P1 P2 P3 / (N_tr kf^3)), the estimator's Gaussian standard deviation), in FP32 sums so the
comparison sees the maps only.  Measured: 1.2e-6 to 1.3e-5 on the band-limited grids; 9.4e-6,
3.4e-6, 1.8e-5 and 5.4e-5 on the full 128^3, 128^3, 192^3 and 256^3 grids, where the sums run over
up to 1.7e7 points of strongly cancelling triple products (the FP32 sums of two exact transforms'
roundings).  Sizes without a transform of the library's own must give bitwise the rocFFT result.
Against the FP64 oracle the default routes are held to 1e-4 by the parity and configuration
tests, the full-grid route by the cfg4 test's 128^3, 192^3 and 384^3 requests.
"""
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import pbi4_oracle as O  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from pbi4_amd.core import Context
    return Context(0)


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _run(ctx, mode, fn):
    old = os.environ.get("PBI4_MAPS_C2R")
    try:
        if mode is None:
            os.environ.pop("PBI4_MAPS_C2R", None)
        else:
            os.environ["PBI4_MAPS_C2R"] = mode
        return fn().cpu().numpy()
    finally:
        if old is None:
            os.environ.pop("PBI4_MAPS_C2R", None)
        else:
            os.environ["PBI4_MAPS_C2R"] = old


# (grid, shells, kbin, kcenter, iopen, multipoles, maps grid)
CASES = [
    (128, 16, 2.0, 2.0, True, True, 100),    # the 256^3 serving shape's 16 shells x 3
    (128, 16, 2.0, 2.0, True, False, 100),
    (256, 48, 1.0, 1.0, True, False, 0),     # auto: 150 (LDS beyond 64 KB)
    (256, 20, 1.0, 1.0, False, True, 150),
    (256, 32, 2.0, 2.0, True, True, 0),      # auto: 200 (LDS ~160 KB)
    (32, 5, 1.0, 1.0, True, False, 30),      # rocFFT 2D
    (64, 8, 2.0, 2.0, True, True, 0),        # auto: 50, rocFFT 2D
    (128, 21, 2.0, 2.0, True, True, 0),      # the reference's default shells: full grid, 2 passes
    (128, 21, 2.0, 2.0, True, False, 0),
    (192, 32, 2.0, 2.0, True, True, 0),
    (256, 42, 2.0, 2.0, True, True, 0),
    (256, 16, 2.0, 2.0, True, True, 256),    # full grid forced on band-limited shells
    (96, 16, 2.0, 2.0, True, True, 0),       # full grid, 96: rocFFT 3D
]
BAND_OWN = (100, 150, 200)
FULL_OWN = (128, 192, 256, 384)


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"N{c[0]}_nb{c[1]}_ell{int(c[5])}_mg{c[6]}")
def test_own_maps_transform_matches_rocfft(ctx, case):
    from pbi4_amd.catalogue import zeldovich
    N, nb, kb, kc, iopen, mult, mg = case
    n_side = min(N, 128)
    print(f"maps c2r LDS vs rocFFT: N={N} nb={nb} kb={kb} multipoles={mult} maps_grid={mg}",
          flush=True)
    p, v = zeldovich(n_side, box=1000.0, seed=70 + nb, device="cuda")
    n = p.shape[1]
    rsd = (0.0, 0.0, 0.011462) if mult else None
    g = ctx.assign(p, N, 4, True, vel=v if mult else None, box=1000.0, rsd=rsd, normalised=False)
    delta = ctx.density(g, N, 4, True, 1000.0, n)
    itot = O.itot_count(nb, kb, kc, iopen)
    ntr = ctx.count_triangles(nb, N, kb, kc, iopen, False, itot)
    counts = _dev((ntr.astype(np.float64) * N ** 3).astype(np.float32))
    kf = np.float32(2 * np.pi / 1000.0)
    desc = ctx.bk_desc(nb, N, kb, kc, iopen, kf, itot, n, mult, (0.0, 0.0, 1.0), 0, mg)
    Np = ctx.maps_grid(desc)
    own = Np in BAND_OWN if Np < N else Np in FULL_OWN
    b_own = _run(ctx, None, lambda: ctx.bk(delta, counts, desc))
    b_roc = _run(ctx, "rocfft", lambda: ctx.bk(delta, counts, desc))
    brows = [6, 9, 10] if mult else [6]
    other = [r for r in range(b_own.shape[1]) if r not in brows]
    assert np.array_equal(b_own[:, other], b_roc[:, other])
    assert np.all(np.isfinite(b_own))
    if not own:
        assert np.array_equal(b_own, b_roc), "a maps grid without a transform of the library's own"
        return
    assert not np.array_equal(b_own[:, 6], b_roc[:, 6]), "the library's transform was not taken"
    sigma_b = np.sqrt(np.abs(b_roc[:, 3] * b_roc[:, 4] * b_roc[:, 5])
                      / (np.maximum(b_roc[:, 8], 1) * float(kf) ** 3))
    b0 = b_roc[:, 6] - b_roc[:, 7]
    worst = 0.0
    for r in brows:
        ref = b0 if r == 6 else b_roc[:, r]
        mine = b_own[:, 6] - b_own[:, 7] if r == 6 else b_own[:, r]
        scale = np.maximum(np.maximum(np.abs(ref), np.abs(b0)), sigma_b)
        e = np.abs(mine - ref) / scale
        worst = max(worst, float(e.max()))
    print(f"  N'={Np} ({'band-limited' if Np < N else 'full grid'}): max |B_own - B_rocfft| / scale"
          f" = {worst:.3g}", flush=True)
    assert worst <= (3e-5 if Np < N else 1e-4)
