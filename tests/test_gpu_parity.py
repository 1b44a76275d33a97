"""GPU parity of each stage of the HIP path against the FP64 oracle and the golden fixtures.

Tolerances (FP32 device arithmetic against FP64):
  grids: 2e-6 of the largest cell; delta: 2e-5 of max|delta|; P_l: 1e-4 relative of P0 per bin;
  B: 1e-4 relative with an absolute floor of 1e-4 x the bin's Gaussian scale (see _bk_close);
  mode counts and N_tr: bit-exact.
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


def _pos(n, seed, clustered=True):
    rng = np.random.default_rng(seed)
    if not clustered:
        return rng.random((3, n)).astype(np.float32)
    c = rng.random((3, 6))
    k = rng.integers(0, 6, n)
    p = (c[:, k] + 0.06 * rng.standard_normal((3, n))) % 1.0
    return np.minimum(p.astype(np.float32), np.float32(1.0 - 2 ** -24))


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _grid_to_xyz(t):
    """[z][y][x] device grid -> numpy indexed [x, y, z]."""
    return t.cpu().numpy().transpose(2, 1, 0)


def _oracle_density(pos, N, order, inter, box=1000.0):
    n = pos.shape[1]
    if inter:
        g = O.assign_double_grid(N, order, n, pos.astype(np.float64))
    else:
        g = O.assign_single_grid(N, order, n, pos.astype(np.float64)).astype(np.complex128)
    return g, O.compute_density(n, g, N, order, inter, box)


# ------------------------------------------------------------------ assignment


@pytest.mark.parametrize("order", [1, 2, 3, 4])
@pytest.mark.parametrize("inter", [0, 1])
@pytest.mark.parametrize("N", [16, 32])
def test_assign(ctx, order, inter, N):
    pos = _pos(20000, 100 + order + N)
    g = ctx.assign(_dev(pos), N, order, inter)
    torch.cuda.synchronize()
    ref = (O.assign_double_grid if inter else O.assign_single_grid)(N, order, pos.shape[1],
                                                                     pos.astype(np.float64))
    mine = _grid_to_xyz(g[0])
    exp = ref.real if inter else ref
    scale = np.abs(exp).max()
    assert np.abs(mine - exp).max() <= 2e-6 * scale
    if inter:
        assert np.abs(_grid_to_xyz(g[1]) - ref.imag).max() <= 2e-6 * scale
    assert np.allclose(g.double().sum(dim=(1, 2, 3)).cpu().numpy(), pos.shape[1], rtol=1e-5)


def test_assign_deterministic_and_box_mode(ctx):
    N = 32
    rng = np.random.default_rng(7)
    box = 1000.0
    x = (rng.random((3, 30000)) * box).astype(np.float32)
    v = (rng.standard_normal((3, 30000)) * 300).astype(np.float32)
    rsd = (0.0, 0.0, np.float32(0.011462))
    a = ctx.assign(_dev(x), N, 4, True, vel=_dev(v), box=box, rsd=rsd, normalised=False)
    b = ctx.assign(_dev(x), N, 4, True, vel=_dev(v), box=box, rsd=rsd, normalised=False)
    assert torch.equal(a, b)
    # reference normalisation: (s + box) % box / box in float32 (pbi4.py:154, 242-243)
    s = x.astype(np.float64)
    s[2] += v[2].astype(np.float64) * np.float64(rsd[2])
    s = s.astype(np.float32)
    p = ((s + np.float32(box)) % np.float32(box)) / np.float32(box)
    ref = O.assign_double_grid(N, 4, p.shape[1], p.astype(np.float64))
    scale = np.abs(ref.real).max()
    assert np.abs(_grid_to_xyz(a[0]) - ref.real).max() <= 2e-5 * scale


@pytest.mark.parametrize("N", [20, 24, 36])
def test_assign_and_density_tiles_other_than_16(ctx, N):
    # tile sizes 10, 12, 12 (the largest divisor of N in [4, 16]): regions and their gather
    n = 30000
    pos = _pos(n, 300 + N)
    for order, inter in ((4, 1), (3, 1), (2, 0), (1, 1)):
        g = ctx.assign(_dev(pos), N, order, inter)
        ref = (O.assign_double_grid if inter else O.assign_single_grid)(N, order, n, pos.astype(np.float64))
        exp = ref.real if inter else ref
        scale = np.abs(exp).max()
        assert np.abs(_grid_to_xyz(g[0]) - exp).max() <= 2e-6 * scale, (order, inter)
        if inter:
            assert np.abs(_grid_to_xyz(g[1]) - ref.imag).max() <= 2e-6 * scale
    delta = ctx.density(ctx.assign(_dev(pos), N, 4, True), N, 4, True, 1000.0, n)
    _, dcl = _oracle_density(pos, N, 4, True)
    exp = dcl[:N // 2 + 1]
    assert np.abs(delta.cpu().numpy().transpose(2, 1, 0) - exp).max() <= 2e-5 * np.abs(exp).max()


@pytest.mark.parametrize("N", [92, 148])
@pytest.mark.parametrize("ordered", [False, True])
def test_assign_hashed_tile_counters(ctx, N, ordered):
    # 4-cell tiles: 23^3 = 12167 and 37^3 = 50653 tiles, counted in the per-work-group hash table;
    # a random catalogue refills and flushes it every pass, a spatially ordered one makes long runs
    n = 60000
    pos = _pos(n, 900 + N)
    if ordered:
        pos = np.ascontiguousarray(pos[:, np.lexsort((pos[0], pos[1], pos[2]))])
    per_cell = O.assign_single_grid(N, 1, n, pos.astype(np.float64)).max()
    for order, inter in ((4, 1), (2, 0)):
        g = ctx.assign(_dev(pos), N, order, inter)
        ref = (O.assign_double_grid if inter else O.assign_single_grid)(N, order, n, pos.astype(np.float64))
        exp = ref.real if inter else ref
        # FP32 and the fixed-point rounding bound of the edge-position test; a particle lost or
        # counted twice by the sort moves a cell by its weight, orders of magnitude more
        tol = 2e-6 * np.abs(exp).max() + 2.0 ** -19 * order ** 3 * per_cell
        assert np.abs(_grid_to_xyz(g[0]) - exp).max() <= tol, (order, inter)
        if inter:
            assert np.abs(_grid_to_xyz(g[1]) - ref.imag).max() <= tol


def _fixed_point_tolerance(pos, N, order, inter):
    """Per-cell bound on the deposit's fixed-point rounding: half a unit (2^-20 for TSC/PCS, 2^-19
    for CIC/NGP, 2^-18 for single-grid CIC, 32- or 64-bit cells alike) per contribution, one
    contribution per particle whose home cell lies within order + 1 cells."""
    smax = 20 if order >= 3 else (18 if order == 2 and not inter else 19)
    home = np.floor(pos.astype(np.float64) * N).astype(np.int64) % N     # positions < 1
    counts = np.zeros((N, N, N))
    np.add.at(counts, (home[0], home[1], home[2]), 1)
    box = np.zeros_like(counts)
    r = order + 1
    for dx in range(-r, r + 1):
        for dy in range(-r, r + 1):
            for dz in range(-r, r + 1):
                box += np.roll(counts, (dx, dy, dz), (0, 1, 2))
    return 0.5 * 2.0 ** -smax * box


@pytest.mark.parametrize("case", ["clustered32", "clustered64", "pileup32"])
@pytest.mark.parametrize("order,inter", [(4, 1), (3, 0), (2, 0), (1, 1)])
def test_assign_dense_tiles_order_independent(ctx, case, order, inter):
    # 2^20 particles: clustered on 32^3 and 64^3 grids (up to ~1500 particles in a cell: tiles in
    # 64-bit LDS cells, alone or beside 32-bit ones), or half of them piled into one cell; the
    # grids are bitwise the same for any order of the catalogue, conserve mass and match the oracle
    n = 1 << 20
    N = 64 if case == "clustered64" else 32
    pos = _pos(n, 4242 + N)
    if case == "pileup32":
        rng = np.random.default_rng(5)
        pos[:, : n // 2] = ((np.array([[9], [20], [3]]) + rng.random((3, n // 2))) / N).astype(np.float32)
    perm = np.random.default_rng(N).permutation(n)
    a = ctx.assign(_dev(pos), N, order, inter)
    b = ctx.assign(_dev(pos[:, perm]), N, order, inter)
    assert torch.equal(a, b)
    # mass: NGP weights are 1, so its fixed-point grids are exact integers summing to n; other
    # orders round each contribution to the nearest unit, which loses ~3e-7 n net (many
    # contributions fall below half a unit)
    tot = a.double().sum(dim=(1, 2, 3)).cpu().numpy()
    if order == 1:
        assert np.all(tot == n), tot
    else:
        assert np.abs(tot - n).max() <= 2e-6 * n, tot
    ref = (O.assign_double_grid if inter else O.assign_single_grid)(N, order, n, pos.astype(np.float64))
    exp = ref.real if inter else ref
    tol = 2e-6 * np.abs(exp).max() + _fixed_point_tolerance(pos, N, order, inter)
    assert np.all(np.abs(_grid_to_xyz(a[0]) - exp) <= tol)
    if inter:
        assert np.all(np.abs(_grid_to_xyz(a[1]) - ref.imag) <= tol)


@pytest.mark.parametrize("order,inter", [(1, 1), (2, 0), (2, 1), (3, 1), (4, 0), (4, 1)])
def test_assign_edge_positions(ctx, order, inter):
    # coordinates at 0, 2^-24, 1 - 2^-24, exactly 1.0 (the reference's normalisation can return
    # it; it wraps to cell 0), on cell and tile boundaries and one ulp either side of them
    N = 32
    vals = np.array([0.0, 2.0 ** -24, 1.0 - 2.0 ** -24, 1.0, 0.5, 1 / N, 15 / N, 16 / N, 17 / N,
                     16 / N - 2.0 ** -24, 16 / N + 2.0 ** -23, 0.25 + 2.0 ** -20], dtype=np.float32)
    g = np.stack(np.meshgrid(vals, vals, vals, indexing="ij")).reshape(3, -1)
    pos = np.concatenate([g, _pos(2000, 77)], axis=1).astype(np.float32)
    n = pos.shape[1]
    out = ctx.assign(_dev(pos), N, order, inter)
    ref = (O.assign_double_grid if inter else O.assign_single_grid)(N, order, n, pos.astype(np.float64))
    exp = ref.real if inter else ref
    scale = np.abs(exp).max()
    # the lattice piles up to 64 particles near one corner: beside the FP32 term allow the
    # fixed-point rounding of the deposit, at most 2^-19 per contribution (the LDS unit is never
    # coarser than 2^-18), with order^3 x (particles per cell) contributions at most
    per_cell = O.assign_single_grid(N, 1, n, pos.astype(np.float64)).max()
    tol = 2e-6 * scale + 2.0 ** -19 * order ** 3 * per_cell
    assert np.abs(_grid_to_xyz(out[0]) - exp).max() <= tol
    if inter:
        assert np.abs(_grid_to_xyz(out[1]) - ref.imag).max() <= tol


def test_assign_empty_catalogue(ctx):
    g = ctx.assign(torch.empty((3, 0), dtype=torch.float32, device="cuda"), 32, 4, True)
    torch.cuda.synchronize()
    assert g.shape == (2, 32, 32, 32) and not bool(g.any())


@pytest.mark.parametrize("N", [512, 1024])
def test_assign_large_tile_counts(ctx, N):
    # 32768 and 262144 tiles (hashed tile counters, a random catalogue of 2^22 particles): mass
    # conservation, and shifting every particle by exactly one cell
    # along each axis rolls both grids by one cell (up to the per-tile fixed-point rounding)
    n = 1 << 22
    rng = np.random.default_rng(N)
    q = rng.integers(0, 1 << 20, size=(3, n))                 # positions on a 2^-20 lattice
    pos = (q * 2.0 ** -20).astype(np.float32)
    shifted = ((q + (1 << 20) // N) % (1 << 20) * 2.0 ** -20).astype(np.float32)
    a = ctx.assign(_dev(pos), N, 4, True)
    sums = a.double().sum(dim=(1, 2, 3)).cpu().numpy()
    assert np.allclose(sums, n, rtol=1e-5)
    b = ctx.assign(_dev(shifted), N, 4, True)
    rolled = torch.roll(a, shifts=(1, 1, 1), dims=(1, 2, 3))
    assert float((b - rolled).abs().max()) <= 1e-5 * float(a.abs().max())
    del a, b, rolled
    torch.cuda.empty_cache()


# ------------------------------------------------------------------ density


@pytest.mark.parametrize("order,inter", [(4, 1), (3, 1), (2, 1), (3, 0), (4, 0)])
def test_density(ctx, order, inter):
    N, n = 32, 40000
    pos = _pos(n, 5 + order)
    g = ctx.assign(_dev(pos), N, order, inter)
    delta = ctx.density(g, N, order, inter, 1000.0, n)
    torch.cuda.synchronize()
    _, dcl = _oracle_density(pos, N, order, inter)
    h = N // 2 + 1
    mine = delta.cpu().numpy().transpose(2, 1, 0)
    exp = dcl[:h]
    assert np.abs(mine - exp).max() <= 2e-5 * np.abs(exp).max()


# ------------------------------------------------------------------ P(k)


@pytest.mark.parametrize("kb,kc", [(1.0, 1.0), (2.0, 2.0), (1.5, 0.5)])
def test_pk_multipoles(ctx, kb, kc):
    N, n, box = 32, 40000, 1000.0
    pos = _pos(n, 31)
    g = ctx.assign(_dev(pos), N, 4, True)
    delta = ctx.density(g, N, 4, True, box, n)
    los = np.array([0.0, 0.0, 1.0])
    nb = int((N / 2 + kb - kc) // kb)
    pout = ctx.pk(delta, N, box, kb, kc, nb, True, los).cpu().numpy().T
    _, dcl = _oracle_density(pos, N, 4, True, box)
    d2, d4 = O.compute_density_multipoles(dcl, N, box, los)
    ref = O.measure_pk_multipoles(N, box, kb, kc, nb, dcl, d2, d4)
    assert np.array_equal(pout[4], ref[4])
    ok = ref[4] > 0
    assert np.allclose(pout[0][ok], ref[0][ok], rtol=1e-6)
    for r in (1, 2, 3):
        assert np.all(np.abs(pout[r][ok] - ref[r][ok]) <= 1e-4 * np.abs(ref[1][ok])), r
    # real-space form
    pr = ctx.pk(delta, N, box, kb, kc, nb, False).cpu().numpy().T
    rr = O.measure_pk(N, box, kb, kc, nb, dcl, dcl)
    assert np.array_equal(pr[2], rr[2])
    assert np.all(np.abs(pr[1][ok] - rr[1][ok]) <= 1e-4 * np.abs(rr[1][ok]))


def test_pk_golden_nmodes(ctx, golden_dir):
    a = np.loadtxt(os.path.join(golden_dir, "ps_N64_kb1_kc1.txt"))
    N = 64
    delta = torch.ones((N, N, N // 2 + 1), dtype=torch.complex64, device="cuda")
    pout = ctx.pk(delta, N, 1000.0, 1.0, 1.0, 32, True).cpu().numpy().T
    assert np.array_equal(pout[4].astype(np.int64), a[:, 1].astype(np.int64))


# ------------------------------------------------------------------ triangle counts


def test_counts_golden(ctx, golden_dir):
    bs = np.loadtxt(os.path.join(golden_dir, "bs_N64_kb1_kc1.txt")).astype(np.int64)
    ntr = ctx.count_triangles(21, 64, 1.0, 1.0, True, False, 1156)
    assert np.array_equal(ntr, bs[:, 3])


@pytest.mark.parametrize("N,nb,kb,kc,io,cross", [(12, 4, 1.0, 1.0, True, False),
                                                 (16, 4, 2.0, 2.0, True, False),
                                                 (24, 8, 1.0, 1.0, False, False),
                                                 (16, 5, 1.0, 1.0, True, True)])
def test_counts_vs_oracle(ctx, N, nb, kb, kc, io, cross):
    itot = O.itot_count(nb, kb, kc, io, cross)
    ntr = ctx.count_triangles(nb, N, kb, kc, io, cross, itot)
    assert np.array_equal(ntr, O.count_triangles_exact(nb, N, kb, kc, io, cross))


# ------------------------------------------------------------------ triangle sums (GEMM)


@pytest.mark.parametrize("prec,tol", [(0, 2e-6), (1, 3e-5)])
@pytest.mark.parametrize("nb,nell,K", [(3, 1, 1000), (5, 3, 4096), (16, 3, 20000), (20, 1, 7000)])
def test_triangle_sums(ctx, prec, tol, nb, nell, K):
    rng = np.random.default_rng(nb * 100 + nell)
    Ks = ((K + 63) // 64) * 64
    maps = np.zeros((nb * nell, Ks), dtype=np.float32)
    maps[:, :K] = rng.standard_normal((nb * nell, K)).astype(np.float32)
    C = ctx.triangle_sums(_dev(maps), nb, nell, K, prec).cpu().numpy()
    m = maps[:, :K].astype(np.float64)
    p = 0
    for j in range(nb):
        for i in range(j + 1):
            prod = m[i] * m[j]
            exp = m @ prod
            scale = np.sqrt(np.sum(prod[None, :] ** 2 * m ** 2, axis=1))
            assert np.all(np.abs(C[p] - exp) <= tol * scale * 5 + 1e-12), (i, j)
            p += 1


# ------------------------------------------------------------------ bispectrum end to end


def _bk_close(mine, ref, rows, rtol=1e-4):
    """Rows 0-5, 7, 8: k's, pow, BSN, Ntr; rows 6, 9, 10: B.  B0 is compared relative to
    max(|B0_ref|, sqrt(P1 P2 P3 / Ntr)); B2 and B4 relative to max(|B_l|, |B0|, that scale), as
    P2 and P4 are relative to P0 (B_l crosses zero where B0 does not).  The Gaussian standard
    deviation of B in a bin is sqrt(s_123 P1 P2 P3 / (Ntr kf^3)); the floor used here leaves out
    the 1/kf^3 (kf = 2 pi / L < 1), so it is the stricter of the two."""
    for r in (0, 1, 2, 8):
        assert np.array_equal(mine[r], ref[r]), r
    for r in (3, 4, 5, 7):
        assert np.allclose(mine[r], ref[r], rtol=rtol), r
    gauss = np.sqrt(np.abs(ref[3] * ref[4] * ref[5]) / np.maximum(ref[8], 1))
    b0 = np.abs(ref[6] - ref[7])
    for r in [6] + ([9, 10] if rows == 11 else []):
        bref = ref[r] - (ref[7] if r == 6 else 0)
        bmine = mine[r] - (mine[7] if r == 6 else 0)
        scale = np.maximum(np.maximum(np.abs(bref), gauss), b0)
        err = np.abs(bmine - bref) / scale
        assert err.max() <= rtol, (r, err.max(), np.argmax(err))


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("multipoles", [False, True])
@pytest.mark.parametrize("maps_grid", [0, 30, 32])
def test_bispectrum(ctx, prec, multipoles, maps_grid):
    """maps_grid 0: the band-limited grid (z transform fused into the shell filter, one 2D c2r
    over all maps); 30: also fused, but 30^3 is not a multiple of 64, so the maps' stride differs
    from 30^3 and the 2D c2r runs per map; 32: the full grid (3D c2r)."""
    N, n, box, nb, kb, kc = 32, 60000, 1000.0, 6, 1.0, 1.0
    pos = _pos(n, 77)
    g = ctx.assign(_dev(pos), N, 4, True)
    delta = ctx.density(g, N, 4, True, box, n)
    itot = O.itot_count(nb, kb, kc, True)
    ntr = ctx.count_triangles(nb, N, kb, kc, True, False, itot)
    counts = (ntr.astype(np.float64) * N ** 3).astype(np.float32)
    kf = np.float32(2 * np.pi / box)
    los = (0.0, 0.0, 1.0)
    desc = ctx.bk_desc(nb, N, kb, kc, True, kf, itot, n, multipoles, los, prec, maps_grid)
    if maps_grid == 0:
        assert ctx.maps_grid(desc) < N          # band-limited grid chosen
    bout = ctx.bk(delta, _dev(counts), desc).cpu().numpy().T
    _, dcl = _oracle_density(pos, N, 4, True, box)
    if multipoles:
        d2, d4 = O.compute_density_multipoles(dcl, N, box, np.array(los))
        ref = O.measure_bk(nb, N, kb, kc, True, kf, itot, n, counts, dcl, d2, d4)
    else:
        ref = O.measure_bk(nb, N, kb, kc, True, kf, itot, n, counts, dcl)
    _bk_close(bout, ref, 11 if multipoles else 9)


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("multipoles", [False, True])
def test_bispectrum_aliased_closure(ctx, prec, multipoles):
    """8 shells of dk = k_F on 24^3 reach |q| = 8.5, so 3 R >= N: the reference's mod-N closure
    counts wrapped triangles (bispectrumI4.f90:884-908), the maps stay on the full grid and the
    sums must include the wrapped terms."""
    N, n, box, nb, kb, kc = 24, 50000, 1000.0, 8, 1.0, 1.0
    pos = _pos(n, 91)
    delta = ctx.density(ctx.assign(_dev(pos), N, 4, True), N, 4, True, box, n)
    itot = O.itot_count(nb, kb, kc, True)
    ntr = ctx.count_triangles(nb, N, kb, kc, True, False, itot)
    assert np.array_equal(ntr, O.count_triangles_exact(nb, N, kb, kc, True))
    # some triangles exist only through the wrap: the exact closure on a grid above 3 R has fewer
    assert np.any(ntr > O.count_triangles_exact(nb, 64, kb, kc, True))
    counts = (ntr.astype(np.float64) * N ** 3).astype(np.float32)
    kf = np.float32(2 * np.pi / box)
    los = (0.0, 0.0, 1.0)
    desc = ctx.bk_desc(nb, N, kb, kc, True, kf, itot, n, multipoles, los, prec, 0)
    assert ctx.maps_grid(desc) == N
    bout = ctx.bk(delta, _dev(counts), desc).cpu().numpy().T
    _, dcl = _oracle_density(pos, N, 4, True, box)
    extra = O.compute_density_multipoles(dcl, N, box, np.array(los)) if multipoles else ()
    ref = O.measure_bk(nb, N, kb, kc, True, kf, itot, n, counts, dcl, *extra)
    _bk_close(bout, ref, 11 if multipoles else 9)


@pytest.mark.parametrize("prec", [0, 1])
def test_bispectrum_explicit_multipole_fields_fused_vs_full_grid(ctx, prec):
    """Explicit delta_2 / delta_4 fields (here the densities of two other catalogues) through the
    fused z filter on the band-limited grid agree with the shell filter + 3D c2r on the full grid."""
    N, n, box, nb, kb, kc = 32, 60000, 1000.0, 6, 1.0, 1.0
    fields = [ctx.density(ctx.assign(_dev(_pos(n, seed)), N, 4, True), N, 4, True, box, n)
              for seed in (77, 78, 79)]
    itot = O.itot_count(nb, kb, kc, True)
    ntr = ctx.count_triangles(nb, N, kb, kc, True, False, itot)
    counts = _dev((ntr.astype(np.float64) * N ** 3).astype(np.float32))
    kf = np.float32(2 * np.pi / box)
    out = {}
    for mg in (0, N):
        desc = ctx.bk_desc(nb, N, kb, kc, True, kf, itot, n, True, (0.0, 0.0, 1.0), prec, mg)
        if mg == 0:
            assert ctx.maps_grid(desc) < N
        out[mg] = ctx.bk(fields[0], counts, desc, delta2=fields[1], delta4=fields[2]).cpu().numpy().T
    _bk_close(out[0], out[N], 11)


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("maps_grid", [0, 32])
def test_bispectrum_cross(ctx, prec, maps_grid):
    # two different clustered fields; B = kf^3 sum B_i A_j A_l / N_tr over the cross loop
    N, n, box, nb, kb, kc = 32, 50000, 1000.0, 6, 1.0, 1.0
    # two tracers of one clustered field: 3/4 of the particles, each moved by a small offset
    pa = _pos(n, 77)
    rng = np.random.default_rng(78)
    keep = rng.random(n) < 0.75
    pb = np.minimum(((pa[:, keep] + 0.004 * rng.standard_normal((3, int(keep.sum())))) % 1.0)
                    .astype(np.float32), np.float32(1.0 - 2 ** -24))
    nb_part = pb.shape[1]
    da = ctx.density(ctx.assign(_dev(pa), N, 3, True), N, 3, True, box, n)
    db = ctx.density(ctx.assign(_dev(pb), N, 3, True), N, 3, True, box, nb_part)
    itot = O.itot_count(nb, kb, kc, True, True)
    ntr = ctx.count_triangles(nb, N, kb, kc, True, True, itot)
    assert np.array_equal(ntr, O.count_triangles_exact(nb, N, kb, kc, True, cross=True))
    counts = (ntr.astype(np.float64) * N ** 3).astype(np.float32)
    kf = np.float32(2 * np.pi / box)
    desc = ctx.bk_desc(nb, N, kb, kc, True, kf, itot, n, False, (0.0, 0.0, 1.0), prec, maps_grid)
    bout = ctx.bk_cross(da, db, _dev(counts), desc).cpu().numpy().T
    _, dca = _oracle_density(pa, N, 3, True, box)
    _, dcb = _oracle_density(pb, N, 3, True, box)   # npart taken from pb
    ref = O.measure_bk_cross(nb, N, kb, kc, True, kf, itot, n, counts, dca, dcb)
    for r in (0, 1, 2, 4):
        assert np.array_equal(bout[r], ref[r]), r
    kf3 = O.kf3_bk(kf)
    pw = []
    for dcl in (dca, dcb):
        maps, nmodes = O.shell_maps(dcl, N, nb, kb, kc)
        pw.append(kf3 * np.sum(maps * maps, axis=1) / float(N) ** 3 / nmodes)
    l, j, i = (ref[0] - kc).astype(int), (ref[1] - kc).astype(int), (ref[2] - kc).astype(int)
    gauss = np.sqrt(np.abs(pw[0][j] * pw[0][l] * pw[1][i]) / np.maximum(ref[4], 1))
    scale = np.maximum(np.abs(ref[3]), gauss)
    err = np.abs(bout[3] - ref[3]) / scale
    assert err.max() <= 1e-4, (err.max(), np.argmax(err))
