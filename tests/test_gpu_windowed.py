"""Block-sparse triangle GEMM against the dense one: every bispectrum entry bitwise equal.

The block-sparse kernel multiplies, per wave of pair rows, only the column fragments its
triangles read, with the same products, bf16 splits and MFMA sequence per K chunk as the dense
kernel and the same FP64 split-K combine, so B (and B_2, B_4, and the cross bispectrum) must
match bit for bit.  PBI4_TRIANGLE_GEMM=dense selects the dense kernel; =block_sparse makes the
library refuse a shape that the block-sparse tiles would not cover, so a pass shows that the
block-sparse kernel ran.  Its partial rows are never wider than the dense ones.
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


def _field(ctx, N, n, seed, box=1000.0):
    rng = np.random.default_rng(seed)
    c = rng.random((3, 8))
    k = rng.integers(0, 8, n)
    p = (c[:, k] + 0.05 * rng.standard_normal((3, n))) % 1.0
    p = np.minimum(p.astype(np.float32), np.float32(1.0 - 2 ** -24))
    return ctx.density(ctx.assign(_dev(p), N, 4, True), N, 4, True, box, n)


def _run(ctx, mode, fn, ws_fn):
    old = os.environ.get("PBI4_TRIANGLE_GEMM")
    try:
        if mode is None:
            os.environ.pop("PBI4_TRIANGLE_GEMM", None)
        else:
            os.environ["PBI4_TRIANGLE_GEMM"] = mode
        ws = ws_fn()
        out = fn().cpu().numpy()
    finally:
        if old is None:
            os.environ.pop("PBI4_TRIANGLE_GEMM", None)
        else:
            os.environ["PBI4_TRIANGLE_GEMM"] = old
    return ws, out


# (grid, shells, kbin, kcenter, iopen, multipoles): the tile shapes of the real-space and
# redshift-space workloads (16 x 3, 48 x 1, 48 x 3, 64 x 3) and small ones with one row fragment
CASES = [
    (64, 20, 1.0, 1.0, True, False),
    (64, 16, 2.0, 2.0, True, True),
    (64, 22, 1.0, 1.0, False, True),
    (96, 48, 1.0, 1.0, True, False),
    (96, 48, 1.0, 1.0, True, True),
    (128, 64, 1.0, 1.0, True, True),
]


@pytest.mark.parametrize("prec", [1, 0])
@pytest.mark.parametrize("case", CASES, ids=lambda c: f"N{c[0]}_nb{c[1]}_kb{c[2]:g}_open{int(c[4])}_ell{int(c[5])}")
def test_block_sparse_equals_dense(ctx, case, prec):
    N, nb, kb, kc, iopen, mult = case
    n = 200000
    print(f"block-sparse vs dense: N={N} nb={nb} kb={kb} iopen={iopen} multipoles={mult} prec={prec}",
          flush=True)
    delta = _field(ctx, N, n, 5 + nb)
    itot = O.itot_count(nb, kb, kc, iopen)
    ntr = ctx.count_triangles(nb, N, kb, kc, iopen, False, itot)
    counts = _dev((ntr.astype(np.float64) * N ** 3).astype(np.float32))
    kf = np.float32(2 * np.pi / 1000.0)
    desc = ctx.bk_desc(nb, N, kb, kc, iopen, kf, itot, n, mult, (0.0, 0.0, 1.0), prec, 0)
    ws_w, b_w = _run(ctx, "block_sparse", lambda: ctx.bk(delta, counts, desc),
                     lambda: ctx.bk_workspace(desc))
    ws_d, b_d = _run(ctx, "dense", lambda: ctx.bk(delta, counts, desc),
                     lambda: ctx.bk_workspace(desc))
    assert ws_w <= ws_d
    assert np.all(np.isfinite(b_w))
    bad = np.argwhere(b_w != b_d)
    assert bad.size == 0, f"{len(bad)} entries differ, first {bad[:5].tolist()}"


@pytest.mark.parametrize("prec", [1, 0])
@pytest.mark.parametrize("case", [(64, 24, 1.0, 1.0, True), (96, 40, 1.0, 1.0, False)],
                         ids=lambda c: f"N{c[0]}_nb{c[1]}_open{int(c[4])}")
def test_block_sparse_cross_equals_dense(ctx, case, prec):
    N, nb, kb, kc, iopen = case
    n = 150000
    print(f"block-sparse vs dense cross: N={N} nb={nb} iopen={iopen} prec={prec}", flush=True)
    d1, d2 = _field(ctx, N, n, 31), _field(ctx, N, n, 32)
    itot = O.itot_count(nb, kb, kc, iopen, cross=True)
    ntr = ctx.count_triangles(nb, N, kb, kc, iopen, True, itot)
    counts = _dev((ntr.astype(np.float64) * N ** 3).astype(np.float32))
    kf = np.float32(2 * np.pi / 1000.0)
    desc = ctx.bk_desc(nb, N, kb, kc, iopen, kf, itot, n, False, (0.0, 0.0, 1.0), prec, 0)
    ws_w, b_w = _run(ctx, "block_sparse", lambda: ctx.bk_cross(d1, d2, counts, desc),
                     lambda: ctx.bk_cross_workspace(desc))
    ws_d, b_d = _run(ctx, "dense", lambda: ctx.bk_cross(d1, d2, counts, desc),
                     lambda: ctx.bk_cross_workspace(desc))
    assert ws_w <= ws_d
    assert np.all(np.isfinite(b_w))
    bad = np.argwhere(b_w != b_d)
    assert bad.size == 0, f"{len(bad)} entries differ, first {bad[:5].tolist()}"


@pytest.mark.parametrize("prec", [1, 0])
@pytest.mark.parametrize("case", [(64, 8, 2.0, 2.0, True), (64, 20, 1.0, 1.0, True),
                                  (64, 22, 1.0, 1.0, False), (96, 48, 1.0, 1.0, True),
                                  (128, 64, 1.0, 1.0, True)],
                         ids=lambda c: f"N{c[0]}_nb{c[1]}_kb{c[2]:g}_open{int(c[4])}")
def test_real_space_rows_jl_match_dense(ctx, case, prec):
    """Real space takes the rows (j, l) of the triangle loop with the leg i as the column (about
    half the (i <= j) pair rows): the same triple sums paired and ordered differently, so some B
    differ in the last bits from the (i <= j) block-sparse ones (bitwise the dense ones), which
    shows the (j, l) form was taken.  Rows 0-5, 7, 8 (k's, powers, shot noise, N_tr) do not
    involve the sums and are bitwise equal.  B is measured per bin against the dense kernel's
    FP32 result, relative to max(|B|, sigma_B), sigma_B = sqrt(P1 P2 P3 / (N_tr kf^3)) the
    estimator's Gaussian standard deviation: 1e-4, plus four times the bin's own FP32 spread
    (|B_FP32 (j, l) - B_FP32 (i, j)| over the same scale: how far FP32 rounding alone moves
    that sum).  The FP32 pairings differ by up to 2.6e-5 (64 shells on 128^3): a different
    rounding of the pair products and of the FP32 chunk sums of strongly cancelling sums.  These
    small fields put many bins at noise level, where split-BF16 lands 4-8 times the FP32 spread
    from FP32: at 128^3, 64 shells the largest split-BF16 error is 1.25e-4 at triangle
    (59, 29, 29), |B| = 0.17 sigma_B, FP32 spread 2.6e-5 (8.7e-5 there when the density comes from
    rocFFT's 3D transform instead of the fused z transform, which differs by 2e-7 of the largest
    mode).  Well-conditioned bins keep the 1e-4 bound; against the FP64 oracle the real-space
    configuration tests hold both forms to 1e-4 of the stricter floor sqrt(P1 P2 P3 / N_tr)."""
    from pbi4_amd.catalogue import zeldovich
    N, nb, kb, kc, iopen = case
    print(f"real-space (j, l) rows vs dense: N={N} nb={nb} kb={kb} iopen={iopen} prec={prec}",
          flush=True)
    p, _ = zeldovich(N, box=1000.0, seed=11 + nb, device="cuda", velocities=False)
    n = p.shape[1]
    delta = ctx.density(ctx.assign(p, N, 4, True, box=1000.0, normalised=False), N, 4, True,
                        1000.0, n)
    itot = O.itot_count(nb, kb, kc, iopen)
    ntr = ctx.count_triangles(nb, N, kb, kc, iopen, False, itot)
    counts = _dev((ntr.astype(np.float64) * N ** 3).astype(np.float32))
    kf = np.float32(2 * np.pi / 1000.0)

    def desc(pr):
        return ctx.bk_desc(nb, N, kb, kc, iopen, kf, itot, n, False, (0.0, 0.0, 1.0), pr, 0)

    d = desc(prec)
    _, b_jl = _run(ctx, None, lambda: ctx.bk(delta, counts, d), lambda: ctx.bk_workspace(d))
    _, b_ij = _run(ctx, "block_sparse", lambda: ctx.bk(delta, counts, d),
                   lambda: ctx.bk_workspace(d))
    _, b_d = _run(ctx, "dense", lambda: ctx.bk(delta, counts, d), lambda: ctx.bk_workspace(d))
    d32 = desc(0)
    _, ref = _run(ctx, "dense", lambda: ctx.bk(delta, counts, d32), lambda: ctx.bk_workspace(d32))
    assert not np.array_equal(b_jl[:, 6], b_ij[:, 6]), "the (j, l) row form was not taken"
    assert np.array_equal(b_jl[:, [0, 1, 2, 3, 4, 5, 7, 8]], b_d[:, [0, 1, 2, 3, 4, 5, 7, 8]])
    sigma_b = np.sqrt(np.abs(ref[:, 3] * ref[:, 4] * ref[:, 5])
                      / (np.maximum(ref[:, 8], 1) * float(kf) ** 3))
    bref = ref[:, 6] - ref[:, 7]                                 # B without the shot noise
    scale = np.maximum(np.abs(bref), sigma_b)
    e_jl = np.abs(b_jl[:, 6] - b_jl[:, 7] - bref) / scale
    e_d = np.abs(b_d[:, 6] - b_d[:, 7] - bref) / scale
    _, ref_jl = _run(ctx, None, lambda: ctx.bk(delta, counts, d32), lambda: ctx.bk_workspace(d32))
    spread = np.abs(ref_jl[:, 6] - ref_jl[:, 7] - bref) / scale
    print(f"  max error vs dense FP32: (j, l) {e_jl.max():.3g}, (i, j) {e_d.max():.3g}; FP32 "
          f"spread {spread.max():.3g}", flush=True)
    bound = 1e-4 + 4.0 * spread
    assert np.all(e_jl <= bound), f"(j, l) at triangle {int((e_jl - bound).argmax())}"
    assert np.all(e_d <= bound), f"(i, j) at triangle {int((e_d - bound).argmax())}"


@pytest.mark.parametrize("prec", [1, 0])
@pytest.mark.parametrize("case", [(96, 48, 1.0, 1.0, True, True), (128, 64, 1.0, 1.0, True, True),
                                  (96, 48, 1.0, 1.0, True, False), (128, 64, 1.0, 1.0, True, False)],
                         ids=lambda c: f"N{c[0]}_nb{c[1]}_ell{int(c[5])}")
def test_width_classes_equal_one_launch(ctx, case, prec):
    """The block-sparse plan may launch each class of window widths on its own compiled body (the
    narrowest that holds its row blocks); every row block still runs the MFMA body of its own
    width, so the sums are bitwise those of the single launch.  With multipoles these shapes have
    5 and 17 row blocks of pair rows with windows of 6, 5, 4, 2, 1 and 1 to 7 tiles; in real
    space (j, l) rows with windows of 1 to 4 tiles.  PBI4_TRIANGLE_GEMM_SPLIT=1 launches each
    width class on its own body; =0 keeps one launch."""
    N, nb, kb, kc, iopen, mult = case
    n = 200000
    print(f"width classes vs one launch: N={N} nb={nb} multipoles={mult} prec={prec}", flush=True)
    delta = _field(ctx, N, n, 9 + nb)
    itot = O.itot_count(nb, kb, kc, iopen)
    ntr = ctx.count_triangles(nb, N, kb, kc, iopen, False, itot)
    counts = _dev((ntr.astype(np.float64) * N ** 3).astype(np.float32))
    kf = np.float32(2 * np.pi / 1000.0)
    desc = ctx.bk_desc(nb, N, kb, kc, iopen, kf, itot, n, mult, (0.0, 0.0, 1.0), prec, 0)
    out = {}
    old = os.environ.get("PBI4_TRIANGLE_GEMM_SPLIT")
    try:
        for flag in ("1", "0"):
            os.environ["PBI4_TRIANGLE_GEMM_SPLIT"] = flag
            out[flag] = ctx.bk(delta, counts, desc).cpu().numpy()
    finally:
        if old is None:
            os.environ.pop("PBI4_TRIANGLE_GEMM_SPLIT", None)
        else:
            os.environ["PBI4_TRIANGLE_GEMM_SPLIT"] = old
    assert np.all(np.isfinite(out["1"]))
    bad = np.argwhere(out["1"] != out["0"])
    assert bad.size == 0, f"{len(bad)} entries differ, first {bad[:5].tolist()}"
