"""GPU parity at the workloads of the five benchmark configurations (seeded Zel'dovich
catalogues generated on the GPU, L = 1000 Mpc/h, P(k) = 2e4 (k/0.1)^-1.5 cut at the lattice
Nyquist, float32 positions and velocities):

  cfg1  128^3 particles, grid 64^3 TSC interlaced, real space: P with kbin = kcenter = 1 (32 bins),
        B with kbin = kcenter = 2 over 8 shells (98 triangles); everything against the FP64
        oracle, mode counts and N_tr bit-exact (N_tr against the exact integer counter).
  cfg2  256^3 particles, grid 256^3 PCS interlaced, real space and redshift space along z, through
        the serving pipeline (fused P_l + B pass) in both precisions (split-BF16 on the
        band-limited maps grid, FP32 on the full grid): P_l on all 64 bins, B0 B2 B4 on all 564
        triangles against the FP64 oracle; N_tr against the oracle's FFT-form count on the smallest
        grid above 3 R (no aliasing at 256, so the closures agree).
  cfg3  512^3 particles, grid 512^3 PCS interlaced, real space: P on 256 bins against the oracle,
        B over 48 shells of dk = k_F (11 228 triangles) in split-BF16 and FP32; N_tr for every
        triangle against the oracle's FFT form on 150^3; B against the FP64 oracle on the
        triangles whose three shells lie in a fixed set of ten (maps of those shells only).
  cfg4  a mixed stream (grids 128, 192, 256, 384; CIC/TSC/PCS; 1e6-5e7 particles; amplitudes
        1e4-4e4; redshift space; the reference's default shell count per grid, which makes every
        request's closure aliased, 3 R >= N): each result bitwise the same alone and in the mix;
        the 128^3 and 192^3 requests against the FP64 oracle.
  cfg5  1024^3 particles, grid 1024^3 PCS interlaced, redshift space along x, y and z, 32 shells:
        mode counts against the oracle's bin index, N_tr against the FFT form on 200^3, every row
        finite and the k / N_tr rows as the reference writes them.

Tolerances: mode counts and N_tr exact; P_l within 1e-4 of P0 per bin; B within 1e-4 of
max(|B_l|, |B0|, sqrt(P1 P2 P3 / N_tr)) per bin (the Gaussian scale without its 1/kf^3, the
stricter floor) for FP32 sums, and for split-BF16 sums on the band-limited maps grid.
Split-BF16 on the full grid (the aliased closure, 3 R >= N): within BF16_FULL_GRID_TOL.  There the
triple sum runs over all N^3 map points and cancels down to the N_tr closed triangles, so the
split's per-term error (~2^-16.5) is amplified by the ratio of sum_x |M_i M_j M_l| to the sum;
the FP32 sums of the same requests are held to 1e-4.  Per-bin errors are written to
$PBI4_REPORT_DIR/<test>.json when it is set.
"""
import dataclasses
import json
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import pbi4_oracle as O  # noqa: E402

pytestmark = pytest.mark.gpu
BOX = 1000.0
KF = np.float32(2 * np.pi / BOX)
BF16_FULL_GRID_TOL = 3e-4


@pytest.fixture(scope="module")
def ctx():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from pbi4_amd.core import Context
    return Context(0)


def _report(name, obj):
    d = os.environ.get("PBI4_REPORT_DIR")
    if d:
        os.makedirs(d, exist_ok=True)
        with open(os.path.join(d, name + ".json"), "w") as f:
            json.dump(obj, f)


def _say(*a):
    print(*a, flush=True)


def _oracle_positions(p, v, rsd, box=BOX):
    """IO.read_catalogues on a device catalogue (pbi4/pbi4.py:151-154, 241-243): the
    redshift-space shift s = x + v rsd_d in float64 (rsd_d the float32 coefficient the device
    path uses), stored as float32, then ((s + box) % box) / box in float32 and min(., 1)."""
    pos = p.cpu().numpy().astype(np.float64)
    if rsd is not None:
        vel = v.cpu().numpy()
        for d in range(3):
            if rsd[d] != 0.0:
                pos[d] = pos[d] + vel[d].astype(np.float64) * np.float64(np.float32(rsd[d]))
    pos = pos.astype(np.float32)
    pos = ((pos + np.float32(box)) % np.float32(box)) / np.float32(box)
    return np.minimum(pos, np.float32(1.0)).astype(np.float64)


def _rsd_of(f, coef):
    return tuple(float(np.float32(coef * l)) for l in f.los) if f.rsd else None


def _oracle_field(p, v, f, coef, N, order, npart, box=BOX):
    pos = _oracle_positions(p, v, _rsd_of(f, coef), box)
    g = O.assign_double_grid(N, order, npart, pos)
    del pos
    dcl = O.compute_density(npart, g, N, order, True, box)
    del g
    if not f.rsd:
        return dcl, None, None
    d2, d4 = O.compute_density_multipoles(dcl, N, box, np.array(f.los))
    return dcl, d2, d4


def _pk_errors(mine, ref, multipoles):
    """mine, ref: (5, nbins) pout.  Asserts counts and avgk, returns the per-bin errors of the
    P rows relative to |P0|."""
    cnt = 4 if multipoles else 2
    assert np.array_equal(mine[cnt], ref[cnt]), "mode counts differ"
    ok = ref[cnt] > 0
    assert np.allclose(mine[0][ok], ref[0][ok], rtol=1e-6), "avgk differs"
    rows = (1, 2, 3) if multipoles else (1,)
    return {r: (np.abs(mine[r][ok] - ref[r][ok]) / np.abs(ref[1][ok])) for r in rows}


def _bk_errors(mine, ref):
    """mine, ref: (rows, itot) bout.  Asserts the k, N_tr rows exactly and the powers / BSN to
    1e-4, returns the per-bin errors of B0 (and B2, B4) relative to the bin's scale."""
    for r in (0, 1, 2, 8):
        assert np.array_equal(mine[r], ref[r]), r
    for r in (3, 4, 5, 7):
        assert np.allclose(mine[r], ref[r], rtol=1e-4), r
    gauss = np.sqrt(np.abs(ref[3] * ref[4] * ref[5]) / np.maximum(ref[8], 1))
    b0 = np.abs(ref[6] - ref[7])
    out = {}
    for r in [6] + ([9, 10] if ref.shape[0] == 11 else []):
        bref = ref[r] - (ref[7] if r == 6 else 0)
        bm = mine[r] - (mine[7] if r == 6 else 0)
        out[r] = np.abs(bm - bref) / np.maximum(np.maximum(np.abs(bref), gauss), b0)
    return out


def _check(errs, what, tol=1e-4):
    for r, e in errs.items():
        assert e.max() <= tol, (what, r, float(e.max()), int(np.argmax(e)))


def _summary(errs):
    return {str(r): {"max": float(e.max()), "median": float(np.median(e)),
                     "per_bin": [float(x) for x in e]} for r, e in errs.items()}


def _ntr_fft(nb, Ng, kb, kc):
    """N_tr from the oracle's FFT form of count_triangles on an Ng^3 grid, as integers (the
    FP64 sums are within 1e-6 of integers times Ng^3 at these sizes)."""
    c = O.count_triangles(nb, Ng, kb, kc, True) / float(Ng) ** 3
    r = np.rint(c)
    assert np.max(np.abs(c - r)) < 1e-3
    return r.astype(np.int64)


# ------------------------------------------------------------------ cfg1


@pytest.mark.timeout(900)
def test_cfg1_tsc_64_real_space(ctx):
    from pbi4_amd.catalogue import zeldovich
    N, n = 64, 128
    npart = n ** 3
    _say(f"[cfg1] {n}^3 Zel'dovich particles, grid {N}^3 TSC interlaced")
    p, _ = zeldovich(n, box=BOX, seed=101, device="cuda", velocities=False)
    g = ctx.assign(p, N, 3, True, box=BOX, normalised=False)
    d = ctx.density(g, N, 3, True, BOX, npart)
    pout = ctx.pk(d, N, BOX, 1.0, 1.0, 32, False).cpu().numpy().T
    nb, itot = 8, O.itot_count(8, 2.0, 2.0, True)
    assert itot == 98
    ntr = ctx.count_triangles(nb, N, 2.0, 2.0, True, False, itot)
    counts = (ntr.astype(np.float64) * N ** 3).astype(np.float32)
    bouts = {}
    for prec in (0, 1):
        for mg in (0, N):
            desc = ctx.bk_desc(nb, N, 2.0, 2.0, True, KF, itot, npart, False, (0.0, 0.0, 1.0),
                               prec, mg)
            bouts[(prec, mg)] = ctx.bk(d, torch.from_numpy(counts).cuda(), desc).cpu().numpy().T
    from pbi4_amd.pipeline import Field
    dcl, _, _ = _oracle_field(p, None, Field(False), 0.0, N, 3, npart)
    assert np.array_equal(ntr, O.count_triangles_exact(nb, N, 2.0, 2.0, True))
    pref = O.measure_pk(N, BOX, 1.0, 1.0, 32, dcl, dcl)
    perr = _pk_errors(pout, pref, False)
    _check(perr, "cfg1 P")
    bref = O.measure_bk(nb, N, 2.0, 2.0, True, KF, itot, npart, counts, dcl)
    rep = {"P": _summary(perr)}
    for key, bo in bouts.items():
        e = _bk_errors(bo, bref)
        rep[f"B prec{key[0]} maps{key[1]}"] = _summary(e)
        _check(e, ("cfg1 B", key))
    _report("cfg1", rep)
    _say(f"[cfg1] P max {max(e.max() for e in perr.values()):.2e}; B max "
         f"{max(v['max'] for k, r in rep.items() if k != 'P' for v in r.values()):.2e}")


# ------------------------------------------------------------------ cfg2


@pytest.mark.timeout(900)
def test_cfg2_pcs_256_real_and_redshift_space(ctx):
    from pbi4_amd import _lib
    from pbi4_amd.catalogue import zeldovich
    from pbi4_amd.pipeline import CataloguePlan, Field, Measurement
    N, n = 256, 256
    npart = n ** 3
    _say(f"[cfg2] {n}^3 Zel'dovich particles, grid {N}^3 PCS interlaced, real + z-space")
    p, v = zeldovich(n, box=BOX, seed=1000, device="cuda")
    fields = (Field(False), Field(True, (0.0, 0.0, 1.0)))
    res, plans = {}, {}
    for prec, mg in ((_lib.PREC_BF16X3, 0), (_lib.PREC_FP32, N)):
        m = Measurement(grid=N, order=4, interlacing=True, box=BOX, kbin=2.0, kcenter=2.0,
                        nbins_ps=0, nbins_bisp=16, iopen=True, fields=fields, precision=prec,
                        maps_grid=mg)
        plan = plans[prec] = CataloguePlan(ctx, m)
        out = plan.run(p, v, npart)
        torch.cuda.synchronize()
        res[(prec, mg)] = {f: (po.cpu().numpy().T, bo.cpu().numpy().T) for f, (po, bo) in out.items()}
    assert plan.itot == 564 and m.nb_ps() == 64
    assert plans[_lib.PREC_BF16X3].maps_grid == 100 and plans[_lib.PREC_FP32].maps_grid == N
    assert np.array_equal(plan.ntr, _ntr_fft(16, 100, 2.0, 2.0))
    counts = plan.counts.cpu().numpy()
    rep = {}
    for f in fields:
        _say(f"[cfg2] oracle, field rsd={f.rsd}")
        dcl, d2, d4 = _oracle_field(p, v, f, plan.rsd_coef, N, 4, npart)
        if f.rsd:
            pref = O.measure_pk_multipoles(N, BOX, 2.0, 2.0, 64, dcl, d2, d4)
        else:
            pref = O.measure_pk(N, BOX, 2.0, 2.0, 64, dcl, dcl)
        bref = O.measure_bk(16, N, 2.0, 2.0, True, KF, 564, npart, counts, dcl, d2, d4)
        del dcl, d2, d4
        for key, r in res.items():
            pe = _pk_errors(r[f][0], pref, f.rsd)
            be = _bk_errors(r[f][1], bref)
            name = f"{'zspace' if f.rsd else 'real'} prec{key[0]} maps{key[1]}"
            rep[name] = {"P": _summary(pe), "B": _summary(be)}
            _say(f"[cfg2] {name}: P max {max(e.max() for e in pe.values()):.2e}, "
                 f"B max {max(e.max() for e in be.values()):.2e}")
            _check(pe, ("cfg2 P", name))
            _check(be, ("cfg2 B", name))
    _report("cfg2", rep)


# ------------------------------------------------------------------ cfg3


@pytest.mark.timeout(900)
def test_cfg3_pcs_512_precision_form(ctx):
    from pbi4_amd import _lib
    from pbi4_amd.catalogue import zeldovich
    from pbi4_amd.pipeline import CataloguePlan, Field, Measurement
    N, n, nb = 512, 512, 48
    npart = n ** 3
    _say(f"[cfg3] {n}^3 Zel'dovich particles, grid {N}^3 PCS interlaced, 48 shells of k_F")
    p, _ = zeldovich(n, box=BOX, seed=3000, device="cuda", velocities=False)
    res, plan = {}, None
    for prec in (_lib.PREC_BF16X3, _lib.PREC_FP32):
        m = Measurement(grid=N, order=4, interlacing=True, box=BOX, kbin=1.0, kcenter=1.0,
                        nbins_ps=0, nbins_bisp=nb, iopen=True, fields=(Field(False),),
                        precision=prec, maps_grid=0)
        plan = CataloguePlan(ctx, m)
        out = plan.run(p, None, npart)
        torch.cuda.synchronize()
        po, bo = out[Field(False)]
        res[prec] = (po.cpu().numpy().T, bo.cpu().numpy().T)
        del out, po, bo
    assert plan.itot == 11228 and plan.maps_grid == 150
    _say("[cfg3] N_tr against the FFT form on 150^3")
    assert np.array_equal(plan.ntr, _ntr_fft(nb, 150, 1.0, 1.0))
    counts = plan.counts.cpu().numpy()
    # split-BF16 against FP32 on every triangle (the per-bin precision statement)
    rep = {"bf16x3_vs_fp32_all_triangles": _summary(_bk_errors(res[1][1], res[0][1]))}
    _say("[cfg3] oracle density")
    dcl, _, _ = _oracle_field(p, None, Field(False), 0.0, N, 4, npart)
    del p
    pref = O.measure_pk(N, BOX, 1.0, 1.0, 256, dcl, dcl)
    for prec in res:
        pe = _pk_errors(res[prec][0], pref, False)
        rep[f"P prec{prec}"] = _summary(pe)
        _check(pe, ("cfg3 P", prec))
    shells = {1, 2, 4, 8, 12, 16, 24, 32, 47, 48}
    tl = O.triangle_list(nb, 1.0, 1.0, True)
    rows_m = [m for m, t in enumerate(tl) if set(int(x) for x in t) <= shells]
    assert len(rows_m) >= 30
    _say(f"[cfg3] oracle B on {len(rows_m)} triangles of shells {sorted(shells)}")
    bref = O.measure_bk_subset(nb, N, 1.0, 1.0, True, KF, len(tl), npart, counts, dcl, rows_m)
    rep["triangles"] = [[int(x) for x in tl[m]] for m in rows_m]
    for prec in res:
        be = _bk_errors(res[prec][1][:, rows_m], bref)
        rep[f"B prec{prec} vs fp64"] = _summary(be)
        _say(f"[cfg3] B prec{prec} vs FP64: max {max(e.max() for e in be.values()):.2e}")
        _check(be, ("cfg3 B", prec))
    _report("cfg3", rep)


# ------------------------------------------------------------------ cfg4


@pytest.mark.timeout(900)
def test_cfg4_mixed_stream_default_shells(ctx):
    from pbi4_amd import _lib
    from pbi4_amd.catalogue import zeldovich
    from pbi4_amd.pipeline import Field, Measurement
    from pbi4_amd.serving import Request, ServingEngine
    # (grid, order, particles per side, amplitude): counts 2.0e6 ... 4.98e7, log-spread
    draws = [(128, 2, 126, 1.5e4), (192, 3, 171, 3.1e4), (256, 4, 368, 2.2e4),
             (384, 2, 215, 3.8e4), (128, 4, 100, 1.1e4), (192, 2, 292, 2.6e4),
             (256, 3, 160, 3.4e4), (384, 4, 271, 1.8e4)]
    f = Field(True, (0.0, 0.0, 1.0))

    def meas(grid, order):
        return Measurement(grid=grid, order=order, interlacing=True, box=BOX, kbin=2.0,
                           kcenter=2.0, nbins_ps=0, nbins_bisp=0, iopen=True, fields=(f,),
                           precision=_lib.PREC_BF16X3, maps_grid=0)

    cats = [zeldovich(n, box=BOX, amp=a, seed=4000 + k, device="cuda")
            for k, (_, _, n, a) in enumerate(draws)]
    ms = [meas(g, o) for g, o, _, _ in draws]
    assert [m.nb_bisp() for m in ms[:4]] == [21, 32, 42, 64]
    one = ServingEngine(ctx, slots=1)
    mix = ServingEngine(ctx, slots=4)
    try:
        _cfg4_body(ctx, f, cats, ms, one, mix, Request)
    finally:
        # release the arenas (up to ~100 GB each) and catalogues even when a check fails, so
        # the frames a failure report keeps do not hold the device memory
        one.release()
        mix.release()
        one.plans.clear()
        mix.plans.clear()
        cats.clear()
        torch.cuda.empty_cache()


def _cfg4_body(ctx, f, cats, ms, one, mix, Request):
    from pbi4_amd import _lib
    alone = []
    for k, m in enumerate(ms):
        _say(f"[cfg4] alone: request {k}: grid {m.grid}, order {m.order}, "
             f"{cats[k][0].shape[1]} particles, {m.nb_bisp()} shells")
        r = one.run([Request(k, m, *cats[k], arrival=0.0)])[0]
        alone.append(r.result[f])
        plan = one.plan(m)
        assert plan.maps_grid == m.grid           # 3 R >= N: full grid, aliased closure
    one.release()
    mix.plans = one.plans
    reqs = [Request(k, m, *cats[k], arrival=0.002 * k) for k, m in enumerate(ms)]
    _say("[cfg4] the same requests as one stream through 4 slots")
    done = mix.run(reqs)
    assert len(done) == len(ms)
    differ = []
    for r in sorted(done, key=lambda r: r.rid):
        a = alone[r.rid]
        for what, x, y in (("pout", r.result[f][0], a[0]), ("bout", r.result[f][1], a[1])):
            if not np.array_equal(x, y):
                rel = np.abs(x - y) / np.maximum(np.abs(y), 1e-300)
                rows = sorted(set(np.nonzero(x != y)[1].tolist()))
                differ.append((r.rid, what, rows, float(np.nanmax(rel)), int((x != y).sum())))
    _say(f"[cfg4] alone vs mix: {'bitwise equal' if not differ else differ}")
    mix.release()
    rep = {"alone_vs_mix_differences": differ}
    # the plain FP32 sums of requests 1 and 7 (same maps, FP32 GEMM), one slot
    fp32 = {}
    for k in (1, 7):
        m32 = dataclasses.replace(ms[k], precision=_lib.PREC_FP32)
        r = one.run([Request(100 + k, m32, *cats[k], arrival=0.0)])[0]
        fp32[k] = r.result[f]
    one.release()
    for k in (0, 1):
        m = ms[k]
        N, npart = m.grid, cats[k][0].shape[1]
        plan = one.plan(m)
        nb = m.nb_bisp()
        _say(f"[cfg4] oracle for request {k} (grid {N}, {nb} shells)")
        assert np.array_equal(plan.ntr, _ntr_fft(nb, N, 2.0, 2.0))
        dcl, d2, d4 = _oracle_field(cats[k][0], cats[k][1], f, plan.rsd_coef, N, m.order, npart)
        pref = O.measure_pk_multipoles(N, BOX, 2.0, 2.0, m.nb_ps(), dcl, d2, d4)
        bref = O.measure_bk(nb, N, 2.0, 2.0, True, KF, m.itot(), npart,
                            plan.counts.cpu().numpy(), dcl, d2, d4)
        pout, bout = alone[k]
        pe = _pk_errors(pout.T, pref, True)
        be = _bk_errors(bout.T, bref)
        rep[f"request{k}_grid{N}"] = {"P": _summary(pe), "B split-bf16": _summary(be)}
        _say(f"[cfg4] request {k}: P max {max(e.max() for e in pe.values()):.2e}, "
             f"B split-BF16 max {max(e.max() for e in be.values()):.2e}")
        _check(pe, ("cfg4 P", k))
        _check(be, ("cfg4 B split-bf16", k), BF16_FULL_GRID_TOL)
        if k in fp32:
            be32 = _bk_errors(fp32[k][1].T, bref)
            rep[f"request{k}_grid{N}"]["B fp32"] = _summary(be32)
            _say(f"[cfg4] request {k}: B FP32 max {max(e.max() for e in be32.values()):.2e}")
            _check(be32, ("cfg4 B fp32", k))
        del dcl, d2, d4
    # the 384^3 request with 64 shells x 3 maps (the largest full-grid transform, 1.09e10 map
    # values): B0 B2 B4 against the FP64 oracle on the triangles of six shells across the range
    k = 7
    m = ms[k]
    plan = one.plan(m)
    N, npart, nb = m.grid, cats[k][0].shape[1], m.nb_bisp()
    shells = {1, 2, 21, 22, 63, 64}
    tl = O.triangle_list(nb, 2.0, 2.0, True)
    rows_m = [q for q, t in enumerate(tl) if set(int(x) for x in t) <= shells]
    _say(f"[cfg4] oracle for request {k} (grid {N}, {nb} shells) on {len(rows_m)} triangles")
    dcl, d2, d4 = _oracle_field(cats[k][0], cats[k][1], f, plan.rsd_coef, N, m.order, npart)
    bref = O.measure_bk_subset(nb, N, 2.0, 2.0, True, KF, m.itot(), npart,
                               plan.counts.cpu().numpy(), dcl, rows_m, d2, d4)
    del dcl, d2, d4
    be = _bk_errors(alone[k][1].T[:, rows_m], bref)
    be32 = _bk_errors(fp32[k][1].T[:, rows_m], bref)
    rep[f"request{k}_grid{N}_subset"] = {"B split-bf16": _summary(be), "B fp32": _summary(be32),
                                         "triangles": [[int(x) for x in tl[q]] for q in rows_m]}
    _say(f"[cfg4] request {k}: B split-BF16 max {max(e.max() for e in be.values()):.2e}, "
         f"FP32 max {max(e.max() for e in be32.values()):.2e}")
    _check(be, ("cfg4 B split-bf16", k), BF16_FULL_GRID_TOL)
    _check(be32, ("cfg4 B fp32", k))
    _report("cfg4", rep)
    assert not differ, differ


# ------------------------------------------------------------------ cfg5


@pytest.mark.timeout(900)
def test_cfg5_large_box_properties(ctx):
    from pbi4_amd import _lib
    from pbi4_amd.catalogue import zeldovich
    from pbi4_amd.pipeline import CataloguePlan, Field, Measurement
    import gc
    gc.collect()
    torch.cuda.empty_cache()
    N = n = 1024
    npart = n ** 3
    _say(f"[cfg5] {n}^3 Zel'dovich particles, grid {N}^3 PCS interlaced, LOS x, y, z")
    p, v = zeldovich(n, box=BOX, seed=4242, device="cuda")
    los3 = ((1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0))
    m = Measurement(grid=N, order=4, interlacing=True, box=BOX, kbin=2.0, kcenter=2.0,
                    nbins_ps=0, nbins_bisp=32, iopen=True,
                    fields=tuple(Field(True, l) for l in los3), precision=_lib.PREC_BF16X3,
                    maps_grid=0)
    plan = CataloguePlan(ctx, m)
    out = plan.run(p, v, npart)
    torch.cuda.synchronize()
    res = {fl: (po.cpu().numpy().T, bo.cpu().numpy().T) for fl, (po, bo) in out.items()}
    del out, p, v
    torch.cuda.empty_cache()
    nb_ps = m.nb_ps()
    assert plan.itot == 3624 and plan.maps_grid == 200
    _say("[cfg5] N_tr against the FFT form on 200^3 (3 R = 195: exact closure on both grids)")
    assert np.array_equal(plan.ntr, _ntr_fft(32, 200, 2.0, 2.0))
    _say("[cfg5] mode counts against the oracle's bin index")
    co = O.pk_mode_counts(N, BOX, 2.0, 2.0, nb_ps).astype(np.float64)
    tl = O.triangle_list(32, 2.0, 2.0, True)
    tk = 2.0 + (tl - 1) * 2.0
    # bout(9) = real(counts / N^3) with the REAL(4) counts (bispectrumI4.f90:182)
    counts = plan.counts.cpu().numpy().astype(np.float64)
    ntr32 = (counts / float(N) ** 3).astype(np.float32).astype(np.float64)
    psn = 1.0 / npart / float(O.kf3_bk(KF))
    for fl, (pout, bout) in res.items():
        assert np.array_equal(pout[4], co), fl
        assert np.all(np.isfinite(pout)) and np.all(np.isfinite(bout)), fl
        assert np.all(pout[1] > 0), fl
        assert bout.shape == (11, 3624)
        for r, c in ((0, 0), (1, 1), (2, 2)):
            assert np.array_equal(bout[r], tk[:, c]), (fl, r)
        assert np.array_equal(bout[8], ntr32), fl
        assert np.all(bout[3] > -psn) and np.all(bout[7] > 0), fl
    # the monopole does not depend on which axis is the line of sight, up to sample variance
    p0 = np.array([res[fl][0][1] for fl in res])
    assert np.all(np.abs(p0 - p0.mean(axis=0)) <= 0.2 * p0.mean(axis=0))
