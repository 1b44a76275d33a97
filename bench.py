"""Benchmark: P+B catalogues/sec at 256^3 (split-BF16 serving path vs plain FP32), p99 latency.

Headline workload (one "step" = one batch of `--slots` catalogues through the serving engine):
  256^3 Zel'dovich particles (16.8M, float32 positions + velocities) in L = 1000 Mpc/h,
  grid 256^3, PCS with interlacing, real space and redshift space (LOS z); P_l (l = 0, 2, 4 in
  redshift space) with kbin = kcenter = 2; bispectrum with kbin = kcenter = 2 up to 32 k_F
  (16 shells, 564 triangles, open triangles), B_0 in real space and B_0,2,4 in redshift space.
Inputs are resident in HBM before the timed region.  `value` = catalogues/sec of the serving
path (split-BF16x3 triangle GEMM, band-limited maps grid, concurrent slots).

Also reported, each as its own key of the one JSON line (stdout; progress goes to stderr):
  speedup     the same workload one catalogue at a time with a host synchronisation after every
              call: plain FP32 on the full maps grid (the reference's structure), plain FP32 on
              the band-limited grid, split-BF16 on the band-limited grid; the ratios between
              them and the serving value (band-limiting, split-BF16, slot overlap, like for like)
  latency     p50/p99 of a Poisson stream of --latency-requests requests at 80% of the capacity
  precision   per-bin B errors of the split-BF16 and FP32 results against the FP64 oracle and
              against each other, for one catalogue (both fields, all 564 triangles)
  cpu_baseline  the FP64 oracle timed in full on the host cores at this workload (and at the
              128^3 / 64^3 TSC workload), no extrapolation
  cfg4        a Poisson stream of --cfg4-requests requests over grids 128..384, CIC/TSC/PCS,
              1e6-5e7 particles, with the reference's default shell count per grid, against the
              plain FP32 path one request at a time
  cfg3        512^3 particles, grid 512^3, 48 shells of k_F: split-BF16 vs FP32 throughput and
              per-bin error
  cfg5        1024^3 particles, grid 1024^3, three lines of sight, 32 shells: seconds per
              catalogue, deposit GB/s, GEMM MFMA fraction
  triangle_gemm, permuted_assign, stage_ms_rsd_field
A deadline (--deadline seconds from process start) bounds the whole run: an optional leg that
would not fit is shortened or skipped, and says so in its key.
"""
import argparse
import json
import math
import os
import sys
import time

T_START = time.perf_counter()

import numpy as np  # noqa: E402

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import pbi4_amd  # noqa: E402,F401  (sets how rocFFT builds its kernels before torch is used)

METRIC = "P+B catalogues/sec at 256³ grid (split-BF16 vs plain FP32), p99 latency, 1e-4"
HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0
BOX = 1000.0


def log(*a):
    print(f"[bench {time.perf_counter() - T_START:6.1f}s]", *a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--slots", type=int, default=6,
                    help="concurrent slots of the headline engine (measured 4: 391.6, 6: 405.2, "
                         "8: 405.1 catalogues/s)")
    ap.add_argument("--pool", type=int, default=4, help="distinct resident catalogues")
    ap.add_argument("--lattice", type=int, default=256)
    ap.add_argument("--grid", type=int, default=256)
    ap.add_argument("--plain", type=int, default=3,
                    help="catalogues per one-at-a-time point of the speedup ladder (0 = skip)")
    ap.add_argument("--latency-requests", type=int, default=1000)
    ap.add_argument("--roofline-kernel", default="pbi4_deposit")
    ap.add_argument("--roofline-steps", type=int, default=4)
    ap.add_argument("--precision", type=int, default=1, help="per-bin B errors vs FP64")
    ap.add_argument("--cpu-baseline", type=int, default=1)
    ap.add_argument("--stages", type=int, default=1)
    ap.add_argument("--cfg4-requests", type=int, default=2000)
    ap.add_argument("--cfg4-pool", type=int, default=24)
    ap.add_argument("--cfg4-slots", type=int, default=4)
    ap.add_argument("--cfg4-fp32", type=int, default=100,
                    help="requests of the cfg4 stream timed through the plain FP32 path")
    ap.add_argument("--cfg3", type=int, default=1)
    ap.add_argument("--cfg5", type=int, default=1024, help="particles per side and grid (0 = skip)")
    ap.add_argument("--gemm", type=int, default=1)
    ap.add_argument("--permuted", type=int, default=1)
    ap.add_argument("--deadline", type=float, default=530.0,
                    help="seconds from process start by which the run must end")
    ap.add_argument("--gemm-only", type=int, default=0,
                    help="run only the triangle GEMM timing (for profiling), print it and exit: "
                         "1 = 48 shells on 150^3, 2 = the full-grid 384^3 / 256^3 shapes")
    ap.add_argument("--cfg4-only", type=int, default=0,
                    help="run only the cfg4 stream, print it and exit")
    ap.add_argument("--large-only", type=int, default=0,
                    help="run only the cfg5 large box with this many particles per side and exit")
    return ap.parse_args()


def remaining(args):
    return args.deadline - (time.perf_counter() - T_START)


def pmc_traffic(kernel):
    """HBM bytes per launch of `kernel` (calibrated FETCH_SIZE + WRITE_SIZE) from the rocprofv3
    summary committed in profiles/ for this workload, or None if it is absent."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_256cube_pcs.txt")
    try:
        for line in open(path):
            # fixed-width table: a 60-character kernel name column (names contain spaces)
            f = line[60:].split()
            if line[:60].strip() == kernel and len(f) >= 2:
                return round((float(f[0]) + float(f[1])) * 1e6)
    except (OSError, ValueError):
        pass
    return None


def sigma_b(bout_t, kf):
    """Gaussian standard deviation of B in each bin, sqrt(P1 P2 P3 / (N_tr kf^3)) (the estimator
    B = kf^3 / N_tr sum delta1 delta2 delta3 with P = kf^3 <|delta|^2>; the symmetry factor
    s_123 in {1, 2, 6} is left out, so this is a lower bound).  bout_t: rows of a bout array."""
    return np.sqrt(np.abs(bout_t[3] * bout_t[4] * bout_t[5]) / np.maximum(bout_t[8], 1)
                   / float(kf) ** 3)


def b_errors(ref, mine, kf):
    """Per-bin |B - B_ref| of the B rows (B0 from rows 6 - 7, then B2, B4) of two bout arrays
    (rows x itot), relative to max(|B_l|, |B0|, sigma_B), sigma_B the bin's Gaussian standard
    deviation.  Returns {row name: array}."""
    fl = sigma_b(ref, kf)
    b0 = np.abs(ref[6] - ref[7])
    out = {}
    for r, name in ((6, "B0"), (9, "B2"), (10, "B4")):
        if r >= ref.shape[0]:
            continue
        br = ref[r] - (ref[7] if r == 6 else 0)
        bm = mine[r] - (mine[7] if r == 6 else 0)
        out[name] = np.abs(bm - br) / np.maximum(np.maximum(np.abs(br), fl), b0)
    return out


def tenths_of_decades(err):
    """Compact per-bin encoding for the JSON line: one two-digit field per bin, v =
    round(-10 log10(err)) in 00..99 (so err = 10^(-v/10); 99 also stands for anything below
    1.3e-10), concatenated in the reference's triangle order."""
    e = np.maximum(np.asarray(err, dtype=np.float64), 1e-300)
    return "".join(f"{int(v):02d}" for v in np.clip(np.rint(-10.0 * np.log10(e)), 0, 99))


def stats(err):
    e = np.concatenate([np.ravel(x) for x in err]) if isinstance(err, (list, tuple)) else np.ravel(err)
    return {"max": float(e.max()), "p99": float(np.percentile(e, 99)), "median": float(np.median(e)),
            "bins": int(e.size)}


def triangle_list(nb, kbin, kcenter, iopen):
    """(l, j, i) of the reference's triangle loop (pbi4/pbi4.py:105-117)."""
    ishift = int(kcenter / kbin) + (1 if iopen else 0)
    return [(l, j, i) for l in range(1, nb + 1) for j in range(1, l + 1)
            for i in range(max(1, l - j + 1 - ishift), j + 1)]


def gemm_blocks(nb, nell, kbin, kcenter, iopen, pair_rows=False):
    """16-row x 16-column MFMA tiles the triangle GEMM issues, dense and block-sparse, and the
    number of sums the finalize reads.  Block-sparse (the library's rule): a wave of 16 rf GEMM
    rows spans the column fragments between the first and last column its triangles read, and
    all 4 waves of a work-group multiply the widest span of the four.  Real space (nell = 1)
    takes the rows (j, l) of the triangle loop, l-major from j = max(1, ceil((l + 1 - ishift) /
    2)), with the leg i as the column; redshift space the pairs (i <= j) with columns (l, ell),
    kept dense where the spans would be more than 80 % of the dense tiles.  pair_rows: the pair
    rows in real space too, without the 80 % rule (PBI4_TRIANGLE_GEMM=block_sparse)."""
    ishift = int(kcenter / kbin) + (1 if iopen else 0)

    def rf_of(rows, cf):
        rf = 4 if rows >= 900 else 2 if rows >= 300 else 1
        if cf >= 12 and rf > 2:
            rf = 2
        if rows <= 256 and cf >= 3:
            rf = 4
        return rf

    npair = nb * (nb + 1) // 2
    ncol = nb * nell
    cf = next(c for c in (1, 2, 3, 4, 6, 8, 9, 12) if c >= -(-ncol // 16))
    rf = rf_of(npair, cf)
    dense = 4 * (-(-npair // (64 * rf))) * rf * cf
    tl = triangle_list(nb, kbin, kcenter, iopen)
    if nell == 1 and not pair_rows:
        j0 = [0] + [max(1, (l + 2 - ishift) // 2) for l in range(1, nb + 1)]
        off = [0, 0]
        for l in range(1, nb + 1):
            off.append(off[-1] + l - j0[l] + 1)
        rows = off[-1]
        rf = rf_of(rows, cf)

        def entry(l, j, i):
            return off[l] + j - j0[l], i - 1, i - 1
    else:
        rows = npair

        def entry(l, j, i):
            return (j - 1) * j // 2 + (i - 1), (l - 1) * nell, (l - 1) * nell + nell - 1
    nblocks = -(-rows // (64 * rf))
    lo, hi = [1 << 30] * (4 * nblocks), [-1] * (4 * nblocks)
    for l, j, i in tl:
        r, c0, c1 = entry(l, j, i)
        w = r // (16 * rf)
        lo[w] = min(lo[w], c0 // 16)
        hi[w] = max(hi[w], c1 // 16)
    sparse = 0
    for b in range(nblocks):
        spans = [hi[w] - lo[w] + 1 for w in range(4 * b, 4 * b + 4) if hi[w] >= 0]
        sparse += 4 * rf * max(spans, default=0)
    if nell > 1 and not pair_rows and 5 * sparse > 4 * dense:   # the library keeps dense there
        sparse = dense
    return {"dense": dense, "block_sparse": sparse, "useful": len(tl) * nell}


def guarded(name, fn, out, args, need):
    """Run leg `fn` if at least `need` seconds remain before the deadline; record its result or
    why it did not run under out[name]."""
    left = remaining(args)
    if left < need:
        out[name] = {"skipped": f"{left:.0f} s left before the deadline, the leg needs ~{need:.0f} s"}
        log(f"{name}: skipped ({left:.0f} s left)")
        return
    t0 = time.perf_counter()
    try:
        out[name] = fn()
    except Exception as e:  # the other legs still report
        out[name] = {"error": f"{type(e).__name__}: {e}"}
        log(f"{name}: error {e}")
    log(f"{name}: {time.perf_counter() - t0:.1f} s")


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); every box here has "
              "one GPU and the path does not shard across GPUs", file=sys.stderr)
        return 2
    import torch
    from pbi4_amd import _lib
    from pbi4_amd.catalogue import zeldovich
    from pbi4_amd.core import Context
    from pbi4_amd.pipeline import CataloguePlan, Field, Measurement
    from pbi4_amd.serving import Request, ServingEngine, latency_summary

    torch.cuda.init()
    ctx = Context(0)
    if args.gemm_only == 2:
        print(json.dumps({"triangle_gemm_full_grid": gemm_full_grid(ctx, torch)}), flush=True)
        return 0
    if args.gemm_only:
        print(json.dumps({"triangle_gemm": gemm_bench(ctx, torch)}), flush=True)
        return 0
    if args.cfg4_only:
        print(json.dumps({"cfg4": cfg4_stream(ctx, torch, args, 1e9)}), flush=True)
        return 0
    if args.large_only:
        print(json.dumps(large_box(ctx, torch, n=args.large_only)), flush=True)
        return 0
    N, n = args.grid, args.lattice
    fields = (Field(False), Field(True, (0.0, 0.0, 1.0)))

    def meas(prec, mg):
        return Measurement(grid=N, order=4, interlacing=True, box=BOX, kbin=2.0, kcenter=2.0,
                           nbins_ps=0, nbins_bisp=16, iopen=True, fields=fields, precision=prec,
                           maps_grid=mg)

    m = meas(_lib.PREC_BF16X3, 0)
    log(f"generating {args.pool} catalogues of {n}^3 particles on the GPU")
    pool = [zeldovich(n, box=BOX, seed=1000 + i, device="cuda") for i in range(args.pool)]
    torch.cuda.synchronize()
    npart = n ** 3
    plan = CataloguePlan(ctx, m)
    torch.cuda.synchronize()
    log(f"plan ready: itot={plan.itot} maps grid {plan.maps_grid}^3")

    eng = ServingEngine(ctx, slots=args.slots)
    eng.plans[m] = plan
    rid = [0]

    def batch(k, mm=m):
        reqs = []
        for _ in range(k):
            p, v = pool[rid[0] % len(pool)]
            reqs.append(Request(rid[0], mm, p, v, arrival=0.0))
            rid[0] += 1
        return reqs

    for _ in range(args.warmup):
        eng.run(batch(args.slots))
    torch.cuda.synchronize()
    log("warmup done")

    # ---------------------------------------------------------------- headline (timed region)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        eng.run(batch(args.slots))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    n_cat = args.steps * args.slots
    value = n_cat / elapsed
    log(f"serving: {n_cat} catalogues in {elapsed:.3f}s -> {value:.2f} catalogues/s")

    # ---------------------------------------------------------------- roofline kernel timing
    # the same workload through a one-slot engine: launches do not overlap, so the HIP event pair
    # the library records around each launch on its stream brackets exactly one kernel
    eng1 = ServingEngine(ctx, slots=1)
    eng1.plans[m] = plan
    eng1.run(batch(1))
    torch.cuda.synchronize()
    ctx.timing_enable(args.roofline_kernel)
    for _ in range(args.roofline_steps):
        eng1.run(batch(1))
    torch.cuda.synchronize()
    kern_ms, launches = ctx.timing_read()
    ctx.timing_enable(None)
    eng1.release()
    del eng1
    avg_ms = kern_ms / max(launches, 1)
    V = N ** 3
    # per launch: read the tile-sorted particles (12 B each) and write both interlaced grids
    # once (2 x 4 B per cell): the deposit's algorithmic bytes
    alg_bytes = 12.0 * npart + 8.0 * V
    achieved = alg_bytes / (avg_ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": pmc_traffic("k_deposit<4, 2, 20>"),
            "kernel": args.roofline_kernel, "avg_launch_ms": round(avg_ms, 5), "launches": launches,
            "timing": "HIP events around each launch on its stream, one-slot pass of the workload",
            "algorithmic_bytes_per_launch": alg_bytes}

    out = {}

    # ---------------------------------------------------------------- per-stage breakdown
    if args.stages:
        s = torch.cuda.current_stream()
        p, v = pool[0]
        rsd = tuple(float(np.float32(plan.rsd_coef * l)) for l in (0.0, 0.0, 1.0))
        names = ["assign", "density", "pk", "bk"]
        for _ in range(2):
            ev = [torch.cuda.Event(enable_timing=True) for _ in range(5)]
            ev[0].record(s)
            g = ctx.assign(p, N, 4, True, vel=v, box=BOX, rsd=rsd, normalised=False)
            ev[1].record(s)
            d = ctx.density(g, N, 4, True, BOX, npart)
            ev[2].record(s)
            ctx.pk(d, N, BOX, 2.0, 2.0, m.nb_ps(), True)
            ev[3].record(s)
            desc = plan.desc[fields[1]]
            desc.npart = npart
            ctx.bk(d, plan.counts, desc)
            ev[4].record(s)
            torch.cuda.synchronize()
        out["stage_ms_rsd_field"] = {nm: round(ev[i].elapsed_time(ev[i + 1]), 4)
                                     for i, nm in enumerate(names)}
        # the serving pipeline's route: assignment and density in one call (fused gather and
        # x transform, y pass, z transform + combine)
        for _ in range(2):
            ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
            ev[0].record(s)
            ctx.assign_density(p, N, 4, True, BOX, npart, vel=v, rsd=rsd, normalised=False)
            ev[1].record(s)
            torch.cuda.synchronize()
        out["stage_ms_rsd_field"]["assign_density_one_call"] = round(ev[0].elapsed_time(ev[1]), 4)
        log(f"redshift-space field stages (ms): {out['stage_ms_rsd_field']}")

    # ---------------------------------------------------------------- speedup ladder
    results = {}
    if args.plain > 0:
        def one_at_a_time(prec, mg):
            mm = meas(prec, mg)
            pp = CataloguePlan(ctx, mm)
            e1 = ServingEngine(ctx, slots=1)
            e1.plans[mm] = pp
            e1.run(batch(1, mm))                     # warm: plans, workspace
            torch.cuda.synchronize()
            rid[0] = 0                               # the timed catalogues start at pool[0]
            t = time.perf_counter()
            done = []
            for _ in range(args.plain):
                done += e1.run(batch(1, mm))         # returns after the host has the result
            rate = args.plain / (time.perf_counter() - t)
            e1.release()
            res = {r.rid % len(pool): r.result for r in done}
            return rate, res, pp.maps_grid

        ladder = {}
        for key, prec, mg in (("plain_fp32_full_grid", _lib.PREC_FP32, N),
                              ("plain_fp32_band_limited", _lib.PREC_FP32, 0),
                              ("split_bf16_band_limited", _lib.PREC_BF16X3, 0)):
            rate, res, mgrid = one_at_a_time(prec, mg)
            ladder[key] = {"catalogues_per_s": round(rate, 3), "maps_grid": mgrid}
            results[key] = res
            log(f"one at a time, {key}: {rate:.2f} catalogues/s (maps grid {mgrid}^3)")
        a = ladder["plain_fp32_full_grid"]["catalogues_per_s"]
        b = ladder["plain_fp32_band_limited"]["catalogues_per_s"]
        c = ladder["split_bf16_band_limited"]["catalogues_per_s"]
        ladder["split_bf16_serving_slots"] = {"catalogues_per_s": round(value, 3), "slots": args.slots}
        lfl = value / b
        ladder["ratios"] = {"band_limiting": round(b / a, 3), "split_bf16": round(c / b, 3),
                            "slot_overlap": round(value / c, 3),
                            "like_for_like_vs_plain_fp32_band_limited": round(lfl, 3),
                            "vs_plain_fp32_reference_structure": round(value / a, 3)}
        ladder["note"] = ("like for like (the same band-limited maps grid) the split-BF16 serving "
                          f"path is {lfl:.2f}x the plain FP32 one-at-a-time path" +
                          ("" if lfl >= 3 else ", below the 3x target"))
        out["speedup"] = ladder

    # ---------------------------------------------------------------- Poisson latency
    if args.latency_requests > 0:
        rate = 0.8 * value
        rng = np.random.default_rng(5)
        t = np.cumsum(rng.exponential(1.0 / rate, args.latency_requests))
        reqs = batch(args.latency_requests)
        for r, a_ in zip(reqs, t):
            r.arrival = float(a_)
        t0 = time.perf_counter()
        done = eng.run(reqs)
        lat = latency_summary(done)
        lat["arrival_rate_per_s"] = round(rate, 3)
        lat["served_per_s"] = round(len(done) / (time.perf_counter() - t0), 3)
        out["latency"] = lat
        log(f"Poisson stream at {rate:.1f}/s: {lat}")
    eng.release()
    del eng

    # ---------------------------------------------------------------- cfg4 stream
    if args.cfg4_requests:
        guarded("cfg4", lambda: cfg4_stream(ctx, torch, args, remaining(args) - 170), out, args, 120)

    # ---------------------------------------------------------------- precision + CPU baseline
    cpu = None
    if args.cpu_baseline or args.precision:
        guarded("cpu_cfg2", lambda: cpu_cfg2(pool[0], plan, m, N, results), out, args, 90)
        c2 = out.pop("cpu_cfg2", {})
        if "precision" in c2:
            out["precision"] = c2.pop("precision")
        if "value" in c2:
            cpu = c2
        else:
            out["cpu_baseline_cfg2"] = c2
    guarded("cfg1_cpu", lambda: cpu_cfg1(ctx, torch), out, args, 20)
    if cpu is not None and isinstance(out.get("cfg1_cpu"), dict):
        cpu["cfg1"] = out.pop("cfg1_cpu")

    # ---------------------------------------------------------------- cfg3, cfg5, GEMM, sort
    if args.cfg3:
        guarded("cfg3", lambda: cfg3_leg(ctx, torch), out, args, 45)
    if args.cfg5:
        guarded("cfg5", lambda: large_box(ctx, torch, n=args.cfg5), out, args, 60)
    if args.gemm:
        guarded("triangle_gemm", lambda: gemm_bench(ctx, torch), out, args, 30)
    if args.permuted:
        guarded("permuted_assign", lambda: permuted_assign(ctx, torch, pool[0], plan), out, args, 10)

    res = {
        "metric": METRIC, "value": round(value, 4), "unit": "catalogues/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+bf16x3",
        "data": "synthetic (seeded Zel'dovich catalogues generated on the GPU)",
        "config": {"workload": "256^3 Zel'dovich particles, grid 256^3 PCS interlaced, real + "
                               "redshift space (LOS z), P_l l=0,2,4 (kbin=kcenter=2), B with "
                               "kbin=kcenter=2 up to 32 kF (16 shells, 564 triangles)",
                   "grid": N, "npart": npart, "slots": args.slots, "maps_grid": plan.maps_grid,
                   "precision": "split-bf16x3", "parallelism": "single GPU, concurrent slots"},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    res.update(out)
    res["run_seconds"] = round(time.perf_counter() - T_START, 1)
    res["child_processes_at_exit"] = children_at_exit()
    line = json.dumps(res, separators=(",", ":"))
    log(f"JSON line: {len(line)} characters")
    print(line, flush=True)
    return 0


def children_at_exit():
    """Child processes still alive (pid:name), after waiting up to 5 s for them to end."""
    try:
        import psutil
    except ImportError:
        return None
    kids = psutil.Process().children(recursive=True)
    _, alive = psutil.wait_procs(kids, timeout=5)
    return [f"{k.pid}:{k.name()}" for k in alive]


def cfg4_stream(ctx, torch, args, budget):
    """Serving configuration: a Poisson stream over --cfg4-pool resident Zel'dovich catalogues
    (particle counts log-uniform in [1e6, 5e7], lattice n = round(count^(1/3)), amplitudes
    uniform in [1e4, 4e4]); each request draws its grid from {128, 192, 256, 384}^3 and its
    assignment from CIC/TSC/PCS (interlaced) and asks for redshift-space P_l (l = 0, 2, 4) and the
    bispectrum with kbin = kcenter = 2 over the reference's default shell count for the grid,
    N/(3 kbin) = 21, 32, 42, 64 (pbi4/pbi4.py:99).  A capacity pass (the first 200 requests all
    queued at t = 0) sets the Poisson rate at 80% of it; the stream's served rate and p50/p99
    latency are reported, and the plain FP32 path (one slot, one request at a time) over the first
    --cfg4-fp32 requests.  `budget` seconds bound the stream: fewer requests are sent if the full
    count would not fit."""
    from pbi4_amd import _lib
    from pbi4_amd.catalogue import zeldovich
    from pbi4_amd.pipeline import Field, Measurement
    from pbi4_amd.serving import Request, ServingEngine, latency_summary
    f = Field(True, (0.0, 0.0, 1.0))
    rng = np.random.default_rng(2024)
    pool = []
    for i in range(args.cfg4_pool):
        n = int(round(math.exp(rng.uniform(math.log(1e6), math.log(5e7))) ** (1.0 / 3.0)))
        amp = float(rng.uniform(1e4, 4e4))
        pool.append(zeldovich(n, box=BOX, amp=amp, seed=7000 + i, device="cuda"))
    torch.cuda.synchronize()
    nreq = args.cfg4_requests
    draws = [(i % len(pool), int(rng.choice((128, 192, 256, 384))), int(rng.choice((2, 3, 4))))
             for i in range(nreq)]

    def meas(grid, order, prec):
        return Measurement(grid=grid, order=order, interlacing=True, box=BOX, kbin=2.0,
                           kcenter=2.0, nbins_ps=0, nbins_bisp=0, iopen=True, fields=(f,),
                           precision=prec, maps_grid=0)

    def requests(prec, ids):
        return [Request(k, meas(draws[k][1], draws[k][2], prec), *pool[draws[k][0]], arrival=0.0)
                for k in ids]

    t_start = time.perf_counter()
    eng = ServingEngine(ctx, slots=args.cfg4_slots)
    first = {}
    for k, (c, g, o) in enumerate(draws):
        first.setdefault((g, o), k)
    eng.run(requests(_lib.PREC_BF16X3, sorted(first.values())))       # plans, one pass per shape
    torch.cuda.synchronize()
    ncap = min(200, nreq)
    t0 = time.perf_counter()
    eng.run(requests(_lib.PREC_BF16X3, range(ncap)))
    torch.cuda.synchronize()
    cap = ncap / (time.perf_counter() - t0)
    rate = 0.8 * cap
    log(f"cfg4: capacity {cap:.2f}/s over {ncap} queued requests")
    # the plain path: FP32 sums, one slot, the host waits for each result (the serving slots'
    # workspaces are released first: a 384^3 request with 64 shells needs ~100 GB)
    eng.release()
    torch.cuda.empty_cache()
    one = ServingEngine(ctx, slots=1)
    nfp = min(args.cfg4_fp32, nreq)
    one.run(requests(_lib.PREC_FP32, sorted(first.values())))         # its plans
    t0 = time.perf_counter()
    for k in range(nfp):
        one.run(requests(_lib.PREC_FP32, [k]))
    plain = nfp / (time.perf_counter() - t0)
    one.release()
    del one
    torch.cuda.empty_cache()
    log(f"cfg4: plain FP32 one at a time {plain:.2f}/s over {nfp} requests")
    left = budget - (time.perf_counter() - t_start)
    nstream = int(min(nreq, max(ncap, rate * left * 0.85)))
    reqs = requests(_lib.PREC_BF16X3, range(nstream))
    t = np.cumsum(np.random.default_rng(6).exponential(1.0 / rate, nstream))
    for r, a in zip(reqs, t):
        r.arrival = float(a)
    t0 = time.perf_counter()
    mark = [0]

    def progress(k):
        if k >= mark[0] + 200:
            mark[0] = k - k % 200
            log(f"cfg4: {k} of {nstream} requests served")

    done = eng.run(reqs, progress=progress)
    torch.cuda.synchronize()
    served = nstream / (time.perf_counter() - t0)
    lat = latency_summary(done)
    eng.release()
    del eng
    counts = [int(pool[c][0].shape[1]) for c, _, _ in draws[:nstream]]
    res = {"requests": nstream, "requests_asked": nreq, "distinct_catalogues": len(pool),
           "shells": {g: meas(g, 4, 1).nb_bisp() for g in (128, 192, 256, 384)},
           "triangles": {g: meas(g, 4, 1).itot() for g in (128, 192, 256, 384)},
           "npart_min": min(counts), "npart_max": max(counts),
           "grids": {g: sum(1 for d in draws[:nstream] if d[1] == g) for g in (128, 192, 256, 384)},
           "orders": {o: sum(1 for d in draws[:nstream] if d[2] == o) for o in (2, 3, 4)},
           "capacity_catalogues_per_s": round(cap, 3),
           "poisson_arrival_rate_per_s": round(rate, 3),
           "poisson_served_per_s": round(served, 3), "latency": lat,
           "plain_fp32_one_at_a_time_per_s": round(plain, 3), "plain_fp32_requests": nfp,
           "speedup_vs_plain": round(served / plain, 3) if plain else None}
    if nstream < nreq:
        res["note"] = f"{nstream} of {nreq} requests: the full stream would not fit the deadline"
    del pool
    torch.cuda.empty_cache()
    log(f"cfg4: {res}")
    return res


def cpu_cfg2(cat, plan, m, N, results):
    """The FP64 oracle (the reference algorithm restated, oracle/) timed in full on the host
    cores for one catalogue of the headline workload: both fields (real space, redshift space
    along z), assignment, FFT, fcomb, the 22-FFT multipoles, P_l, all shell maps and all 564
    triangle sums per field; nothing is extrapolated.  Its B is the FP64 reference of the
    per-bin errors of the GPU results of the same catalogue (the speedup ladder's split-BF16 and
    FP32 band-limited results and the FP32 full-grid result)."""
    from oracle import pbi4_oracle as O
    cores = O.nthreads()
    p, v = cat
    npart = p.shape[1]
    nb = m.nb_bisp()
    itot = O.itot_count(nb, m.kbin, m.kcenter, m.iopen)
    counts = plan.counts.cpu().numpy()
    T = {}
    ref = {}
    t_all = time.perf_counter()
    for f in m.fields:
        pos = p.cpu().numpy().astype(np.float64)
        if f.rsd:
            pos[2] = pos[2] + v[2].cpu().numpy().astype(np.float64) * np.float64(
                np.float32(plan.rsd_coef))                    # pbi4.py:154
        pos = pos.astype(np.float32)
        pos = ((pos + np.float32(BOX)) % np.float32(BOX)) / np.float32(BOX)
        pos = np.minimum(pos, np.float32(1.0)).astype(np.float64)
        tag = "z" if f.rsd else "real"
        t = time.perf_counter()
        g = O.assign_double_grid(N, 4, npart, pos)
        T[f"assign_{tag}"] = time.perf_counter() - t
        del pos
        t = time.perf_counter()
        dcl = O.compute_density(npart, g, N, 4, True, BOX)
        T[f"density_{tag}"] = time.perf_counter() - t
        del g
        d2 = d4 = None
        if f.rsd:
            t = time.perf_counter()
            d2, d4 = O.compute_density_multipoles(dcl, N, BOX, np.array(f.los))
            T["multipoles_z"] = time.perf_counter() - t
        t = time.perf_counter()
        if f.rsd:
            O.measure_pk_multipoles(N, BOX, m.kbin, m.kcenter, m.nb_ps(), dcl, d2, d4)
        else:
            O.measure_pk(N, BOX, m.kbin, m.kcenter, m.nb_ps(), dcl, dcl)
        T[f"pk_{tag}"] = time.perf_counter() - t
        t = time.perf_counter()
        ref[f] = O.measure_bk(nb, N, m.kbin, m.kcenter, m.iopen, m.kf(), itot, npart, counts,
                              dcl, d2, d4)
        T[f"bk_{tag}"] = time.perf_counter() - t
        del dcl, d2, d4
    total = time.perf_counter() - t_all
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    log(f"CPU FP64 oracle, one full catalogue: {total:.1f} s on {cores} cores ({model})")
    out = {"value": round(1.0 / total, 6), "unit": "catalogues/s", "cores": cores, "kind": "port",
           "sample": "one full catalogue of the headline workload: both fields, assignment, FFTs, "
                     "fcomb, the 22-FFT multipoles, P_l, every shell map and all 564 triangle sums "
                     "per field, timed without extrapolation",
           "seconds_per_catalogue": round(total, 2), "cpu_model": model,
           "stage_seconds": {k: round(x, 3) for k, x in T.items()}}
    # per-bin B errors of the GPU results of this catalogue (pool[0]) against FP64
    kf = m.kf()
    prec = {}
    per_bin = {}
    want = (("split_bf16_band_limited", "split_bf16"), ("plain_fp32_band_limited", "fp32"),
            ("plain_fp32_full_grid", "fp32_full_grid"))
    for key, name in want:
        if key not in results or 0 not in results[key]:
            continue
        res = results[key][0]
        errs = {}
        for f in m.fields:
            tag = "z" if f.rsd else "real"
            for row, e in b_errors(ref[f], res[f][1].T, kf).items():
                errs[f"{row}_{tag}"] = e
        prec[f"{name}_vs_fp64"] = stats(list(errs.values()))
        if name == "split_bf16":
            per_bin[f"{name}_vs_fp64"] = {k: tenths_of_decades(e) for k, e in errs.items()}
    if "split_bf16_band_limited" in results and "plain_fp32_band_limited" in results:
        errs = {}
        for f in m.fields:
            tag = "z" if f.rsd else "real"
            a = results["plain_fp32_band_limited"][0][f][1].T
            b = results["split_bf16_band_limited"][0][f][1].T
            for row, e in b_errors(a, b, kf).items():
                errs[f"{row}_{tag}"] = e
        prec["split_bf16_vs_fp32"] = stats(list(errs.values()))
        per_bin["split_bf16_vs_fp32"] = {k: tenths_of_decades(e) for k, e in errs.items()}
    out["precision"] = {
        "scale": "|B - B_ref| / max(|B_l|, |B0|, sigma_B), sigma_B = sqrt(P1 P2 P3 / (N_tr kf^3)) "
                 "the bin's Gaussian standard deviation",
        "summary": prec,
        "per_bin_encoding": "per row a string of two-digit fields, one per triangle in the "
                            "reference's loop order: v = round(-10 log10(error)), error = "
                            "10^(-v/10), 99 = below 1.3e-10",
        "per_bin": per_bin}
    log(f"precision vs FP64: {prec}")
    return out


def cpu_cfg1(ctx, torch):
    """The small CPU-runnable workload timed in full with the FP64 oracle and on the GPU: 128^3
    Zel'dovich particles, grid 64^3 TSC interlaced, real space, P with kbin = kcenter = 1 (32
    bins), B with kbin = kcenter = 2 over 8 shells (98 triangles)."""
    from oracle import pbi4_oracle as O
    from pbi4_amd.catalogue import zeldovich
    N, n = 64, 128
    npart = n ** 3
    p, _ = zeldovich(n, box=BOX, seed=101, device="cuda", velocities=False)
    itot = O.itot_count(8, 2.0, 2.0, True)
    ntr = ctx.count_triangles(8, N, 2.0, 2.0, True, False, itot)
    counts = torch.from_numpy((ntr.astype(np.float64) * N ** 3).astype(np.float32)).cuda()
    kf = np.float32(2 * np.pi / BOX)
    desc = ctx.bk_desc(8, N, 2.0, 2.0, True, kf, itot, npart, False, (0.0, 0.0, 1.0), 1, 0)

    def gpu_one():
        g = ctx.assign(p, N, 3, True, box=BOX, normalised=False)
        d = ctx.density(g, N, 3, True, BOX, npart)
        po = ctx.pk(d, N, BOX, 1.0, 1.0, 32, False).cpu()
        bo = ctx.bk(d, counts, desc).cpu()
        return po, bo

    gpu_one()
    torch.cuda.synchronize()
    t = time.perf_counter()
    reps = 10
    for _ in range(reps):
        gpu_one()
    gpu_s = (time.perf_counter() - t) / reps
    pos = np.minimum(((p.cpu().numpy() + np.float32(BOX)) % np.float32(BOX)) / np.float32(BOX),
                     np.float32(1.0)).astype(np.float64)
    t = time.perf_counter()
    dcl = O.compute_density(npart, O.assign_double_grid(N, 3, npart, pos), N, 3, True, BOX)
    O.measure_pk(N, BOX, 1.0, 1.0, 32, dcl, dcl)
    O.measure_bk(8, N, 2.0, 2.0, True, kf, itot, npart, counts.cpu().numpy(), dcl)
    cpu_s = time.perf_counter() - t
    return {"workload": "128^3 Zel'dovich particles, grid 64^3 TSC interlaced, real space, P "
                        "(32 bins of k_F), B (8 shells of 2 k_F, 98 triangles)",
            "cpu_fp64_seconds_per_catalogue": round(cpu_s, 3), "cores": O.nthreads(),
            "gpu_seconds_per_catalogue": round(gpu_s, 5),
            "gpu_catalogues_per_s": round(1.0 / gpu_s, 2)}


def cfg3_leg(ctx, torch):
    """Precision form: 512^3 Zel'dovich particles, grid 512^3 PCS interlaced, real space, P with
    kbin = kcenter = 1 and the bispectrum over 48 shells of dk = k_F (11 228 triangles) on the
    band-limited 150^3 maps grid: catalogues per second one at a time with split-BF16 and with
    FP32 triangle sums, and the per-bin difference of the two (the FP64 comparison of sampled
    triangles is in the GPU tests' cfg3 report)."""
    from pbi4_amd import _lib
    from pbi4_amd.catalogue import zeldovich
    from pbi4_amd.pipeline import CataloguePlan, Field, Measurement
    n = N = 512
    p, _ = zeldovich(n, box=BOX, seed=3000, device="cuda", velocities=False)
    torch.cuda.synchronize()
    out, res = {}, {}
    for prec, name in ((_lib.PREC_BF16X3, "split_bf16"), (_lib.PREC_FP32, "fp32")):
        m = Measurement(grid=N, order=4, interlacing=True, box=BOX, kbin=1.0, kcenter=1.0,
                        nbins_ps=0, nbins_bisp=48, iopen=True, fields=(Field(False),),
                        precision=prec, maps_grid=0)
        plan = CataloguePlan(ctx, m)
        r = plan.run(p, None, n ** 3)
        torch.cuda.synchronize()
        reps = 3
        t = time.perf_counter()
        for _ in range(reps):
            r = plan.run(p, None, n ** 3)
            torch.cuda.synchronize()
        out[f"{name}_catalogues_per_s"] = round(reps / (time.perf_counter() - t), 3)
        res[name] = r[Field(False)][1].cpu().numpy().T
        out["maps_grid"] = plan.maps_grid
        out["triangles"] = plan.itot
    out["split_bf16_speedup_vs_fp32"] = round(out["split_bf16_catalogues_per_s"] /
                                              out["fp32_catalogues_per_s"], 3)
    e = b_errors(res["fp32"], res["split_bf16"], np.float32(2 * np.pi / BOX))["B0"]
    out["split_bf16_vs_fp32"] = stats(e)
    del p
    torch.cuda.empty_cache()
    log(f"cfg3: {out}")
    return out


def large_box(ctx, torch, n=1024, reps=1):
    """The large single-box configuration: n^3 Zel'dovich particles (float32 positions and
    velocities, 26 GB at n = 1024), grid n^3, PCS with interlacing, redshift space along x, y and
    z, P_l (l = 0, 2, 4, kbin = kcenter = 2) and the split-BF16x3 bispectrum over 32 shells of
    dk = 2 k_F.  Inputs resident in HBM; per line of sight the stages are timed with events on
    the stream, the deposit and the triangle GEMM with the library's launch timers."""
    from pbi4_amd import _lib
    from pbi4_amd.catalogue import rsd_factor, zeldovich
    from pbi4_amd.pipeline import CataloguePlan, Field, Measurement
    N = n
    log(f"large: generating {n}^3 Zel'dovich particles on the GPU")
    p, v = zeldovich(n, box=BOX, seed=4242, device="cuda")
    torch.cuda.synchronize()
    npart = n ** 3
    los3 = ((1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0))
    m = Measurement(grid=N, order=4, interlacing=True, box=BOX, kbin=2.0, kcenter=2.0,
                    nbins_ps=0, nbins_bisp=32, iopen=True, fields=tuple(Field(True, l) for l in los3),
                    precision=_lib.PREC_BF16X3, maps_grid=0)
    plan = CataloguePlan(ctx, m)
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    coef = float(rsd_factor())
    acc = {"dep": [0.0, 0], "gemm": [0.0, 0]}

    def field(l, timed):
        rsd = tuple(float(np.float32(coef * c)) for c in l)
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(5)]
        if timed:
            ctx.timing_enable("pbi4_deposit")
        ev[0].record(s)
        g = ctx.assign(p, N, 4, True, vel=v, box=BOX, rsd=rsd, normalised=False)
        ev[1].record(s)
        if timed:
            torch.cuda.synchronize()
            ms, k = ctx.timing_read()
            acc["dep"][0] += ms
            acc["dep"][1] += k
            ctx.timing_enable(None)
        d = ctx.density(g, N, 4, True, BOX, npart)
        del g
        ev[2].record(s)
        po = ctx.pk(d, N, BOX, 2.0, 2.0, m.nb_ps(), True, l)
        ev[3].record(s)
        desc = plan.desc[Field(True, l)]
        desc.npart = npart
        if timed:
            ctx.timing_enable("pbi4_triangle_gemm")
        bo = ctx.bk(d, plan.counts, desc)
        ev[4].record(s)
        torch.cuda.synchronize()
        if timed:
            ms, k = ctx.timing_read()
            acc["gemm"][0] += ms
            acc["gemm"][1] += k
            ctx.timing_enable(None)
        del d
        return [ev[i].elapsed_time(ev[i + 1]) for i in range(4)], po, bo

    field(los3[2], False)                           # warm-up (plans, workspaces)
    t0 = time.perf_counter()
    for _ in range(reps):
        for l in los3:
            field(l, False)
    torch.cuda.synchronize()
    per_cat = (time.perf_counter() - t0) / reps
    stages = np.zeros(4)
    for l in los3:                                  # a timed pass for the per-stage and kernel times
        st, po, bo = field(l, True)
        stages += np.array(st)
    stages /= 3
    dep_ms = acc["dep"][0] / max(acc["dep"][1], 1)
    gemm_ms = acc["gemm"][0] / max(acc["gemm"][1], 1)
    V = N ** 3
    dep_gbs = (12.0 * npart + 8.0 * V) / (dep_ms * 1e-3) / 1e9
    nb = m.nb_bisp()
    dsc = plan.desc[Field(True, los3[0])]
    tiles = gemm_blocks(nb, 3, dsc.kbin, dsc.kcenter, bool(dsc.iopen))
    gemm_tf = 3 * 2.0 * 256 * tiles["block_sparse"] * plan.maps_grid ** 3 / (gemm_ms * 1e-3) / 1e12
    out = {"workload": f"{n}^3 Zel'dovich particles, grid {N}^3 PCS interlaced, redshift space "
                       "along x, y, z, P_l l=0,2,4, B with 32 shells of 2 k_F (split-BF16x3)",
           "seconds_per_catalogue": round(per_cat, 4), "npart": npart, "itot": plan.itot,
           "maps_grid": plan.maps_grid,
           "stage_ms_per_field": {k: round(float(x), 3) for k, x in
                                  zip(("assign", "density", "pk", "bk"), stages)},
           "deposit": {"avg_launch_ms": round(dep_ms, 3), "algorithmic_GBps": round(dep_gbs, 1),
                       "frac_of_hbm": round(dep_gbs / HBM_PEAK_GBS, 4)},
           "triangle_gemm": {"avg_launch_ms": round(gemm_ms, 3), "executed_tflops": round(gemm_tf, 1),
                             "mfma_frac": round(gemm_tf / BF16_PEAK_TFLOPS, 4),
                             "tiles_issued_of_dense": round(tiles["block_sparse"] / tiles["dense"], 4)},
           "finite": bool(torch.isfinite(po).all() and torch.isfinite(bo).all())}
    log(f"large: {out}")
    del p, v
    torch.cuda.empty_cache()
    return out


def gemm_bench(ctx, torch):
    """MFMA utilisation of the triangle GEMM at the precision-config shape: 48 shells (kbin =
    kcenter = 1 at grid 512, max |q| = 48) on the band-limited maps grid 150^3, real space (48
    columns) and redshift space (144 columns).
      all_sums:            every (pair, column) sum, through pbi4_triangle_sums on random
                           unit-variance maps (dense kernel);
      in_bispectrum:       the bispectrum's own GEMM (block-sparse: the column windows its
                           triangles read; real space on the (j, l) rows), timed inside
                           pbi4_measure_bk on a random 512^3 half spectrum;
      in_bispectrum_dense: the same with PBI4_TRIANGLE_GEMM=dense;
      in_bispectrum_pair_rows (real space): PBI4_TRIANGLE_GEMM=block_sparse, the (i <= j) pair
                           rows.
    executed_tflops counts the MFMA work issued (x3 for split-BF16), useful_tflops only the sums
    that are read; mfma_frac = executed / peak."""
    from pbi4_amd import _lib
    out = {}
    nb, Np, N = 48, 150, 512
    K = Np ** 3
    Ks = ((K + 63) // 64) * 64
    npair = nb * (nb + 1) // 2
    g = torch.Generator(device="cuda")
    g.manual_seed(7)
    itot = len(triangle_list(nb, 1.0, 1.0, True))
    ntr = ctx.count_triangles(nb, N, 1.0, 1.0, True, False, itot)
    counts = torch.from_numpy((ntr.astype(np.float64) * float(N) ** 3).astype(np.float32)).cuda()
    delta = torch.randn((N, N, N // 2 + 1), generator=g, device="cuda", dtype=torch.complex64)
    kf = np.float32(2 * np.pi / BOX)

    def timed(fn, reps=5):
        fn()
        torch.cuda.synchronize()
        ctx.timing_enable("pbi4_triangle_gemm")
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        ms, k = ctx.timing_read()
        ctx.timing_enable(None)
        return ms / max(k, 1)

    for nell in (1, 3):
        maps = torch.randn((nb * nell, Ks), generator=g, device="cuda", dtype=torch.float32)
        maps[:, K:] = 0
        tiles = gemm_blocks(nb, nell, 1.0, 1.0, True)
        pair_tiles = gemm_blocks(nb, nell, 1.0, 1.0, True, pair_rows=True)["block_sparse"]
        modes = [("dense", "_in_bispectrum_dense", tiles["dense"]),
                 (None, "_in_bispectrum", tiles["block_sparse"])]
        if nell == 1:
            modes.append(("block_sparse", "_in_bispectrum_pair_rows", pair_tiles))
        res = {}
        for prec, name, peak, x in ((_lib.PREC_BF16X3, "bf16x3", BF16_PEAK_TFLOPS, 3),
                                    (_lib.PREC_FP32, "fp32", 157.3, 1)):
            def rates(avg, issued_tiles):
                ex = x * 2.0 * 256 * issued_tiles * K / (avg * 1e-3) / 1e12
                return {"avg_ms": round(avg, 4), "executed_tflops": round(ex, 1),
                        "mfma_frac": round(ex / peak, 4),
                        "useful_tflops": round(x * 2.0 * tiles["useful"] * K / (avg * 1e-3) / 1e12, 1)}

            avg = timed(lambda: ctx.triangle_sums(maps, nb, nell, K, prec))
            res[name + "_all_sums"] = rates(avg, tiles["dense"])
            desc = ctx.bk_desc(nb, N, 1.0, 1.0, True, kf, itot, 10 ** 8, nell == 3,
                               (0.0, 0.0, 1.0), prec, 0)
            assert ctx.maps_grid(desc) == Np
            for mode, key, issued in modes:
                if mode:
                    os.environ["PBI4_TRIANGLE_GEMM"] = mode
                try:
                    ws = ctx._ws(ctx.bk_workspace(desc))
                    avg = timed(lambda: ctx.bk(delta, counts, desc, ws=ws))
                finally:
                    os.environ.pop("PBI4_TRIANGLE_GEMM", None)
                res[name + key] = rates(avg, issued)
                del ws
            res[name + "_block_sparse_speedup"] = round(
                res[name + "_in_bispectrum_dense"]["avg_ms"] / res[name + "_in_bispectrum"]["avg_ms"], 3)
        res["tiles_issued_of_dense"] = round(tiles["block_sparse"] / tiles["dense"], 4)
        out[f"nb48_nell{nell}_K150^3"] = res
        del maps
    log(f"triangle GEMM alone: {out}")
    return out


def gemm_full_grid(ctx, torch):
    """The bispectrum's triangle GEMM on the full-grid aliasing shapes of the mixed-grid stream
    at the reference's shell counts (kbin = kcenter = 2, open triangles, redshift space, three
    multipoles, split-BF16x3): 384^3 with 64 shells (25 424 triangles) and 256^3 with 42 (7 714),
    timed inside pbi4_measure_bk on a random half spectrum (the GEMM does not depend on the
    values).  mfma_frac = issued MFMA work over the bf16 peak."""
    from pbi4_amd import _lib
    out = {}
    g = torch.Generator(device="cuda")
    g.manual_seed(11)
    kf = np.float32(2 * np.pi / BOX)
    for N, nb in ((384, 64), (256, 42)):
        itot = len(triangle_list(nb, 2.0, 2.0, True))
        ntr = ctx.count_triangles(nb, N, 2.0, 2.0, True, False, itot)
        counts = torch.from_numpy((ntr.astype(np.float64) * float(N) ** 3).astype(np.float32)).cuda()
        delta = torch.randn((N, N, N // 2 + 1), generator=g, device="cuda", dtype=torch.complex64)
        desc = ctx.bk_desc(nb, N, 2.0, 2.0, True, kf, itot, 10 ** 8, True, (0.0, 0.0, 1.0),
                           _lib.PREC_BF16X3, 0)
        assert ctx.maps_grid(desc) == N
        ws = ctx._ws(ctx.bk_workspace(desc))
        ctx.bk(delta, counts, desc, ws=ws)
        torch.cuda.synchronize()
        ctx.timing_enable("pbi4_triangle_gemm")
        reps = 3
        for _ in range(reps):
            ctx.bk(delta, counts, desc, ws=ws)
        torch.cuda.synchronize()
        ms, k = ctx.timing_read()
        ctx.timing_enable(None)
        avg = ms / max(k, 1)
        tiles = gemm_blocks(nb, 3, 2.0, 2.0, True)
        ex = 3 * 2.0 * 256 * tiles["block_sparse"] * N ** 3 / (avg * 1e-3) / 1e12
        out[f"{N}^3_nb{nb}_nell3"] = {"avg_ms": round(avg, 3), "executed_tflops": round(ex, 1),
                                      "mfma_frac": round(ex / BF16_PEAK_TFLOPS, 4),
                                      "tiles_issued_of_dense": round(tiles["block_sparse"] / tiles["dense"], 4)}
        log(f"full-grid GEMM {N}^3 nb={nb}: {out[f'{N}^3_nb{nb}_nell3']}")
        del ws, delta, counts
        torch.cuda.empty_cache()
    return out


def permuted_assign(ctx, torch, cat, plan):
    """Assignment (sort + deposit + gather) and deposit-kernel time of the headline catalogue in
    its lattice order and randomly permuted (the sort's run aggregation and the deposit's lane
    hand-out are tuned on spatially ordered input; the grids are bitwise the same either way)."""
    p, v = cat
    npart = p.shape[1]
    rsd = tuple(float(np.float32(plan.rsd_coef * l)) for l in (0.0, 0.0, 1.0))
    gen = torch.Generator(device=p.device)
    gen.manual_seed(3)
    perm = torch.randperm(npart, device=p.device, generator=gen)
    pp, vp = p[:, perm].contiguous(), v[:, perm].contiguous()
    out, grids = {}, {}
    for name, (a, b) in (("lattice_order", (p, v)), ("permuted", (pp, vp))):
        s = torch.cuda.current_stream()
        ctx.assign(a, 256, 4, True, vel=b, box=BOX, rsd=rsd, normalised=False)
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ctx.timing_enable("pbi4_deposit")
        ev[0].record(s)
        reps = 5
        for _ in range(reps):
            g = ctx.assign(a, 256, 4, True, vel=b, box=BOX, rsd=rsd, normalised=False)
        ev[1].record(s)
        torch.cuda.synchronize()
        dep_ms, k = ctx.timing_read()
        ctx.timing_enable(None)
        out[name] = {"assign_ms": round(ev[0].elapsed_time(ev[1]) / reps, 4),
                     "deposit_ms": round(dep_ms / max(k, 1), 4)}
        grids[name] = g
    out["grids_bitwise_equal"] = bool(torch.equal(grids["lattice_order"], grids["permuted"]))
    log(f"permuted assign: {out}")
    return out


if __name__ == "__main__":
    sys.exit(main())
