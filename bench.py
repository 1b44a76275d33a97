"""Benchmark: P+B catalogues/sec at 256^3 (split-BF16 serving path vs plain FP32), p99 latency.

Workload (one "step" = one batch of `--slots` catalogues through the serving engine):
  256^3 Zel'dovich particles (16.8M, float32 positions + velocities) in L = 1000 Mpc/h,
  grid 256^3, PCS with interlacing, real space and redshift space (LOS z); P_l (l = 0, 2, 4 in
  redshift space) with kbin = kcenter = 2; bispectrum with kbin = kcenter = 2 up to 32 k_F
  (16 shells, 564 triangles, open triangles), B_0 in real space and B_0,2,4 in redshift space.
Inputs are resident in HBM before the timed region.  `value` = catalogues/sec of the serving
path (split-BF16x3 triangle GEMM, band-limited maps grid, concurrent slots).  Also reported:
the plain one-at-a-time FP32 path (reference structure: maps on the full 256^3 grid, FP32
triangle sums, host synchronisation after every call), p50/p99 latency of a Poisson stream at
80% of the measured capacity, the per-bin split-BF16 error, the roofline of one kernel timed with
HIP events inside the timed region, and the FP64 CPU oracle on a bounded sample.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import pbi4_amd  # noqa: E402,F401  (sets how rocFFT builds its kernels before torch is used)

METRIC = "P+B catalogues/sec at 256³ grid (split-BF16 vs plain FP32), p99 latency, 1e-4"
HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--slots", type=int, default=4)
    ap.add_argument("--pool", type=int, default=4, help="distinct resident catalogues")
    ap.add_argument("--lattice", type=int, default=256)
    ap.add_argument("--grid", type=int, default=256)
    ap.add_argument("--plain", type=int, default=3, help="catalogues for the plain path (0 = skip)")
    ap.add_argument("--latency-requests", type=int, default=24)
    ap.add_argument("--mixed-requests", type=int, default=200,
                    help="requests of the mixed-grid Poisson stream (0 = skip)")
    ap.add_argument("--mixed-pool", type=int, default=24, help="distinct catalogues it cycles")
    ap.add_argument("--mixed-shells", type=int, default=16,
                    help="bispectrum shells of every mixed-stream request; 0 = the reference's "
                         "default per grid (21, 32, 42, 64 at 128, 192, 256, 384)")
    ap.add_argument("--mixed-slots", type=int, default=4, help="serving slots of the mixed stream")
    ap.add_argument("--mixed-only", type=int, default=0,
                    help="run only the mixed stream, print it and exit")
    ap.add_argument("--roofline-kernel", default="pbi4_deposit")
    ap.add_argument("--roofline-steps", type=int, default=4,
                    help="one-slot catalogues of the timed pass that times the roofline kernel")
    ap.add_argument("--cpu-baseline", type=int, default=1)
    ap.add_argument("--stages", type=int, default=1, help="print a per-stage breakdown")
    ap.add_argument("--gemm", type=int, default=1,
                    help="time the triangle GEMM alone at the 512^3 / 48-shell shape")
    ap.add_argument("--large", type=int, default=0,
                    help="measure the large single box with this many particles per side and grid "
                         "(1024 for the full configuration; three lines of sight, 32 shells) "
                         "instead of the serving workload, print it and exit")
    ap.add_argument("--gemm-only", type=int, default=0,
                    help="run only the triangle GEMM timing (for profiling), print it and exit")
    return ap.parse_args()


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


def bf16_error(ref_out, out, kf):
    """Max / median / p99 per-bin difference of the B rows of `out` from `ref_out` ({field:
    (pout, bout)} host arrays), relative to max(|B_l|, |B_0|, Gaussian sigma of the bin)."""
    worst = 0.0
    per_bin = []
    for f in ref_out:
        rb, mb = ref_out[f][1].T, out[f][1].T
        gauss = sigma_b(rb, kf)
        b0 = np.abs(rb[6] - rb[7])
        rows = [6] + ([9, 10] if rb.shape[0] == 11 else [])
        for r in rows:
            bref = rb[r] - (rb[7] if r == 6 else 0)
            bm = mb[r] - (mb[7] if r == 6 else 0)
            err = np.abs(bm - bref) / np.maximum(np.maximum(np.abs(bref), gauss), b0)
            per_bin.append(err)
            worst = max(worst, float(err.max()))
    allb = np.concatenate(per_bin)
    return worst, float(np.median(allb)), float(np.percentile(allb, 99))


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
    if args.gemm_only:
        print(json.dumps({"triangle_gemm": gemm_bench(ctx, torch)}), flush=True)
        return 0
    if args.large:
        print(json.dumps(large_box(ctx, torch, n=args.large)), flush=True)
        return 0
    if args.mixed_only:
        print(json.dumps({"mixed_stream": mixed_stream(ctx, torch, args.mixed_requests,
                                                       args.mixed_pool, 1000.0,
                                                       args.mixed_shells, args.mixed_slots)}),
              flush=True)
        return 0
    N, n = args.grid, args.lattice
    box = 1000.0
    m = Measurement(grid=N, order=4, interlacing=True, box=box, kbin=2.0, kcenter=2.0,
                    nbins_ps=0, nbins_bisp=16, iopen=True,
                    fields=(Field(False), Field(True, (0.0, 0.0, 1.0))),
                    precision=_lib.PREC_BF16X3, maps_grid=0)
    log(f"[bench] generating {args.pool} catalogues of {n}^3 particles on the GPU")
    pool = [zeldovich(n, box=box, seed=1000 + i, device="cuda") for i in range(args.pool)]
    torch.cuda.synchronize()
    npart = n ** 3
    t0 = time.perf_counter()
    plan = CataloguePlan(ctx, m)
    torch.cuda.synchronize()
    log(f"[bench] plan ready in {time.perf_counter() - t0:.2f}s: itot={plan.itot} "
        f"maps grid {plan.maps_grid}^3")

    eng = ServingEngine(ctx, slots=args.slots)
    eng.plans[m] = plan
    rid = [0]

    def batch(k):
        reqs = []
        for _ in range(k):
            p, v = pool[rid[0] % len(pool)]
            reqs.append(Request(rid[0], m, p, v, arrival=0.0))
            rid[0] += 1
        return reqs

    def step():
        return eng.run(batch(args.slots))

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    log("[bench] warmup done")

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    n_cat = args.steps * args.slots
    value = n_cat / elapsed
    log(f"[bench] serving: {n_cat} catalogues in {elapsed:.3f}s -> {value:.2f} catalogues/s")

    # ---------------------------------------------------------------- timed pass for the roofline
    # the same workload through a one-slot engine: launches do not overlap, so the HIP event pair
    # the library records around each launch on its stream brackets exactly one kernel (with
    # several slots the pair also holds the time the launch queues behind other slots' kernels)
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

    # ---------------------------------------------------------------- roofline of one kernel
    avg_ms = kern_ms / max(launches, 1)
    V = N ** 3
    ntiles = (N // 16) ** 3
    if args.roofline_kernel == "pbi4_deposit":
        # per launch: read the tile-sorted particles (12 B each), write both interlaced grids'
        # halo regions once (algorithmic = the 2 x 4 B x V grid cells)
        alg_bytes = 12.0 * npart + 8.0 * V
        roof = {"bound": "hbm", "peak": HBM_PEAK_GBS, "unit": "GB/s"}
        achieved = alg_bytes / (avg_ms * 1e-3) / 1e9
    elif args.roofline_kernel == "pbi4_triangle_gemm":
        # executed bf16 MFMA flops per launch (averaged over the real-space launch with nb
        # columns and the redshift-space launch with 3 nb columns): 3 x 2 x pairs x cols x K
        nb = m.nb_bisp()
        npair = nb * (nb + 1) // 2
        K = plan.maps_grid ** 3
        flops = 3 * 2.0 * npair * nb * (1 + 3) * K / 2
        roof = {"bound": "mfma", "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s"}
        achieved = flops / (avg_ms * 1e-3) / 1e12
        alg_bytes = None
    else:
        alg_bytes = None
        roof = {"bound": "hbm", "peak": HBM_PEAK_GBS, "unit": "GB/s"}
        achieved = float("nan")
    roof.update({"kernel": args.roofline_kernel, "achieved": round(achieved, 2),
                 "frac": round(achieved / roof["peak"], 4),
                 "traffic": pmc_traffic("k_deposit<4, 2, 20>") if args.roofline_kernel == "pbi4_deposit" else None,
                 "avg_launch_ms": round(avg_ms, 5), "launches": launches,
                 "timing": "HIP events around each launch on its stream, one-slot timed pass of "
                           "the same workload",
                 "algorithmic_bytes_per_launch": alg_bytes})

    # ---------------------------------------------------------------- per-stage breakdown
    stages = {}
    if args.stages:
        s = torch.cuda.current_stream()
        p, v = pool[0]
        rsd = tuple(float(np.float32(plan.rsd_coef * l)) for l in (0.0, 0.0, 1.0))
        names = ["assign", "density", "pk", "bk"]
        for rep in range(2):
            ev = [torch.cuda.Event(enable_timing=True) for _ in range(5)]
            ev[0].record(s)
            g = ctx.assign(p, N, 4, True, vel=v, box=box, rsd=rsd, normalised=False)
            ev[1].record(s)
            d = ctx.density(g, N, 4, True, box, npart)
            ev[2].record(s)
            ctx.pk(d, N, box, 2.0, 2.0, m.nb_ps(), True)
            ev[3].record(s)
            desc = plan.desc[Field(True, (0.0, 0.0, 1.0))]
            desc.npart = npart
            ctx.bk(d, plan.counts, desc)
            ev[4].record(s)
            torch.cuda.synchronize()
        stages = {nm: round(ev[i].elapsed_time(ev[i + 1]), 4) for i, nm in enumerate(names)}
        log(f"[bench] redshift-space field stages (ms): {stages}")

    # ---------------------------------------------------------------- plain one-at-a-time FP32
    plain_value = None
    errs = None
    if args.plain > 0:
        mp = Measurement(**{**m.__dict__, "precision": _lib.PREC_FP32, "maps_grid": N})
        pplan = CataloguePlan(ctx, mp)

        def plain_one(p, v):
            outs = {}
            for f in mp.fields:
                rsd = tuple(float(np.float32(pplan.rsd_coef * l)) for l in f.los) if f.rsd else None
                g = ctx.assign(p, N, 4, True, vel=v if f.rsd else None, box=box, rsd=rsd,
                               normalised=False)
                torch.cuda.synchronize()
                d = ctx.density(g, N, 4, True, box, npart)
                torch.cuda.synchronize()
                po = ctx.pk(d, N, box, 2.0, 2.0, mp.nb_ps(), f.rsd, f.los).cpu().numpy()
                desc = pplan.desc[f]
                desc.npart = npart
                bo = ctx.bk(d, pplan.counts, desc).cpu().numpy()
                outs[f] = (po, bo)
            return outs

        plain_one(*pool[0])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ref_outs = []
        for i in range(args.plain):
            ref_outs.append(plain_one(*pool[i % len(pool)]))
        plain_value = args.plain / (time.perf_counter() - t0)
        log(f"[bench] plain FP32 one-at-a-time: {plain_value:.2f} catalogues/s")
        # per-bin error of the serving (split-BF16, band-limited) result vs the plain FP32 one
        res = eng.run(batch(1))[0].result
        idx = (rid[0] - 1) % len(pool)
        plain_same = ref_outs[idx] if idx < len(ref_outs) else plain_one(*pool[idx])
        errs = bf16_error(plain_same, res, m.kf())
        log(f"[bench] serving vs plain FP32 B error: max {errs[0]:.2e} median {errs[1]:.2e} "
            f"p99 {errs[2]:.2e}")

    # ---------------------------------------------------------------- Poisson latency
    lat = {}
    if args.latency_requests > 0:
        rate = 0.8 * value
        rng = np.random.default_rng(5)
        t = np.cumsum(rng.exponential(1.0 / rate, args.latency_requests))
        reqs = batch(args.latency_requests)
        for r, a in zip(reqs, t):
            r.arrival = float(a)
        done = eng.run(reqs)
        lat = latency_summary(done)
        lat["arrival_rate_per_s"] = round(rate, 3)
        log(f"[bench] Poisson stream at {rate:.1f}/s: {lat}")

    # ---------------------------------------------------------------- mixed-grid stream
    mixed = None
    eng.release()
    if args.mixed_requests > 0:
        mixed = mixed_stream(ctx, torch, args.mixed_requests, args.mixed_pool, box,
                             args.mixed_shells, args.mixed_slots)

    # ---------------------------------------------------------------- triangle GEMM alone
    gemm = gemm_bench(ctx, torch) if args.gemm else None

    # ---------------------------------------------------------------- CPU baseline
    cpu = None
    if args.cpu_baseline:
        cpu = cpu_baseline(pool[0], plan, m, N, box, ctx)

    out = {
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
        "plain_fp32_catalogues_per_s": None if plain_value is None else round(plain_value, 4),
        "speedup_vs_plain": None if plain_value is None else round(value / plain_value, 3),
        "latency": lat,
        "mixed_stream": mixed,
        "split_bf16_vs_fp32_err": None if errs is None else
        {"max": errs[0], "median": errs[1], "p99": errs[2]},
        "stage_ms_rsd_field": stages,
        "triangle_gemm": gemm,
    }
    print(json.dumps(out), flush=True)
    return 0


def mixed_stream(ctx, torch, nreq, npool, box, shells=16, slots=4):
    """Serving a mixed stream: `nreq` requests cycling `npool` resident Zel'dovich catalogues
    whose particle counts are log-uniform in [1e6, 5e7] (lattice n = round(count^(1/3))) with
    amplitudes uniform in [1e4, 4e4]; each request draws its grid from {128, 192, 256, 384}^3 and
    its assignment from CIC/TSC/PCS (interlaced) and asks for redshift-space P_l (l = 0, 2, 4)
    plus the bispectrum with kbin = kcenter = 2 over `shells` shells (16: up to 32 k_F; 0: the
    reference's default count for the grid, N / (3 kbin)).  Reported: capacity (all requests
    queued at t = 0), p50/p99 latency of a Poisson stream at 80% of it, and the single-slot FP32
    path (full-grid shell maps, FP32 triangle sums) over the same requests."""
    from pbi4_amd import _lib
    from pbi4_amd.catalogue import zeldovich
    from pbi4_amd.pipeline import Field, Measurement
    from pbi4_amd.serving import Request, ServingEngine, latency_summary
    rng = np.random.default_rng(2024)
    pool = []
    for i in range(npool):
        n = int(round(math.exp(rng.uniform(math.log(1e6), math.log(5e7))) ** (1.0 / 3.0)))
        amp = float(rng.uniform(1e4, 4e4))
        pool.append(zeldovich(n, box=box, amp=amp, seed=7000 + i, device="cuda"))
    torch.cuda.synchronize()
    draws = [(i % npool, int(rng.choice((128, 192, 256, 384))), int(rng.choice((2, 3, 4))))
             for i in range(nreq)]

    def meas(grid, order, prec, full):
        return Measurement(grid=grid, order=order, interlacing=True, box=box, kbin=2.0,
                           kcenter=2.0, nbins_ps=0, nbins_bisp=shells, iopen=True,
                           fields=(Field(True, (0.0, 0.0, 1.0)),), precision=prec,
                           maps_grid=grid if full else 0)

    def requests(prec, full):
        out = []
        for rid, (c, grid, order) in enumerate(draws):
            p, v = pool[c]
            out.append(Request(rid, meas(grid, order, prec, full), p, v, arrival=0.0))
        return out

    def warm(eng, prec, full):
        # plans (rocFFT plans, exact triangle counts) and one pass per shape, untimed
        seen = {}
        for c, grid, order in draws:
            seen.setdefault((grid, order), c)
        eng.run([Request(-1, meas(g, o, prec, full), *pool[c], arrival=0.0)
                 for (g, o), c in seen.items()])
        torch.cuda.synchronize()

    eng = ServingEngine(ctx, slots=slots)
    warm(eng, _lib.PREC_BF16X3, False)
    log("[bench] mixed stream: plans and one pass per shape done")
    t0 = time.perf_counter()
    eng.run(requests(_lib.PREC_BF16X3, False))
    torch.cuda.synchronize()
    cap = nreq / (time.perf_counter() - t0)
    log(f"[bench] mixed stream: capacity {cap:.2f}/s")
    rate = 0.8 * cap
    reqs = requests(_lib.PREC_BF16X3, False)
    t = np.cumsum(np.random.default_rng(6).exponential(1.0 / rate, nreq))
    for r, a in zip(reqs, t):
        r.arrival = float(a)
    t0 = time.perf_counter()
    done = eng.run(reqs)
    torch.cuda.synchronize()
    poisson_rate = nreq / (time.perf_counter() - t0)
    lat = latency_summary(done)
    eng.release()
    del eng
    log(f"[bench] mixed stream: Poisson at {rate:.2f}/s served {poisson_rate:.2f}/s")
    plain = ServingEngine(ctx, slots=1)
    warm(plain, _lib.PREC_FP32, True)
    log("[bench] mixed stream: FP32 path warmed")
    t0 = time.perf_counter()
    plain.run(requests(_lib.PREC_FP32, True))
    torch.cuda.synchronize()
    plain_rate = nreq / (time.perf_counter() - t0)
    plain.release()
    del plain
    counts = [int(pool[c][0].shape[1]) for c, _, _ in draws]
    res = {"requests": nreq, "distinct_catalogues": npool,
           "shells": {g: meas(g, 4, _lib.PREC_BF16X3, False).nb_bisp() for g in (128, 192, 256, 384)},
           "triangles": {g: meas(g, 4, _lib.PREC_BF16X3, False).itot() for g in (128, 192, 256, 384)},
           "npart_min": min(counts), "npart_max": max(counts),
           "grids": {g: sum(1 for d in draws if d[1] == g) for g in (128, 192, 256, 384)},
           "orders": {o: sum(1 for d in draws if d[2] == o) for o in (2, 3, 4)},
           "capacity_catalogues_per_s": round(cap, 3),
           "poisson_arrival_rate_per_s": round(rate, 3),
           "poisson_served_per_s": round(poisson_rate, 3),
           "latency": lat,
           "plain_fp32_single_slot_catalogues_per_s": round(plain_rate, 3),
           "speedup_vs_plain": round(cap / plain_rate, 3)}
    log(f"[bench] mixed stream: {res}")
    del pool
    torch.cuda.empty_cache()
    return res


def large_box(ctx, torch, n=1024, box=1000.0, reps=2):
    """The large single-box configuration: n^3 Zel'dovich particles (float32 positions and
    velocities, 26 GB at n = 1024), grid n^3, PCS with interlacing, redshift space along x, y and
    z, P_l (l = 0, 2, 4, kbin = kcenter = 2) and the split-BF16x3 bispectrum over 32 shells of
    dk = 2 k_F.  Inputs resident in HBM; per line of sight the stages are timed with events on
    the stream, the deposit and the triangle GEMM with the library's launch timers."""
    from pbi4_amd import _lib
    from pbi4_amd.catalogue import rsd_factor, zeldovich
    from pbi4_amd.pipeline import CataloguePlan, Field, Measurement
    N = n
    log(f"[large] generating {n}^3 Zel'dovich particles on the GPU")
    p, v = zeldovich(n, box=box, seed=4242, device="cuda")
    torch.cuda.synchronize()
    npart = n ** 3
    los3 = ((1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0))
    m = Measurement(grid=N, order=4, interlacing=True, box=box, kbin=2.0, kcenter=2.0,
                    nbins_ps=0, nbins_bisp=32, iopen=True, fields=tuple(Field(True, l) for l in los3),
                    precision=_lib.PREC_BF16X3, maps_grid=0)
    t0 = time.perf_counter()
    plan = CataloguePlan(ctx, m)
    torch.cuda.synchronize()
    log(f"[large] plan (bin tables, rocFFT plans, exact triangle counts) in "
        f"{time.perf_counter() - t0:.2f} s: itot={plan.itot}, maps grid {plan.maps_grid}^3")
    s = torch.cuda.current_stream()
    coef = float(rsd_factor())

    acc = {"dep": [0.0, 0], "gemm": [0.0, 0]}

    def field(l, timed):
        rsd = tuple(float(np.float32(coef * c)) for c in l)
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(5)]
        if timed:
            ctx.timing_enable("pbi4_deposit")
        ev[0].record(s)
        g = ctx.assign(p, N, 4, True, vel=v, box=box, rsd=rsd, normalised=False)
        ev[1].record(s)
        if timed:
            torch.cuda.synchronize()
            ms, k = ctx.timing_read()
            acc["dep"][0] += ms
            acc["dep"][1] += k
            ctx.timing_enable(None)
        d = ctx.density(g, N, 4, True, box, npart)
        del g
        ev[2].record(s)
        po = ctx.pk(d, N, box, 2.0, 2.0, m.nb_ps(), True, l)
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
    for rep in range(reps):
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
    gemm_tf = 3 * 2.0 * (nb * (nb + 1) // 2) * 3 * nb * plan.maps_grid ** 3 / (gemm_ms * 1e-3) / 1e12
    out = {"metric": "large box: seconds per catalogue (3 lines of sight, P_l + B)",
           "value": round(per_cat, 4), "unit": "s", "higher_is_better": False,
           "config": {"workload": f"{n}^3 Zel'dovich particles, grid {N}^3 PCS interlaced, "
                                  "redshift space along x, y, z, P_l l=0,2,4, B with 32 shells of "
                                  "2 k_F (split-BF16x3)", "npart": npart, "itot": plan.itot,
                      "maps_grid": plan.maps_grid},
           "stage_ms_per_field": {k: round(float(x), 3) for k, x in
                                  zip(("assign", "density", "pk", "bk"), stages)},
           "deposit": {"avg_launch_ms": round(dep_ms, 3), "algorithmic_GBps": round(dep_gbs, 1),
                       "frac_of_hbm": round(dep_gbs / HBM_PEAK_GBS, 4)},
           "triangle_gemm": {"avg_launch_ms": round(gemm_ms, 3), "executed_tflops": round(gemm_tf, 1),
                             "mfma_frac": round(gemm_tf / BF16_PEAK_TFLOPS, 4)},
           "pk_rows": int(po.shape[0]), "bk_rows": int(bo.shape[0]),
           "finite": bool(torch.isfinite(po).all() and torch.isfinite(bo).all())}
    log(f"[large] {out}")
    del p, v
    torch.cuda.empty_cache()
    return out


def triangle_list(nb, kbin, kcenter, iopen):
    """(l, j, i) of the reference's triangle loop (pbi4/pbi4.py:105-117)."""
    ishift = int(kcenter / kbin) + (1 if iopen else 0)
    return [(l, j, i) for l in range(1, nb + 1) for j in range(1, l + 1)
            for i in range(max(1, l - j + 1 - ishift), j + 1)]


def gemm_blocks(nb, nell, kbin, kcenter, iopen):
    """16 x 16 output blocks of the triangle GEMM that the masked kernel computes (per 16-row
    fragment of pair rows, the column tiles holding a sum the finalize reads) and the number of
    sums it reads (the useful entries)."""
    tiles = {}
    tl = triangle_list(nb, kbin, kcenter, iopen)
    for l, j, i in tl:
        p = (j - 1) * j // 2 + (i - 1)
        for e in range(nell):
            tiles.setdefault(p // 16, set()).add((e * nb + l - 1) // 16)
    return sum(len(t) for t in tiles.values()), len(tl) * nell


def gemm_bench(ctx, torch):
    """MFMA utilisation of the triangle GEMM at the precision-config shape: 48 shells (kbin =
    kcenter = 1 at grid 512, max |q| = 48) on the band-limited maps grid 150^3, real space (48
    columns) and redshift space (144 columns).
      dense:         every (pair, column) sum, through pbi4_triangle_sums on random
                     unit-variance maps;
      in_bispectrum: the bispectrum's own GEMM, timed inside pbi4_measure_bk on a random 512^3
                     half spectrum (executed = the 16 x 16 blocks the kernel issues).
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
    kf = np.float32(2 * np.pi / 1000.0)

    def timed(fn, reps=5):
        fn()
        torch.cuda.synchronize()
        ctx.timing_enable("pbi4_triangle_gemm")
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        ms, n = ctx.timing_read()
        ctx.timing_enable(None)
        return ms / max(n, 1)

    for nell in (1, 3):
        maps = torch.randn((nb * nell, Ks), generator=g, device="cuda", dtype=torch.float32)
        maps[:, K:] = 0
        blocks, useful = gemm_blocks(nb, nell, 1.0, 1.0, True)
        res = {}
        for prec, name, peak, x in ((_lib.PREC_BF16X3, "bf16x3", BF16_PEAK_TFLOPS, 3),
                                    (_lib.PREC_FP32, "fp32", 157.3, 1)):
            avg = timed(lambda: ctx.triangle_sums(maps, nb, nell, K, prec))
            executed = x * 2.0 * npair * nb * nell * K
            tf = executed / (avg * 1e-3) / 1e12
            res[name + "_dense"] = {"avg_ms": round(avg, 4), "executed_tflops": round(tf, 1),
                                    "mfma_frac": round(tf / peak, 4),
                                    "useful_tflops": round(x * 2.0 * useful * K / (avg * 1e-3) / 1e12, 1)}
            desc = ctx.bk_desc(nb, N, 1.0, 1.0, True, kf, itot, 10 ** 8, nell == 3,
                               (0.0, 0.0, 1.0), prec, 0)
            assert ctx.maps_grid(desc) == Np
            ws = ctx._ws(ctx.bk_workspace(desc))
            avg = timed(lambda: ctx.bk(delta, counts, desc, ws=ws))
            executed = x * 2.0 * npair * nb * nell * K
            tf = executed / (avg * 1e-3) / 1e12
            res[name + "_in_bispectrum"] = {
                "avg_ms": round(avg, 4), "executed_tflops": round(tf, 1),
                "mfma_frac": round(tf / peak, 4),
                "useful_tflops": round(x * 2.0 * useful * K / (avg * 1e-3) / 1e12, 1),
                "blocks_needed_of_dense": round(blocks / (-(-npair // 16) * -(-nb * nell // 16)), 4)}
            del ws
        out[f"nb48_nell{nell}_K150^3"] = res
        del maps
    log(f"[bench] triangle GEMM alone: {out}")
    return out


def cpu_baseline(cat, plan, m, N, box, ctx):
    """FP64 oracle (the reference algorithm restated) on the host cores, bounded sample: the
    redshift-space field of one catalogue in full except the triangle sums, which run on every
    16th triangle; the real-space field and the other triangles are extrapolated from the
    measured per-stage times.  Also checks the GPU bispectrum on the sampled triangles."""
    import torch
    from oracle import pbi4_oracle as O
    cores = O.nthreads()
    p, v = cat
    npart = p.shape[1]
    pos = p.cpu().numpy().astype(np.float64)
    vz = v[2].cpu().numpy().astype(np.float64)
    pos[2] = pos[2] + vz * (plan.rsd_coef * 1.0)          # pbi4.py:154 (float64 then float32)
    pos = pos.astype(np.float32)
    pos = ((pos + np.float32(box)) % np.float32(box)) / np.float32(box)
    pos = np.minimum(pos, np.float32(1.0)).astype(np.float64)
    los = np.array([0.0, 0.0, 1.0])
    T = {}
    t = time.perf_counter()
    g = O.assign_double_grid(N, 4, npart, pos)
    T["assign"] = time.perf_counter() - t
    del pos
    t = time.perf_counter()
    dcl = O.compute_density(npart, g, N, 4, True, box)
    T["density"] = time.perf_counter() - t
    del g
    t = time.perf_counter()
    d2, d4 = O.compute_density_multipoles(dcl, N, box, los)
    T["multipoles"] = time.perf_counter() - t
    t = time.perf_counter()
    O.measure_pk_multipoles(N, box, m.kbin, m.kcenter, m.nb_ps(), dcl, d2, d4)
    T["pk"] = time.perf_counter() - t
    nb = m.nb_bisp()
    tl, idx = O._triangle_map_index(nb, m.kbin, m.kcenter, m.iopen)
    sample = np.arange(0, len(tl), 16)
    t = time.perf_counter()
    maps, nmodes = O.shell_maps(dcl, N, nb, m.kbin, m.kcenter)
    m2, _ = O.shell_maps(d2, N, nb, m.kbin, m.kcenter)
    m4, _ = O.shell_maps(d4, N, nb, m.kbin, m.kcenter)
    T["maps"] = time.perf_counter() - t
    del d2, d4
    t = time.perf_counter()
    s0 = O.triple_sums(maps, maps, idx[sample])
    s2 = O.triple_sums(maps, m2, idx[sample])
    s4 = O.triple_sums(maps, m4, idx[sample])
    T["sums_sample"] = time.perf_counter() - t
    full_sums = T["sums_sample"] * len(tl) / len(sample)
    t_rsd = T["assign"] + T["density"] + T["multipoles"] + T["pk"] + T["maps"] + full_sums
    t_real = T["assign"] + T["density"] + T["pk"] + T["maps"] / 3 + full_sums / 3
    total = t_rsd + t_real
    sampled = sum(T.values())
    log(f"[bench] CPU oracle stages (s, {cores} cores): {({k: round(x, 2) for k, x in T.items()})}"
        f" -> {total:.1f} s per catalogue (extrapolated)")
    # FP64 parity of the GPU path at full size on every triangle of the redshift-space field
    # (the sums beyond the timed sample are not timed)
    rest = np.setdiff1d(np.arange(len(tl)), sample)
    order = np.concatenate([sample, rest])
    s0 = np.concatenate([s0, O.triple_sums(maps, maps, idx[rest])])[np.argsort(order)]
    s2 = np.concatenate([s2, O.triple_sums(maps, m2, idx[rest])])[np.argsort(order)]
    s4 = np.concatenate([s4, O.triple_sums(maps, m4, idx[rest])])[np.argsort(order)]
    del m2, m4
    counts = plan.counts.cpu().numpy().astype(np.float64)
    kf3 = O.kf3_bk(m.kf())
    f = [k for k in plan.desc if k.rsd][0]
    check = {}
    for name, prec, mg in (("split_bf16_bandlimited", 1, 0), ("fp32_full_grid", 0, N)):
        d = plan.desc[f]
        gdesc = ctx.bk_desc(nb, N, m.kbin, m.kcenter, m.iopen, m.kf(), plan.itot, npart, True,
                            (0.0, 0.0, 1.0), prec, mg)
        rsd = (0.0, 0.0, float(np.float32(plan.rsd_coef)))
        gg = ctx.assign(p, N, 4, True, vel=v, box=box, rsd=rsd, normalised=False)
        dd = ctx.density(gg, N, 4, True, box, npart)
        bo = ctx.bk(dd, plan.counts, gdesc).cpu().numpy()
        del d
        ref = {6: kf3 * s0 / counts, 9: 5 * kf3 * s2 / counts, 10: 9 * kf3 * s4 / counts}
        b0 = np.abs(ref[6])
        gauss = sigma_b(bo.T, m.kf())
        worst, worst_plain = 0.0, 0.0
        for r in (6, 9, 10):
            mine = bo[:, r] - (bo[:, 7] if r == 6 else 0)
            err = np.abs(mine - ref[r])
            worst = max(worst, float((err / np.maximum(np.maximum(np.abs(ref[r]), b0), gauss)).max()))
            worst_plain = max(worst_plain, float((err / np.maximum(np.abs(ref[r]), b0)).max()))
        check[name] = {"vs_max_B_sigma": worst, "vs_max_B": worst_plain}
        del gg, dd
        torch.cuda.synchronize()
    log(f"[bench] GPU vs FP64 oracle at full size, all {len(tl)} triangles (B0, B2, B4), max "
        f"|B - B_oracle| / max(|B_l|, |B0|, sigma_B) and without sigma_B: {check}")
    return {"value": round(1.0 / total, 6), "unit": "catalogues/s", "cores": cores, "kind": "port",
            "sample": f"redshift-space field of one 256^3 catalogue (assignment, FFTs, fcomb, 22-FFT "
                      f"multipoles, P_l, 48 shell maps) in full; triangle sums on {len(sample)} of "
                      f"{len(tl)} triangles x3; real-space field and other triangles extrapolated; "
                      f"{sampled:.1f} s measured",
            "seconds_per_catalogue": round(total, 2), "stage_seconds": {k: round(x, 3) for k, x in T.items()},
            "gpu_vs_fp64_oracle_max_rel_err_all_triangles": check}


if __name__ == "__main__":
    sys.exit(main())
