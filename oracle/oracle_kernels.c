/*
 * FP64 CPU restatement of the PBI4 hot loops -- TEST INFRASTRUCTURE ONLY.
 *
 * This file is the checker for the HIP path: tests, the smoke check and the CPU-baseline leg of
 * the benchmark call it; the product library never links or loads it.  Every function follows
 * one reference routine literally (same loop order, same FP64 operation order, Fortran NINT /
 * INT / MOD semantics) so that its sums can be compared to the reference bit for bit where the
 * reference itself can be compiled (powerI4.f90 via oracle/Makefile) and to the committed golden
 * fixtures elsewhere.
 *
 * Array conventions (same memory layout as the f2py-returned Fortran arrays):
 *   pos      : numpy (3, npart) C-order, pos[d*npart + i]   (pbi4/pbi4.py:241-243)
 *   grids    : Fortran order, cell (ix,iy,iz) 0-based at ix + N*(iy + N*iz)
 *   complex  : interleaved (re, im) doubles
 *
 * Build: gcc -O2 -ffp-contract=off -fopenmp -shared -fPIC (oracle/Makefile).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#include <omp.h>

/* Fortran NINT for REAL(8): round half away from zero (gfortran lowers it to lround). */
static inline long fnint(double x) { return lround(x); }
/* Fortran MOD for the non-negative arguments used below. */
static inline long fmodi(long a, long n) { return a % n; }

/* First 0-based cell index (before the periodic wrap) and width of the stencil of one axis
 * for the coordinate v = N pos + 1 (or v - 1/2 for the interlaced grid), as each order's
 * branch below computes its indices; used only to skip particles that cannot touch a slab. */
static inline void stencil_span(int nintp, double v, long *first, int *width)
{
    switch (nintp) {
        case 4: *first = (long)v - 2; *width = 4; break;
        case 3: *first = fnint(v) - 2; *width = 3; break;
        case 2: *first = (long)v; *width = 2; break;
        default: *first = fnint(v) - 1; *width = 1; break;
    }
}

static inline int span_touches(long first, int width, long N, long zlo, long zhi)
{
    for (int k = 0; k < width; ++k) {
        const long z = (first + k + N) % N;
        if (z >= zlo && z < zhi) return 1;
    }
    return 0;
}

/* The particle loop of assign_single_grid / assign_double_grid restricted to the output
 * z-planes [zlo, zhi): every particle is visited in catalogue order and only its contributions
 * to those planes are added, so each cell receives exactly the additions of the serial loop,
 * in the same order. */
static void assign_slab(int grid, int nintp, int interlacing, int64_t npart, const double *pos,
                        const double *weight, double *out, long zlo, long zhi)
{
    const long N = grid;
    const int nintph = (nintp >= 3) ? 3 : 1;
    double w1[3][4] = {{0}}, w2[3][4] = {{0}};
    long iv[3][4], jv[3][4];
    for (int64_t i = 0; i < npart; ++i) {
        double vr[3], vt[3];
        long ivr[3], ivt[3];
        for (int d = 0; d < 3; ++d) {
            vr[d] = (double)N * pos[(int64_t)d * npart + i] + 1.0;
            vt[d] = vr[d] - 0.5;
        }
        {
            long f;
            int w;
            stencil_span(nintp, vr[2], &f, &w);
            int hit = span_touches(f, w, N, zlo, zhi);
            if (!hit && interlacing) {
                stencil_span(nintp, vt[2], &f, &w);
                hit = span_touches(f, w, N, zlo, zhi);
            }
            if (!hit) continue;
        }
        if (!interlacing) {
            if (nintp == 4) { /* PCS  powerI4.f90:87-100 */
                for (int d = 0; d < 3; ++d) ivr[d] = (long)vr[d];
                for (int j = 1; j <= 4; ++j)
                    for (int d = 0; d < 3; ++d) iv[d][j - 1] = fmodi(ivr[d] - nintph + j + N, N);
                for (int d = 0; d < 3; ++d) {
                    double h = vr[d] - (double)(long)vr[d];
                    double h2 = h * h;
                    double t = 1.0 - h;
                    w1[d][0] = (t * t) * t / 6.0;
                    w1[d][1] = 4.0 / 6.0 + (0.5 * h - 1.0) * h2;
                    w1[d][3] = h2 * h / 6.0;
                    w1[d][2] = 1.0 - w1[d][0] - w1[d][1] - w1[d][3];
                }
            } else if (nintp == 3) { /* TSC  :101-113 */
                for (int d = 0; d < 3; ++d) ivr[d] = fnint(vr[d]);
                for (int j = 1; j <= 3; ++j)
                    for (int d = 0; d < 3; ++d) iv[d][j - 1] = fmodi(ivr[d] - nintph + j + N, N);
                for (int d = 0; d < 3; ++d) {
                    double h = vr[d] - (double)fnint(vr[d]);
                    double h2 = h * h;
                    double t = 0.5 - h;
                    w1[d][0] = 0.5 * (t * t);
                    w1[d][1] = 0.75 - h2;
                    w1[d][2] = 1.0 - w1[d][0] - w1[d][1];
                }
            } else if (nintp == 2) { /* CIC  :114-124, note h = vr - nint(vr) */
                for (int d = 0; d < 3; ++d) ivr[d] = (long)vr[d];
                for (int j = 1; j <= 2; ++j)
                    for (int d = 0; d < 3; ++d) iv[d][j - 1] = fmodi(ivr[d] - nintph + j + N, N);
                for (int d = 0; d < 3; ++d) {
                    double h = vr[d] - (double)fnint(vr[d]);
                    w1[d][0] = 1.0 - h;
                    w1[d][1] = h;
                }
            } else { /* NGP :125-130 */
                for (int d = 0; d < 3; ++d) {
                    for (int j = 0; j < 4; ++j) w1[d][j] = 1.0;
                    ivr[d] = fnint(vr[d]);
                    iv[d][0] = fmodi(ivr[d] - 1 + N, N);
                }
            }
            const double wt = weight ? weight[i] : 1.0;
            for (int ix = 0; ix < nintp; ++ix)
                for (int iy = 0; iy < nintp; ++iy)
                    for (int iz = 0; iz < nintp; ++iz) {
                        if (iv[2][iz] < zlo || iv[2][iz] >= zhi) continue;
                        long c = iv[0][ix] + N * (iv[1][iy] + N * iv[2][iz]);
                        out[c] = out[c] + w1[0][ix] * w1[1][iy] * w1[2][iz] * wt;
                    }
            continue;
        }
        if (nintp == 4) { /* PCS  powerI4.f90:164-186 */
            for (int d = 0; d < 3; ++d) { ivr[d] = (long)vr[d]; ivt[d] = (long)vt[d]; }
            for (int j = 1; j <= 4; ++j)
                for (int d = 0; d < 3; ++d) {
                    iv[d][j - 1] = fmodi(ivr[d] - nintph + j + N, N);
                    jv[d][j - 1] = fmodi(ivt[d] - nintph + j + N, N);
                }
            for (int d = 0; d < 3; ++d) {
                double h1 = vr[d] - (double)(long)vr[d];
                double h2 = vt[d] - (double)(long)vt[d];
                double h1sq = h1 * h1, h2sq = h2 * h2;
                double t1 = 1.0 - h1, t2 = 1.0 - h2;
                w1[d][0] = (t1 * t1) * t1 / 6.0;
                w1[d][1] = 4.0 / 6.0 + (0.5 * h1 - 1.0) * h1sq;
                w1[d][3] = h1sq * h1 / 6.0;
                w1[d][2] = 1.0 - w1[d][0] - w1[d][1] - w1[d][3];
                w2[d][0] = (t2 * t2) * t2 / 6.0;
                w2[d][1] = 4.0 / 6.0 + (0.5 * h2 - 1.0) * h2sq;
                w2[d][3] = h2sq * h2 / 6.0;
                w2[d][2] = 1.0 - w2[d][0] - w2[d][1] - w2[d][3];
            }
        } else if (nintp == 3) { /* TSC  :187-207 */
            for (int d = 0; d < 3; ++d) { ivr[d] = fnint(vr[d]); ivt[d] = fnint(vt[d]); }
            for (int j = 1; j <= 3; ++j)
                for (int d = 0; d < 3; ++d) {
                    iv[d][j - 1] = fmodi(ivr[d] - nintph + j + N, N);
                    jv[d][j - 1] = fmodi(ivt[d] - nintph + j + N, N);
                }
            for (int d = 0; d < 3; ++d) {
                double h1 = vr[d] - (double)fnint(vr[d]);
                double h2 = vt[d] - (double)fnint(vt[d]);
                double h1sq = h1 * h1, h2sq = h2 * h2;
                double t1 = 0.5 - h1, t2 = 0.5 - h2;
                w1[d][0] = 0.5 * (t1 * t1);
                w1[d][1] = 0.75 - h1sq;
                w1[d][2] = 1.0 - w1[d][0] - w1[d][1];
                w2[d][0] = 0.5 * (t2 * t2);
                w2[d][1] = 0.75 - h2sq;
                w2[d][2] = 1.0 - w2[d][0] - w2[d][1];
            }
        } else if (nintp == 2) { /* CIC  :208-224 */
            for (int d = 0; d < 3; ++d) { ivr[d] = (long)vr[d]; ivt[d] = (long)vt[d]; }
            for (int j = 1; j <= 2; ++j)
                for (int d = 0; d < 3; ++d) {
                    iv[d][j - 1] = fmodi(ivr[d] - nintph + j + N, N);
                    jv[d][j - 1] = fmodi(ivt[d] - nintph + j + N, N);
                }
            for (int d = 0; d < 3; ++d) {
                double h1 = vr[d] - (double)(long)vr[d];
                double h2 = vt[d] - (double)(long)vt[d];
                w1[d][0] = 1.0 - h1; w1[d][1] = h1;
                w2[d][0] = 1.0 - h2; w2[d][1] = h2;
            }
        } else { /* NGP :225-233 */
            for (int d = 0; d < 3; ++d) {
                for (int j = 0; j < 4; ++j) { w1[d][j] = 1.0; w2[d][j] = 1.0; }
                ivr[d] = fnint(vr[d]);
                ivt[d] = fnint(vt[d]);
                iv[d][0] = fmodi(ivr[d] - 1 + N, N);
                jv[d][0] = fmodi(ivt[d] - 1 + N, N);
            }
        }
        const double wt = weight ? weight[i] : 1.0;
        for (int ix = 0; ix < nintp; ++ix)
            for (int iy = 0; iy < nintp; ++iy)
                for (int iz = 0; iz < nintp; ++iz) {
                    if (iv[2][iz] >= zlo && iv[2][iz] < zhi) {
                        long c = iv[0][ix] + N * (iv[1][iy] + N * iv[2][iz]);
                        out[2 * c] = out[2 * c] + w1[0][ix] * w1[1][iy] * w1[2][iz] * wt;
                    }
                    if (jv[2][iz] >= zlo && jv[2][iz] < zhi) {
                        long e = jv[0][ix] + N * (jv[1][iy] + N * jv[2][iz]);
                        out[2 * e + 1] = out[2 * e + 1] + w2[0][ix] * w2[1][iy] * w2[2][iz] * wt;
                    }
                }
    }
}

/*
 * assign_single_grid (powerI4.f90:68-140) when interlacing == 0, out = real N^3
 * assign_double_grid (powerI4.f90:142-249) when interlacing != 0, out = complex N^3
 * Zero-fills `out` first, as the f2py intent(out) wrapper does.  The output z-planes are split
 * over nthreads threads (0 = OpenMP default); each thread runs the reference's particle loop
 * in order for its own planes, so the result is bit-identical to the serial loop for any
 * thread count.
 */
void orc_assign(int grid, int nintp, int interlacing, int64_t npart, const double *pos,
                const double *weight, double *out, int nthreads)
{
    const long N = grid;
    const long V = N * N * N;
    memset(out, 0, sizeof(double) * (size_t)V * (interlacing ? 2 : 1));
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads)
    {
        const long T = omp_get_num_threads(), t = omp_get_thread_num();
        const long zlo = N * t / T, zhi = N * (t + 1) / T;
        if (zlo < zhi) assign_slab(grid, nintp, interlacing, npart, pos, weight, out, zlo, zhi);
    }
}

/*
 * measure_pk (powerI4.f90:510-594) when mode == 0: P(k) of <dcl1* dcl2>, pout rows avgk, P, co
 * measure_pk_multipoles (powerI4.f90:596-693) when mode == 1: rows avgk, P0, P2, P4, co
 * box, kbin, kcenter are the REAL(4) arguments.  pout is Fortran (5, nbins): pout[r + 5*b].
 */
void orc_measure_pk(int mode, int grid, float box, float kbin, float kcenter, int nbins,
                    const double *d1, const double *d2, const double *d4, double *pout)
{
    const long N = grid;
    const float twopi = 6.28319f;
    const double kF = (double)(twopi / box);
    const double kF3 = kF * kF * kF;
    const double dk = (double)kbin * kF;
    const double kcb = (double)(kcenter / kbin);
    double *avgk = calloc((size_t)nbins, sizeof(double));
    double *avgP = calloc((size_t)nbins, sizeof(double));
    double *avgP2 = calloc((size_t)nbins, sizeof(double));
    double *avgP4 = calloc((size_t)nbins, sizeof(double));
    double *co = calloc((size_t)nbins, sizeof(double));
    for (long iz = 1; iz <= N; ++iz) {
        long ikz = fmodi(iz + N / 2 - 2, N) - N / 2 + 1;
        double rkz = kF * (double)(float)ikz;
        for (long iy = 1; iy <= N; ++iy) {
            long iky = fmodi(iy + N / 2 - 2, N) - N / 2 + 1;
            double rky = kF * (double)(float)iky;
            for (long ix = 1; ix <= N / 2 + 1; ++ix) {
                long ikx = fmodi(ix + N / 2 - 2, N) - N / 2 + 1;
                double rkx = kF * (double)(float)ikx;
                if (ix != 1 || (ix == 1 && iky > 0) || (ix == 1 && iky == 0 && ikz > 0)) {
                    double rk = sqrt(rkx * rkx + rky * rky + rkz * rkz);
                    long imk = fnint(rk / dk - kcb - DBL_MIN) + 1;
                    if (imk <= nbins && imk > 0) {
                        long b = imk - 1;
                        long c = (ix - 1) + N * ((iy - 1) + N * (iz - 1));
                        co[b] = co[b] + 1.0;
                        avgk[b] = avgk[b] + rk;
                        double ar = d1[2 * c], ai = d1[2 * c + 1];
                        if (mode == 0) {
                            /* pk = real(conjg(ct1)*ct2): REAL() without KIND is REAL(4), so
                             * every term is rounded to single precision (powerI4.f90:580-582) */
                            double br = d2[2 * c], bi = d2[2 * c + 1];
                            avgP[b] = avgP[b] + (double)(float)(ar * br - (-ai) * bi);
                        } else {
                            double h = hypot(ar, ai);
                            avgP[b] = avgP[b] + h * h;
                            double br = d2[2 * c], bi = d2[2 * c + 1];
                            avgP2[b] = avgP2[b] + (ar * br - (-ai) * bi);
                            double cr = d4[2 * c], ci = d4[2 * c + 1];
                            avgP4[b] = avgP4[b] + (ar * cr - (-ai) * ci);
                        }
                    }
                }
            }
        }
    }
    for (int b = 0; b < nbins; ++b) {
        double *p = pout + 5 * (size_t)b;
        p[0] = avgk[b] / co[b];
        p[1] = avgP[b] * kF3 / co[b];
        if (mode == 0) {
            p[2] = co[b]; p[3] = 0.0; p[4] = 0.0;
        } else {
            p[2] = 5.0 * avgP2[b] * kF3 / co[b];
            p[3] = 9.0 * avgP4[b] * kF3 / co[b];
            p[4] = co[b];
        }
    }
    free(avgk); free(avgP); free(avgP2); free(avgP4); free(co);
}

/*
 * Triangle triple sums over the map grid, the inner loop of measure_bk / measure_bk_multipoles /
 * count_triangles (bispectrumI4.f90:164-193, :476-511, :884-908):
 *   out[t] = sum_n a[n] * b[n] * c[n],  a = maps[tri[3t]], b = maps[tri[3t+1]], c = maps[tri[3t+2]]
 * summed sequentially in n, product evaluated (a*b)*c.  Triangles run in parallel (OpenMP).
 */
void orc_triple_sums(int64_t V, const double *maps, int ntri, const int *tri, double *out,
                     int nthreads)
{
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
    for (int t = 0; t < ntri; ++t) {
        const double *a = maps + (int64_t)tri[3 * t] * V;
        const double *b = maps + (int64_t)tri[3 * t + 1] * V;
        const double *c = maps + (int64_t)tri[3 * t + 2] * V;
        double s = 0.0;
        for (int64_t n = 0; n < V; ++n) s = s + a[n] * b[n] * c[n];
        out[t] = s;
    }
}

/*
 * Exact triangle count with modular closure, the integer meaning of count_triangles
 * (bispectrumI4.f90:763-915):  T[i][j][l] = #{q1 in S_i, q2 in S_j : shell(-(q1+q2) mod N) = l},
 * for 1 <= i <= j <= nb and every l.  shell_of_d2[d2] gives the shell (1..nb, 0 = none) of a
 * grid point whose per-axis distances d = min(c, N-c) satisfy dx^2+dy^2+dz^2 = d2.
 * modes: (nm, 3) signed components in (-N/2, N/2], mode_shell: shell of each mode (1..nb),
 * sorted by shell.  Brute force over all pairs; used at small sizes to pin the GPU counter.
 * T is (nb+1)^3 int64, indexed [i][j][l].
 */
void orc_count_pairs_exact(int grid, int nb, int64_t nm, const int *modes, const int *mode_shell,
                           const int *shell_of_d2, int64_t max_d2, int64_t *T, int nthreads)
{
    const long N = grid;
    const long S = nb + 1;
    memset(T, 0, sizeof(int64_t) * (size_t)(S * S * S));
#pragma omp parallel num_threads(nthreads)
    {
        int64_t *loc = calloc((size_t)(S * S * S), sizeof(int64_t));
#pragma omp for schedule(dynamic, 64)
        for (int64_t a = 0; a < nm; ++a) {
            const int i = mode_shell[a];
            for (int64_t b = 0; b < nm; ++b) {
                const int j = mode_shell[b];
                if (j < i) continue;
                long d2 = 0;
                for (int d = 0; d < 3; ++d) {
                    long c = -(long)(modes[3 * a + d] + modes[3 * b + d]);
                    c = ((c % N) + N) % N;
                    long e = c < N - c ? c : N - c;
                    d2 += e * e;
                }
                int l = d2 <= max_d2 ? shell_of_d2[d2] : 0;
                loc[(i * S + j) * S + l] += 1;
            }
        }
#pragma omp critical
        for (long k = 0; k < S * S * S; ++k) T[k] += loc[k];
        free(loc);
    }
}
