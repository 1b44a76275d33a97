/*
 * pbi4_amd -- MI355X (gfx950) implementation of the PBI4 measurement path.
 *
 * C ABI of libpbi4hip.so.  Plain pointers and sizes only: device buffers are caller-owned
 * (any allocator: torch tensors, hipMalloc), host outputs are caller-allocated.  Every entry
 * point returns 0 on success or a negative PBI4_E* code; pbi4_last_error() gives the message.
 * Calls that take a stream are asynchronous on that stream (hipStream_t passed as void*;
 * NULL = the null stream) unless stated otherwise.
 *
 * Grid layout: real grids are [z][y][x] with x contiguous (the memory order of the reference's
 * Fortran arrays dcl(ix,iy,iz)); half spectra are [z][y][x'] complex float2 with
 * x' = 0..N/2 (rocFFT r2c layout).  Frequencies of index i along y or z are i (i <= N/2) or
 * i - N (i > N/2), as in the reference (powerI4.f90:549-560).
 */
#ifndef PBI4_AMD_H
#define PBI4_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PBI4_OK 0
#define PBI4_EINVAL (-1)       /* bad argument (the reference's SystemExit / stop cases)   */
#define PBI4_EHIP (-2)         /* HIP runtime error                                        */
#define PBI4_EFFT (-3)         /* rocFFT error                                             */
#define PBI4_ENOMEM (-4)       /* workspace too small or allocation failed                 */
#define PBI4_EUNSUPPORTED (-5) /* configuration this build does not implement              */

#define PBI4_PREC_FP32 0   /* triangle sums on FP32 MFMA (the plain path)                  */
#define PBI4_PREC_BF16X3 1 /* split-BF16x3 MFMA with FP32 accumulation                      */

#define PBI4_POS_NORMALISED 0 /* positions already in [0,1) (output of pbi4.py:241-243)      */
#define PBI4_POS_BOX 1        /* positions in box units: RSD shift, wrap, /box in the kernel */

typedef struct pbi4_ctx pbi4_ctx;

/* Context: owns rocFFT plans and the integer tables (P(k) bin table, shell table, triangle
 * lists, exact triangle counts) cached per configuration.  One context per host thread.
 * Replaces the module state of the f2py extensions imported at pbi4/pbi4.py:12-14. */
int pbi4_ctx_create(int device, pbi4_ctx **out);
int pbi4_ctx_destroy(pbi4_ctx *ctx);
/* ctx = NULL: the message of the calling thread's last failed call made without a context
 * (pbi4_assign_workspace, pbi4_triangle_sums_workspace return 0 when they refuse a shape). */
const char *pbi4_last_error(const pbi4_ctx *ctx);
int pbi4_version(void);

/* ------------------------------------------------------------------ mass assignment
 * Replaces powerI4.assign_double_grid (powerI4.f90:142-249, interlacing = 1) and
 * powerI4.assign_single_grid (powerI4.f90:68-140, interlacing = 0), called from
 * DENSITIES.assign_grid (pbi4/pbi4.py:290-302), with the catalogue normalisation and the
 * redshift-space shift of IO.read_catalogues (pbi4/pbi4.py:151-154, 241-243) fused in when
 * pos_mode == PBI4_POS_BOX:  s_d = pos_d + rsd[d]*vel_d;  s_d wrapped into [0, box);  s_d/box.
 * order: 1 NGP, 2 CIC, 3 TSC, 4 PCS.  Output grids[g][z][y][x] (g = 0, and g = 1 for the
 * half-cell-shifted interlacing grid) hold the raw sum of weights (weight = 1 per particle).
 * Accumulation is in 64-bit fixed point inside each tile, so the grids are deterministic.
 */
typedef struct {
    int grid;          /* N                                                               */
    int order;         /* nintp, 1..4                                                     */
    int interlacing;   /* 0 or 1                                                          */
    int pos_mode;      /* PBI4_POS_NORMALISED or PBI4_POS_BOX                             */
    float box;         /* box size (PBI4_POS_BOX only)                                    */
    float rsd[3];      /* per-axis velocity coefficient los_d*(1+z)/H(z); ignored if vel NULL */
} pbi4_assign_desc;

size_t pbi4_assign_workspace(const pbi4_assign_desc *d, int64_t npart);
int pbi4_assign(pbi4_ctx *ctx, const pbi4_assign_desc *d, const float *pos, const float *vel,
                int64_t npart, float *grids, void *ws, size_t ws_bytes, void *stream);

/* ------------------------------------------------------------------ density
 * Replaces DENSITIES.compute_density (pbi4/pbi4.py:304-341): dcl /= npart, FFT, and
 * powerI4.fcomb (powerI4.f90:251-378): window deconvolution sinc^p and, with interlacing, the
 * combination of the two grids with the ordered-store semantics of powerI4.f90:335-342.
 * grids: output of pbi4_assign; delta: half spectrum [N][N][N/2+1] float2 (the ix <= N/2+1
 * half of the reference's dcl, which is the only part any later routine reads).
 */
typedef struct {
    int grid, order, interlacing;
    float box;
    double npart;
} pbi4_density_desc;

size_t pbi4_density_workspace(pbi4_ctx *ctx, const pbi4_density_desc *d);
int pbi4_compute_density(pbi4_ctx *ctx, const pbi4_density_desc *d, const float *grids,
                         void *delta, void *ws, size_t ws_bytes, void *stream);

/* pbi4_assign followed by pbi4_compute_density in one call: the two steps the reference runs
 * back to back (pbi4/pbi4.py:304-341, DENSITIES.assign_* then compute_density).  On interlaced
 * grids 128, 192, 256, 384 and 512 the gather of the tile regions is fused with the FFT's x axis
 * and the FFT runs in grids (no rocFFT pass); elsewhere, or with PBI4_DENSITY_FFT set, it is
 * exactly pbi4_assign + pbi4_compute_density.  grids (2 N^3 floats, or N^3 single grid) is
 * scratch here and holds no defined values afterwards; ws must hold max(pbi4_assign_workspace,
 * pbi4_density_workspace) bytes (a ws that is not 16-byte aligned takes the two-step route).
 * ad and dd must name the same grid, order and interlacing.
 */
int pbi4_assign_density(pbi4_ctx *ctx, const pbi4_assign_desc *ad, const pbi4_density_desc *dd,
                        const float *pos, const float *vel, int64_t npart, float *grids,
                        void *delta, void *ws, size_t ws_bytes, void *stream);

/* Full-array forms of the reference routines for callers that hold complete N^3 arrays (the
 * f2py-named Python functions).  Arrays are complex float2 [z][y][x] (Fortran (x,y,z) order).
 * pbi4_fcomb: powerI4.fcomb (powerI4.f90:251-378) on fdcl = FFT(grid1 + i grid2), all N^3
 * outputs written with the reference's ordered-store semantics (dcl /= npart is the caller's).
 * pbi4_compute_qk: powerI4.compute_q2k (norder 2, idx[0..1]) and compute_q4k (norder 4,
 * idx[0..3]) (powerI4.f90:409-508); idx are 1-based axes as in the reference.
 */
int pbi4_fcomb(pbi4_ctx *ctx, int grid, int nintp, int interlacing, float box, const void *fdcl,
               void *dcl, void *stream);
int pbi4_compute_qk(pbi4_ctx *ctx, int grid, int norder, const int *idx, const void *deltak,
                    void *out, void *stream);

/* Catalogue front-end: the redshift-space shift, periodic wrap and normalisation of
 * IO.read_catalogues (pbi4/pbi4.py:151-154, 162-163, 179-180, 213-216 and 241-243) on positions
 * a host reader produced, with numpy's exact arithmetic.
 * pos: device (n, 3) rows of float (pos_f64 = 0) or double (1), box units (Mpc/h).
 * vel: device (n, 3) rows (vel_f64 as for pos), or NULL with rsd = PBI4_RSD_NONE.
 * rsd: PBI4_RSD_NONE; PBI4_RSD_HZ: s = p + ((v los_d) / hz) (1 + z) (Gadget, Rockstar and
 *      Quijote readers, zf = 1 + z); PBI4_RSD_VEL: s = p + v los_d (the Minerva reader);
 *      the shift is FP64 and s is stored in the output type.
 * out: device (3, n) of float (out_f64 = 0) or double (1): min(1, ((s + box) % box) / box) in the
 *      output type, numpy's remainder and correctly rounded division. */
#define PBI4_RSD_NONE 0
#define PBI4_RSD_HZ 1
#define PBI4_RSD_VEL 2
int pbi4_catalogue_positions(pbi4_ctx *ctx, const void *pos, int pos_f64, const void *vel, int vel_f64,
                             int64_t n, int rsd, const double los[3], double hz, double zf, double box,
                             void *out, int out_f64, void *stream);

/* Lowest-frequency sub-array of a density stored at a larger grid: powerI4.reduce_array
 * (powerI4.f90:380-407), as DENSITIES.read_density uses it (pbi4/pbi4.py:257-288) to bring a
 * density saved at grid L to the requested grid N <= L.
 * din: device [z][y][x] array of L x L x lx complex elements, lx >= N/2+1 (only x <= N/2 is read).
 * dout: device [z][y][x] array of N x N x nx elements, N/2+1 <= nx <= N.
 * in_bytes / out_bytes: 8 = complex float, 16 = complex double (16 -> 8 rounds to nearest).
 * Output (x, y, z) with x <= N/2 copies source (x, m(y), m(z)), m(i) = i for i < N - N/2 and
 * L - N + i above: the reference's store order, in which the negative-frequency source stays on
 * the Nyquist index of an even N.  Output x > N/2 is zero (the reference's zero-filled
 * intent(out) array). */
int pbi4_reduce_array(pbi4_ctx *ctx, int L, int lx, int N, int nx, int in_bytes, int out_bytes,
                      const void *din, void *dout, void *stream);

/* ------------------------------------------------------------------ power spectrum
 * Replaces powerI4.measure_pk (powerI4.f90:510-594, multipoles = 0) and
 * powerI4.measure_pk_multipoles (powerI4.f90:596-693, multipoles = 1) together with
 * DENSITIES.compute_density_multipoles (pbi4/pbi4.py:343-388): for a fixed line of sight the
 * reference's delta_2, delta_4 are exactly L_2(mu) delta and L_4(mu) delta with
 * mu^2 = (l.k)^2/k^2, l = los/|los|^2 (pbi4/pbi4.py:103), so they are applied per mode.
 * Mode-to-bin assignment uses a table built once per (N, box, kbin, kcenter) with the
 * reference's FP64 formula (powerI4.f90:565-566), so mode counts are bit-exact.
 * pout: device double, Fortran order (5, nbins): rows avgk, P0(+PSN), P2, P4, Nmodes
 * (multipoles) or avgk, P, Nmodes, 0, 0 (real space), as the reference's pout.
 * delta_b: real space only, the second field of measure_pk(dcl1, dcl2) (cross spectrum,
 * pbi4/pbi4.py:445-454); NULL = delta.  delta2, delta4: multipoles only, explicit delta_2 and
 * delta_4 half spectra as measure_pk_multipoles receives them; NULL = L_l(mu) delta.
 */
typedef struct {
    int grid;
    float box, kbin, kcenter;
    int nbins;
    int multipoles;
    double los[3];
} pbi4_pk_desc;

size_t pbi4_pk_workspace(pbi4_ctx *ctx, const pbi4_pk_desc *d);
int pbi4_measure_pk(pbi4_ctx *ctx, const pbi4_pk_desc *d, const void *delta, const void *delta_b,
                    const void *delta2, const void *delta4, double *pout, void *ws,
                    size_t ws_bytes, void *stream);

/* ------------------------------------------------------------------ triangle counts
 * Replaces bispectrumI4.count_triangles (bispectrumI4.f90:763-915), called from
 * MEASUREMENTS.triangle_counts (pbi4/pbi4.py:463-476).  Exact integer count with the
 * reference's modular closure: ntr[m] = #{q1 in S_i, q2 in S_j, q3 in S_l : q1+q2+q3 = 0 mod N}
 * for the m-th triangle (l, j, i) of the reference loop order; the reference's counts(m) is
 * ntr[m] * N^3.  Synchronous (runs on the GPU, waits, caches the result in the context).
 */
int pbi4_count_triangles(pbi4_ctx *ctx, int nbins, int grid, float kbin, float kcenter,
                         int iopen, int cross, int64_t itot, int64_t *ntr);

/* ------------------------------------------------------------------ bispectrum
 * Replaces bispectrumI4.measure_bk (bispectrumI4.f90:1-198, multipoles = 0) and
 * bispectrumI4.measure_bk_multipoles (bispectrumI4.f90:200-516, multipoles = 1), called from
 * MEASUREMENTS.bispectrum (pbi4/pbi4.py:478-553).
 * Shell maps are the c2r transforms of the shell-filtered spectrum (ell = 0, and the L_2, L_4
 * weighted spectra in redshift space); the triangle sums sum_x M_i M_j M_l are one
 * MFMA GEMM C[(i<=j),(ell,l)] = sum_x (M_i M_j)(x) M_{ell,l}(x) with split-K partials
 * combined in FP64.  maps_grid = 0 picks the smallest FFT size N' > 3*max|q| of the shells
 * (N' = N when the reference's mod-N closure aliases), on which the band-limited triple sums
 * equal the reference's N^3 sums times (N'/N)^3 exactly; maps_grid = N forces the full grid.
 * counts: device float, the reference's REAL(4) counts(itot).  pow(s) uses Parseval.
 * delta2, delta4: multipoles only, explicit delta_2 / delta_4 half spectra as
 * measure_bk_multipoles receives them; NULL = L_l(mu) delta (the fused path).
 * bout: device double, Fortran order (9 or 11, itot) as the reference's bout.
 */
typedef struct {
    int nbins, grid;
    float kbin, kcenter;
    int iopen;
    float kf;
    int64_t itot, npart;
    int multipoles;
    double los[3];
    int precision;     /* PBI4_PREC_FP32 or PBI4_PREC_BF16X3 */
    int maps_grid;     /* 0 = automatic band-limited grid, else the grid to build maps on */
} pbi4_bk_desc;

size_t pbi4_bk_workspace(pbi4_ctx *ctx, const pbi4_bk_desc *d);
int pbi4_bk_maps_grid(pbi4_ctx *ctx, const pbi4_bk_desc *d);
int pbi4_measure_bk(pbi4_ctx *ctx, const pbi4_bk_desc *d, const float *counts, const void *delta,
                    const void *delta2, const void *delta4, double *bout, void *ws,
                    size_t ws_bytes, void *stream);

/* ------------------------------------------------------------------ cross bispectrum
 * Replaces bispectrumI4.measure_bk_cross (bispectrumI4.f90:518-761), called from
 * MEASUREMENTS.bispectrum when params['cross'] is set (pbi4/pbi4.py:481-496).
 * B(k_l, k_j, k_i) = kf^3 sum_x B_i(x) A_j(x) A_l(x) / N_tr with A the shell maps of delta1 and
 * B those of delta2, over the cross loop's triangles (i up to min(nbins, j + l - 1 + ishift),
 * :733-735).  Same shells, maps grid, precision and GEMM as pbi4_measure_bk, which runs with
 * the nb field-1 maps as pair rows and the 2 nb maps (field 1, then field 2) as columns.
 * d->multipoles must be 0; d->itot is the cross loop's count; counts from
 * pbi4_count_triangles with cross = 1.  bout: device double, Fortran order (5, itot):
 * k_l, k_j, k_i, B, N_tr / N^3 (REAL(4)-rounded), no shot noise (:751-755).
 */
size_t pbi4_bk_cross_workspace(pbi4_ctx *ctx, const pbi4_bk_desc *d);
int pbi4_measure_bk_cross(pbi4_ctx *ctx, const pbi4_bk_desc *d, const float *counts,
                          const void *delta1, const void *delta2, double *bout, void *ws,
                          size_t ws_bytes, void *stream);

/* ------------------------------------------------------------------ fused P_l + B_l
 * pbi4_measure_pk followed by pbi4_measure_bk on the same delta, with one binning pass that
 * produces the P_l bins and the shell powers together (the serving path).  Same outputs as
 * the two calls.  The descriptors must agree on grid and multipoles; L_l(mu) is applied.
 */
size_t pbi4_pk_bk_workspace(pbi4_ctx *ctx, const pbi4_pk_desc *pd, const pbi4_bk_desc *d);
int pbi4_measure_pk_bk(pbi4_ctx *ctx, const pbi4_pk_desc *pd, const pbi4_bk_desc *d,
                       const float *counts, const void *delta, double *pout, double *bout,
                       void *ws, size_t ws_bytes, void *stream);

/* ------------------------------------------------------------------ triangle sums
 * The triple-sum loops of bispectrumI4.f90:164-193 (measure_bk), :476-511
 * (measure_bk_multipoles) and :884-908 (count_triangles) over caller-provided real maps:
 * C[p][c] = sum_{x < K} M_i(x) M_j(x) B_c(x) for every pair p = j(j+1)/2 + i (0 <= i <= j < nb)
 * and column c < nb*nell, where M_s = maps[s] (s < nb) and B_c = maps[c].
 * maps: device float [nb*nell][Kstride]; C: device double [nb(nb+1)/2][nb*nell].
 * K and Kstride must be multiples of 4.
 */
size_t pbi4_triangle_sums_workspace(int nb, int nell, int64_t K, int64_t Kstride, int precision);
int pbi4_triangle_sums(pbi4_ctx *ctx, int nb, int nell, int64_t K, int64_t Kstride, int precision,
                       const float *maps, double *C, void *ws, size_t ws_bytes, void *stream);

/* ------------------------------------------------------------------ triangle sums, no allocation
 * No counterpart in the reference.  pbi4_triangle_sums for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc):
 * same arguments without the workspace, same plan and kernels, so C is bitwise the same as
 * pbi4_triangle_sums gives.  The split-K partials live in one static device array of
 * PBI4_NOALLOC_SCRATCH_BYTES per GPU, held from library load until the process ends; shapes
 * whose partials need more (pbi4_triangle_sums_workspace(...) - 256 bytes above the limit) are
 * rejected with PBI4_ENOMEM.  Every call on a GPU shares that GPU's array: calls that overlap in
 * time on one GPU (two streams, two threads, two contexts) race and must be serialised by the
 * caller.  The context must have been created on the device (pbi4_ctx_create records the
 * device's array address).  Nothing is allocated or freed by this call.
 */
#define PBI4_NOALLOC_SCRATCH_BYTES (64u << 20)
int pbi4_triangle_sums_noalloc(pbi4_ctx *ctx, int nb, int nell, int64_t K, int64_t Kstride,
                               int precision, const float *maps, double *C, void *stream);

/* ------------------------------------------------------------------ introspection
 * Kernel timing support for the benchmark: records HIP events around the named kernel
 * (0 = none) on every launch and returns the accumulated milliseconds and launch count.
 * No counterpart in the reference.
 */
int pbi4_timing_enable(pbi4_ctx *ctx, const char *kernel_name);
int pbi4_timing_read(pbi4_ctx *ctx, double *total_ms, int64_t *launches);

#ifdef __cplusplus
}
#endif
#endif
