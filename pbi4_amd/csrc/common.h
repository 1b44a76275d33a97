// Internal definitions shared by the HIP translation units of libpbi4hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <string>
#include <utility>
#include <vector>

#include "../../include/pbi4_amd.h"

namespace pbi4 {

struct KernelTimer;

// REAL(4), parameter :: twopi = 6.28319 of the reference (powerI4.f90:72 and every routine).
constexpr float kTwoPiF32 = 6.28319f;

struct Error {
    int code;
    std::string msg;
};

// Throwing helpers: every C entry point catches and converts to a status code.
[[noreturn]] void fail(int code, const std::string &msg);

#define PBI4_HIP(call)                                                                     \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            ::pbi4::fail(PBI4_EHIP, std::string(#call) + ": " + hipGetErrorString(e_));   \
    } while (0)

#define PBI4_CHECK(cond, code, msg)                                                        \
    do {                                                                                   \
        if (!(cond)) ::pbi4::fail((code), (msg));                                          \
    } while (0)

__host__ __device__ inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }
inline size_t align_up(size_t a, size_t b) { return (a + b - 1) / b * b; }

// Bump allocator over a caller-provided device workspace.
struct Arena {
    char *base;
    size_t size, used;
    Arena(void *b, size_t s) : base(static_cast<char *>(b)), size(s), used(0) {}
    template <class T>
    T *take(size_t count) {
        size_t off = align_up(used, 256);
        size_t bytes = align_up(count * sizeof(T), 256);
        if (base != nullptr && off + bytes > size)
            fail(PBI4_ENOMEM, "workspace too small: need " + std::to_string(off + bytes) +
                                  " bytes, have " + std::to_string(size));
        used = off + bytes;
        return base ? reinterpret_cast<T *>(base + off) : nullptr;
    }
};

// ---------------------------------------------------------------- assignment (assign.hip)
struct AssignPlan {
    int N, T, NT, R, order, interlacing, ngrids, pos_mode;
    int64_t ntiles;
    float box, rsd[3];
};
AssignPlan make_assign_plan(const pbi4_assign_desc &d, int64_t npart);
size_t assign_workspace(const AssignPlan &p, int64_t npart);
// Sorts, deposits and (grids != NULL) gathers the grids; returns the tile regions, which stay in
// ws for launch_gather_xfft.
const float *assign_run(const AssignPlan &p, const float *pos, const float *vel, int64_t npart,
                        float *grids, Arena &ws, hipStream_t s, KernelTimer *timer);
// Interlaced grids 128, 192, 256, 384, 512: the gather fused with the density FFT's x axis, out[z][y][k]
// = forward DFT over x of grid1 + i grid2 (cells bitwise the gathered ones); tw[k] = e^{2 pi i k / N}.
bool gather_xfft_supported(const AssignPlan &p);
void launch_gather_xfft(const AssignPlan &p, const float *regions, const float2 *tw, float2 *out,
                        hipStream_t s);

// ---------------------------------------------------------------- spectra (spectrum.hip)
// Interlaced: G1 is the full N^3 transform of grid1 + i grid2 and ph[s + N] = e^{-i theta s},
// theta = pi / N, s in [-N, 3N/2]; single grid: G1 is the r2c half spectrum and ph is unused.
void launch_combine(int N, int interlacing, const float *wtab, const float2 *ph, float cf,
                    const float2 *G1, float2 *delta, hipStream_t s);

// Interlaced density with the FFT's z axis fused into the combine: G holds every z plane of
// grid1 + i grid2 transformed over (y, x); tw[k] = e^{2 pi i k / N}.  Grids 128, 192, 256, 384, 512.
bool zfft_combine_supported(int N);
void launch_zfft_combine(int N, const float *wtab, const float2 *ph, const float2 *tw, float cf,
                         const float2 *G, float2 *delta, hipStream_t s);
// In place, every (z, x) column of G[z][y][x] (N^3 complex) replaced by its forward DFT over y;
// grids 128, 192, 256, 384, 512.
void launch_yfft(int N, const float2 *tw, float2 *G, hipStream_t s);

void launch_fcomb_full(int N, int interlacing, const float *wtab, float cf, const float2 *F,
                       float2 *out, hipStream_t s);
void launch_qk(int N, int norder, const int *idx, const float2 *F, float2 *out, hipStream_t s);
void launch_catalogue_positions(const void *pos, int pos_f64, const void *vel, int vel_f64,
                                int64_t n, int rsd, const double *los, double hz, double zf,
                                double box, void *out, int out_f64, hipStream_t s);
void launch_reduce_array(int L, int lx, int N, int nx, int in_bytes, int out_bytes, const void *din,
                         void *dout, hipStream_t s);

struct BinPlan {
    int N;                    // grid
    int nbins_pk;             // P(k) bins (0 = none)
    int nshell;               // bispectrum shells (0 = none)
    int multipoles;
    float los[3];
    float kF;                 // |k| scale for avgk
    const uint16_t *pk_table; // (N/2+1)^3 bin ids [|kz|][|ky|][kx], 0 = excluded
    const float2 *cross;      // second field of a cross spectrum (real space), or null
    const float2 *f2, *f4;    // explicit delta_2, delta_4 (multipoles), or null for L_l(mu)
    const int16_t *shell_table; // d2 -> shell id, 0 = none, size max_d2+1
    int max_d2;
    int nwg;                  // work-groups of the binning kernel
};
size_t bin_partials_floats(const BinPlan &b);
void launch_bin(const BinPlan &b, const float2 *delta, float *partials, hipStream_t s);
// Reduces the per-work-group partials in FP64.  pk_out: (5, nbins) Fortran, shell_out: nshell.
void launch_bin_finalize(const BinPlan &b, const float *partials, double kF3, double *pk_out,
                         double *shell_sum, hipStream_t s);

struct FilterPlan {
    int N, Np;        // delta grid, maps grid
    int R;            // max |component| of shell modes
    int nshell, nell; // shells, 1 or 3
    float los[3];
    const int16_t *shell_table;
    int max_d2;
    const int32_t *shell_d2;  // per shell: smallest and largest d2 (the fused z filter)
    int ld;                   // the fused z filter's output row stride (>= N'/2 + 1)
    const float2 *f2, *f4;    // explicit delta_2, delta_4 fields, or null for L_l(mu) delta
    int cols;                 // fused z filter: columns x' < cols written (0 = all N'/2 + 1)
    int disk_only;            // fused z filter: lines (x', q_y) with x'^2 + q_y^2 above the shell's
                              // largest d2 (zero) are not written; their readers take them as zero
};
void launch_filter(const FilterPlan &f, const float2 *delta, const double *shell_sum,
                   const double *nmodes, float2 *spectra, float *scale_exp, hipStream_t s);
// Band-limited maps grid (3R < N' < N): shell filter fused with the inverse transform over z.
void launch_filter_z(const FilterPlan &f, const float2 *delta, const double *shell_sum,
                     const float2 *tw, float2 *spectra, float *scale_exp, hipStream_t s);

// Inverse 2D transform over (y, x) of the band-limited maps' planes in one pass through LDS:
// per plane, the half spectrum's columns x' < C (the rest is zero) go through inverse DFTs along
// y, then each row's c2r as an inverse complex DFT of length N'/2 (the even / odd split of the
// real output), written as real [y][x]; four-step transforms with compile-time factors for the
// maps grids 100, 150 and 200.  Unnormalised, as a c2r of the same half spectrum.
struct C2r2dPlan {
    int Np, C, ld, yc;        // N', columns carried, input row stride, rows per LDS chunk
    int64_t in_plane;         // complex entries per input plane (N' ld)
    int64_t Kstride;          // floats per output map
    size_t lds;               // dynamic LDS bytes
    int nb;                   // shells: map m holds shell m mod nb
    const int32_t *shell_d2;  // per shell: smallest and largest d2; lines (x', q_y) with
                              // x'^2 + q_y^2 above the largest are zero and are not read
};
// False when N' has no compiled transform or the planes do not fit LDS.
bool make_c2r2d_plan(int Np, int C, int ld, int64_t Kstride, C2r2dPlan &p);
void launch_c2r2d(const C2r2dPlan &p, const float2 *spectra, const float2 *tw, float *maps,
                  int nmaps, hipStream_t s);
// Full-grid maps (N' = N in 128, 192, 256, 384): after the fused z filter, the inverse DFTs along
// y of the columns x' < C in place (row stride ld), then each row's c2r into the maps.
bool maps_two_pass_supported(int Np);
void launch_maps_two_pass(int Np, int C, int ld, int nmaps, int nb, const int32_t *shell_d2,
                          int64_t Kstride, const float2 *tw, float2 *spectra, float *maps,
                          hipStream_t s);

// ---------------------------------------------------------------- triangles (triangle.hip)
struct GemmPlan {
    int nb, nell, ncol, npair;
    int64_t K, Kstride; // map length and per-map stride (floats)
    int precision;
    int cf, rf;         // column fragments (of 16), row fragments per wave (of 16)
    int kchunk;         // x-values per work-group
    int nkchunks, nkchunks_pad, nrowblocks;
    // Block-sparse mode, set by gemm_set_windows from a triangle list: logical column c is map
    // moff + (c % mdiv) * nb + c / mdiv, and wave w (16 rf pair rows) multiplies only the
    // logical column fragments [wstart[w], wstart[w] + wwidth[w]) that its triangles read.  C is
    // then npair x ncol_l in logical columns and holds only those entries.
    bool win = false;
    int cfw = 0;        // widest window, rounded up to a compiled width
    int maxw = 0;       // widest window (column fragments): partial rows are 16 maxw long
    int ucap = 0;       // most column fragments one work-group stages (union of its 4 waves)
    // launches, one per compiled window width: row blocks rbs[cls_off[k] .. cls_off[k + 1]) on
    // the cls_w[k]-tile body, staging at most cls_ucap[k] column fragments
    std::vector<uint8_t> rbs;
    std::vector<int> cls_w, cls_off, cls_ucap;
    int moff = 0, mdiv = 1, ncol_l = 0;
    // rows: 0 = pairs (i <= j) at j (j + 1) / 2 + i; 1 = real-space rows (j, l) of the triangle
    // loop, ordered by l then j from max(1, ceil((l + 1 - ish) / 2)), columns = the leg i
    int rowmode = 0, ish = 0;
    std::vector<uint8_t> wstart, wwidth;
    // finalize indexing: auto triangle (l, ell) reads C column (l - 1) * cs_l + ell * cs_e; cross
    // triangle i reads column coff + i - 1; cld is the C row length
    int cs_l = 1, cs_e = 0, coff = 0, cld = 0;
};
// nrows > 0: that many GEMM rows instead of the nb (nb + 1) / 2 pairs (block-sparse row forms)
GemmPlan make_gemm_plan(int nb, int nell, int64_t K, int64_t Kstride, int precision,
                        int nrows = 0);
// Switch plan g to block-sparse mode for the triangles in tri ((l, j, i) triples, 1-based, as
// triangle_list writes them, ishift = int(kcenter / kbin) + iopen); cross: the cross-bispectrum
// layout (pair rows (j <= l) of field 1, field-2 map i); real space (nell = 1) uses the (j, l)
// rows.  Keeps the dense kernel where the windows do not fit the compiled tiles.
void gemm_set_windows(GemmPlan &g, const std::vector<int32_t> &tri, bool cross, int ishift);
size_t gemm_partials_floats(const GemmPlan &g);
void launch_gemm(const GemmPlan &g, const float *maps, float *partials, hipStream_t s,
                 KernelTimer *timer);
void launch_gemm_reduce(const GemmPlan &g, const float *partials, double *C, hipStream_t s);

void launch_bk_finalize(int rows, int64_t itot, const int32_t *tri, const double *C,
                        const GemmPlan &g, const float *scale_exp, const double *shell_sum,
                        const double *nmodes, const float *counts, double kf3, double powSN,
                        double gridD, double sum_scale, double kcenter, double kbin, double *bout,
                        hipStream_t s);

void launch_bk_cross_finalize(int64_t itot, const int32_t *tri, const double *C, const GemmPlan &g,
                              const float *scale_exp, const float *counts, double kf3,
                              double gridD, double sum_scale, double kcenter, double kbin,
                              double *bout, hipStream_t s);

void count_pairs_gpu(int N, int nb, const int32_t *modes_d, int64_t nmodes, const int64_t *shell_off,
                     const int32_t *canon_d, const int32_t *canon_w, int64_t ncanon,
                     const int16_t *shell_table_d, int max_d2, unsigned long long *T_d,
                     hipStream_t s);

// ---------------------------------------------------------------- timing
// HIP events around one named kernel on the stream it is launched on (benchmark support).
struct KernelTimer {
    std::string name;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
};
struct TimedScope {
    KernelTimer *t;
    hipStream_t s;
    hipEvent_t b = nullptr;
    TimedScope(KernelTimer *t_, const char *k, hipStream_t s_);
    ~TimedScope();
};

void launch_fill_u32(uint32_t *p, uint32_t v, int64_t n, hipStream_t s);

}  // namespace pbi4
