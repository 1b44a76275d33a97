// Bispectrum triangle sums as one MFMA GEMM, FP64 finalisation, exact triangle counts.
//
// C[p][c] = sum_x A[p][x] B[x][c] with A[(i<=j)][x] = M_i(x) M_j(x) (formed on the fly from the
// ell = 0 shell maps staged in LDS) and B[x][(ell,l)] = M_{ell,l}(x) -- the triple sums
// sum_x M_i M_j M_{ell,l} of bispectrumI4.f90:164-193 and :476-511 for every (i<=j, ell, l).
// Split-K: each work-group owns 64 RF pair-rows x all columns x one K chunk and writes an FP32
// partial; the partials are summed over chunks in FP64 in a fixed order.
//   PBI4_PREC_FP32  : v_mfma_f32_16x16x4_f32 (the plain path).
//   PBI4_PREC_BF16X3: a = a_hi + a_lo, b = b_hi + b_lo in bf16 (round-to-nearest), and
//                     a b ~ a_hi b_hi + a_hi b_lo + a_lo b_hi on v_mfma_f32_16x16x32_bf16 with
//                     FP32 accumulation (relative representation error ~2^-17 per product).
#include "common.h"

namespace pbi4 {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kGemmThreads = 256;     // 4 waves
constexpr int kMaxCF = 12;            // column fragments (ncol <= 192)
constexpr int kXcds = 8;              // MI355X: work-groups are dealt round-robin to 8 XCDs
constexpr int kKT = 32;               // x-values per LDS stage: 8 quads (128 B) per map row
constexpr int kGemmLds = 80 * 1024;   // LDS ring per work-group (two work-groups per CU)

typedef __attribute__((address_space(3))) void lds_void_t;
typedef __attribute__((address_space(1))) void gbl_void_t;

__device__ __forceinline__ void pair_of(int p, int &i, int &j) {
    // p = j(j+1)/2 + i with 0 <= i <= j
    int jj = (int)((sqrtf(8.0f * (float)p + 1.0f) - 1.0f) * 0.5f);
    while ((jj + 1) * (jj + 2) / 2 <= p) ++jj;
    while (jj * (jj + 1) / 2 > p) --jj;
    j = jj;
    i = p - jj * (jj + 1) / 2;
}

// Ring geometry: each stage holds NRP map rows x kKT floats; a wave's load instruction fills
// 8 rows (64 lanes x 16 B), so NRP is the column count rounded up to 32 (4 waves).
template <int CF>
struct GemmRing {
    static constexpr int NRP = (CF * 16 + 31) / 32 * 32;
    static constexpr int LPW = NRP / 32;                        // load instructions per wave
    static constexpr int STAGE = NRP * kKT;                     // floats
    static constexpr int S0 = kGemmLds / (STAGE * 4);
    static constexpr int S = S0 > 8 ? 8 : S0;                   // stages in the ring
};

// Wait until at most n of this wave's stage loads are outstanding, then barrier.  The loads
// write LDS directly (no registers), so the compiler does not track them: the counts are ours,
// and the memory clobber keeps LDS reads below the barrier.
template <int LPW>
__device__ __forceinline__ void wait_stages_barrier(int n) {
#define PBI4_WAIT(N)                                                         \
    case N:                                                                  \
        if constexpr (N * LPW <= 63)                                         \
            asm volatile("s_waitcnt vmcnt(%0)\n\ts_barrier" ::"n"(N * LPW) : "memory"); \
        break;
    switch (n) {
        PBI4_WAIT(0)
        PBI4_WAIT(1)
        PBI4_WAIT(2)
        PBI4_WAIT(3)
        PBI4_WAIT(4)
        PBI4_WAIT(5)
        PBI4_WAIT(6)
        default: asm volatile("s_waitcnt vmcnt(0)\n\ts_barrier" ::: "memory"); break;
    }
#undef PBI4_WAIT
}

// Quad q (4 floats) of ring row r is stored at quad position q ^ (r & 7): the 16 rows one MFMA
// fragment reads at one quad fall in distinct LDS banks.
__device__ __forceinline__ float4 ring_quad(const float *stage, int r, int q) {
    return *reinterpret_cast<const float4 *>(stage + r * kKT + 4 * (q ^ (r & 7)));
}

__device__ __forceinline__ void split8(const float (&v)[8], bf16x8 &hi, bf16x8 &lo) {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const __bf16 h = (__bf16)v[e];
        hi[e] = h;
        lo[e] = (__bf16)(v[e] - (float)h);
    }
}

// Work-group tile: 4 waves x RF row fragments x 16 pair-rows, all CF column fragments, one K
// chunk.  The chunk streams through an S-stage LDS ring filled by direct global->LDS loads
// (S-1 stages in flight, one barrier per stage); A = M_i M_j is formed from the ring's ell = 0
// rows and split into bf16 hi/lo in registers, as is B.  x beyond the chunk end loads clamped
// (finite) data and its A products are zeroed.  The linear work-group id is mapped so that all
// row blocks of one K chunk land on the same XCD (ids = chunk mod 8): the chunk's maps come from
// HBM once and from that XCD's L2 for the other row blocks.
template <int PREC, int CF, int RF>
__global__ __launch_bounds__(kGemmThreads) void k_triangle_gemm(const float *maps, int64_t K, int64_t Kstride, int nb, int ncol, int npair,
                     int kchunk, int nrowblocks, int nchunks, float *partials) {
    using G = GemmRing<CF>;
    constexpr int S = G::S, LPW = G::LPW, STAGE = G::STAGE;
    constexpr int NCP = CF * 16;
    extern __shared__ __align__(16) float ring[];

    const int w = blockIdx.x;
    const int qd = w / kXcds;
    const int rowblock = qd % nrowblocks;
    const int chunk = (w % kXcds) + kXcds * (qd / nrowblocks);
    if (chunk >= nchunks) return;

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, grp = lane >> 4;
    const int64_t kbeg = (int64_t)chunk * kchunk;
    const int64_t kend = min(K, kbeg + kchunk);
    const int nst = (int)((kend - kbeg + kKT - 1) / kKT);

    // loader: instruction i = wave + 4 j fills ring rows 8 i .. 8 i + 7; lane L writes row
    // 8 i + L / 8 = row0 + 32 j at quad position L % 8, i.e. quad (L % 8) ^ (row0 & 7) of that
    // map row (rows past the last map re-read it; the ring never reads them)
    const int row0 = 8 * wave + lane / 8;
    const int lq = (lane % 8) ^ (row0 & 7);
    auto issue = [&](int st) {
        const int64_t x = min(kbeg + (int64_t)st * kKT + 4 * lq, K - 4);
        float *dst = ring + (st % S) * STAGE;
#pragma unroll
        for (int j = 0; j < LPW; ++j) {
            const float *src = maps + (int64_t)min(row0 + 32 * j, ncol - 1) * Kstride + x;
            __builtin_amdgcn_global_load_lds((gbl_void_t *)src,
                                             (lds_void_t *)(dst + (wave + 4 * j) * 256), 16, 0, 0);
        }
    };

    int pi[RF], pj[RF];
#pragma unroll
    for (int r = 0; r < RF; ++r) {
        const int prow = (rowblock * 4 + wave) * RF * 16 + r * 16 + (lane & 15);
        pi[r] = pj[r] = 0;
        if (prow < npair) pair_of(prow, pi[r], pj[r]);
    }
    f32x4 acc[RF][CF];
#pragma unroll
    for (int r = 0; r < RF; ++r)
#pragma unroll
        for (int c = 0; c < CF; ++c) acc[r][c] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};

    auto compute = [&](int st) {
        const float *b = ring + (st % S) * STAGE;
        // x-values of this stage still inside the chunk, relative to the stage start
        const int64_t left = kend - (kbeg + (int64_t)st * kKT);
        if (PREC == PBI4_PREC_BF16X3) {
            // lane: k = 8 grp .. 8 grp + 7 = quads 2 grp, 2 grp + 1
            const int q0 = 2 * grp;
            bf16x8 ahi[RF], alo[RF];
#pragma unroll
            for (int r = 0; r < RF; ++r) {
                const float4 a0 = ring_quad(b, pi[r], q0), a1 = ring_quad(b, pi[r], q0 + 1);
                const float4 c0 = ring_quad(b, pj[r], q0), c1 = ring_quad(b, pj[r], q0 + 1);
                float prod[8] = {a0.x * c0.x, a0.y * c0.y, a0.z * c0.z, a0.w * c0.w,
                                 a1.x * c1.x, a1.y * c1.y, a1.z * c1.z, a1.w * c1.w};
                if (left < kKT) {
#pragma unroll
                    for (int e = 0; e < 8; ++e)
                        if (8 * grp + e >= left) prod[e] = 0.0f;
                }
                split8(prod, ahi[r], alo[r]);
            }
#pragma unroll
            for (int c = 0; c < CF; ++c) {
                const int col = c * 16 + (lane & 15);
                const float4 b0 = ring_quad(b, col, q0), b1 = ring_quad(b, col, q0 + 1);
                const float bv[8] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w};
                bf16x8 bh, bl;
                split8(bv, bh, bl);
#pragma unroll
                for (int r = 0; r < RF; ++r) {
                    acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi[r], bh, acc[r][c], 0, 0, 0);
                    acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi[r], bl, acc[r][c], 0, 0, 0);
                    acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo[r], bh, acc[r][c], 0, 0, 0);
                }
            }
        } else {
            // two k-steps of 16: lane supplies k = 4 grp + e of each MFMA e
#pragma unroll
            for (int ks = 0; ks < 2; ++ks) {
                const int q = 4 * ks + grp;
                float pr[RF][4];
#pragma unroll
                for (int r = 0; r < RF; ++r) {
                    const float4 a = ring_quad(b, pi[r], q), c4 = ring_quad(b, pj[r], q);
                    pr[r][0] = a.x * c4.x;
                    pr[r][1] = a.y * c4.y;
                    pr[r][2] = a.z * c4.z;
                    pr[r][3] = a.w * c4.w;
                    if (left < kKT) {
#pragma unroll
                        for (int e = 0; e < 4; ++e)
                            if (4 * q + e >= left) pr[r][e] = 0.0f;
                    }
                }
#pragma unroll
                for (int c = 0; c < CF; ++c) {
                    const int col = c * 16 + (lane & 15);
                    const float4 bv = ring_quad(b, col, q);
#pragma unroll
                    for (int r = 0; r < RF; ++r) {
                        acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(pr[r][0], bv.x, acc[r][c], 0, 0, 0);
                        acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(pr[r][1], bv.y, acc[r][c], 0, 0, 0);
                        acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(pr[r][2], bv.z, acc[r][c], 0, 0, 0);
                        acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(pr[r][3], bv.w, acc[r][c], 0, 0, 0);
                    }
                }
            }
        }
    };

    for (int t = 0; t < S - 1; ++t)
        if (t < nst) issue(t);
    for (int st = 0; st < nst; ++st) {
        // stages issued after st: st+1 .. min(nst-1, st+S-2); they may stay in flight
        wait_stages_barrier<LPW>(min(S - 2, nst - 1 - st));
        // every wave is past compute(st - 1): its buffer is free for stage st + S - 1
        if (st + S - 1 < nst) issue(st + S - 1);
        compute(st);
    }
    // D layout of the 16x16 MFMAs: lane holds rows 4*(lane/16) + i, column lane%16
    const int64_t npair_pad = (int64_t)nrowblocks * 4 * RF * 16;
    float *out = partials + (int64_t)chunk * npair_pad * NCP;
#pragma unroll
    for (int r = 0; r < RF; ++r)
#pragma unroll
        for (int c = 0; c < CF; ++c)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int64_t row = ((int64_t)rowblock * 4 + wave) * RF * 16 + r * 16 + 4 * grp + i;
                out[row * NCP + c * 16 + (lane & 15)] = acc[r][c][i];
            }
}

// Split-K combine in FP64: each block owns 32 consecutive C entries; the chunk range of every
// entry is split over 8 threads in fixed sub-ranges, then summed in a fixed order in LDS.
__global__ __launch_bounds__(256) void k_gemm_reduce(const float *partials, int nchunks,
                                                     int64_t npair_pad, int ncol_pad, int npair,
                                                     int ncol, double *C) {
    __shared__ double red[8][32];
    const int lo = threadIdx.x & 31, part = threadIdx.x >> 5;
    const int64_t t = (int64_t)blockIdx.x * 32 + lo;
    const bool ok = t < (int64_t)npair * ncol;
    const int64_t p = ok ? t / ncol : 0, c = ok ? t % ncol : 0;
    const int per = (nchunks + 7) / 8;
    const int k0 = part * per, k1 = min(nchunks, k0 + per);
    double s = 0.0;
    for (int k = k0; k < k1; ++k) s += (double)partials[((int64_t)k * npair_pad + p) * ncol_pad + c];
    red[part][lo] = s;
    __syncthreads();
    if (part == 0 && ok) {
        double r = red[0][lo];
        for (int q = 1; q < 8; ++q) r += red[q][lo];
        C[t] = r;
    }
}

// One thread per triangle m (reference order): bout(:, m) of bispectrumI4.f90:173-182 / :489-500
// with pow(s) from Parseval and the triple sums unscaled from the band-limited maps grid.
__global__ void k_bk_finalize(int rows, int64_t itot, const int32_t *tri, const double *C,
                              int ncol, int nb, const float *scale_exp, const double *shell_sum,
                              const double *nmodes, const float *counts, double kf3, double powSN,
                              double gridD, double sum_scale, double kcenter, double kbin,
                              double *bout) {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= itot) return;
    const int l = tri[3 * m], j = tri[3 * m + 1], i = tri[3 * m + 2];   // 1-based
    auto pw = [&](int s) { return kf3 * shell_sum[s - 1] / nmodes[s - 1] - powSN; };
    const double pl = pw(l), pj = pw(j), pi_ = pw(i);
    const double powSN2 = powSN * powSN;
    const double bispSN = ((pi_ + pj) + pl) * powSN + powSN2;
    const int64_t p = (int64_t)(j - 1) * j / 2 + (i - 1);
    const double unscale =
        sum_scale * ldexp(1.0, -(int)(scale_exp[i - 1] + scale_exp[j - 1] + scale_exp[l - 1]));
    const double cnt = (double)counts[m];
    double *b = bout + (int64_t)rows * m;
    const double sum = C[p * ncol + (l - 1)] * unscale;
    b[0] = kcenter + (double)(l - 1) * kbin;
    b[1] = kcenter + (double)(j - 1) * kbin;
    b[2] = kcenter + (double)(i - 1) * kbin;
    b[3] = pl;
    b[4] = pj;
    b[5] = pi_;
    b[6] = kf3 * sum / cnt + bispSN;
    b[7] = bispSN;
    b[8] = (double)(float)(cnt / gridD);
    if (rows == 11) {
        const double s2 = C[p * ncol + nb + (l - 1)] * unscale;
        const double s4 = C[p * ncol + 2 * nb + (l - 1)] * unscale;
        b[9] = 5.0 * kf3 * s2 / cnt;
        b[10] = 9.0 * kf3 * s4 / cnt;
    }
}

// ------------------------------------------------------------------ exact triangle counts
// One work-group per canonical q1 (|q1x| >= |q1y| >= |q1z| >= 0 up to the lattice symmetry
// group, weighted by its orbit size): for every q2 in shells j >= shell(q1), the shell l of
// q3 = -(q1 + q2) mod N is histogrammed in LDS; the histogram times the weight is added to
// T[i][j][l] (uint64).  Integer arithmetic only: the counts are exact.
constexpr int kCountThreads = 256;

__global__ __launch_bounds__(kCountThreads) void k_count_pairs(int N, int nb,
                                                               const int32_t *modes,
                                                               const int64_t *shell_off,
                                                               const int32_t *canon,
                                                               const int32_t *canon_w,
                                                               const int16_t *shell_table,
                                                               int max_d2,
                                                               unsigned long long *T) {
    extern __shared__ uint32_t hist[];   // (nb+1) x (nb+1): [j][l]
    const int S = nb + 1;
    for (int e = threadIdx.x; e < S * S; e += blockDim.x) hist[e] = 0;
    __syncthreads();
    const int qx = canon[4 * blockIdx.x], qy = canon[4 * blockIdx.x + 1];
    const int qz = canon[4 * blockIdx.x + 2], si = canon[4 * blockIdx.x + 3];
    const int64_t b = shell_off[si - 1], e = shell_off[nb];
    for (int64_t q = b + threadIdx.x; q < e; q += blockDim.x) {
        const int mx = modes[4 * q], my = modes[4 * q + 1], mz = modes[4 * q + 2];
        const int sj = modes[4 * q + 3];
        int cx = -(qx + mx), cy = -(qy + my), cz = -(qz + mz);
        cx = ((cx % N) + N) % N;
        cy = ((cy % N) + N) % N;
        cz = ((cz % N) + N) % N;
        const int dx = min(cx, N - cx), dy = min(cy, N - cy), dz = min(cz, N - cz);
        const int d2 = dx * dx + dy * dy + dz * dz;
        const int sl = d2 <= max_d2 ? shell_table[d2] : 0;
        if (sl > 0 && sl <= nb) atomicAdd(&hist[sj * S + sl], 1u);
    }
    __syncthreads();
    const unsigned long long w = (unsigned long long)canon_w[blockIdx.x];
    for (int t = threadIdx.x; t < S * S; t += blockDim.x) {
        const uint32_t h = hist[t];
        if (h) atomicAdd(&T[(int64_t)si * S * S + t], w * h);
    }
}

}  // namespace

static int gemm_cf(int ncol) {
    const int cf = (int)ceil_div(ncol, 16);
    for (int c : {1, 2, 3, 4, 6, 8, 9, 12})
        if (c >= cf) return c;
    return -1;
}

GemmPlan make_gemm_plan(int nb, int nell, int64_t K, int64_t Kstride, int precision) {
    GemmPlan g{};
    g.nb = nb;
    g.nell = nell;
    g.ncol = nb * nell;
    g.npair = nb * (nb + 1) / 2;
    g.K = K;
    g.Kstride = Kstride;
    g.precision = precision;
    g.cf = gemm_cf(g.ncol);
    PBI4_CHECK(g.cf > 0, PBI4_EUNSUPPORTED,
               "too many shells x multipoles for the triangle GEMM (max 192 columns)");
    PBI4_CHECK(K >= 4 && K % 4 == 0 && Kstride % 4 == 0, PBI4_EINVAL,
               "maps length must be a positive multiple of 4");
    g.rf = g.npair >= 900 ? 4 : g.npair >= 300 ? 2 : 1;
    g.nrowblocks = (int)ceil_div(g.npair, 64 * g.rf);
    // ~2048 work-groups, K chunks a multiple of the LDS stage
    int64_t target = std::max<int64_t>(kXcds, 2048 / g.nrowblocks);
    int64_t kc = std::max<int64_t>(kKT * 8, align_up(ceil_div(K, target), kKT));
    g.kchunk = (int)kc;
    g.nkchunks = (int)ceil_div(K, kc);
    g.nkchunks_pad = (int)align_up(g.nkchunks, kXcds);
    return g;
}

size_t gemm_partials_floats(const GemmPlan &g) {
    return (size_t)g.nkchunks * g.nrowblocks * 64 * g.rf * g.cf * 16;
}

template <int PREC, int CF, int RF>
static void launch_one(const GemmPlan &g, const float *maps, float *partials, hipStream_t s) {
    using G = GemmRing<CF>;
    const size_t lds = sizeof(float) * G::S * G::STAGE;
    auto kern = k_triangle_gemm<PREC, CF, RF>;
    if (lds > 64 * 1024)
        PBI4_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)lds));
    const unsigned nwg = (unsigned)((int64_t)g.nrowblocks * g.nkchunks_pad);
    hipLaunchKernelGGL(kern, dim3(nwg), dim3(kGemmThreads), lds, s, maps, g.K, g.Kstride, g.nb,
                       g.ncol, g.npair, g.kchunk, g.nrowblocks, g.nkchunks, partials);
}

template <int PREC, int CF>
static void launch_rf(const GemmPlan &g, const float *maps, float *partials, hipStream_t s) {
    switch (g.rf) {
        case 1: launch_one<PREC, CF, 1>(g, maps, partials, s); break;
        case 2: launch_one<PREC, CF, 2>(g, maps, partials, s); break;
        default: launch_one<PREC, CF, 4>(g, maps, partials, s); break;
    }
}

template <int PREC>
static void launch_cf(const GemmPlan &g, const float *maps, float *partials, hipStream_t s) {
    switch (g.cf) {
        case 1: launch_rf<PREC, 1>(g, maps, partials, s); break;
        case 2: launch_rf<PREC, 2>(g, maps, partials, s); break;
        case 3: launch_rf<PREC, 3>(g, maps, partials, s); break;
        case 4: launch_rf<PREC, 4>(g, maps, partials, s); break;
        case 6: launch_rf<PREC, 6>(g, maps, partials, s); break;
        case 8: launch_rf<PREC, 8>(g, maps, partials, s); break;
        case 9: launch_rf<PREC, 9>(g, maps, partials, s); break;
        default: launch_rf<PREC, 12>(g, maps, partials, s); break;
    }
}

void launch_gemm(const GemmPlan &g, const float *maps, float *partials, hipStream_t s,
                 KernelTimer *timer) {
    TimedScope ts(timer, "pbi4_triangle_gemm", s);
    if (g.precision == PBI4_PREC_BF16X3)
        launch_cf<PBI4_PREC_BF16X3>(g, maps, partials, s);
    else
        launch_cf<PBI4_PREC_FP32>(g, maps, partials, s);
    PBI4_HIP(hipGetLastError());
}

void launch_gemm_reduce(const GemmPlan &g, const float *partials, double *C, hipStream_t s) {
    const int cf = g.cf;
    const int64_t n = (int64_t)g.npair * g.ncol;
    hipLaunchKernelGGL(k_gemm_reduce, dim3((unsigned)ceil_div(n, 32)), dim3(256), 0, s, partials,
                       g.nkchunks, (int64_t)g.nrowblocks * 64 * g.rf, cf * 16, g.npair, g.ncol, C);
    PBI4_HIP(hipGetLastError());
}

void launch_bk_finalize(int rows, int64_t itot, const int32_t *tri, const double *C, int npair,
                        int ncol, int nb, const float *scale_exp, const double *shell_sum,
                        const double *nmodes, const float *counts, double kf3, double powSN,
                        double gridD, double sum_scale, double kcenter, double kbin, double *bout,
                        hipStream_t s) {
    (void)npair;
    hipLaunchKernelGGL(k_bk_finalize, dim3((unsigned)ceil_div(itot, 128)), dim3(128), 0, s, rows,
                       itot, tri, C, ncol, nb, scale_exp, shell_sum, nmodes, counts, kf3, powSN,
                       gridD, sum_scale, kcenter, kbin, bout);
    PBI4_HIP(hipGetLastError());
}

void count_pairs_gpu(int N, int nb, const int32_t *modes_d, int64_t nmodes, const int64_t *shell_off,
                     const int32_t *canon_d, const int32_t *canon_w, int64_t ncanon,
                     const int16_t *shell_table_d, int max_d2, unsigned long long *T_d,
                     hipStream_t s) {
    (void)nmodes;
    const size_t lds = sizeof(uint32_t) * (nb + 1) * (nb + 1);
    PBI4_CHECK(lds <= 64 * 1024, PBI4_EUNSUPPORTED, "too many shells for the count kernel");
    if (ncanon > 0)
        hipLaunchKernelGGL(k_count_pairs, dim3((unsigned)ncanon), dim3(kCountThreads), lds, s, N, nb,
                           modes_d, shell_off, canon_d, canon_w, shell_table_d, max_d2, T_d);
    PBI4_HIP(hipGetLastError());
}

}  // namespace pbi4
