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
#include <cstdlib>
#include <cstring>

#include "common.h"

namespace pbi4 {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kGemmThreads = 256;     // 4 waves
constexpr int kPadF = 4;              // float padding per LDS row
constexpr int kPadB = 8;              // bf16 padding per LDS row
constexpr int kMaxCF = 12;            // column fragments (ncol <= 192)
constexpr int kXcds = 8;              // MI355X: work-groups are dealt round-robin to 8 XCDs


__device__ __forceinline__ void pair_of(int p, int &i, int &j) {
    // p = j(j+1)/2 + i with 0 <= i <= j
    int jj = (int)((sqrtf(8.0f * (float)p + 1.0f) - 1.0f) * 0.5f);
    while ((jj + 1) * (jj + 2) / 2 <= p) ++jj;
    while (jj * (jj + 1) / 2 > p) --jj;
    j = jj;
    i = p - jj * (jj + 1) / 2;
}

// First row (1-based j) of l in the real-space row order: j >= max(1, ceil((l + 1 - ishift) / 2)),
// the j of the triangle loop that have an i with max(1, l - j + 1 - ishift) <= i <= j.
__host__ __device__ __forceinline__ int jl_first(int l, int ish) {
    const int j = (l + 2 - ish) / 2;
    return j > 1 ? j : 1;
}

// Maps (a <= b, 0-based) whose product is GEMM row p.  Mode 0: the pair (i <= j) at
// p = j (j + 1) / 2 + i (rows of the dense kernel and of the multipole form).  Mode 1 (real
// space): the rows (j, l) of the triangle loop in (l, j) order, j from jl_first(l) to l, whose
// column is the smallest leg i; rows past the last are padding (maps 0, 0).
__device__ __forceinline__ void row_pair(int p, int mode, int nb, int ish, int &a, int &b) {
    if (mode == 0) {
        pair_of(p, a, b);
        return;
    }
    a = b = 0;
    int off = 0;
    for (int l = 1; l <= nb; ++l) {
        const int j0 = jl_first(l, ish), n = l - j0 + 1;
        if (p < off + n) {
            a = j0 + (p - off) - 1;
            b = l - 1;
            return;
        }
        off += n;
    }
}

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

// (a, b) -> packed bf16 hi and lo with a = hi + lo + O(2^-16 |a|): one conversion for two values
// (v_cvt_pk_bf16_f32), the hi halves widened by shifts, one more conversion for the residuals.
__device__ __forceinline__ void split2(float a, float b, uint32_t &hi, uint32_t &lo) {
    const f32x2 v = {a, b};
    hi = __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2));
    const f32x2 r = {a - __uint_as_float(hi << 16), b - __uint_as_float(hi & 0xffff0000u)};
    lo = __builtin_bit_cast(uint32_t, __builtin_convertvector(r, bf16x2));
}

template <int CF>
struct GemmTile {
    static constexpr int KT = CF <= 6 ? 64 : 32;                 // x-values per LDS stage
    static constexpr int NV = (CF * 16 * KT / 4 + kGemmThreads - 1) / kGemmThreads;  // float4/thread
};

template <int PREC>
__host__ __device__ constexpr size_t gemm_stage_bytes(int nb, int ncol_pad, int kt) {
    return PREC == PBI4_PREC_BF16X3
               ? sizeof(float) * nb * (kt + kPadF) + 2 * sizeof(__bf16) * ncol_pad * (kt + kPadB)
               : sizeof(float) * ncol_pad * (kt + kPadF);
}

// Work-group tile: 4 waves x RF row fragments x 16 pair-rows, all CF column fragments, one K
// chunk.  LDS is double buffered: the global loads of stage s+1 are in flight while stage s is
// multiplied; B is split into bf16 hi/lo once, when stored to LDS.  The linear work-group id is
// mapped so that all row blocks of one K chunk land on the same XCD (ids = chunk mod 8): the
// chunk's maps come from HBM once and from that XCD's L2 for the other row blocks.
template <int PREC, int CF, int RF>
__global__ __launch_bounds__(kGemmThreads) void k_triangle_gemm(const float *maps, int64_t K,
                                                                int64_t Kstride, int nb, int ncol,
                                                                int npair, int kchunk,
                                                                int nrowblocks, int nchunks,
                                                                float *partials) {
    constexpr int KT = GemmTile<CF>::KT, NV = GemmTile<CF>::NV;
    constexpr int LDF = KT + kPadF, LDB = KT + kPadB;
    constexpr int NCP = CF * 16;
    extern __shared__ __align__(16) unsigned char smem[];
    const size_t stage = gemm_stage_bytes<PREC>(nb, NCP, KT);

    const int w = blockIdx.x;
    const int q = w / kXcds;
    const int rowblock = q % nrowblocks;
    const int chunk = (w % kXcds) + kXcds * (q / nrowblocks);
    if (chunk >= nchunks) return;

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, grp = lane >> 4;
    const int64_t kbeg = (int64_t)chunk * kchunk;
    const int64_t kend = min(K, kbeg + kchunk);

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

    float4 reg[NV];
    auto load = [&](int64_t x0) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const int e = threadIdx.x + v * kGemmThreads;
            const int c = e / (KT / 4), xv = e % (KT / 4);
            const int64_t x = x0 + 4 * xv;
            reg[v] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (c < ncol && x < kend) reg[v] = *reinterpret_cast<const float4 *>(maps + (int64_t)c * Kstride + x);
        }
    };
    auto store = [&](int buf) {
        unsigned char *base = smem + buf * stage;
        float *tf = reinterpret_cast<float *>(base);
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const int e = threadIdx.x + v * kGemmThreads;
            const int c = e / (KT / 4), xv = e % (KT / 4);
            if (c >= NCP) continue;
            if (PREC == PBI4_PREC_BF16X3) {
                if (c < nb) *reinterpret_cast<float4 *>(tf + c * LDF + 4 * xv) = reg[v];
                __bf16 *bhi = reinterpret_cast<__bf16 *>(base + sizeof(float) * nb * LDF);
                __bf16 *blo = bhi + NCP * LDB;
                uint32_t h[2], l[2];
                split2(reg[v].x, reg[v].y, h[0], l[0]);
                split2(reg[v].z, reg[v].w, h[1], l[1]);
                *reinterpret_cast<uint2 *>(bhi + c * LDB + 4 * xv) = make_uint2(h[0], h[1]);
                *reinterpret_cast<uint2 *>(blo + c * LDB + 4 * xv) = make_uint2(l[0], l[1]);
            } else {
                *reinterpret_cast<float4 *>(tf + c * LDF + 4 * xv) = reg[v];
            }
        }
    };
    auto compute = [&](int buf) {
        const unsigned char *base = smem + buf * stage;
        const float *tf = reinterpret_cast<const float *>(base);
        if (PREC == PBI4_PREC_BF16X3) {
            const __bf16 *bhi = reinterpret_cast<const __bf16 *>(base + sizeof(float) * nb * LDF);
            const __bf16 *blo = bhi + NCP * LDB;
#pragma unroll
            for (int ks = 0; ks < KT; ks += 32) {
                const int xo = ks + 8 * grp;
                bf16x8 ahi[RF], alo[RF];
#pragma unroll
                for (int r = 0; r < RF; ++r) {
                    const float4 *ai = reinterpret_cast<const float4 *>(tf + pi[r] * LDF + xo);
                    const float4 *aj = reinterpret_cast<const float4 *>(tf + pj[r] * LDF + xo);
                    const float4 a0 = ai[0], a1 = ai[1], b0 = aj[0], b1 = aj[1];
                    uint32_t h[4], l[4];
                    split2(a0.x * b0.x, a0.y * b0.y, h[0], l[0]);
                    split2(a0.z * b0.z, a0.w * b0.w, h[1], l[1]);
                    split2(a1.x * b1.x, a1.y * b1.y, h[2], l[2]);
                    split2(a1.z * b1.z, a1.w * b1.w, h[3], l[3]);
                    ahi[r] = __builtin_bit_cast(bf16x8, h);
                    alo[r] = __builtin_bit_cast(bf16x8, l);
                }
                // Wide tiles: raise the wave's priority over its MFMA block so that the other wave on
                // the SIMD issues its split VALU work in the MFMA shadow (measured faster for CF >= 6,
                // slower for the VALU-heavier CF = 3 tiles).
                if (CF >= 6) __builtin_amdgcn_s_setprio(1);
#pragma unroll
                for (int c = 0; c < CF; ++c) {
                    const int col = c * 16 + (lane & 15);
                    const bf16x8 bh = *reinterpret_cast<const bf16x8 *>(bhi + col * LDB + xo);
                    const bf16x8 bl = *reinterpret_cast<const bf16x8 *>(blo + col * LDB + xo);
#pragma unroll
                    for (int r = 0; r < RF; ++r) {
                        acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi[r], bh, acc[r][c], 0, 0, 0);
                        acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi[r], bl, acc[r][c], 0, 0, 0);
                        acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo[r], bh, acc[r][c], 0, 0, 0);
                    }
                }
                if (CF >= 6) __builtin_amdgcn_s_setprio(0);
            }
        } else {
#pragma unroll
            for (int ks = 0; ks < KT; ks += 16) {
                const int xo = ks + 4 * grp;
                float pr[RF][4];
#pragma unroll
                for (int r = 0; r < RF; ++r) {
                    const float4 a = *reinterpret_cast<const float4 *>(tf + pi[r] * LDF + xo);
                    const float4 b = *reinterpret_cast<const float4 *>(tf + pj[r] * LDF + xo);
                    pr[r][0] = a.x * b.x;
                    pr[r][1] = a.y * b.y;
                    pr[r][2] = a.z * b.z;
                    pr[r][3] = a.w * b.w;
                }
#pragma unroll
                for (int c = 0; c < CF; ++c) {
                    const int col = c * 16 + (lane & 15);
                    const float4 bv = *reinterpret_cast<const float4 *>(tf + col * LDF + xo);
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

    // padded columns (ncol <= c < NCP) are zero in both buffers for the whole kernel: the loads
    // give zeros there and store() writes them
    int buf = 0;
    load(kbeg);
    store(0);
    __syncthreads();
    for (int64_t x0 = kbeg; x0 < kend; x0 += KT) {
        const bool more = x0 + KT < kend;
        if (more) load(x0 + KT);
        compute(buf);
        if (more) store(buf ^ 1);
        __syncthreads();
        buf ^= 1;
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

// ------------------------------------------------------------------ block-sparse triangle GEMM
// Triangle (l, j, i) reads only columns l >= j of pair (i, j), and with the columns in (l, ell)
// order those form one short run per wave of pair rows: about a third of the dense tile at 48
// shells x 3 multipoles.  Each work-group (4 waves, 64 RF consecutive pair rows) multiplies W
// column fragments per wave, W the widest of its waves' windows, each wave from its own first
// fragment.  Kernel argument, filled on the host from the triangle list:
constexpr int kMaxWinWaves = 1024;
struct GemmWindows {
    uint8_t start[kMaxWinWaves];       // first logical column fragment of each wave's window
    uint8_t width[kMaxWinWaves / 4];   // W of each work-group (0: no triangle reads its rows)
    uint8_t rbs[kMaxWinWaves / 4];     // row blocks in launch order (launch k: a range of them)
};

constexpr int kWinKT = 32;         // x-values per LDS stage (one bf16 MFMA k-step)
constexpr int kWinMaxMaps = 64;    // A maps staged per work-group: nb <= 64 (2 float4 per thread)

template <int CFW>
struct WinTile {
    static constexpr int NVB = CFW / 2 + 2;    // B float4 per thread: staged union <= 2 NVB tiles
    static constexpr int UCAP = 2 * NVB;
};

template <int PREC>
__host__ __device__ constexpr size_t win_stage_bytes(int nb, int ucap) {
    return sizeof(float) * nb * (kWinKT + kPadF) +
           (PREC == PBI4_PREC_BF16X3 ? 2 * sizeof(__bf16) * ucap * 16 * (kWinKT + kPadB)
                                     : sizeof(float) * ucap * 16 * (kWinKT + kPadF));
}

// One work-group of the block-sparse GEMM with W column fragments per wave (straight-line MFMA
// code).  Same products and MFMA sequence per C entry as k_triangle_gemm (A = M_i M_j split
// once per x, B split when staged, k-steps of 32 in x order), so every entry is bitwise the
// dense kernel's.  The work-group stages the FP32 maps 0 .. jmax-1 its pair rows take (the A
// sources) and the union of its waves' windows (B).  Partials: [chunk][row][nw], the column
// offset relative to the wave's window.
template <int PREC, int W, int NVB, int RF>
__device__ __forceinline__ void win_body(const float *maps, int64_t K, int64_t Kstride, int nb,
                                         int moff, int mdiv, int ncol, int npair, int kchunk,
                                         int nrowblocks, int rowblock, int chunk, int ucap, int nw,
                                         int rowmode, int ish, const GemmWindows &win,
                                         float *partials) {
    constexpr int KT = kWinKT, LDF = KT + kPadF, LDB = KT + kPadB;
    constexpr int NVA = kWinMaxMaps * (KT / 4) / kGemmThreads;
    extern __shared__ __align__(16) unsigned char smem[];
    const size_t abytes = sizeof(float) * nb * LDF;
    const size_t stage = win_stage_bytes<PREC>(nb, ucap);

    const int lane = threadIdx.x & 63, grp = lane >> 4;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wid = rowblock * 4 + wave;
    int u0 = 255, u1 = 0;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        u0 = min(u0, (int)win.start[rowblock * 4 + v]);
        u1 = max(u1, (int)win.start[rowblock * 4 + v] + W);
    }
    const int nstage = (u1 - u0) * 16;          // staged B columns
    const int t0 = win.start[wid] - u0;         // the wave's first fragment in the staged tile
    const int64_t kbeg = (int64_t)chunk * kchunk;
    const int64_t kend = min(K, kbeg + kchunk);
    int alast, blast;
    row_pair(min(npair, (rowblock + 1) * 4 * RF * 16) - 1, rowmode, nb, ish, alast, blast);
    const int amax = blast + 1;                 // rows (a <= b, b non-decreasing) read maps 0 .. blast

    int pi[RF], pj[RF];
#pragma unroll
    for (int r = 0; r < RF; ++r) {
        const int prow = wid * RF * 16 + r * 16 + (lane & 15);
        pi[r] = pj[r] = 0;
        if (prow < npair) row_pair(prow, rowmode, nb, ish, pi[r], pj[r]);
    }
    f32x4 acc[RF][W];
#pragma unroll
    for (int r = 0; r < RF; ++r)
#pragma unroll
        for (int c = 0; c < W; ++c) acc[r][c] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};

    // staging: thread t moves x-quad t % 8 of A map (t / 8) + 32 v and of staged column
    // (t / 8) + 32 v; the map behind a staged column is fixed for the whole kernel
    const int xq = threadIdx.x & 7, row0 = threadIdx.x >> 3;
    const float *asrc[NVA], *bsrc[NVB];
#pragma unroll
    for (int v = 0; v < NVA; ++v) {
        const int m = row0 + 32 * v;
        asrc[v] = m < amax ? maps + (int64_t)m * Kstride + 4 * xq : nullptr;
    }
#pragma unroll
    for (int v = 0; v < NVB; ++v) {
        const int u = row0 + 32 * v, c = u0 * 16 + u;
        bsrc[v] = (u < nstage && c < ncol)
                      ? maps + (int64_t)(moff + (c % mdiv) * nb + c / mdiv) * Kstride + 4 * xq
                      : nullptr;
    }
    float4 ra[NVA], rb[NVB];
    auto load = [&](int64_t x0) {
        const bool in = x0 + 4 * xq < kend;
#pragma unroll
        for (int v = 0; v < NVA; ++v) {
            ra[v] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (asrc[v] && in) ra[v] = *reinterpret_cast<const float4 *>(asrc[v] + x0);
        }
#pragma unroll
        for (int v = 0; v < NVB; ++v) {
            rb[v] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (bsrc[v] && in) rb[v] = *reinterpret_cast<const float4 *>(bsrc[v] + x0);
        }
    };
    auto store = [&](int buf) {
        unsigned char *base = smem + buf * stage;
        float *tf = reinterpret_cast<float *>(base);
#pragma unroll
        for (int v = 0; v < NVA; ++v)
            if (row0 + 32 * v < amax) *reinterpret_cast<float4 *>(tf + (row0 + 32 * v) * LDF + 4 * xq) = ra[v];
#pragma unroll
        for (int v = 0; v < NVB; ++v) {
            const int u = row0 + 32 * v;
            if (u >= nstage) continue;
            if (PREC == PBI4_PREC_BF16X3) {
                __bf16 *bhi = reinterpret_cast<__bf16 *>(base + abytes);
                __bf16 *blo = bhi + ucap * 16 * LDB;
                uint32_t h[2], l[2];
                split2(rb[v].x, rb[v].y, h[0], l[0]);
                split2(rb[v].z, rb[v].w, h[1], l[1]);
                *reinterpret_cast<uint2 *>(bhi + u * LDB + 4 * xq) = make_uint2(h[0], h[1]);
                *reinterpret_cast<uint2 *>(blo + u * LDB + 4 * xq) = make_uint2(l[0], l[1]);
            } else {
                float *tb = reinterpret_cast<float *>(base + abytes);
                *reinterpret_cast<float4 *>(tb + u * LDF + 4 * xq) = rb[v];
            }
        }
    };
    auto compute = [&](int buf) {
        const unsigned char *base = smem + buf * stage;
        const float *tf = reinterpret_cast<const float *>(base);
        if (PREC == PBI4_PREC_BF16X3) {
            const __bf16 *bhi = reinterpret_cast<const __bf16 *>(base + abytes);
            const __bf16 *blo = bhi + ucap * 16 * LDB;
            const int xo = 8 * grp;
            // the wave's W B fragments in registers, then per row fragment: form A (VALU) and
            // issue its 3 W MFMAs, so that the next fragment's products fill the MFMA shadow
            bf16x8 bh[W], bl[W];
            const __bf16 *bh0 = bhi + (t0 * 16 + (lane & 15)) * LDB + xo;
            const __bf16 *bl0 = blo + (t0 * 16 + (lane & 15)) * LDB + xo;
#pragma unroll
            for (int c = 0; c < W; ++c) {
                bh[c] = *reinterpret_cast<const bf16x8 *>(bh0 + c * 16 * LDB);
                bl[c] = *reinterpret_cast<const bf16x8 *>(bl0 + c * 16 * LDB);
            }
#pragma unroll
            for (int r = 0; r < RF; ++r) {
                const float4 *ai = reinterpret_cast<const float4 *>(tf + pi[r] * LDF + xo);
                const float4 *aj = reinterpret_cast<const float4 *>(tf + pj[r] * LDF + xo);
                const float4 a0 = ai[0], a1 = ai[1], b0 = aj[0], b1 = aj[1];
                uint32_t h[4], l[4];
                split2(a0.x * b0.x, a0.y * b0.y, h[0], l[0]);
                split2(a0.z * b0.z, a0.w * b0.w, h[1], l[1]);
                split2(a1.x * b1.x, a1.y * b1.y, h[2], l[2]);
                split2(a1.z * b1.z, a1.w * b1.w, h[3], l[3]);
                const bf16x8 ahi = __builtin_bit_cast(bf16x8, h);
                const bf16x8 alo = __builtin_bit_cast(bf16x8, l);
#pragma unroll
                for (int c = 0; c < W; ++c) {
                    acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi, bh[c], acc[r][c], 0, 0, 0);
                    acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi, bl[c], acc[r][c], 0, 0, 0);
                    acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo, bh[c], acc[r][c], 0, 0, 0);
                }
            }
        } else {
            const float *tb = reinterpret_cast<const float *>(base + abytes);
#pragma unroll
            for (int ks = 0; ks < KT; ks += 16) {
                const int xo = ks + 4 * grp;
                float pr[RF][4];
#pragma unroll
                for (int r = 0; r < RF; ++r) {
                    const float4 a = *reinterpret_cast<const float4 *>(tf + pi[r] * LDF + xo);
                    const float4 b = *reinterpret_cast<const float4 *>(tf + pj[r] * LDF + xo);
                    pr[r][0] = a.x * b.x;
                    pr[r][1] = a.y * b.y;
                    pr[r][2] = a.z * b.z;
                    pr[r][3] = a.w * b.w;
                }
                const float *b0 = tb + (t0 * 16 + (lane & 15)) * LDF + xo;
#pragma unroll
                for (int c = 0; c < W; ++c) {
                    const float4 bv = *reinterpret_cast<const float4 *>(b0 + c * 16 * LDF);
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

    int buf = 0;
    load(kbeg);
    store(0);
    __syncthreads();
    for (int64_t x0 = kbeg; x0 < kend; x0 += KT) {
        const bool more = x0 + KT < kend;
        if (more) load(x0 + KT);
        compute(buf);
        if (more) store(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }
    const int64_t npair_pad = (int64_t)nrowblocks * 4 * RF * 16;
    float *out = partials + (int64_t)chunk * npair_pad * nw;
#pragma unroll
    for (int r = 0; r < RF; ++r)
#pragma unroll
        for (int c = 0; c < W; ++c)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int64_t row = (int64_t)wid * RF * 16 + r * 16 + 4 * grp + i;
                out[row * nw + c * 16 + (lane & 15)] = acc[r][c][i];
            }
}

// Dispatches each work-group on its window width W (1 .. CFW), so that the MFMA loop of every
// width is straight-line code; the kernel's registers are those of the widest case, so the host
// launches the row blocks of each width class on the narrowest body that holds them.  This
// launch covers the row blocks win.rbs[rb_off .. rb_off + nrb_l).
template <int PREC, int CFW, int RF>
__global__ __launch_bounds__(kGemmThreads) void k_triangle_gemm_win(
    const float *maps, int64_t K, int64_t Kstride, int nb, int moff, int mdiv, int ncol, int npair,
    int kchunk, int nrowblocks, int nchunks, int ucap, int nw, int rowmode, int ish, int rb_off,
    int nrb_l, GemmWindows win, float *partials) {
    constexpr int NVB = WinTile<CFW>::NVB;
    const int w = blockIdx.x;
    const int q = w / kXcds;
    const int rowblock = win.rbs[rb_off + q % nrb_l];
    const int chunk = (w % kXcds) + kXcds * (q / nrb_l);
    if (chunk >= nchunks) return;
    const int W = win.width[rowblock];
#define PBI4_WIN_CASE(n)                                                                       \
    case n:                                                                                    \
        if constexpr (n <= CFW)                                                                \
            win_body<PREC, n, NVB, RF>(maps, K, Kstride, nb, moff, mdiv, ncol, npair, kchunk,  \
                                       nrowblocks, rowblock, chunk, ucap, nw, rowmode, ish, win, \
                                       partials);                                             \
        break;
    switch (W) {
        PBI4_WIN_CASE(1)
        PBI4_WIN_CASE(2)
        PBI4_WIN_CASE(3)
        PBI4_WIN_CASE(4)
        PBI4_WIN_CASE(5)
        PBI4_WIN_CASE(6)
        PBI4_WIN_CASE(7)
        PBI4_WIN_CASE(8)
        default: break;
    }
#undef PBI4_WIN_CASE
}

// Split-K combine of the block-sparse partials, entries in the waves' windows only: C[p][col]
// for col = 16 start(wave) + offset, same chunk split and summation order as k_gemm_reduce.
__global__ __launch_bounds__(256) void k_gemm_reduce_win(const float *partials, int nchunks,
                                                         int64_t npair_pad, int nw, int wave_rows,
                                                         int npair, int ncol, GemmWindows win,
                                                         double *C) {
    __shared__ double red[8][32];
    const int lo = threadIdx.x & 31, part = threadIdx.x >> 5;
    const int64_t t = (int64_t)blockIdx.x * 32 + lo;
    const int64_t p = t / nw;
    const int cr = (int)(t % nw);
    bool ok = p < npair;
    int col = 0;
    if (ok) {
        const int wv = (int)(p / wave_rows);
        col = 16 * win.start[wv] + cr;
        ok = cr < 16 * win.width[wv / 4] && col < ncol;
    }
    double s = 0.0;
    if (ok) {
        const int per = (nchunks + 7) / 8;
        const int k0 = part * per, k1 = min(nchunks, k0 + per);
        int k = k0;
        for (; k + 8 <= k1; k += 8) {
            float v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = partials[((int64_t)(k + u) * npair_pad + p) * nw + cr];
#pragma unroll
            for (int u = 0; u < 8; ++u) s += (double)v[u];
        }
        for (; k < k1; ++k) s += (double)partials[((int64_t)k * npair_pad + p) * nw + cr];
    }
    red[part][lo] = s;
    __syncthreads();
    if (part == 0 && ok) {
        double r = red[0][lo];
        for (int q = 1; q < 8; ++q) r += red[q][lo];
        C[p * ncol + col] = r;
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
    // loads issued 8 at a time, added in chunk order (the sum stays in a fixed order)
    int k = k0;
    for (; k + 8 <= k1; k += 8) {
        float v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = partials[((int64_t)(k + u) * npair_pad + p) * ncol_pad + c];
#pragma unroll
        for (int u = 0; u < 8; ++u) s += (double)v[u];
    }
    for (; k < k1; ++k) s += (double)partials[((int64_t)k * npair_pad + p) * ncol_pad + c];
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
                              int cld, int cs_l, int cs_e, int rowmode, int ish,
                              const float *scale_exp,
                              const double *shell_sum,
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
    // C entry of the triangle: row (i <= j), column (l, ell); or (real-space rows) row (j, l),
    // column i
    const double *cp;
    if (rowmode == 0) {
        cp = C + ((int64_t)(j - 1) * j / 2 + (i - 1)) * cld + (int64_t)(l - 1) * cs_l;
    } else {
        int64_t row = j - jl_first(l, ish);
        for (int ll = 1; ll < l; ++ll) row += ll - jl_first(ll, ish) + 1;
        cp = C + row * cld + (i - 1);
    }
    const double unscale =
        sum_scale * ldexp(1.0, -(int)(scale_exp[i - 1] + scale_exp[j - 1] + scale_exp[l - 1]));
    const double cnt = (double)counts[m];
    double *b = bout + (int64_t)rows * m;
    const double sum = cp[0] * unscale;
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
        const double s2 = cp[cs_e] * unscale;
        const double s4 = cp[2 * cs_e] * unscale;
        b[9] = 5.0 * kf3 * s2 / cnt;
        b[10] = 9.0 * kf3 * s4 / cnt;
    }
}

// Cross bispectrum (bispectrumI4.f90:733-755): triangle m = (l, j, i) takes the pair (j <= l)
// of field-1 maps (GEMM rows) and field-2 map i (GEMM column nb + i - 1); rows k_l, k_j, k_i,
// kf3 sum / N_tr, N_tr / N^3.  scale_exp holds the field-1 scales, then the field-2 scales.
__global__ void k_bk_cross_finalize(int64_t itot, const int32_t *tri, const double *C, int cld,
                                    int coff, int nb, const float *scale_exp, const float *counts,
                                    double kf3, double gridD, double sum_scale, double kcenter,
                                    double kbin, double *bout) {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= itot) return;
    const int l = tri[3 * m], j = tri[3 * m + 1], i = tri[3 * m + 2];   // 1-based
    const int64_t p = (int64_t)(l - 1) * l / 2 + (j - 1);
    const double unscale =
        sum_scale * ldexp(1.0, -(int)(scale_exp[j - 1] + scale_exp[l - 1] + scale_exp[nb + i - 1]));
    const double cnt = (double)counts[m];
    const double sum = C[p * cld + coff + (i - 1)] * unscale;
    double *b = bout + 5 * m;
    b[0] = kcenter + (double)(l - 1) * kbin;
    b[1] = kcenter + (double)(j - 1) * kbin;
    b[2] = kcenter + (double)(i - 1) * kbin;
    b[3] = kf3 * sum / cnt;
    b[4] = (double)(float)(cnt / gridD);
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
    for (int c : {1, 2, 3, 4, 6, 8, 9, kMaxCF})
        if (c >= cf) return c;
    return -1;
}

GemmPlan make_gemm_plan(int nb, int nell, int64_t K, int64_t Kstride, int precision, int nrows) {
    GemmPlan g{};
    g.nb = nb;
    g.nell = nell;
    g.ncol = nb * nell;
    g.npair = nrows > 0 ? nrows : nb * (nb + 1) / 2;
    g.K = K;
    g.Kstride = Kstride;
    g.precision = precision;
    g.cf = gemm_cf(g.ncol);
    PBI4_CHECK(g.cf > 0, PBI4_EUNSUPPORTED,
               "too many shells x multipoles for the triangle GEMM (max 192 columns)");
    PBI4_CHECK(K >= 4 && K % 4 == 0 && Kstride % 4 == 0, PBI4_EINVAL,
               "maps length must be a positive multiple of 4");
    g.rf = g.npair >= 900 ? 4 : g.npair >= 300 ? 2 : 1;
    // 12 column tiles with 4 row tiles hold 192 accumulator registers: one wave per SIMD; two
    // row tiles keep two (192-column shapes: 161 -> 144 ms at 64 shells on 384^3)
    if (g.cf >= 12 && g.rf > 2) g.rf = 2;
    // few pairs and several column tiles: one row block reads the maps once instead of once per
    // block (16 shells x 3 multipoles on 100^3: 98 -> 89 us; 21 shells on 128^3: 336 -> 242 us)
    if (g.npair <= 256 && g.cf >= 3) g.rf = 4;
    g.nrowblocks = (int)ceil_div(g.npair, 64 * g.rf);
    // ~2048 work-groups, K chunks a multiple of the LDS stage and at least 2048 long: each chunk
    // writes a partial tile that the FP64 reduce reads back, so short chunks on small problems
    // cost more in partials than they gain in parallelism (16 shells x 3 on 100^3: GEMM + reduce
    // 197 -> 118 us)
    // and at most kMaxChunk long: each chunk is accumulated in FP32 on the MFMA before the FP64
    // combine, and on full grids (K = N^3 up to 10^8) chunks of ~5e5 points left 2.7e-4 relative
    // errors on sums that cancel (384^3, 64 shells x 3, FP32 and split-BF16 alike)
    constexpr int64_t kMaxChunk = 32768;
    const int kt = g.cf <= 6 ? 64 : 32;
    int64_t target = std::max<int64_t>(kXcds, 2048 / g.nrowblocks);
    int64_t kc = std::max<int64_t>(2048, align_up(ceil_div(K, target), kt));
    kc = std::min<int64_t>(kc, kMaxChunk);
    g.kchunk = (int)kc;
    g.nkchunks = (int)ceil_div(K, kc);
    g.nkchunks_pad = (int)align_up(g.nkchunks, kXcds);
    g.ncol_l = g.cld = g.ncol;
    g.cs_l = 1;
    g.cs_e = nb;
    return g;
}

size_t gemm_partials_floats(const GemmPlan &g) {
    return (size_t)g.nkchunks * g.nrowblocks * 64 * g.rf * (g.win ? g.maxw : g.cf) * 16;
}

// Per-wave windows of plan g for the C entries (row, logical column) of the triangles; false
// (with the reason) if they do not fit the compiled block-sparse tiles.
template <class Entry>
static bool plan_windows(GemmPlan &g, const std::vector<int32_t> &tri, const Entry &entry,
                         std::string &why) {
    const int nwaves = g.nrowblocks * 4, wave_rows = 16 * g.rf;
    if (g.nb > kWinMaxMaps) return why = "more than 64 shells", false;
    if (nwaves > kMaxWinWaves) return why = "too many pair rows", false;
    std::vector<int> lo(nwaves, 1 << 30), hi(nwaves, -1);
    for (size_t m = 0; m + 2 < tri.size(); m += 3) {
        int64_t row, c0, c1;
        entry(tri[m], tri[m + 1], tri[m + 2], row, c0, c1);
        const int w = (int)(row / wave_rows);
        lo[w] = std::min<int>(lo[w], (int)(c0 / 16));
        hi[w] = std::max<int>(hi[w], (int)(c1 / 16));
    }
    // work-group width W = its widest wave window; every wave multiplies W fragments from its
    // own first one, moved left where the window would pass the last logical fragment
    const int cfl = (int)ceil_div(g.ncol_l, 16);
    std::vector<uint8_t> start(nwaves, 0), width(g.nrowblocks, 0);
    std::vector<int> span(g.nrowblocks, 0);
    int maxw = 0, ucap = 0;
    for (int b = 0; b < g.nrowblocks; ++b) {
        int W = 0, first = -1;
        for (int w = 4 * b; w < 4 * b + 4; ++w)
            if (hi[w] >= 0) W = std::max(W, hi[w] - lo[w] + 1);
        if (W == 0) continue;
        for (int w = 4 * b; w < 4 * b + 4; ++w)
            if (hi[w] >= 0) {
                start[w] = (uint8_t)std::min(lo[w], cfl - W);
                if (first < 0) first = start[w];
            }
        int u0 = 1 << 30, u1 = -1;
        for (int w = 4 * b; w < 4 * b + 4; ++w) {
            if (hi[w] < 0) start[w] = (uint8_t)first;     // rows without triangles: no new tiles
            u0 = std::min<int>(u0, start[w]);
            u1 = std::max<int>(u1, start[w] + W);
        }
        width[b] = (uint8_t)W;
        span[b] = u1 - u0;
        maxw = std::max(maxw, W);
        ucap = std::max(ucap, u1 - u0);
    }
    int cfw = 0;
    for (int c : {2, 3, 4, 6, 8})
        if (c >= maxw) {
            cfw = c;
            break;
        }
    if (maxw == 0) return why = "no triangles", false;
    if (cfw == 0) return why = "a window wider than 8 column tiles", false;
    auto ucap_of = [](int c) {
        return c == 2   ? WinTile<2>::UCAP
               : c == 3 ? WinTile<3>::UCAP
               : c == 4 ? WinTile<4>::UCAP
               : c == 6 ? WinTile<6>::UCAP
                        : WinTile<8>::UCAP;
    };
    if (ucap > ucap_of(cfw)) return why = "a work-group's windows span too many column tiles", false;
    g.cfw = cfw;
    g.maxw = maxw;
    g.ucap = ucap;
    g.wstart = std::move(start);
    g.wwidth = std::move(width);
    // launch classes: each row block on the narrowest compiled body that holds its window and
    // staged span (the 2- and 4-tile bodies hold 103 and 142 registers against 220 for the 8-tile
    // one: 4 and 3 waves per SIMD instead of 2, and less LDS).  Separate launches also stop the
    // row blocks of one K chunk from sharing its maps in L2, which outweighs the occupancy from
    // five row blocks on (split / one launch: 256^3 with 42 shells x 3, 4 row blocks: 7.8 / 9.5
    // ms; 192^3 with 32 x 3, 5: 2.48 / 2.61; 150^3 with 48 x 3, 5: 2.21 / 2.03; 384^3 with
    // 64 x 3, 17: 78.0 / 74.1), so only plans of at most four row blocks on 2^21 or more points
    // split.  PBI4_TRIANGLE_GEMM_SPLIT=1 / =0 splits always / never (comparisons and tests).
    const int widths[5] = {2, 3, 4, 6, 8};
    std::vector<std::vector<int>> by(5);
    const char *force_split = std::getenv("PBI4_TRIANGLE_GEMM_SPLIT");
    const bool split = force_split ? std::strcmp(force_split, "1") == 0
                                   : g.K >= (int64_t)1 << 21 && g.nrowblocks <= 4;
    for (int b = 0; b < g.nrowblocks; ++b) {
        int k = 4;
        while (widths[k] != cfw) --k;                  // the widest body: holds every row block
        if (split)
            for (int c = 0; c < 5 && widths[c] <= cfw; ++c)
                if (widths[c] >= g.wwidth[b] && ucap_of(widths[c]) >= span[b]) {
                    k = c;
                    break;
                }
        by[k].push_back(b);
    }
    g.rbs.clear();
    g.cls_w.clear();
    g.cls_off.assign(1, 0);
    g.cls_ucap.clear();
    for (int k = 0; k < 5; ++k) {
        if (by[k].empty()) continue;
        int u = 1;
        for (int b : by[k]) {
            g.rbs.push_back((uint8_t)b);
            u = std::max(u, span[b]);
        }
        g.cls_w.push_back(widths[k]);
        g.cls_ucap.push_back(u);
        g.cls_off.push_back((int)g.rbs.size());
    }
    return true;
}

void gemm_set_windows(GemmPlan &g, const std::vector<int32_t> &tri, bool cross, int ishift) {
    g.win = false;
    g.rowmode = 0;
    g.ish = ishift;
    g.ncol_l = cross ? g.nb : g.nb * g.nell;
    g.moff = cross ? g.nb : 0;
    g.mdiv = cross ? 1 : g.nell;
    // dense indexing (maps in ell-major order; the cross field-2 maps after the field-1 maps)
    g.cld = g.ncol;
    g.cs_l = 1;
    g.cs_e = g.nb;
    g.coff = cross ? g.nb : 0;
    // PBI4_TRIANGLE_GEMM=dense keeps the dense kernel, =block_sparse the block-sparse kernel on
    // (i <= j) rows and refuses shapes its tiles do not cover (comparisons and timing)
    const char *mode = std::getenv("PBI4_TRIANGLE_GEMM");
    if (mode && std::strcmp(mode, "dense") == 0) return;
    const bool force = mode && std::strcmp(mode, "block_sparse") == 0;
    std::string why;
    // real space: rows (j, l) of the triangle loop with the smallest leg i as the column -- about
    // half the (i <= j) pair rows (624 of 1176 at 48 shells of k_F) and the forming of the pair
    // products is what bounds the block-sparse kernel
    if (!cross && g.nell == 1 && !force) {
        std::vector<int64_t> off(g.nb + 2, 0);
        for (int l = 1; l <= g.nb; ++l) off[l + 1] = off[l] + (l - jl_first(l, ishift) + 1);
        bool rows_ok = true;
        for (size_t m = 0; m + 2 < tri.size(); m += 3)
            if (tri[m + 1] < jl_first(tri[m], ishift) || tri[m + 1] > tri[m]) rows_ok = false;
        GemmPlan h = make_gemm_plan(g.nb, 1, g.K, g.Kstride, g.precision, (int)off[g.nb + 1]);
        h.rowmode = 1;
        h.ish = ishift;
        h.ncol_l = g.nb;
        h.moff = 0;
        h.mdiv = 1;
        auto entry = [&](int l, int j, int i, int64_t &row, int64_t &c0, int64_t &c1) {
            row = off[l] + (j - jl_first(l, ishift));
            c0 = c1 = i - 1;
        };
        if (rows_ok && plan_windows(h, tri, entry, why)) {
            h.win = true;
            h.cld = g.nb;
            h.cs_l = 1;
            h.cs_e = 0;
            h.coff = 0;
            g = std::move(h);
            return;
        }
    }
    auto entry = [&](int l, int j, int i, int64_t &row, int64_t &c0, int64_t &c1) {
        if (cross) {            // pair (j <= l) of field 1, field-2 column i
            row = (int64_t)(l - 1) * l / 2 + (j - 1);
            c0 = c1 = i - 1;
        } else {                // pair (i <= j), columns (l, ell) for every ell
            row = (int64_t)(j - 1) * j / 2 + (i - 1);
            c0 = (int64_t)(l - 1) * g.nell;
            c1 = c0 + g.nell - 1;
        }
    };
    GemmPlan h = g;
    if (!plan_windows(h, tri, entry, why)) {
        PBI4_CHECK(!force, PBI4_EUNSUPPORTED, "PBI4_TRIANGLE_GEMM=block_sparse: " + why);
        return;
    }
    // a work-group with every column tile gains nothing over the dense kernel's wider x-stages
    // (64 x-values for up to 6 tiles) and single staging of each map
    int64_t tiles = 0;
    for (int b = 0; b < h.nrowblocks; ++b) tiles += h.wwidth[b];
    if (!force && 5 * tiles > 4 * (int64_t)h.nrowblocks * h.cf) return;
    h.win = true;
    h.cld = h.ncol_l;
    h.cs_l = cross ? 1 : h.nell;
    h.cs_e = 1;
    h.coff = 0;
    g = std::move(h);
}

static GemmWindows windows_arg(const GemmPlan &g) {
    GemmWindows w{};
    std::copy(g.wstart.begin(), g.wstart.end(), w.start);
    std::copy(g.wwidth.begin(), g.wwidth.end(), w.width);
    std::copy(g.rbs.begin(), g.rbs.end(), w.rbs);
    return w;
}

// Launch k of the block-sparse plan (its row blocks on the CFW-tile body).
template <int PREC, int CFW, int RF>
static void launch_win(const GemmPlan &g, int k, const GemmWindows &w, const float *maps,
                       float *partials, hipStream_t s) {
    const int ucap = g.cls_ucap[k], rb_off = g.cls_off[k], nrb = g.cls_off[k + 1] - rb_off;
    PBI4_CHECK(ucap <= WinTile<CFW>::UCAP, PBI4_EUNSUPPORTED, "triangle GEMM window class");
    const size_t lds = 2 * win_stage_bytes<PREC>(g.nb, ucap);
    PBI4_CHECK(lds <= 160 * 1024, PBI4_EUNSUPPORTED, "triangle GEMM tile does not fit in LDS");
    auto kern = k_triangle_gemm_win<PREC, CFW, RF>;
    if (lds > 64 * 1024)
        PBI4_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)lds));
    const unsigned nwg = (unsigned)((int64_t)nrb * g.nkchunks_pad);
    hipLaunchKernelGGL(kern, dim3(nwg), dim3(kGemmThreads), lds, s, maps, g.K, g.Kstride, g.nb,
                       g.moff, g.mdiv, g.ncol_l, g.npair, g.kchunk, g.nrowblocks, g.nkchunks,
                       ucap, g.maxw * 16, g.rowmode, g.ish, rb_off, nrb, w, partials);
}

template <int PREC, int CFW>
static void launch_win_rf(const GemmPlan &g, int k, const GemmWindows &w, const float *maps,
                          float *partials, hipStream_t s) {
    switch (g.rf) {
        case 1: launch_win<PREC, CFW, 1>(g, k, w, maps, partials, s); break;
        case 2: launch_win<PREC, CFW, 2>(g, k, w, maps, partials, s); break;
        default: launch_win<PREC, CFW, 4>(g, k, w, maps, partials, s); break;
    }
}

template <int PREC>
static void launch_win_cfw(const GemmPlan &g, const float *maps, float *partials, hipStream_t s) {
    const GemmWindows w = windows_arg(g);
    for (size_t k = 0; k < g.cls_w.size(); ++k) {
        switch (g.cls_w[k]) {
            case 2: launch_win_rf<PREC, 2>(g, (int)k, w, maps, partials, s); break;
            case 3: launch_win_rf<PREC, 3>(g, (int)k, w, maps, partials, s); break;
            case 4: launch_win_rf<PREC, 4>(g, (int)k, w, maps, partials, s); break;
            case 6: launch_win_rf<PREC, 6>(g, (int)k, w, maps, partials, s); break;
            default: launch_win_rf<PREC, 8>(g, (int)k, w, maps, partials, s); break;
        }
    }
}

template <int PREC, int CF, int RF>
static void launch_one(const GemmPlan &g, const float *maps, float *partials, hipStream_t s) {
    constexpr int KT = GemmTile<CF>::KT;
    const size_t lds = 2 * gemm_stage_bytes<PREC>(g.nb, CF * 16, KT);
    PBI4_CHECK(lds <= 160 * 1024, PBI4_EUNSUPPORTED, "triangle GEMM tile does not fit in LDS");
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
    if (g.win) {
        if (g.precision == PBI4_PREC_BF16X3)
            launch_win_cfw<PBI4_PREC_BF16X3>(g, maps, partials, s);
        else
            launch_win_cfw<PBI4_PREC_FP32>(g, maps, partials, s);
    } else if (g.precision == PBI4_PREC_BF16X3) {
        launch_cf<PBI4_PREC_BF16X3>(g, maps, partials, s);
    } else {
        launch_cf<PBI4_PREC_FP32>(g, maps, partials, s);
    }
    PBI4_HIP(hipGetLastError());
}

void launch_gemm_reduce(const GemmPlan &g, const float *partials, double *C, hipStream_t s) {
    const int64_t npair_pad = (int64_t)g.nrowblocks * 64 * g.rf;
    if (g.win) {
        const int nw = g.maxw * 16;
        const int64_t n = (int64_t)g.npair * nw;
        hipLaunchKernelGGL(k_gemm_reduce_win, dim3((unsigned)ceil_div(n, 32)), dim3(256), 0, s,
                           partials, g.nkchunks, npair_pad, nw, 16 * g.rf, g.npair, g.ncol_l,
                           windows_arg(g), C);
    } else {
        const int64_t n = (int64_t)g.npair * g.ncol;
        hipLaunchKernelGGL(k_gemm_reduce, dim3((unsigned)ceil_div(n, 32)), dim3(256), 0, s,
                           partials, g.nkchunks, npair_pad, g.cf * 16, g.npair, g.ncol, C);
    }
    PBI4_HIP(hipGetLastError());
}

void launch_bk_finalize(int rows, int64_t itot, const int32_t *tri, const double *C,
                        const GemmPlan &g, const float *scale_exp, const double *shell_sum,
                        const double *nmodes, const float *counts, double kf3, double powSN,
                        double gridD, double sum_scale, double kcenter, double kbin, double *bout,
                        hipStream_t s) {
    hipLaunchKernelGGL(k_bk_finalize, dim3((unsigned)ceil_div(itot, 128)), dim3(128), 0, s, rows,
                       itot, tri, C, g.cld, g.cs_l, g.cs_e, g.rowmode, g.ish, scale_exp,
                       shell_sum, nmodes, counts,
                       kf3, powSN, gridD, sum_scale, kcenter, kbin, bout);
    PBI4_HIP(hipGetLastError());
}

void launch_bk_cross_finalize(int64_t itot, const int32_t *tri, const double *C, const GemmPlan &g,
                              const float *scale_exp, const float *counts, double kf3,
                              double gridD, double sum_scale, double kcenter, double kbin,
                              double *bout, hipStream_t s) {
    if (itot <= 0) return;
    hipLaunchKernelGGL(k_bk_cross_finalize, dim3((unsigned)ceil_div(itot, 256)), dim3(256), 0, s,
                       itot, tri, C, g.cld, g.coff, g.nb, scale_exp, counts, kf3, gridD, sum_scale,
                       kcenter, kbin, bout);
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
