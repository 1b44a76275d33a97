// Fourier-space kernels: interlacing combine (fcomb), P_l / shell-power binning, shell filter.
#include "common.h"
#include "dft.h"

namespace pbi4 {

namespace {

// ------------------------------------------------------------------ fcomb (powerI4.f90:251-378)

// The octet (ix, iy, iz) of the interlaced combine from its 8 values orig[k] = F(X, Y, Z), X in
// {ix, icx}, Y in {iy, icy}, Z in {iz, icz} (code bits 1 = icx, 2 = icy, 4 = icz): the cells in
// the x <= N/2 half are written, replaying the eight ordered stores where cells coincide.
__device__ __forceinline__ void combine_octet(int N, int ix, int iy, int iz, const float2 *orig,
                                              const float2 *ph, const float *wtab, float cf,
                                              float2 *delta) {
    const int h = N / 2 + 1;
    const int icx = (N - ix) % N, icy = (N - iy) % N, icz = (N - iz) % N;
    const int mask = (ix == icx ? 1 : 0) | (iy == icy ? 2 : 0) | (iz == icz ? 4 : 0);
    auto cm = [&](int sidx) {  // i * e^{-i theta s}
        const float2 v = ph[sidx + N];
        return make_float2(-v.y, v.x);
    };
    auto comb = [&](float2 a, float2 b, float2 c) {  // a (1 - c) + conj(b) (1 + c)
        return cadd(cmul(a, make_float2(1.0f - c.x, -c.y)), cmul(conjf2(b), make_float2(1.0f + c.x, c.y)));
    };
    const float cfac = cf / (wtab[ix] * wtab[iy] * wtab[iz]);
    const float2 c000 = cscale(comb(orig[0], orig[7], cm(ix + iy + iz)), cfac);
    const float2 c001 = cscale(comb(orig[4], orig[3], cm(ix + iy - iz)), cfac);
    const float2 c010 = cscale(comb(orig[2], orig[5], cm(ix - iy + iz)), cfac);
    const float2 c011 = cscale(comb(orig[6], orig[1], cm(ix - iy - iz)), cfac);
    if (!(mask & 1)) {
        // ix != icx: the odd codes' stores land in other cells, and of the even codes, stored in
        // the order 0, 4, 2, 6, the last one that reaches cell k is code k | mask (mask: y, z)
        auto val = [&](int code) {
            return code == 0 ? c000 : code == 2 ? c010 : code == 4 ? c001 : c011;
        };
#pragma unroll
        for (int k = 0; k < 8; k += 2) {
            const int Y = (k & 2) ? icy : iy, Z = (k & 4) ? icz : iz;
            delta[((int64_t)Z * N + Y) * h + ix] = val(k | mask);
        }
        return;
    }

    float2 mem[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) mem[k] = make_float2(0.0f, 0.0f);
    auto put = [&](int code, float2 v) {
        const int cc = code & ~mask;
#pragma unroll
        for (int j = 0; j < 8; ++j)
            if (j == cc) mem[j] = v;
    };
    auto get = [&](int code) {
        const int cc = code & ~mask;
        float2 v = mem[0];
#pragma unroll
        for (int j = 1; j < 8; ++j)
            if (j == cc) v = mem[j];
        return v;
    };
    put(0, c000);
    put(4, c001);
    put(2, c010);
    put(6, c011);
    put(1, conjf2(get(6)));
    put(5, conjf2(get(2)));
    put(3, conjf2(get(4)));
    put(7, conjf2(get(0)));
#pragma unroll
    for (int k = 0; k < 8; k += 2) {
        const int Y = (k & 2) ? icy : iy, Z = (k & 4) ? icz : iz;
        delta[((int64_t)Z * N + Y) * h + ix] = get(k);
    }
}

// One thread per iteration (ix, iy, iz) in [0, N/2]^3 of the reference's interlaced loop.  The
// iteration owns the <= 8 cells {ix,icx} x {iy,icy} x {iz,icz}; their eight ordered stores are
// replayed in registers (a later store to the same cell wins, later stores read earlier ones),
// then the cells in the x <= N/2 half are written.  Code bits: 1 = icx, 2 = icy, 4 = icz.
// F is the full N^3 transform of grid1 + i grid2 (one complex FFT).
// Threads: first the iterations with x in [1, N/2) (x runs fastest), then those with x in
// {0, N/2}, so the ordered-store replay, needed only where ix = icx, runs in waves of its own.
// ph[s + N] = e^{-i theta s} for s in [-N, 3N/2]: cm(s) = i e^{-i theta s}, s = ix +- iy +- iz,
// one table entry (rounded once from FP64) per phase.
__global__ void k_combine_interlaced(int N, const float2 *ph, const float *wtab, float cf,
                                     const float2 *F, float2 *delta) {
    const int h = N / 2 + 1, hin = N / 2 - 1;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nin = (int64_t)hin * h * h;
    int ix, iy, iz;
    if (t < nin) {
        ix = 1 + (int)(t % hin);
        const int64_t r = t / hin;
        iy = (int)(r % h);
        iz = (int)(r / h);
    } else {
        const int64_t u = t - nin;
        if (u >= 2 * (int64_t)h * h) return;
        ix = (u & 1) ? N / 2 : 0;
        iy = (int)((u >> 1) % h);
        iz = (int)((u >> 1) / h);
    }
    const int icx = (N - ix) % N, icy = (N - iy) % N, icz = (N - iz) % N;
    // The 8 values F(X, Y, Z), X in {ix, icx}, Y in {iy, icy}, Z in {iz, icz}, loaded as one batch.
    float2 orig[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int X = (k & 1) ? icx : ix, Y = (k & 2) ? icy : iy, Z = (k & 4) ? icz : iz;
        orig[k] = F[((int64_t)Z * N + Y) * N + X];
    }
    combine_octet(N, ix, iy, iz, orig, ph, wtab, cf, delta);
}

// Non-interlaced fcomb: only the window deconvolution and the 1/kF^3 normalisation of the
// ix <= N/2 half (powerI4.f90:351-374).
__global__ void k_combine_single(int N, const float *wtab, float cf, const float2 *G1,
                                 float2 *delta) {
    const int h = N / 2 + 1;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)N * N * h) return;
    const int x = (int)(t % h), y = (int)((t / h) % N), z = (int)(t / ((int64_t)h * N));
    const int ay = y <= N / 2 ? y : N - y, az = z <= N / 2 ? z : N - z;
    delta[t] = cscale(G1[t], cf / (wtab[x] * wtab[ay] * wtab[az]));
}

// ------------------------------------------------------------------ binning

constexpr int kBinThreads = 256;
constexpr int kBinWaves = kBinThreads / 64;

// DPP lane moves (VALU, no LDS): CTRL 0x110 + n = row_shr:n inside 16-lane rows, 0x142 =
// row_bcast:15 (lane 15 of each row to the next row), 0x143 = row_bcast:31.
template <int CTRL>
__device__ __forceinline__ int dpp_i(int v) { return __builtin_amdgcn_mov_dpp(v, CTRL, 0xf, 0xf, false); }

template <int CTRL, int NV>
__device__ __forceinline__ void seg_step(int key, float (&v)[NV], bool ok) {
    const int k2 = dpp_i<CTRL>(key);
    float t[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) t[i] = __int_as_float(dpp_i<CTRL>(__float_as_int(v[i])));
    if (ok && k2 == key) {
#pragma unroll
        for (int i = 0; i < NV; ++i) v[i] += t[i];
    }
}

// Segmented inclusive prefix sum across a wave for runs of equal key, keys non-decreasing in
// lane order; afterwards the last lane of each run holds the run total.  A source lane with the
// same key implies every lane between has it, so each step adds exactly the run's lanes.
template <int NV>
__device__ __forceinline__ void seg_prefix_sum(int key, float (&v)[NV], int lane) {
    const int rl = lane & 15;
    seg_step<0x111>(key, v, rl >= 1);
    seg_step<0x112>(key, v, rl >= 2);
    seg_step<0x114>(key, v, rl >= 4);
    seg_step<0x118>(key, v, rl >= 8);
    seg_step<0x142>(key, v, (lane & 31) >= 16);
    seg_step<0x143>(key, v, lane >= 32);
}

// Raw inputs of one mode, loaded unconditionally (indices clamped into range; with some optional
// field present the absent ones read delta, with none present they are not loaded) so a step's
// loads issue as one batch; a guarded load becomes a branch with a wait after it.
struct ModeLoads {
    float2 d, e2, e4, ec;
    int bin, shell;
};

// EXTRA: some of the optional fields (explicit delta_2, delta_4, cross field) are present.
template <bool EXTRA>
__device__ __forceinline__ ModeLoads mode_loads(const BinPlan &b, const float2 *delta, int64_t row,
                                                int h, int x, int aky, int akz) {
    ModeLoads m;
    const int64_t o = row * h + x;
    m.d = delta[o];
    if (EXTRA) {
        m.e2 = (b.f2 ? b.f2 : delta)[o];
        m.e4 = (b.f4 ? b.f4 : delta)[o];
        m.ec = (b.cross ? b.cross : delta)[o];
    } else {
        m.e2 = m.e4 = m.ec = m.d;
    }
    m.bin = b.nbins_pk > 0 ? (int)b.pk_table[((int64_t)akz * h + aky) * h + x] : 0;
    const int d2 = x * x + aky * aky + akz * akz;
    m.shell = b.nshell > 0 ? (int)b.shell_table[min(d2, b.max_d2)] : 0;
    if (d2 > b.max_d2) m.shell = 0;
    return m;
}

// Contributions of one mode: P(k) bin key with (count, k, P0, P2, P4) and shell key with its
// weighted power.
struct ModeTerms {
    int key, skey;
    float v[5], sv[1];
};

__device__ __forceinline__ ModeTerms mode_terms(const BinPlan &b, const ModeLoads &m, int N, int x,
                                                int y, int ky, int aky, int z, int kz, int akz,
                                                float lt, bool valid) {
    ModeTerms t;
    t.key = -1;
    t.skey = -1;
#pragma unroll
    for (int q = 0; q < 5; ++q) t.v[q] = 0.0f;
    t.sv[0] = 0.0f;
    if (!valid) return t;
    const float2 d = m.d;
    const float p = d.x * d.x + d.y * d.y;
    const int d2 = x * x + aky * aky + akz * akz;
    if (b.nbins_pk > 0) {
        const bool sel = x != 0 || ky > 0 || (ky == 0 && kz > 0);
        const int bin = sel ? m.bin : 0;
        t.key = bin;
        if (bin > 0) {
            auto dot = [&](float2 e) { return d.x * e.x + d.y * e.y; };   // Re(conj(d) e)
            // hardware reciprocal and square root (1 ulp): the IEEE forms cost ~10 instructions
            // each in this VALU-bound loop
            const float k2 = (float)d2;
            const float lk = b.los[0] * (float)x + lt;
            const float mu2 = k2 > 0.0f ? lk * lk * __builtin_amdgcn_rcpf(k2) : 0.0f;
            t.v[0] = 1.0f;
            t.v[1] = b.kF * __builtin_amdgcn_sqrtf(k2);
            t.v[2] = (!b.multipoles && b.cross) ? dot(m.ec) : p;
            if (b.multipoles) {
                t.v[3] = b.f2 ? dot(m.e2) : (-0.5f + 1.5f * mu2) * p;
                t.v[4] = b.f4 ? dot(m.e4) : (0.375f + mu2 * (-3.75f + 4.375f * mu2)) * p;
            }
        }
    }
    if (b.nshell > 0) {
        const int s = m.shell;
        t.skey = s;
        if (s > 0) {
            const bool planex = x == 0 || x == N / 2;
            const bool special = planex && (y == 0 || y == N / 2) && (z == 0 || z == N / 2);
            t.sv[0] = (special ? d.x * d.x : p) * (planex ? 1.0f : 2.0f);
        }
    }
    return t;
}

// One wave per (y, z) row of the half spectrum, 64 consecutive x per step over x < N/2; the
// lane holding x = N/2 - 1 also adds the Nyquist mode x = N/2 after the scan.  Along a row |k|
// grows with x, so P(k) bins and shell ids are non-decreasing in lane order and each run is
// reduced in registers; run tails add into wave-private LDS accumulators (no atomics), the waves
// are summed in a fixed order and each work-group writes one partial: deterministic.
// NV: P(k) quantities scanned per mode, 5 with multipoles (count, k, P0, P2, P4), else 3.
// EXTRA: explicit delta_2 / delta_4 / cross fields are read.
template <int NV, bool EXTRA>
__global__ __launch_bounds__(kBinThreads) void k_bin(BinPlan b, const float2 *delta,
                                                     float *partials) {
    extern __shared__ float acc[];
    const int N = b.N, h = N / 2 + 1, half = N / 2;
    const int nq = 5 * b.nbins_pk + b.nshell;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < kBinWaves * nq; i += blockDim.x) acc[i] = 0.0f;
    __syncthreads();
    float *mine = acc + wave * nq;
    const int64_t nrows = (int64_t)N * N;
    for (int64_t row = (int64_t)blockIdx.x * kBinWaves + wave; row < nrows;
         row += (int64_t)gridDim.x * kBinWaves) {
        const int y = (int)(row % N), z = (int)(row / N);
        const int ky = y <= N / 2 ? y : y - N, kz = z <= N / 2 ? z : z - N;
        const int aky = ky < 0 ? -ky : ky, akz = kz < 0 ? -kz : kz;
        const float lt = b.los[1] * (float)ky + b.los[2] * (float)kz;
        for (int x0 = 0; x0 < half; x0 += 64) {
            const int x = x0 + lane;
            const ModeLoads ml = mode_loads<EXTRA>(b, delta, row, h, min(x, half - 1), aky, akz);
            const ModeLoads mn = mode_loads<EXTRA>(b, delta, row, h, half, aky, akz);   // Nyquist
            ModeTerms t = mode_terms(b, ml, N, x, y, ky, aky, z, kz, akz, lt, x < half);
            const bool nyq_lane = x == half - 1;
            ModeTerms tn;
            if (nyq_lane) tn = mode_terms(b, mn, N, half, y, ky, aky, z, kz, akz, lt, true);
            const int last = min(63, half - 1 - x0);      // highest valid lane of this step
            if (b.nbins_pk > 0) {
                float v[NV];
#pragma unroll
                for (int q = 0; q < NV; ++q) v[q] = t.v[q];
                seg_prefix_sum<NV>(t.key, v, lane);
                const int next = __builtin_amdgcn_mov_dpp(t.key, 0x130, 0xF, 0xF, false);  // lane + 1
                if (t.key > 0 && (lane == last || next != t.key)) {
#pragma unroll
                    for (int q = 0; q < NV; ++q) mine[q * b.nbins_pk + t.key - 1] += v[q];
                }
                if (nyq_lane && tn.key > 0) {
#pragma unroll
                    for (int q = 0; q < NV; ++q) mine[q * b.nbins_pk + tn.key - 1] += tn.v[q];
                }
            }
            if (b.nshell > 0) {
                seg_prefix_sum<1>(t.skey, t.sv, lane);
                const int next = __builtin_amdgcn_mov_dpp(t.skey, 0x130, 0xF, 0xF, false);
                if (t.skey > 0 && (lane == last || next != t.skey))
                    mine[5 * b.nbins_pk + t.skey - 1] += t.sv[0];
                if (nyq_lane && tn.skey > 0) mine[5 * b.nbins_pk + tn.skey - 1] += tn.sv[0];
            }
        }
    }
    __syncthreads();
    // partials are [work-group][quantity]: one contiguous row per work-group
    for (int i = threadIdx.x; i < nq; i += blockDim.x) {
        float s = acc[i];
        for (int w = 1; w < kBinWaves; ++w) s += acc[w * nq + i];
        partials[(int64_t)blockIdx.x * nq + i] = s;
    }
}

// FP64 sums of every quantity over the work-group rows, in kBinSegs segments: block (b, seg) owns
// quantities 64 b .. 64 b + 63 (a wave reads 64 consecutive floats of a row) and rows
// seg + kBinSegs (w + 4 r), r = 0, 1, ... (wave w), loads issued 8 at a time and added in row order;
// the four waves are added in a fixed order into qseg[seg][q] (deterministic).
constexpr int kBinSegs = 16;

__global__ __launch_bounds__(256) void k_bin_reduce(const float *partials, int nwg, int nq,
                                                    double *qseg) {
    __shared__ double red[4][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, seg = blockIdx.y;
    const int q = blockIdx.x * 64 + lane;
    const int step = 4 * kBinSegs;
    double s = 0.0;
    if (q < nq) {
        int g = seg + kBinSegs * w;
        for (; g + 7 * step < nwg; g += 8 * step) {
            float v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = partials[(int64_t)(g + u * step) * nq + q];
#pragma unroll
            for (int u = 0; u < 8; ++u) s += (double)v[u];
        }
        for (; g < nwg; g += step) s += (double)partials[(int64_t)g * nq + q];
    }
    red[w][lane] = s;
    __syncthreads();
    if (w == 0 && q < nq)
        qseg[(int64_t)seg * nq + q] = ((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane];
}

// pout / shell sums from the segment sums (added in segment order): one thread per bin.
__global__ void k_bin_finalize(BinPlan b, const double *qseg, double kF3, double *pk_out,
                               double *shell_sum) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= b.nbins_pk + b.nshell) return;
    const int nq = 5 * b.nbins_pk + b.nshell;
    const bool pk = i < b.nbins_pk;
    const int nv = pk ? 5 : 1;
    // quantity q of bin i: column q * nbins_pk + i (P(k) bins) or 4 nbins_pk + i (shells)
    const int col0 = pk ? i : 4 * b.nbins_pk + i;
    const int cstride = pk ? b.nbins_pk : 0;
    double s[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int q = 0; q < 5; ++q)
        if (q < nv)
            for (int seg = 0; seg < kBinSegs; ++seg) s[q] += qseg[(int64_t)seg * nq + col0 + q * cstride];
    if (pk) {
        double *p = pk_out + 5 * (int64_t)i;
        const double co = s[0];
        p[0] = s[1] / co;
        p[1] = s[2] * kF3 / co;
        if (b.multipoles) {
            p[2] = 5.0 * s[3] * kF3 / co;
            p[3] = 9.0 * s[4] * kF3 / co;
            p[4] = co;
        } else {
            p[2] = co;
            p[3] = 0.0;
            p[4] = 0.0;
        }
    } else {
        shell_sum[i - b.nbins_pk] = s[0];
    }
}

// ------------------------------------------------------------------ shell filter

// One thread per mode of the maps-grid half spectrum (N' = f.Np).  A mode q inside the shells
// takes the reference's full-spectrum value, made exactly Hermitian on the self-conjugate x
// planes (Re of the reference's c2c maps equals the c2r of the Hermitian part), times L_l(mu)
// for the redshift-space maps and the power-of-two scale 2^e_s of its shell; every other map
// gets zero at that mode.  Map m = ell*nshell + (s-1).
__device__ __forceinline__ float2 herm_at(const float2 *f, int N, int x, int Y, int Z) {
    const int h = N / 2 + 1;
    const float2 a = f[((int64_t)Z * N + Y) * h + x];
    if (x == 0 || 2 * x == N) {
        const float2 m = f[((int64_t)((N - Z) % N) * N + (N - Y) % N) * h + x];
        return make_float2(0.5f * (a.x + m.x), 0.5f * (a.y - m.y));
    }
    return a;
}

__global__ void k_filter(FilterPlan f, const float2 *delta, const float *scale_exp,
                         float2 *spectra) {
    const int Np = f.Np, hp = Np / 2 + 1;
    const int64_t per_map = (int64_t)Np * Np * hp;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= per_map) return;
    const int x = (int)(t % hp), y = (int)((t / hp) % Np), z = (int)(t / ((int64_t)hp * Np));
    const int qy = y <= Np / 2 ? y : y - Np, qz = z <= Np / 2 ? z : z - Np;
    int s = 0;
    float2 H = make_float2(0.0f, 0.0f), H2 = H, H4 = H;
    const int R = f.R;
    if (x <= R && qy <= R && qy >= -R && qz <= R && qz >= -R) {
        const int d2 = x * x + qy * qy + qz * qz;
        s = d2 <= f.max_d2 ? f.shell_table[d2] : 0;
        if (s > f.nshell) s = 0;
        if (s > 0) {
            const int N = f.N;
            const int Y = (qy + N) % N, Z = (qz + N) % N;
            const float sc = ldexpf(1.0f, (int)scale_exp[s - 1]);
            H = cscale(herm_at(delta, N, x, Y, Z), sc);
            if (f.nell == 3) {
                const float lk = f.los[0] * x + f.los[1] * qy + f.los[2] * qz;
                const float mu2 = d2 > 0 ? lk * lk / (float)d2 : 0.0f;
                H2 = f.f2 ? cscale(herm_at(f.f2, N, x, Y, Z), sc) : cscale(H, -0.5f + 1.5f * mu2);
                H4 = f.f4 ? cscale(herm_at(f.f4, N, x, Y, Z), sc)
                          : cscale(H, 0.375f + mu2 * (-3.75f + 4.375f * mu2));
            }
        }
    }
    const float2 zero = make_float2(0.0f, 0.0f);
    for (int m = 0; m < f.nshell; ++m) spectra[(int64_t)m * per_map + t] = (m + 1 == s) ? H : zero;
    if (f.nell == 3) {
        for (int m = 0; m < f.nshell; ++m)
            spectra[(int64_t)(f.nshell + m) * per_map + t] = (m + 1 == s) ? H2 : zero;
        for (int m = 0; m < f.nshell; ++m)
            spectra[(int64_t)(2 * f.nshell + m) * per_map + t] = (m + 1 == s) ? H4 : zero;
    }
}

// Shell filter fused with the inverse transform over z, for a maps grid N' below the delta grid:
// every shell mode has |q_c| <= R < N'/3, and along one (x', q_y) line only the q_z of shell s
// with d2 = x'^2 + q_y^2 + q_z^2 inside the shell's d2 range are nonzero: two runs of q_z,
// -qhi..-qlo and qlo..qhi.  One thread owns one line, one shell, its NELL maps and ZC
// consecutive planes z, and sums those modes directly,
// A_l(z) = sum_{q_z in S_s} H_l(q) e^{+2 pi i q_z z / N'}, with H_l exactly as k_filter forms it,
// the modes +q_z and -q_z as one pair sharing one phase.  The phase is stepped by one complex
// product per plane from tw[k] = e^{2 pi i k / N'}.
// Output: the half spectra [m][z][y][x'] after the z transform (m = l * nshell + s - 1), rows of
// f.ld >= N'/2 + 1 entries; a batched 2D c2r over (y, x) completes the maps.  Lines outside the
// band get zeros.
__device__ __forceinline__ int isqrt_floor(int v) {
    int r = (int)sqrtf((float)v);
    while (r * r > v) --r;
    while ((r + 1) * (r + 1) <= v) ++r;
    return r;
}

template <int ZC, int NELL>
__global__ __launch_bounds__(256) void k_filter_z(FilterPlan f, const float2 *delta, const float *scale_exp,
                           const float2 *tw, float2 *out) {
    const int Np = f.Np, hp = f.cols ? f.cols : Np / 2 + 1;   // columns x' < hp written
    const int line = blockIdx.x * blockDim.x + threadIdx.x;  // y * hp + x'
    if (line >= Np * hp) return;
    const int s = blockIdx.y + 1, z0 = blockIdx.z * ZC;
    const int x = line % hp, y = line / hp;
    const int qy = y <= Np / 2 ? y : y - Np;
    const int R = f.R;
    float2 acc[NELL][ZC];
#pragma unroll
    for (int l = 0; l < NELL; ++l)
#pragma unroll
        for (int j = 0; j < ZC; ++j) acc[l][j] = make_float2(0.0f, 0.0f);
    int qlo = 0, qhi = -1;
    const int r2 = x * x + qy * qy;
    if (f.disk_only && r2 > f.shell_d2[2 * (s - 1) + 1]) return;   // zero, and never read
    if (x <= R && qy <= R && qy >= -R) {
        const int a = f.shell_d2[2 * (s - 1)], b = f.shell_d2[2 * (s - 1) + 1];
        if (a <= b && b >= r2) {
            qhi = min(isqrt_floor(b - r2), R);
            if (a > r2) {
                qlo = isqrt_floor(a - r2);
                if (qlo * qlo < a - r2) ++qlo;
            }
        }
    }
    if (qlo <= qhi) {
        const int N = f.N, Y = (qy + N) % N;
        const float sc = ldexpf(1.0f, (int)scale_exp[s - 1]);
        // every d2 of the shell's range belongs to it (shells are |q| intervals; the host checks
        // this before taking this path), so the runs hold exactly the shell's modes
        auto mode = [&](int qz, int d2, float2 *H) {
            const int Z = (qz + N) % N;
            H[0] = cscale(herm_at(delta, N, x, Y, Z), sc);
            if (NELL == 3) {
                const float lk = f.los[0] * x + f.los[1] * qy + f.los[2] * qz;
                const float mu2 = d2 > 0 ? lk * lk * __builtin_amdgcn_rcpf((float)d2) : 0.0f;
                H[1] = f.f2 ? cscale(herm_at(f.f2, N, x, Y, Z), sc)
                            : cscale(H[0], -0.5f + 1.5f * mu2);
                H[2] = f.f4 ? cscale(herm_at(f.f4, N, x, Y, Z), sc)
                            : cscale(H[0], 0.375f + mu2 * (-3.75f + 4.375f * mu2));
            }
        };
        // +q_z and -q_z together: with w = e^{i phi q_z z}, H+ w + H- conj(w) =
        // (S.x w.x - D.y w.y, S.y w.x + D.x w.y), S = H+ + H-, D = H+ - H-
        for (int qz = qlo; qz <= qhi; ++qz) {
            const int d2 = r2 + qz * qz;
            float2 Hp[NELL], Hm[NELL], S[NELL], D[NELL];
            mode(qz, d2, Hp);
            if (qz > 0) {
                mode(-qz, d2, Hm);
            } else {
#pragma unroll
                for (int l = 0; l < NELL; ++l) Hm[l] = make_float2(0.0f, 0.0f);
            }
#pragma unroll
            for (int l = 0; l < NELL; ++l) {
                S[l] = make_float2(Hp[l].x + Hm[l].x, Hp[l].y + Hm[l].y);
                D[l] = make_float2(Hp[l].x - Hm[l].x, Hp[l].y - Hm[l].y);
            }
            float2 w = tw[(int)(((int64_t)qz * z0) % Np)];
            const float2 ws = tw[qz];
#pragma unroll
            for (int j = 0; j < ZC; ++j) {
#pragma unroll
                for (int l = 0; l < NELL; ++l) {
                    acc[l][j].x = fmaf(S[l].x, w.x, fmaf(-D[l].y, w.y, acc[l][j].x));
                    acc[l][j].y = fmaf(S[l].y, w.x, fmaf(D[l].x, w.y, acc[l][j].y));
                }
                w = make_float2(w.x * ws.x - w.y * ws.y, w.x * ws.y + w.y * ws.x);
            }
        }
    }
    const int64_t plane = (int64_t)Np * f.ld;   // rows padded to f.ld entries (never read)
    const int64_t cell = (int64_t)y * f.ld + x;
#pragma unroll
    for (int l = 0; l < NELL; ++l) {
        const int64_t m = (int64_t)l * f.nshell + s - 1;
#pragma unroll
        for (int j = 0; j < ZC; ++j)
            if (z0 + j < Np) out[(m * Np + z0 + j) * plane + cell] = acc[l][j];
    }
}

// Power-of-two shell scales: e_s = -round(log2(sqrt(sum_{q in S_s} |H_q|^2))), so each map has
// O(1) RMS in FP32 / BF16 whatever the catalogue normalisation.
__global__ void k_shell_scales(int nshell, const double *shell_sum, float *scale_exp) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nshell) return;
    const double v = shell_sum[s];
    float e = 0.0f;
    if (v > 0.0 && isfinite(v)) e = (float)(-rint(0.5 * log2(v)));
    scale_exp[s] = e;
}


// ------------------------------------------------------------------ full-array forms
// fcomb on a full N^3 complex array exactly as powerI4.f90:251-378 receives it (the FFT of
// grid1 + i grid2), writing the full N^3 output: all eight ordered stores per iteration.
__global__ void k_fcomb_full_interlaced(int N, float theta, const float *wtab, float cf,
                                        const float2 *F, float2 *out) {
    const int h = N / 2 + 1;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)h * h * h) return;
    const int ix = (int)(t % h), iy = (int)((t / h) % h), iz = (int)(t / ((int64_t)h * h));
    const int icx = (N - ix) % N, icy = (N - iy) % N, icz = (N - iz) % N;
    const int mask = (ix == icx ? 1 : 0) | (iy == icy ? 2 : 0) | (iz == icz ? 4 : 0);
    auto at = [&](int code) {
        const int X = (code & 1) ? icx : ix, Y = (code & 2) ? icy : iy, Z = (code & 4) ? icz : iz;
        return ((int64_t)Z * N + Y) * N + X;
    };
    float2 orig[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) orig[k] = F[at(k)];
    auto cm = [&](int s) {
        float sn, cs;
        sincosf(-theta * (float)s, &sn, &cs);
        return make_float2(-sn, cs);
    };
    auto comb = [&](float2 a, float2 b, float2 c) {
        return cadd(cmul(a, make_float2(1.0f - c.x, -c.y)), cmul(conjf2(b), make_float2(1.0f + c.x, c.y)));
    };
    const float cfac = cf / (wtab[ix] * wtab[iy] * wtab[iz]);
    const float2 c000 = cscale(comb(orig[0], orig[7], cm(ix + iy + iz)), cfac);
    const float2 c001 = cscale(comb(orig[4], orig[3], cm(ix + iy - iz)), cfac);
    const float2 c010 = cscale(comb(orig[2], orig[5], cm(ix - iy + iz)), cfac);
    const float2 c011 = cscale(comb(orig[6], orig[1], cm(ix - iy - iz)), cfac);
    float2 mem[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) mem[k] = make_float2(0.0f, 0.0f);
    auto put = [&](int code, float2 v) {
        const int cc = code & ~mask;
#pragma unroll
        for (int j = 0; j < 8; ++j)
            if (j == cc) mem[j] = v;
    };
    auto get = [&](int code) {
        const int cc = code & ~mask;
        float2 v = mem[0];
#pragma unroll
        for (int j = 1; j < 8; ++j)
            if (j == cc) v = mem[j];
        return v;
    };
    put(0, c000);
    put(4, c001);
    put(2, c010);
    put(6, c011);
    put(1, conjf2(get(6)));
    put(5, conjf2(get(2)));
    put(3, conjf2(get(4)));
    put(7, conjf2(get(0)));
#pragma unroll
    for (int k = 0; k < 8; ++k) out[at(k)] = get(k);
}

__global__ void k_fcomb_full_single(int N, const float *wtab, float cf, const float2 *F,
                                    float2 *out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)N * N * N) return;
    const int x = (int)(t % N), y = (int)((t / N) % N), z = (int)(t / ((int64_t)N * N));
    if (x > N / 2) {
        out[t] = F[t];
        return;
    }
    const int ay = y <= N / 2 ? y : N - y, az = z <= N / 2 ? z : N - z;
    out[t] = cscale(F[t], cf / (wtab[x] * wtab[ay] * wtab[az]));
}

// compute_q2k / compute_q4k (powerI4.f90:409-508): on the ix <= N/2+1 half, Q = F k_i k_j / k^2
// (or k_i k_j k_l k_m / k^4), zero at k = 0; the other half is copied.  kF cancels in the ratio.
__global__ void k_qk(int N, int norder, int4 idx, const float2 *F, float2 *out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)N * N * N) return;
    const int x = (int)(t % N), y = (int)((t / N) % N), z = (int)(t / ((int64_t)N * N));
    if (x > N / 2) {
        out[t] = F[t];
        return;
    }
    if (x == 0 && y == 0 && z == 0) {
        out[t] = make_float2(0.0f, 0.0f);
        return;
    }
    const float v[3] = {(float)x, (float)(y <= N / 2 ? y : y - N), (float)(z <= N / 2 ? z : z - N)};
    const float k2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    float f = v[idx.x - 1] * v[idx.y - 1] / k2;
    if (norder == 4) f = f * (v[idx.z - 1] * v[idx.w - 1] / k2);
    out[t] = cscale(F[t], f);
}


// ------------------------------------------------------------------ maps: 2D c2r through LDS

// One work-group per plane (map m, z); four-step transforms with compile-time factors.
// Columns (length N' = A1 B1, only x' < C): thread (c, j < B1) loads rows q = j + B1 m, m < A1,
// transforms them over m, multiplies by W_N'^{j k1} and stores at k1 B1 + j; thread (c, k1 < A1)
// then transforms positions k1 B1 + j over j in place, so row y's value sits at
// (y mod A1) B1 + y / A1 (W_N'^e = twl[e], W_{N'/2}^e = twl[2e]).  Rows (the c2r of row y as an inverse DFT of length M = N'/2 = A2 B2):
// with X[k] the column outputs (zero for k >= C) and W = e^{2 pi i / N'},
// Z[k] = (X[k] + X*[M-k]) + i W^k (X[k] - X*[M-k]); the same two steps over Z give z, and
// x[2n] = Re z[n], x[2n+1] = Im z[n] are the real outputs.  LDS: twiddles W^k, the columns (stride
// N' + 1 entries) and a chunk of yc rows.
template <int NP, int A1, int B1, int A2, int B2>
__global__ __launch_bounds__(256) void k_c2r2d(C2r2dPlan p, const float2 *spec, const float2 *tw,
                                               float *maps) {
    static_assert(A1 * B1 == NP && 2 * A2 * B2 == NP, "factors of N' and N'/2");
    constexpr int M = NP / 2, CS = NP + 1;
    extern __shared__ float2 sm[];
    float2 *twl = sm, *col = sm + NP, *row = col + p.C * CS;
    const int plane = blockIdx.x;
    const int map = plane / NP, z = plane - map * NP;
    // the shell's disk: lines with x'^2 + q_y^2 > b are zero (the filter did not write them)
    const int b = p.shell_d2[2 * (map % p.nb) + 1];
    const int C = b < 0 ? 0 : min(p.C, isqrt_floor(b) + 1);
    for (int i = threadIdx.x; i < NP; i += blockDim.x) twl[i] = tw[i];
    const float2 *in = spec + (int64_t)plane * p.in_plane;
    for (int t = threadIdx.x; t < C * B1; t += blockDim.x) {
        const int j = t / C, c = t - j * C;
        float2 x[A1];
#pragma unroll
        for (int m = 0; m < A1; ++m) {
            const int q = j + B1 * m, qy = q <= NP / 2 ? q : q - NP;
            x[m] = c * c + qy * qy <= b ? in[(int64_t)q * p.ld + c] : make_float2(0.0f, 0.0f);
        }
        rdft<A1>(x);
        float2 *cc = col + c * CS + j;
#pragma unroll
        for (int k1 = 0; k1 < A1; ++k1) cc[k1 * B1] = k1 ? cmul(x[k1], twl[j * k1]) : x[0];
    }
    __syncthreads();
    for (int t = threadIdx.x; t < C * A1; t += blockDim.x) {
        const int k1 = t / C, c = t - k1 * C;
        float2 *cc = col + c * CS + k1 * B1;
        float2 x[B1];
#pragma unroll
        for (int j = 0; j < B1; ++j) x[j] = cc[j];
        rdft<B1>(x);
#pragma unroll
        for (int j = 0; j < B1; ++j) cc[j] = x[j];
    }
    __syncthreads();
    float *out = maps + (int64_t)map * p.Kstride + (int64_t)z * NP * NP;
    for (int y0 = 0; y0 < NP; y0 += p.yc) {
        const int ny = min(p.yc, NP - y0);
        for (int t = threadIdx.x; t < ny * B2; t += blockDim.x) {
            const int yl = t / B2, j = t - yl * B2, y = y0 + yl;
            const int py = (y % A1) * B1 + y / A1;
            float2 x[A2];
#pragma unroll
            for (int m = 0; m < A2; ++m) {
                const int k = j + B2 * m;
                float2 xk = make_float2(0.0f, 0.0f), xm = xk;
                if (k < C) xk = col[k * CS + py];
                if (M - k < C) xm = conjf2(col[(M - k) * CS + py]);
                x[m] = cadd(cadd(xk, xm), times_i(cmul(csub(xk, xm), twl[k])));
            }
            rdft<A2>(x);
            float2 *rr = row + yl * M + j;
#pragma unroll
            for (int k1 = 0; k1 < A2; ++k1) rr[k1 * B2] = k1 ? cmul(x[k1], twl[2 * j * k1]) : x[0];
        }
        __syncthreads();
        for (int t = threadIdx.x; t < ny * A2; t += blockDim.x) {
            const int yl = t / A2, k1 = t - yl * A2;
            const float2 *rr = row + yl * M + k1 * B2;
            float2 x[B2];
#pragma unroll
            for (int j = 0; j < B2; ++j) x[j] = rr[j];
            rdft<B2>(x);
            float2 *o = reinterpret_cast<float2 *>(out + (int64_t)(y0 + yl) * NP);
#pragma unroll
            for (int k2 = 0; k2 < B2; ++k2) o[k1 + A2 * k2] = x[k2];
        }
        __syncthreads();
    }
}


// Full-grid maps (N' = N, aliasing requests), two passes through HBM for planes too large for LDS.
// Column pass: work-group (plane p = blockIdx.x, block blockIdx.y of kColBlock columns x');
// thread (c, j < B) loads rows q = j + B m, m < A, of its column, transforms them over m,
// multiplies by W_N^{j k1} and stores them to LDS at k1 B + j; thread (c, k1 < A) transforms
// positions k1 B + j over j and writes X[k1 + A k2] back to row k1 + A k2 of the same column (in
// place: a work-group owns whole columns).
constexpr int kColBlock = 16;
template <int NP, int A, int B>
__global__ __launch_bounds__(256) void k_maps_cols(int C, int ld, int nb, const int32_t *shell_d2,
                                                   const float2 *tw, float2 *spec) {
    static_assert(A * B == NP, "factors of N");
    constexpr int CS = NP + 1;
    __shared__ float2 twl[NP];
    __shared__ float2 col[kColBlock * CS];
    const int cb0 = blockIdx.y * kColBlock;
    // the shell's disk: columns x' with x'^2 > b, and lines with x'^2 + q_y^2 > b, are zero (the
    // filter did not write them; the row pass does not read those columns)
    const int b = shell_d2[2 * ((int)(blockIdx.x / NP) % nb) + 1];
    if (b < 0) return;
    C = min(C, isqrt_floor(b) + 1);
    if (cb0 >= C) return;
    float2 *pl = spec + (int64_t)blockIdx.x * NP * ld;
    for (int i = threadIdx.x; i < NP; i += blockDim.x) twl[i] = tw[i];
    __syncthreads();
    for (int t = threadIdx.x; t < kColBlock * B; t += blockDim.x) {
        const int j = t / kColBlock, c = t - j * kColBlock;
        if (cb0 + c >= C) continue;
        const int x2 = (cb0 + c) * (cb0 + c);
        float2 x[A];
#pragma unroll
        for (int m = 0; m < A; ++m) {
            const int q = j + B * m, qy = q <= NP / 2 ? q : q - NP;
            x[m] = x2 + qy * qy <= b ? pl[(int64_t)q * ld + cb0 + c] : make_float2(0.0f, 0.0f);
        }
        rdft<A>(x);
        float2 *cc = col + c * CS + j;
#pragma unroll
        for (int k1 = 0; k1 < A; ++k1) cc[k1 * B] = k1 ? cmul(x[k1], twl[j * k1]) : x[0];
    }
    __syncthreads();
    for (int t = threadIdx.x; t < kColBlock * A; t += blockDim.x) {
        const int k1 = t / kColBlock, c = t - k1 * kColBlock;
        if (cb0 + c >= C) continue;
        const float2 *cc = col + c * CS + k1 * B;
        float2 x[B];
#pragma unroll
        for (int j = 0; j < B; ++j) x[j] = cc[j];
        rdft<B>(x);
#pragma unroll
        for (int k2 = 0; k2 < B; ++k2) pl[(int64_t)(k1 + A * k2) * ld + cb0 + c] = x[k2];
    }
}

// Row pass: work-group of kRowBlock rows (plane p = map N + z, row y); the c2r of a row as one
// inverse DFT of length M = N/2 = A2 B2 of Z[k] = (X[k] + X*[M-k]) + i W^k (X[k] - X*[M-k])
// (X[k] = 0 for k >= C), in the same two steps; x[2n] = Re z[n], x[2n+1] = Im z[n] are written to
// map p / N at z N^2 + y N.  kRowBlock: 32 rows for N = 128, 192, 16 for 256, 384 (one task per
// thread in both steps; 8 rows left half the threads idle, 22.6 -> 18.8 ms for 192 maps of 384^3
// inside the mixed stream).
template <int NP, int A2, int B2, int kRowBlock>
__global__ __launch_bounds__(256) void k_maps_rows(int C, int ld, int64_t nrows, int nb,
                                                   const int32_t *shell_d2, int64_t Kstride,
                                                   const float2 *tw, const float2 *spec,
                                                   float *maps) {
    static_assert(2 * A2 * B2 == NP, "factors of N/2");
    constexpr int M = NP / 2;
    __shared__ float2 twl[NP];
    __shared__ float2 row[kRowBlock * M];
    const int64_t r0 = (int64_t)blockIdx.x * kRowBlock;
    for (int i = threadIdx.x; i < NP; i += blockDim.x) twl[i] = tw[i];
    __syncthreads();
    for (int t = threadIdx.x; t < kRowBlock * B2; t += blockDim.x) {
        const int r = t / B2, j = t - r * B2;
        if (r0 + r >= nrows) continue;
        const float2 *X = spec + (r0 + r) * ld;     // rows of all planes are ld apart
        // columns beyond the shell's disk are zero (the column pass did not write them)
        const int b = shell_d2[2 * ((int)((r0 + r) / ((int64_t)NP * NP)) % nb) + 1];
        const int Cs = b < 0 ? 0 : min(C, isqrt_floor(b) + 1);
        float2 x[A2];
#pragma unroll
        for (int m = 0; m < A2; ++m) {
            const int k = j + B2 * m;
            float2 xk = make_float2(0.0f, 0.0f), xm = xk;
            if (k < Cs) xk = X[k];
            if (M - k < Cs) xm = conjf2(X[M - k]);
            x[m] = cadd(cadd(xk, xm), times_i(cmul(csub(xk, xm), twl[k])));
        }
        rdft<A2>(x);
        float2 *rr = row + r * M + j;
#pragma unroll
        for (int k1 = 0; k1 < A2; ++k1) rr[k1 * B2] = k1 ? cmul(x[k1], twl[2 * j * k1]) : x[0];
    }
    __syncthreads();
    // one task per thread (kRowBlock A2 <= 256): the transforms read the row buffer, then write
    // it back in natural order, and the rows go out with consecutive lanes on consecutive outputs
    static_assert(kRowBlock * A2 <= 256, "one task per thread");
    const int t = threadIdx.x, r = t / A2, k1 = t - r * A2;
    const bool task = t < kRowBlock * A2 && r0 + r < nrows;
    float2 x[B2];
    if (task) {
        const float2 *rr = row + r * M + k1 * B2;
#pragma unroll
        for (int j = 0; j < B2; ++j) x[j] = rr[j];
        rdft<B2>(x);
    }
    __syncthreads();
    if (task) {
        float2 *rr = row + r * M + k1;
#pragma unroll
        for (int k2 = 0; k2 < B2; ++k2) rr[A2 * k2] = x[k2];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < kRowBlock * M; e += blockDim.x) {
        const int rw = e / M, n = e - rw * M;
        const int64_t g = r0 + rw;
        if (g >= nrows) break;                       // rows in order: the rest are past the end
        const int64_t p = g / NP;                    // plane = map N + z
        const int y = (int)(g - p * NP);
        const int64_t map = p / NP, z = p - map * NP;
        // non-temporal: the full-grid maps (0.5 - 43 GB) pass through no cache before the GEMM
        // reads them (192 maps of 384^3: 15.9 -> 15.5 ms)
        typedef float f32x2v __attribute__((ext_vector_type(2)));
        const float2 v = row[rw * M + n];
        __builtin_nontemporal_store(
            f32x2v{v.x, v.y}, reinterpret_cast<f32x2v *>(maps + map * Kstride + (z * NP + y) * NP) + n);
    }
}

// The last axis of the interlaced density's FFT fused with the combine.  rocFFT has transformed
// every z plane of grid1 + i grid2 over (y, x) (G[z][y][x]); a work-group takes the octets
// (ix, iy, iz) of XB consecutive ix, one iy and every iz: it loads the 4 XB z-columns (X, Y),
// X in {ix, N - ix}, Y in {iy, N - iy}, that those octets read (runs of XB consecutive x: 128
// bytes for XB = 16), transforms each over z in LDS (forward DFT of length N = A B: the inverse
// four-step transform of the conjugate, conjugated again when read), then evaluates the octets
// as k_combine_interlaced does.  In a column, z = j + B m sits at m (B + 1) + j and, after the two
// steps, F[k1 + A k2] at k1 (B + 1) + k2: the second step's lanes (one k1 each) read B + 1
// entries apart, on distinct LDS banks.  Threads: 4 XB columns x NT / (4 XB) z rows.
template <int N, int A, int B, int XB, int NT>
__global__ __launch_bounds__(NT) void k_zfft_combine(const float2 *ph, const float *wtab,
                                                          float cf, const float2 *tw,
                                                          const float2 *G, float2 *delta) {
    static_assert(A * B == N, "factors of N");
    static_assert(NT % (4 * XB) == 0 && N % (NT / (4 * XB)) == 0, "threads per column set");
    constexpr int NC = 4 * XB, PB = B + 1, CS = A * PB + 1, H = N / 2 + 1;
    __shared__ float2 col[NC * CS];
    // work-groups are dealt round-robin to the 8 XCDs: hardware order h runs logical block L,
    // consecutive L on one XCD, so the two blocks whose 64-byte runs share a 128-byte line
    // (x0 = 16 k and 16 k + 8, on both sides) read it through one L2
    int bx, iy;
    {
        const int total = gridDim.x * gridDim.y;
        const int h = blockIdx.x + gridDim.x * blockIdx.y;
        const int q = total / 8, r = total % 8, xcd = h % 8;
        const int L = xcd * q + min(xcd, r) + h / 8;
        iy = L / gridDim.x;
        bx = L - iy * gridDim.x;
    }
    const int x0 = bx * XB;
    const int icy = (N - iy) % N;
    // column c = xi + XB (side + 2 ysel): X = side ? N - ix : ix (mod N), Y = ysel ? icy : iy;
    // ix past N / 2 (the last block) loads the cells of ix mod N and writes nothing
    {
        const int xi = threadIdx.x % XB, side = (threadIdx.x / XB) & 1;
        const int ysel = (threadIdx.x / (2 * XB)) & 1, zr = threadIdx.x / (4 * XB);
        const int ix = (x0 + xi) % N;
        const int X = side ? (N - ix) % N : ix, Y = ysel ? icy : iy;
        const int c = xi + XB * (side + 2 * ysel);
        constexpr int ZS = NT / (4 * XB), NZ = N / ZS;
        // all of a thread's loads issue before its LDS stores (NZ in flight per thread)
        float2 v[NZ];
#pragma unroll
        for (int i = 0; i < NZ; ++i) v[i] = G[((int64_t)(zr + ZS * i) * N + Y) * N + X];
#pragma unroll
        for (int i = 0; i < NZ; ++i) {
            const int z = zr + ZS * i;
            col[c * CS + (z / B) * PB + z % B] = make_float2(v[i].x, -v[i].y);
        }
    }
    __syncthreads();
    // step 1: per column and j < B, the A entries z = j + B m over m, times W_N^{j k1}, in place
    for (int t = threadIdx.x; t < NC * B; t += NT) {
        const int c = t / B, j = t - c * B;
        float2 *cc = col + c * CS + j;
        float2 x[A];
#pragma unroll
        for (int m = 0; m < A; ++m) x[m] = cc[m * PB];
        rdft<A>(x);
#pragma unroll
        for (int k1 = 0; k1 < A; ++k1) cc[k1 * PB] = k1 ? cmul(x[k1], tw[j * k1]) : x[0];
    }
    __syncthreads();
    // step 2: per column and k1 < A, the B entries k1 (B + 1) + j over j, in place
    for (int t = threadIdx.x; t < NC * A; t += NT) {
        const int c = t / A, k1 = t - c * A;
        float2 *cc = col + c * CS + k1 * PB;
        float2 x[B];
#pragma unroll
        for (int j = 0; j < B; ++j) x[j] = cc[j];
        rdft<B>(x);
#pragma unroll
        for (int k2 = 0; k2 < B; ++k2) cc[k2] = x[k2];
    }
    __syncthreads();
    auto F = [&](int c, int Z) {     // forward transform of column c at Z
        const float2 v = col[c * CS + (Z % A) * PB + Z / A];
        return make_float2(v.x, -v.y);
    };
    for (int t = threadIdx.x; t < XB * H; t += NT) {
        const int xi = t % XB, iz = t / XB, ix = x0 + xi;
        if (ix > N / 2) continue;
        const int icz = (N - iz) % N;
        float2 orig[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int c = xi + XB * ((k & 1) + 2 * ((k >> 1) & 1));
            orig[k] = F(c, (k & 4) ? icz : iz);
        }
        combine_octet(N, ix, iy, iz, orig, ph, wtab, cf, delta);
    }
}

// The middle axis of the interlaced density's FFT, in place: a work-group loads the 16 y-columns
// (z, x0 .. x0 + 15) of G[z][y][x] (128-byte runs), transforms each over y in LDS (forward DFT of
// length N = A B as the inverse four-step transform of the conjugate; same layout as
// k_zfft_combine) and writes them back.  Threads: 16 x by NT / 16 y rows.
template <int N, int A, int B, int NT>
__global__ __launch_bounds__(NT) void k_yfft(const float2 *tw, float2 *G) {
    static_assert(A * B == N && N % 16 == 0, "factors of N");
    constexpr int XB = 16, RW = NT / XB, PB = B + 1, CS = A * PB + 1, NY = N / RW;
    static_assert(N % RW == 0, "rows per thread");
    __shared__ float2 col[XB * CS];
    const int x0 = blockIdx.x * XB, z = blockIdx.y;
    const int xi = threadIdx.x % XB, yr = threadIdx.x / XB;
    float2 *g = G + (int64_t)z * N * N + x0 + xi;
    {
        float2 v[NY];
#pragma unroll
        for (int i = 0; i < NY; ++i) v[i] = g[(int64_t)(yr + RW * i) * N];
#pragma unroll
        for (int i = 0; i < NY; ++i) {
            const int y = yr + RW * i;
            col[xi * CS + (y / B) * PB + y % B] = make_float2(v[i].x, -v[i].y);
        }
    }
    __syncthreads();
    // step 1: per column and j < B, the A entries y = j + B m over m, times W_N^{j k1}, in place
    for (int t = threadIdx.x; t < XB * B; t += NT) {
        const int c = t / B, j = t - c * B;
        float2 *cc = col + c * CS + j;
        float2 x[A];
#pragma unroll
        for (int m = 0; m < A; ++m) x[m] = cc[m * PB];
        rdft<A>(x);
#pragma unroll
        for (int k1 = 0; k1 < A; ++k1) cc[k1 * PB] = k1 ? cmul(x[k1], tw[j * k1]) : x[0];
    }
    __syncthreads();
    // step 2: per column and k1 < A, the B entries k1 (B + 1) + j over j, in place
    for (int t = threadIdx.x; t < XB * A; t += NT) {
        const int c = t / A, k1 = t - c * A;
        float2 *cc = col + c * CS + k1 * PB;
        float2 x[B];
#pragma unroll
        for (int j = 0; j < B; ++j) x[j] = cc[j];
        rdft<B>(x);
#pragma unroll
        for (int k2 = 0; k2 < B; ++k2) cc[k2] = x[k2];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NY; ++i) {
        const int y = yr + RW * i;
        const float2 f = col[xi * CS + (y % A) * PB + y / A];
        g[(int64_t)y * N] = make_float2(f.x, -f.y);
    }
}

}  // namespace

void launch_yfft(int N, const float2 *tw, float2 *G, hipStream_t s) {
    PBI4_CHECK(zfft_combine_supported(N), PBI4_EUNSUPPORTED, "y transform: grids 128, 192, 256, 384, 512");
    const dim3 grid(N / 16, N);
    switch (N) {
        // 256 threads: with 512 the LDS still holds 4 work-groups per CU and 256^3 took 52.8
        // against 52.1 us
        case 128: hipLaunchKernelGGL((k_yfft<128, 16, 8, 256>), grid, dim3(256), 0, s, tw, G); break;
        case 192: hipLaunchKernelGGL((k_yfft<192, 16, 12, 256>), grid, dim3(256), 0, s, tw, G); break;
        case 256: hipLaunchKernelGGL((k_yfft<256, 16, 16, 256>), grid, dim3(256), 0, s, tw, G); break;
        case 384: hipLaunchKernelGGL((k_yfft<384, 16, 24, 256>), grid, dim3(256), 0, s, tw, G); break;
        default: hipLaunchKernelGGL((k_yfft<512, 16, 32, 256>), grid, dim3(256), 0, s, tw, G); break;
    }
    PBI4_HIP(hipGetLastError());
}

void launch_fcomb_full(int N, int interlacing, const float *wtab, float cf, const float2 *F,
                       float2 *out, hipStream_t s) {
    if (interlacing) {
        const float theta = (float)((double)kTwoPiF32 / 2.0 / (double)N);
        const int64_t n = (int64_t)(N / 2 + 1) * (N / 2 + 1) * (N / 2 + 1);
        hipLaunchKernelGGL(k_fcomb_full_interlaced, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0,
                           s, N, theta, wtab, cf, F, out);
    } else {
        const int64_t n = (int64_t)N * N * N;
        hipLaunchKernelGGL(k_fcomb_full_single, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s, N,
                           wtab, cf, F, out);
    }
    PBI4_HIP(hipGetLastError());
}

// powerI4.f90:380-407 (reduce_array) as a gather: output (x', y', z') of the N grid reads the L
// grid at x = x' and y = m(y'), z = m(z') with m(i) = i below N - N/2 and L - N + i above.  The
// reference writes each output cell from up to four (iy, iz) sources in the order (iy, iz),
// (jcy, iz), (iy, jcz), (jcy, jcz); on the Nyquist index of an even N the later, negative-
// frequency source is the one that stays, which is what m gives.  Rows are contiguous in x, so
// both sides are coalesced.  Output x > N/2 is zero (the reference's intent(out) array is zero-
// filled).
template <typename TI, typename TO>
__device__ __forceinline__ TO convert_complex(TI v) {
    TO o;
    o.x = v.x;
    o.y = v.y;
    return o;
}

template <typename TI, typename TO>
__global__ __launch_bounds__(256) void k_reduce_array(int L, int lx, int N, int nx, const TI *din,
                                                      TO *dout) {
    const int64_t n = (int64_t)N * N * nx;
    const int lo = N - N / 2;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int x = (int)(t % nx);
        const int64_t r = t / nx;
        const int y = (int)(r % N), z = (int)(r / N);
        TO v;
        v.x = 0;
        v.y = 0;
        if (x <= N / 2) {
            const int sy = y < lo ? y : L - N + y, sz = z < lo ? z : L - N + z;
            v = convert_complex<TI, TO>(din[((int64_t)sz * L + sy) * lx + x]);
        }
        dout[t] = v;
    }
}

void launch_reduce_array(int L, int lx, int N, int nx, int in_bytes, int out_bytes, const void *din,
                         void *dout, hipStream_t s) {
    const int64_t n = (int64_t)N * N * nx;
    const unsigned blocks = (unsigned)std::min<int64_t>(ceil_div(n, 256), 1 << 20);
    if (in_bytes == 16 && out_bytes == 16)
        hipLaunchKernelGGL((k_reduce_array<double2, double2>), dim3(blocks), dim3(256), 0, s, L, lx, N,
                           nx, static_cast<const double2 *>(din), static_cast<double2 *>(dout));
    else if (in_bytes == 16)
        hipLaunchKernelGGL((k_reduce_array<double2, float2>), dim3(blocks), dim3(256), 0, s, L, lx, N,
                           nx, static_cast<const double2 *>(din), static_cast<float2 *>(dout));
    else if (out_bytes == 16)
        hipLaunchKernelGGL((k_reduce_array<float2, double2>), dim3(blocks), dim3(256), 0, s, L, lx, N,
                           nx, static_cast<const float2 *>(din), static_cast<double2 *>(dout));
    else
        hipLaunchKernelGGL((k_reduce_array<float2, float2>), dim3(blocks), dim3(256), 0, s, L, lx, N,
                           nx, static_cast<const float2 *>(din), static_cast<float2 *>(dout));
    PBI4_HIP(hipGetLastError());
}

void launch_qk(int N, int norder, const int *idx, const float2 *F, float2 *out, hipStream_t s) {
    const int64_t n = (int64_t)N * N * N;
    const int4 id = make_int4(idx[0], idx[1], norder == 4 ? idx[2] : 1, norder == 4 ? idx[3] : 1);
    hipLaunchKernelGGL(k_qk, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s, N, norder, id, F, out);
    PBI4_HIP(hipGetLastError());
}

void launch_combine(int N, int interlacing, const float *wtab, const float2 *ph, float cf,
                    const float2 *G1, float2 *delta, hipStream_t s) {
    const int h = N / 2 + 1;
    if (interlacing) {
        PBI4_CHECK(ph != nullptr, PBI4_EINVAL, "the interlaced combine needs the phase table");
        const int64_t n = (int64_t)h * h * h;
        hipLaunchKernelGGL(k_combine_interlaced, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s,
                           N, ph, wtab, cf, G1, delta);
    } else {
        const int64_t n = (int64_t)N * N * h;
        hipLaunchKernelGGL(k_combine_single, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s, N,
                           wtab, cf, G1, delta);
    }
    PBI4_HIP(hipGetLastError());
}

bool zfft_combine_supported(int N) { return N == 128 || N == 192 || N == 256 || N == 384 || N == 512; }

void launch_zfft_combine(int N, const float *wtab, const float2 *ph, const float2 *tw, float cf,
                         const float2 *G, float2 *delta, hipStream_t s) {
    PBI4_CHECK(ph != nullptr && tw != nullptr, PBI4_EINVAL, "the fused z transform needs its tables");
    auto run = [&](auto kern, int xb, int nt) {
        const dim3 grid((unsigned)ceil_div(N / 2 + 1, xb), (unsigned)(N / 2 + 1));
        hipLaunchKernelGGL(kern, grid, dim3(nt), 0, s, ph, wtab, cf, tw, G, delta);
    };
    // 8 x-columns per side (64-byte runs, 66 KB of LDS at 256^3, two work-groups per CU): 16
    // (128-byte runs, 140 KB, one work-group per CU) took 78.5 against 90.5 us alone at 256^3 but
    // the six-slot serving stream ran slower with it (435 against 451 catalogues/s), its LDS
    // keeping other slots' work-groups off the CU.  512 threads (256 on 512) rather than 256
    // (128): the LDS caps the work-groups per CU, so more threads per work-group keep more
    // loads in flight (256^3: 61.3 -> 59.0 us; 512^3: 848 -> 708 us)
    switch (N) {
        case 128: run(k_zfft_combine<128, 16, 8, 8, 512>, 8, 512); break;
        case 192: run(k_zfft_combine<192, 16, 12, 8, 512>, 8, 512); break;
        case 256: run(k_zfft_combine<256, 16, 16, 8, 512>, 8, 512); break;
        case 384: run(k_zfft_combine<384, 16, 24, 8, 512>, 8, 512); break;
        // 4 x-columns per side: 16 columns of 512 entries are 68 KB of LDS
        case 512: run(k_zfft_combine<512, 16, 32, 4, 256>, 4, 256); break;
        default: fail(PBI4_EUNSUPPORTED, "no fused z transform for this grid");
    }
    PBI4_HIP(hipGetLastError());
}

// per-work-group rows of FP32 partials, then the FP64 sum of each quantity
static size_t bin_rows_floats(const BinPlan &b) {
    return align_up((size_t)b.nwg * (5 * b.nbins_pk + b.nshell), 2);
}

size_t bin_partials_floats(const BinPlan &b) {
    return bin_rows_floats(b) + 2 * (size_t)kBinSegs * (5 * b.nbins_pk + b.nshell);
}

void launch_bin(const BinPlan &b, const float2 *delta, float *partials, hipStream_t s) {
    const size_t lds = sizeof(float) * kBinWaves * (5 * b.nbins_pk + b.nshell);
    PBI4_CHECK(lds <= 160 * 1024, PBI4_EUNSUPPORTED, "too many bins for the binning kernel");
    const bool extra = b.f2 || b.f4 || b.cross;
    auto launch = [&](auto kern) {
        if (lds > 64 * 1024)
            PBI4_HIP(hipFuncSetAttribute((const void *)kern,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL(kern, dim3(b.nwg), dim3(kBinThreads), lds, s, b, delta, partials);
    };
    if (b.multipoles)
        extra ? launch(k_bin<5, true>) : launch(k_bin<5, false>);
    else
        extra ? launch(k_bin<3, true>) : launch(k_bin<3, false>);
    PBI4_HIP(hipGetLastError());
}

void launch_bin_finalize(const BinPlan &b, const float *partials, double kF3, double *pk_out,
                         double *shell_sum, hipStream_t s) {
    const int n = b.nbins_pk + b.nshell;
    if (n <= 0) return;
    const int nq = 5 * b.nbins_pk + b.nshell;
    // the FP64 segment sums follow the rows in the same buffer (8-byte aligned: the rows are an
    // even count of floats)
    double *qseg = reinterpret_cast<double *>(const_cast<float *>(partials) + bin_rows_floats(b));
    hipLaunchKernelGGL(k_bin_reduce, dim3((unsigned)ceil_div(nq, 64), kBinSegs), dim3(256), 0, s,
                       partials, b.nwg, nq, qseg);
    hipLaunchKernelGGL(k_bin_finalize, dim3((unsigned)ceil_div(n, 64)), dim3(64), 0, s, b, qseg,
                       kF3, pk_out, shell_sum);
    PBI4_HIP(hipGetLastError());
}


// The maps grids with a compiled plane transform: N' = 100 (the 256^3 workload's 16 shells of
// 2 k_F), 150 (48 shells of k_F at 512^3), 200 (32 shells of 2 k_F at 1024^3).
bool make_c2r2d_plan(int Np, int C, int ld, int64_t Kstride, C2r2dPlan &p) {
    if (Np != 100 && Np != 150 && Np != 200) return false;
    if (C < 1 || C > Np / 2 + 1 || ld < C) return false;
    p = C2r2dPlan{};
    p.Np = Np;
    p.C = C;
    p.ld = ld;
    p.in_plane = (int64_t)Np * ld;
    p.Kstride = Kstride;
    // rows per chunk, at least 48 (the row steps then keep the 256 threads busy: 8-row chunks on
    // 150^3 took 2.35 ms for 144 maps, 48-row chunks 1.22 ms): as many as fit 64 KB (several
    // work-groups per CU), else 160 KB, in chunks of equal size
    const size_t fixed = sizeof(float2) * ((size_t)Np + (size_t)C * (Np + 1));
    const size_t per_row = sizeof(float2) * (size_t)(Np / 2);
    const int ymin = std::min(Np, 48);
    for (size_t budget : {(size_t)64 << 10, (size_t)160 << 10}) {
        if (fixed + (size_t)ymin * per_row > budget) continue;
        const int most = (int)std::min<size_t>(Np, (budget - fixed) / per_row);
        const int chunks = (Np + most - 1) / most;
        p.yc = (Np + chunks - 1) / chunks;
        p.lds = fixed + per_row * p.yc;
        return true;
    }
    return false;
}

void launch_c2r2d(const C2r2dPlan &p, const float2 *spectra, const float2 *tw, float *maps,
                  int nmaps, hipStream_t s) {
    auto launch = [&](auto kern) {
        if (p.lds > 64 * 1024)
            PBI4_HIP(hipFuncSetAttribute((const void *)kern,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)p.lds));
        hipLaunchKernelGGL(kern, dim3((unsigned)((int64_t)nmaps * p.Np)), dim3(256), p.lds, s, p,
                           spectra, tw, maps);
    };
    switch (p.Np) {
        case 100: launch(k_c2r2d<100, 10, 10, 10, 5>); break;
        case 150: launch(k_c2r2d<150, 10, 15, 15, 5>); break;
        default: launch(k_c2r2d<200, 10, 20, 10, 10>); break;
    }
    PBI4_HIP(hipGetLastError());
}

bool maps_two_pass_supported(int Np) { return Np == 128 || Np == 192 || Np == 256 || Np == 384; }

void launch_maps_two_pass(int Np, int C, int ld, int nmaps, int nb, const int32_t *shell_d2,
                          int64_t Kstride, const float2 *tw, float2 *spectra, float *maps,
                          hipStream_t s) {
    PBI4_CHECK(maps_two_pass_supported(Np) && C >= 1 && C <= Np / 2 && ld >= C, PBI4_EINVAL,
               "full-grid maps transform: unsupported size");
    const dim3 gc((unsigned)((int64_t)nmaps * Np), (unsigned)ceil_div(C, kColBlock));
    const int64_t nrows = (int64_t)nmaps * Np * Np;
    auto gr = [&](int rb) { return dim3((unsigned)ceil_div(nrows, rb)); };
    auto run = [&](auto kc, auto kr, int rb) {
        hipLaunchKernelGGL(kc, gc, dim3(256), 0, s, C, ld, nb, shell_d2, tw, spectra);
        hipLaunchKernelGGL(kr, gr(rb), dim3(256), 0, s, C, ld, nrows, nb, shell_d2, Kstride, tw,
                           spectra, maps);
    };
    switch (Np) {
        case 128: run(k_maps_cols<128, 8, 16>, k_maps_rows<128, 8, 8, 32>, 32); break;
        case 192: run(k_maps_cols<192, 12, 16>, k_maps_rows<192, 8, 12, 32>, 32); break;
        case 256: run(k_maps_cols<256, 16, 16>, k_maps_rows<256, 8, 16, 16>, 16); break;
        default: run(k_maps_cols<384, 16, 24>, k_maps_rows<384, 12, 16, 16>, 16); break;
    }
    PBI4_HIP(hipGetLastError());
}

void launch_filter_z(const FilterPlan &f, const float2 *delta, const double *shell_sum,
                     const float2 *tw, float2 *spectra, float *scale_exp, hipStream_t s) {
    PBI4_CHECK(f.Np < f.N ? 3 * f.R < f.Np : (f.Np == f.N && 2 * f.R < f.Np), PBI4_EINVAL,
               "the fused z filter needs a band-limited maps grid, or the full grid with 2R < N");
    hipLaunchKernelGGL(k_shell_scales, dim3((unsigned)ceil_div(f.nshell, 64)), dim3(64), 0, s,
                       f.nshell, shell_sum, scale_exp);
    PBI4_HIP(hipGetLastError());
    PBI4_CHECK(f.shell_d2 != nullptr, PBI4_EINVAL, "the fused z filter needs the shell d2 ranges");
    PBI4_CHECK(f.cols >= 0 && f.cols <= f.Np / 2 + 1, PBI4_EINVAL, "bad fused z filter columns");
    PBI4_CHECK(f.ld >= (f.cols ? f.cols : f.Np / 2 + 1), PBI4_EINVAL,
               "the fused z filter's row stride is too small");
    const int lines = f.Np * (f.cols ? f.cols : f.Np / 2 + 1);
    // planes per thread: 8 with three multipole maps, else 25 (fewer planes per thread, more
    // threads in flight: the mode loads of a line's runs are what a thread waits on)
    if (f.nell == 3)
        hipLaunchKernelGGL((k_filter_z<8, 3>),
                           dim3((unsigned)ceil_div(lines, 256), (unsigned)f.nshell,
                                (unsigned)ceil_div(f.Np, 8)),
                           dim3(256), 0, s, f, delta, scale_exp, tw, spectra);
    else
        hipLaunchKernelGGL((k_filter_z<25, 1>),
                           dim3((unsigned)ceil_div(lines, 256), (unsigned)f.nshell,
                                (unsigned)ceil_div(f.Np, 25)),
                           dim3(256), 0, s, f, delta, scale_exp, tw, spectra);
    PBI4_HIP(hipGetLastError());
}

void launch_filter(const FilterPlan &f, const float2 *delta, const double *shell_sum,
                   const double *nmodes, float2 *spectra, float *scale_exp, hipStream_t s) {
    (void)nmodes;
    hipLaunchKernelGGL(k_shell_scales, dim3((unsigned)ceil_div(f.nshell, 64)), dim3(64), 0, s,
                       f.nshell, shell_sum, scale_exp);
    const int64_t per_map = (int64_t)f.Np * f.Np * (f.Np / 2 + 1);
    hipLaunchKernelGGL(k_filter, dim3((unsigned)ceil_div(per_map, 256)), dim3(256), 0, s, f, delta,
                       scale_exp, spectra);
    PBI4_HIP(hipGetLastError());
}

}  // namespace pbi4
