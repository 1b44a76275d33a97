// Mass assignment with interlacing (powerI4.f90:68-249) on gfx950.
//
// Particles are first binned by tile (T^3 cells, T <= 16) with a counting sort, then one
// work-group per tile deposits its particles into LDS copies of the tile plus a 2-cell halo on
// each side ((T+4)^3 cells per grid), accumulating in 32-bit fixed point (2^-18 .. 2^-20 units,
// chosen per tile so no cell can overflow) so the result does not depend on the order of the
// LDS atomics.  Each tile writes its halo-extended block to a region buffer in pieces grouped by
// the tile that owns the cells; a gather pass then sums, for every grid cell, the (1 to 8)
// region entries that cover it in a fixed order.  No global atomics touch grid memory, and the
// output is bitwise reproducible.  When the density follows in the same call, the gather is fused
// with the first axis of the density's FFT (k_gather_xfft: the same cells, transformed along x
// before they leave LDS).
//
// Stencils relative to the cell c = floor(u) containing the particle (u = N * pos), written
// with the fractional part f = u - c, follow the reference exactly:
//   PCS  grid 1: cells c-1..c+2, h = f                         (powerI4.f90:164-186)
//        grid 2: position u - 1/2                              (vt = vr - 0.5, :162)
//   TSC  grid 1: cells nint(u)-1..+1, h = u - nint(u)          (:187-207)
//   CIC  grid 1: cells c+1, c+2, h = f  (a uniform +1 shift)   (:208-224)
//        single grid: h = u - nint(u)                          (:114-124)
//   NGP  cell nint(u)                                          (:225-233)
#include "common.h"
#include "dft.h"

#include <type_traits>

namespace pbi4 {

namespace {

// A sorted particle: per axis 4 bits of cell within the tile and 28 bits of fraction, stored
// as one 12-byte record (the scatter writes it with one store, the deposit loads it with one).
struct Pk3 {
    uint32_t c[3];
};

constexpr int kHalo = 2;          // region cells below the tile origin
constexpr int kMaxR = 16 + 4;     // T <= 16
// LDS plane stride of the deposit is R^2 + kPadZ words: with R^2 = 400 a z step of 4 cells is a
// multiple of 64 words, so the particles of one atomic instruction (4 apart in tile order, which
// for catalogues in lattice order means 4 cells apart along z) would all hit the same LDS banks
constexpr int kPadZ = 9;
constexpr int kDepositThreads = 1024;
constexpr double kTileParticles = 8192.0;   // target mean particles per deposit tile
constexpr int kSortThreads = 256;     // several work-groups per CU overlap their phases
constexpr int kHistPer = 8;       // particles per thread per histogram pass (loads in flight)

struct PosXform {
    int N, pos_mode;
    float box, rsd[3];
    int has_vel;
};

// Grid coordinate of one component: integer cell c in [0, N) and fraction f in [0, 1), with
// u = N * p evaluated with its FMA residual so f is accurate to ~2^-24 cells for any N.
__device__ __forceinline__ void grid_coord(const PosXform &x, float p, int &c, float &f) {
    const float Nf = (float)x.N;
    float u = Nf * p;
    float e = fmaf(Nf, p, -u);
    float fl = floorf(u);
    f = (u - fl) + e;
    int ci = (int)fl;
    if (f < 0.0f) {
        ci -= 1;
        f += 1.0f;
    } else if (f >= 1.0f) {
        ci += 1;
        f -= 1.0f;
    }
    if (f >= 1.0f) f = 0.99999994f;
    // ci is in [-1, N] for components in [0, 1]; the modulo only serves inputs outside
    if (ci >= x.N) ci -= x.N;
    else if (ci < 0) ci += x.N;
    if ((unsigned)ci >= (unsigned)x.N) {
        ci %= x.N;
        if (ci < 0) ci += x.N;
    }
    c = ci;
}

// One particle's raw components.  Loads are not guarded per thread (the caller clamps the index
// into range) so that the loads of several particles issue as one batch; a guarded load becomes
// a branch with a wait after it.
struct RawParticle {
    float p[3], v[3];
};

__device__ __forceinline__ RawParticle load_particle(const PosXform &x, const float *pos,
                                                     const float *vel, int64_t i, int64_t npart) {
    RawParticle r;
#pragma unroll
    for (int d = 0; d < 3; ++d) r.p[d] = pos[(int64_t)d * npart + i];
    // the condition is the same for every thread (kernel arguments): a scalar branch, no
    // divergence, and no second load of the position for components without a shift
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (x.has_vel && x.rsd[d] != 0.0f) r.v[d] = vel[(int64_t)d * npart + i];
        else r.v[d] = 0.0f;
    }
    return r;
}

// Normalised position component as the reference computes it: (s + box) % box / box in
// float32 after the RSD shift (pbi4.py:154, 242-243), or the given value if already in [0,1).
__device__ __forceinline__ float norm_pos(const PosXform &x, const RawParticle &r, int d) {
    const float p = r.p[d];
    if (x.pos_mode == PBI4_POS_NORMALISED) return p;
    const float s = (x.has_vel && x.rsd[d] != 0.0f) ? fmaf(x.rsd[d], r.v[d], p) : p;
    // (s + box) % box: for s + box in [0, 3 box) subtracting box is exact (Sterbenz), so this
    // equals fmod there; fmodf covers anything further out
    float t = s + x.box;
    if (t >= x.box) t -= x.box;
    if (t >= x.box) t -= x.box;
    if (t < 0.0f || t >= x.box) {
        t = fmodf(t, x.box);
        if (t < 0.0f) t += x.box;
    }
    const float q = t / x.box;
    return fminf(q, 1.0f);
}

__device__ __forceinline__ int tile_of(const PosXform &x, int T, int NT, const RawParticle &r,
                                       uint32_t *packed) {
    int c[3];
    float f[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) grid_coord(x, norm_pos(x, r, d), c[d], f[d]);
    int t[3];
    const float invT = 1.0f / (float)T;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        // c / T for 0 <= c < 2^23 from the float reciprocal, corrected by one either way (24-bit
        // products: exact for these operands, and full rate)
        int q = (int)((float)c[d] * invT);
        if (__mul24(q, T) > c[d]) --q;
        else if (__mul24(q + 1, T) <= c[d]) ++q;
        t[d] = q;
        if (packed) {
            uint32_t loc = (uint32_t)(c[d] - __mul24(t[d], T));
            uint32_t fr = (uint32_t)(f[d] * 268435456.0f);   // 2^28
            if (fr > 0x0FFFFFFFu) fr = 0x0FFFFFFFu;
            packed[d] = (loc << 28) | fr;
        }
    }
    // NT^2 < 2^23 for any grid below 46 000 cells, and the tile count fits 31 bits
    return __mul24(__mul24(t[2], NT) + t[1], NT) + t[0];
}

// ------------------------------------------------------------------ counting sort

// DPP lane moves of a 64-lane wave (gfx9 encodings): 0x110 + n row_shr:n (lane i <- i - n inside
// a 16-lane row), 0x142 / 0x143 row_bcast:15 / :31, 0x138 wave_shr:1 (i <- i - 1), 0x130
// wave_shl:1 (i <- i + 1).  VALU only, no LDS traffic.
template <int CTRL>
__device__ __forceinline__ int dpp_mov(int v) {
    return __builtin_amdgcn_mov_dpp(v, CTRL, 0xf, 0xf, false);
}

// Runs of equal keys among consecutive lanes of a wave (catalogues stored in spatial order put
// neighbouring particles in the same tile): rank = position in the run, head / tail = first /
// last lane of the run.  The head below each lane is an inclusive max-scan of head positions.
// One LDS atomic per run instead of one per lane.
__device__ __forceinline__ void run_rank(int key, int &rank, bool &tail) {
    const int lane = threadIdx.x & 63, rl = lane & 15;
    const int prev = dpp_mov<0x138>(key), next = dpp_mov<0x130>(key);
    const bool head = lane == 0 || prev != key;
    tail = lane == 63 || next != key;
    int h = head ? lane : 0, t;
    t = dpp_mov<0x111>(h);
    if (rl >= 1) h = max(h, t);
    t = dpp_mov<0x112>(h);
    if (rl >= 2) h = max(h, t);
    t = dpp_mov<0x114>(h);
    if (rl >= 4) h = max(h, t);
    t = dpp_mov<0x118>(h);
    if (rl >= 8) h = max(h, t);
    t = dpp_mov<0x142>(h);
    if ((lane & 31) >= 16) h = max(h, t);
    t = dpp_mov<0x143>(h);
    if (lane >= 32) h = max(h, t);
    rank = lane - h;
}

// Tile counters of a work-group in LDS.  Direct: one counter per tile (ntiles <= kLdsDirectMax).
// Hashed: kHashSlots (key, value) slots, open addressing from a multiplicative hash, keys set by
// compare-and-swap; a run whose tile finds no slot within kHashProbes counts straight into global
// memory.  A work-group of spatially ordered particles touches a few tens of tiles, so the table
// stays sparse for any grid.
constexpr int64_t kLdsDirectMax = 8192;
constexpr int kHashSlots = 4096;
constexpr int kHashProbes = 32;
constexpr uint32_t kEmpty = 0xFFFFFFFFu;

__device__ __forceinline__ int hash_slot(uint32_t *keys, uint32_t t, uint32_t *nkeys) {
    uint32_t h = (t * 2654435761u) >> 20;                 // 12 bits: kHashSlots = 4096
#pragma unroll 1
    for (int probe = 0; probe < kHashProbes; ++probe) {
        const uint32_t k = keys[h];
        if (k == t) return (int)h;
        if (k == kEmpty) {
            const uint32_t prev = atomicCAS(&keys[h], kEmpty, t);
            if (prev == kEmpty) {
                if (nkeys) atomicAdd(nkeys, 1u);
                return (int)h;
            }
            if (prev == t) return (int)h;
        }
        h = (h + 1) & (kHashSlots - 1);
    }
    return -1;
}

// Direct counters (one per tile): the final flush adds the work-group's count of each tile it
// touched to the tile's total and keeps the total before the add, this work-group's first slot
// inside the tile, in wg_base[work-group][tile] (entries of untouched tiles are not written).
// The scatter pass, over the same particles split the same way, starts its counters there and
// ranks in LDS alone.  Hashed counters (more tiles than LDS holds) leave wg_base untouched.
__global__ __launch_bounds__(kSortThreads) void k_tile_hist(PosXform x, int T, int NT,
                                                            int64_t ntiles, const float *pos,
                                                            const float *vel, int64_t npart,
                                                            uint32_t *counts, uint32_t *wg_base,
                                                            int hashed) {
    // hashed: values lhist[0..S), keys after them, then the number of keys in use; the table is
    // flushed to global memory and cleared whenever it is a quarter full, so a catalogue in any
    // order keeps its probes short
    extern __shared__ uint32_t lhist[];
    uint32_t *keys = lhist + kHashSlots;
    uint32_t *nkeys = lhist + 2 * kHashSlots;
    const int64_t nctr = hashed ? kHashSlots : ntiles;
    const int64_t per_wg = ceil_div(npart, (int64_t)gridDim.x);
    const int64_t beg = (int64_t)blockIdx.x * per_wg;
    const int64_t end = min(npart, beg + per_wg);
    for (int64_t t = threadIdx.x; t < nctr; t += blockDim.x) {
        lhist[t] = 0;
        if (hashed) keys[t] = kEmpty;
    }
    if (hashed && threadIdx.x == 0) *nkeys = 0;
    __syncthreads();
    for (int64_t i0 = beg; i0 < end; i0 += (int64_t)blockDim.x * kHistPer) {
        RawParticle r[kHistPer];
#pragma unroll
        for (int k = 0; k < kHistPer; ++k) {
            const int64_t i = i0 + threadIdx.x + (int64_t)k * blockDim.x;
            r[k] = load_particle(x, pos, vel, i < end ? i : beg, npart);
        }
        int t[kHistPer];
#pragma unroll
        for (int k = 0; k < kHistPer; ++k) {
            const int64_t i = i0 + threadIdx.x + (int64_t)k * blockDim.x;
            t[k] = i < end ? tile_of(x, T, NT, r[k], nullptr) : -1;
        }
#pragma unroll
        for (int k = 0; k < kHistPer; ++k) {
            int rank;
            bool tail;
            run_rank(t[k], rank, tail);
            if (t[k] >= 0 && tail) {
                const int slot = hashed ? hash_slot(keys, (uint32_t)t[k], nkeys) : t[k];
                if (slot >= 0) atomicAdd(&lhist[slot], (uint32_t)(rank + 1));
                else atomicAdd(&counts[t[k]], (uint32_t)(rank + 1));
            }
        }
        if (hashed) {
            __syncthreads();
            if (*nkeys > kHashSlots / 4) {            // uniform: every thread reads one value
                for (int t = threadIdx.x; t < kHashSlots; t += blockDim.x) {
                    if (lhist[t]) atomicAdd(&counts[keys[t]], lhist[t]);
                    lhist[t] = 0;
                    keys[t] = kEmpty;
                }
                __syncthreads();
                if (threadIdx.x == 0) *nkeys = 0;
            }
            __syncthreads();
        }
    }
    __syncthreads();
    if (hashed) {
        for (int64_t t = threadIdx.x; t < nctr; t += blockDim.x)
            if (lhist[t]) atomicAdd(&counts[keys[t]], lhist[t]);
        return;
    }
    // four tiles per thread at a time: the four returning atomics are in flight together
    uint32_t *row = wg_base + (int64_t)blockIdx.x * ntiles;
    for (int64_t t0 = threadIdx.x; t0 < nctr; t0 += 4 * (int64_t)blockDim.x) {
        uint32_t b[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int64_t t = t0 + (int64_t)k * blockDim.x;
            const uint32_t c = t < nctr ? lhist[t] : 0u;
            b[k] = c ? atomicAdd(&counts[t], c) : 0u;
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int64_t t = t0 + (int64_t)k * blockDim.x;
            if (t < nctr && lhist[t]) row[t] = b[k];
        }
    }
}

// Exclusive scan of the tile counts by one work-group; offsets[ntiles] = npart; cursor = offsets.
__global__ __launch_bounds__(1024) void k_tile_scan(const uint32_t *counts, int64_t ntiles,
                                                    uint32_t *offsets, uint32_t *cursor) {
    __shared__ uint32_t part[1024];
    const int64_t per = ceil_div(ntiles, 1024);
    const int64_t b = threadIdx.x * per, e = min(ntiles, b + per);
    uint32_t s = 0;
    for (int64_t t = b; t < e; ++t) s += counts[t];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        uint32_t v = threadIdx.x >= off ? part[threadIdx.x - off] : 0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = threadIdx.x ? part[threadIdx.x - 1] : 0;
    for (int64_t t = b; t < e; ++t) {
        offsets[t] = run;
        cursor[t] = run;
        run += counts[t];
    }
    if (threadIdx.x == 1023) offsets[ntiles] = part[1023];
}

constexpr int kScatterPer = 4;   // particles per thread per LDS round

// Direct counters: one LDS counter per tile starting at the work-group's first slot in the tile
// (offsets + wg_base from the histogram pass); a run's atomic add returns its final slots, so the
// pass needs no global atomics and no barrier per round, whatever the particles' order.
__global__ __launch_bounds__(kSortThreads) void k_tile_scatter_direct(PosXform x, int T, int NT,
                                                                      int64_t ntiles,
                                                                      const float *pos,
                                                                      const float *vel,
                                                                      int64_t npart,
                                                                      const uint32_t *offsets,
                                                                      const uint32_t *wg_base,
                                                                      Pk3 *spk) {
    extern __shared__ uint32_t lhist[];
    const int64_t per_wg = ceil_div(npart, (int64_t)gridDim.x);
    const int64_t beg = (int64_t)blockIdx.x * per_wg;
    const int64_t end = min(npart, beg + per_wg);
    // entries of tiles this work-group does not touch were never written and are never used
    const uint32_t *row = wg_base + (int64_t)blockIdx.x * ntiles;
    for (int64_t t = threadIdx.x; t < ntiles; t += blockDim.x) lhist[t] = offsets[t] + row[t];
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int64_t round = (int64_t)blockDim.x * kScatterPer;
    for (int64_t r0 = beg; r0 < end; r0 += round) {
        RawParticle raw[kScatterPer];
#pragma unroll
        for (int k = 0; k < kScatterPer; ++k) {
            const int64_t i = r0 + threadIdx.x + (int64_t)k * blockDim.x;
            raw[k] = load_particle(x, pos, vel, i < end ? i : beg, npart);
        }
#pragma unroll
        for (int k = 0; k < kScatterPer; ++k) {
            const int64_t i = r0 + threadIdx.x + (int64_t)k * blockDim.x;
            uint32_t pk[3];
            const int tl = i < end ? tile_of(x, T, NT, raw[k], pk) : -1;
            int rr;
            bool tail;
            run_rank(tl, rr, tail);
            uint32_t base = 0;
            if (tl >= 0 && tail) base = atomicAdd(&lhist[tl], (uint32_t)(rr + 1));
            // the tail hands the run's first slot to the head (non-tails push to lane + 1, never
            // a head, so every head receives exactly one value); members read it from the head
            const int dst = tail ? lane - rr : lane + 1;
            const uint32_t b_head = (uint32_t)__builtin_amdgcn_ds_permute(4 * dst, (int)base);
            const uint32_t slot = (uint32_t)__shfl((int)b_head, lane - rr) + (uint32_t)rr;
            if (tl < 0 || slot >= (uint32_t)npart) continue;   // a bad slot cannot happen
            spk[slot] = Pk3{{pk[0], pk[1], pk[2]}};
        }
    }
}

// Hashed counters (more tiles than LDS holds): per round, counters in a hash table, one global
// atomic per touched tile on the cursors, then the writes.
__global__ __launch_bounds__(kSortThreads) void k_tile_scatter(PosXform x, int T, int NT,
                                                               int64_t ntiles, const float *pos,
                                                               const float *vel, int64_t npart,
                                                               uint32_t *cursor, Pk3 *spk,
                                                               int hashed) {
    // LDS: the round's counters (one per tile, or the hashed slots followed by their keys), then
    // the list of counters this round touched and its length; only those are flushed and cleared
    // (a round of spatially ordered particles touches a few tens of tiles)
    extern __shared__ uint32_t lhist[];
    const int64_t nctr = hashed ? kHashSlots : ntiles;
    uint32_t *keys = lhist + kHashSlots;
    uint32_t *touched = lhist + (hashed ? 2 * kHashSlots : ntiles);
    const int64_t round = (int64_t)blockDim.x * kScatterPer;
    uint32_t *ntouched = touched + round;
    const int64_t per_wg = ceil_div(npart, (int64_t)gridDim.x);
    const int64_t beg = (int64_t)blockIdx.x * per_wg;
    const int64_t end = min(npart, beg + per_wg);
    for (int64_t t = threadIdx.x; t < nctr; t += blockDim.x) {
        lhist[t] = 0;
        if (hashed) keys[t] = kEmpty;
    }
    if (threadIdx.x == 0) *ntouched = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63;
    for (int64_t r0 = beg; r0 < end; r0 += round) {
        uint32_t pk[kScatterPer][3];
        int ctr[kScatterPer];          // counter of the particle's run; -1: none; -2: direct
        uint32_t rank[kScatterPer];    // rank in the round's counter, or the final slot if direct
        RawParticle raw[kScatterPer];
#pragma unroll
        for (int k = 0; k < kScatterPer; ++k) {
            const int64_t i = r0 + threadIdx.x + (int64_t)k * blockDim.x;
            raw[k] = load_particle(x, pos, vel, i < end ? i : beg, npart);
        }
#pragma unroll
        for (int k = 0; k < kScatterPer; ++k) {
            int64_t i = r0 + threadIdx.x + (int64_t)k * blockDim.x;
            const int tl = i < end ? tile_of(x, T, NT, raw[k], pk[k]) : -1;
            int rr;
            bool tail;
            run_rank(tl, rr, tail);
            uint32_t base = 0;
            int c = -1;
            if (tl >= 0 && tail) {
                c = hashed ? hash_slot(keys, (uint32_t)tl, nullptr) : tl;
                if (c >= 0) {
                    base = atomicAdd(&lhist[c], (uint32_t)(rr + 1));
                    if (base == 0) touched[atomicAdd(ntouched, 1u)] = (uint32_t)c;
                } else {
                    base = atomicAdd(&cursor[tl], (uint32_t)(rr + 1));   // final slots
                    c = -2;
                }
            }
            // the tail hands the run's base and counter to the head (non-tails push to lane + 1,
            // never a head, so every head receives exactly one value); members read them from
            // the head
            const int dst = tail ? lane - rr : lane + 1;
            const uint32_t b_head = (uint32_t)__builtin_amdgcn_ds_permute(4 * dst, (int)base);
            const int c_head = __builtin_amdgcn_ds_permute(4 * dst, c);
            rank[k] = (uint32_t)__shfl((int)b_head, lane - rr) + (uint32_t)rr;
            ctr[k] = tl >= 0 ? __shfl(c_head, lane - rr) : -1;
        }
        __syncthreads();
        const uint32_t nt = *ntouched;
        for (uint32_t i = threadIdx.x; i < nt; i += blockDim.x) {
            const uint32_t c = touched[i];
            lhist[c] = atomicAdd(&cursor[hashed ? keys[c] : c], lhist[c]);
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kScatterPer; ++k) {
            if (ctr[k] == -1) continue;
            const uint32_t slot = ctr[k] >= 0 ? lhist[ctr[k]] + rank[k] : rank[k];
            if (slot >= (uint32_t)npart) continue;   // cannot happen; keeps a bad rank in bounds
            spk[slot] = Pk3{{pk[k][0], pk[k][1], pk[k][2]}};
        }
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < nt; i += blockDim.x) {
            const uint32_t c = touched[i];
            lhist[c] = 0;
            if (hashed) keys[c] = kEmpty;
        }
        if (threadIdx.x == 0) *ntouched = 0;
        __syncthreads();
    }
}

// ------------------------------------------------------------------ deposit

// Region layout.  A tile deposits into the (T+4)^3 cells [-2, T+2)^3 around it (region
// coordinates a = cell - origin + 2).  Along each axis a region coordinate has a class: 0 for
// a < 2 (cells of the previous tile, its local T-2, T-1), 1 for the T home cells, 2 for
// a >= T+2 (the next tile's local 0, 1).  The region is stored as 27 pieces, one per class
// triple, in (cz, cy, cx) order, each piece z-y-x row-major.  Every piece is read by exactly one
// gather work-group (the tile that owns its cells) as one contiguous block.
__device__ __forceinline__ int region_class(int a, int T) { return a < kHalo ? 0 : a < T + kHalo ? 1 : 2; }
__device__ __forceinline__ int class_start(int c, int T) { return c == 0 ? 0 : c == 1 ? kHalo : T + kHalo; }
__device__ __forceinline__ int class_extent(int c, int T) { return c == 1 ? T : kHalo; }

// Every factor is below R <= 20 and every product below R^3: the 24-bit multiplies (full rate,
// where a 32-bit v_mul_lo issues at a quarter of it) are exact.
__device__ __forceinline__ int region_index(int a, int b, int c, int T, int R) {
    const int ka = region_class(a, T), kb = region_class(b, T), kc = region_class(c, T);
    const int sa = class_start(ka, T), sb = class_start(kb, T), sc = class_start(kc, T);
    const int ea = class_extent(ka, T), eb = class_extent(kb, T), ec = class_extent(kc, T);
    // pieces before (kc, kb, ka): all of the lower z classes, then the lower y classes of this z
    // class, then the lower x classes of this (z, y) class
    const int off = __mul24(__mul24(sc, R), R) + __mul24(ec, __mul24(sb, R) + __mul24(eb, sa));
    return off + __mul24(__mul24(c - sc, eb) + (b - sb), ea) + (a - sa);
}


// Per-dimension stencil start (relative to the tile origin, before the halo offset) and
// weights, for one grid of the given order (see the table at the top of this file).
template <int ORDER>
__device__ __forceinline__ void stencil(int grid_sel, int single, int c, float f, int &start,
                                        float w[4]) {
    if (ORDER == 4) {
        float h;
        if (grid_sel == 0) {
            start = c - 1;
            h = f;
        } else if (f >= 0.5f) {
            start = c - 1;
            h = f - 0.5f;
        } else {
            start = c - 2;
            h = f + 0.5f;
        }
        float t = 1.0f - h, h2 = h * h;
        w[0] = (t * t) * t * (1.0f / 6.0f);
        w[1] = 2.0f / 3.0f + (0.5f * h - 1.0f) * h2;
        w[3] = h2 * h * (1.0f / 6.0f);
        w[2] = 1.0f - w[0] - w[1] - w[3];
    } else if (ORDER == 3) {
        float h;
        if (grid_sel == 0) {
            int up = f >= 0.5f;
            start = c - 1 + up;
            h = f - (float)up;
        } else {
            start = c - 1;
            h = f - 0.5f;
        }
        float t = 0.5f - h;
        w[0] = 0.5f * (t * t);
        w[1] = 0.75f - h * h;
        w[2] = 1.0f - w[0] - w[1];
        w[3] = 0.0f;
    } else if (ORDER == 2) {
        float h;
        if (single) {
            start = c + 1;
            h = f - (f >= 0.5f ? 1.0f : 0.0f);
        } else if (grid_sel == 0) {
            start = c + 1;
            h = f;
        } else {
            int up = f >= 0.5f;
            start = c + up;
            h = f + 0.5f - (float)up;
        }
        w[0] = 1.0f - h;
        w[1] = h;
        w[2] = w[3] = 0.0f;
    } else {
        start = grid_sel == 0 ? c + (f >= 0.5f ? 1 : 0) : c;
        w[0] = 1.0f;
        w[1] = w[2] = w[3] = 0.0f;
    }
}

// The stencil of stencil<ORDER> for ORDER 3 and 4 as its start and local coordinate h, and the
// weight of stencil point k as a polynomial in h: w_k = a_k + b_k h + c_k h^2 + d_k h^3 (PCS: the
// cubic B-spline pieces, h in [0, 1); TSC: the quadratic pieces, h in [-0.5, 0.5), point 3 zero).
template <int ORDER>
__device__ __forceinline__ void stencil_start(int grid_sel, int c, float f, int &start, float &h) {
    if (ORDER == 4) {
        if (grid_sel == 0) { start = c - 1; h = f; }
        else if (f >= 0.5f) { start = c - 1; h = f - 0.5f; }
        else { start = c - 2; h = f + 0.5f; }
    } else {
        if (grid_sel == 0) {
            const int up = f >= 0.5f;
            start = c - 1 + up;
            h = f - (float)up;
        } else {
            start = c - 1;
            h = f - 0.5f;
        }
    }
}
struct WeightPoly {
    float a, b, c, d;
};
template <int ORDER>
__device__ __forceinline__ WeightPoly weight_poly(int k) {
    WeightPoly p;
    if (ORDER == 4) {
        p.a = k == 0 ? 1.0f / 6.0f : k == 1 ? 2.0f / 3.0f : k == 2 ? 1.0f / 6.0f : 0.0f;
        p.b = k == 0 ? -0.5f : k == 2 ? 0.5f : 0.0f;
        p.c = k == 0 ? 0.5f : k == 1 ? -1.0f : k == 2 ? 0.5f : 0.0f;
        p.d = k == 0 ? -1.0f / 6.0f : k == 1 ? 0.5f : k == 2 ? -0.5f : 1.0f / 6.0f;
    } else {
        p.a = k == 0 ? 0.125f : k == 1 ? 0.75f : k == 2 ? 0.125f : 0.0f;
        p.b = k == 0 ? -0.5f : k == 2 ? 0.5f : 0.0f;
        p.c = k == 0 ? 0.5f : k == 1 ? -1.0f : k == 2 ? 0.5f : 0.0f;
        p.d = 0.0f;
    }
    return p;
}
__device__ __forceinline__ float eval_poly(const WeightPoly &p, float h) {
    return __builtin_fmaf(__builtin_fmaf(__builtin_fmaf(p.d, h, p.c), h, p.b), h, p.a);
}
// v of lane k of this lane's quad (DPP quad_perm broadcast)
template <int K>
__device__ __forceinline__ float quad_bcast(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), K * 0x55, 0xF, 0xF, true));
}

// LDS accumulation is fixed point at 2^-s, s = 20 for TSC/PCS, 19 for CIC/NGP, 18 for the
// single-grid CIC of powerI4.f90:114-124; each contribution is rounded to the nearest unit.  A cell
// of the region receives contributions from the particles whose home cells lie in a window of
// ORDER + 1 cells along each axis (either grid), at most B = min(n, (ORDER+1)^2 S) particles for a
// tile of n particles whose fullest window of ORDER + 1 home cells along z holds S, each at most
// w_max = (2/3)^3 (PCS), (3/4)^3 (TSC), 1 (CIC / NGP) or 1.5^3 (single-grid CIC, weights in
// [-0.5, 1.5]) in magnitude.  If B (w_max 2^s + 1/2) < 2^32 (2^31 for the single-grid CIC, whose
// sums can be negative) the tile accumulates in 32-bit LDS cells in one pass; otherwise in 64-bit
// cells, one slab of the region at a time (a grid, or half of the single grid) in the same LDS,
// each slab a pass over the tile's particles.  Either way a tile's cells are exact integer sums
// over all its particles, independent of their order, so the grids are bitwise reproducible
// whatever order the sort leaves the particles in.
template <int ORDER>
__device__ __forceinline__ int deposit_per_thread(int single) {
    return (ORDER >= 3) ? 4 : (ORDER == 2 && single) ? 1 : 2;
}
template <int ORDER>
__device__ __forceinline__ int deposit_scale_bits(int single) {
    return (ORDER >= 3) ? 20 : (ORDER == 2 && single) ? 18 : 19;
}

// Lane layout of the deposit: X lanes share one particle, lane ix owning the stencil's x-offset
// ix and looping over (y, z); a wave deposits P = 64 / X particles per atomic instruction.  The
// wave's particles are preloaded with coalesced loads (lane L holds particles L + 64 m of the
// wave's chunk) and handed out by shuffles so that the P particles of one step are X apart in
// tile order: for particles that arrive in spatial order their x-footprints tile a row instead of
// overlapping, and a cell crowded with particles does not serialise one instruction.
template <int ORDER>
struct DepositLanes {
    static constexpr int X = ORDER == 3 ? 4 : ORDER;
    static constexpr int P = 64 / X;
};
constexpr int kDepositPerMax = 4;

// w[i] for a lane-dependent i, as bit masks: a select of array elements can otherwise be turned
// into a dynamically indexed (scratch) load
template <int ORDER>
__device__ __forceinline__ float pick4(const float w[4], int i) {
    uint32_t r = 0;
#pragma unroll
    for (int k = 0; k < (ORDER == 3 ? 4 : ORDER); ++k) r |= __float_as_uint(w[k]) & (0u - (uint32_t)(i == k));
    return __uint_as_float(r);
}

// Position in the (T+4)^3 region of LDS cell l (l = (c R + b) R + a) and of l + blockDim.x.
struct RegionCursor {
    int g, a, b, c;
    __device__ void init(int l, int R) {
        const int R2 = R * R, R3 = R2 * R;
        g = l / R3;
        l -= g * R3;
        c = l / R2;
        l -= c * R2;
        b = l / R;
        a = l - b * R;
    }
    __device__ void advance(int da, int db, int dc, int R) {
        a += da;
        if (a >= R) { a -= R; ++b; }
        b += db;
        if (b >= R) { b -= R; ++c; }
        c += dc;
        while (c >= R) { c -= R; ++g; }
    }
};

// Nearest integer of a * b for |a * b| < 2^22, in one FMA and one integer subtraction: adding
// 1.5 * 2^23 puts the exact product's rounding to an integer (ties to even) into the FMA's single
// rounding, and the integer is the low mantissa bits.
__device__ __forceinline__ int mul_rint(float a, float b) {
    return __float_as_int(__builtin_fmaf(a, b, 12582912.0f)) - 0x4B400000;
}

// Deposits particles [c0, cend) of a tile into the LDS accumulators (one preload per thread).
// RC: the region edge R when known at compile time (T = 16), so the 16 row offsets of a lane's
// atomics are instruction immediates; 0 = R at run time.
__device__ __forceinline__ void lds_add(int *a, int v) { atomicAdd(a, v); }
__device__ __forceinline__ void lds_add(unsigned long long *a, int v) {
    atomicAdd(a, (unsigned long long)(long long)v);        // two's complement: signed add
}

// The part of the region one deposit pass accumulates: grid g, planes z in [zlo, zhi) (the
// whole region when SLAB is false).
struct Slab {
    int g, zlo, zhi;
};

// The particle of step s for this lane, q of lane j X + s (j = lane / X): a lane of the same
// quad, so a DPP quad permute (a VALU move) instead of an LDS permute on the pipe the atomics use.
// X = 4: broadcast of quad lane s; X = 2: lanes (0, 1) from lane s, (2, 3) from lane 2 + s.
template <int X>
__device__ __forceinline__ uint32_t step_particle(uint32_t v, int s) {
    const int iv = (int)v;
    if constexpr (X == 1) {
        return v;
    } else if constexpr (X == 2) {
        return (uint32_t)(s == 0 ? dpp_mov<0xA0>(iv) : dpp_mov<0xF5>(iv));
    } else {
        switch (s) {   // s is the same on the whole wave
        case 0: return (uint32_t)dpp_mov<0x00>(iv);
        case 1: return (uint32_t)dpp_mov<0x55>(iv);
        case 2: return (uint32_t)dpp_mov<0xAA>(iv);
        default: return (uint32_t)dpp_mov<0xFF>(iv);
        }
    }
}

template <int ORDER, int NG, int RC, bool SLAB, typename Acc>
__device__ __forceinline__ void deposit_chunk(Acc *acc, int R_, int GS, int per, float fix,
                                              Slab sl, uint32_t c0, uint32_t cend,
                                              const Pk3 *spk) {
    constexpr int single = NG == 1;
    constexpr int X = DepositLanes<ORDER>::X;
    const int R = RC ? RC : R_;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int ix = lane % X, j = lane / X;
    const uint32_t cb = c0 + (uint32_t)wave * 64u * (uint32_t)per;   // this wave's part
    uint32_t q[kDepositPerMax][3];
#pragma unroll
    for (int m = 0; m < kDepositPerMax; ++m) {
        const uint32_t p = cb + (uint32_t)(lane + 64 * m);
        const bool ok = m < per && p < cend;
        Pk3 v{{0u, 0u, 0u}};
        if (ok) v = spk[p];
        q[m][0] = v.c[0];
        q[m][1] = v.c[1];
        q[m][2] = v.c[2];
    }
    __syncthreads();   // the accumulators are cleared before the first chunk of a tile
    // orders 3 and 4: the four lanes of a particle are a quad; lane ix evaluates the weight of
    // stencil point ix along each axis and the quad shares them by DPP broadcasts
    const WeightPoly wp = weight_poly<ORDER>(ix);
#pragma unroll
    for (int m = 0; m < kDepositPerMax; ++m) {
        if (m >= per || cb + 64u * m >= cend) continue;
#pragma unroll 1
        for (int s = 0; s < X; ++s) {
            const int src = j * X + s;
            uint32_t qq[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) qq[d] = step_particle<X>(q[m][d], s);
            if constexpr (X == 4) {
                // src is the same on the quad: the quad stays whole for the broadcasts
                if (cb + (uint32_t)(64 * m + src) >= cend) continue;
                int c[3];
                float f[3];
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    c[d] = (int)(qq[d] >> 28);
                    f[d] = (float)(qq[d] & 0x0FFFFFFFu) * 3.7252902984619141e-09f;   // 2^-28
                }
#pragma unroll
                for (int g = 0; g < NG; ++g) {
                    if (SLAB && g != sl.g) continue;
                    int st[3];
                    float wl[3];
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        float h;
                        stencil_start<ORDER>(g, c[d], f[d], st[d], h);
                        wl[d] = eval_poly(wp, h);
                    }
                    const float wx = wl[0] * fix;
                    float wy[4], wxz[4];
                    wy[0] = quad_bcast<0>(wl[1]);
                    wy[1] = quad_bcast<1>(wl[1]);
                    wy[2] = quad_bcast<2>(wl[1]);
                    wy[3] = quad_bcast<3>(wl[1]);
                    wxz[0] = wx * quad_bcast<0>(wl[2]);
                    wxz[1] = wx * quad_bcast<1>(wl[2]);
                    wxz[2] = wx * quad_bcast<2>(wl[2]);
                    wxz[3] = wx * quad_bcast<3>(wl[2]);
                    if (ix >= ORDER) continue;
                    const int PZ = R * R + kPadZ;
                    const int z0 = st[2] + kHalo;
                    const int base = (SLAB ? 0 : g * GS) + (st[1] + kHalo) * R + st[0] + kHalo + ix;
                    // |w_x w_y w_z fix| <= fix <= 2^20: within mul_rint's range
#pragma unroll
                    for (int iz = 0; iz < ORDER; ++iz) {
                        if (SLAB && (z0 + iz < sl.zlo || z0 + iz >= sl.zhi)) continue;
                        Acc *a = acc + base + (z0 + iz - (SLAB ? sl.zlo : 0)) * PZ;
#pragma unroll
                        for (int iy = 0; iy < ORDER; ++iy)
                            lds_add(a + iy * R, mul_rint(wy[iy], wxz[iz]));
                    }
                }
                continue;
            }
            if (cb + (uint32_t)(64 * m + src) >= cend || ix >= ORDER) continue;
            int c[3];
            float f[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                c[d] = (int)(qq[d] >> 28);
                f[d] = (float)(qq[d] & 0x0FFFFFFFu) * 3.7252902984619141e-09f;   // 2^-28
            }
#pragma unroll
            for (int g = 0; g < NG; ++g) {
                if (SLAB && g != sl.g) continue;
                int st[3];
                float w[3][4];
#pragma unroll
                for (int d = 0; d < 3; ++d) stencil<ORDER>(g, single, c[d], f[d], st[d], w[d]);
                const float wx = pick4<ORDER>(w[0], ix) * fix;
                float wxz[ORDER];
#pragma unroll
                for (int iz = 0; iz < ORDER; ++iz) wxz[iz] = wx * w[2][iz];
                const int PZ = R * R + kPadZ;
                const int z0 = st[2] + kHalo;
                const int base = (SLAB ? 0 : g * GS) + (st[1] + kHalo) * R + st[0] + kHalo + ix;
                // |w_x w_y w_z fix| <= max(1, 1.5^3 2^-2) 2^20 (single-grid CIC: weights in
                // [-0.5, 1.5] at scale <= 2^18): within mul_rint's range
#pragma unroll
                for (int iz = 0; iz < ORDER; ++iz) {
                    if (SLAB && (z0 + iz < sl.zlo || z0 + iz >= sl.zhi)) continue;
                    Acc *a = acc + base + (z0 + iz - (SLAB ? sl.zlo : 0)) * PZ;
#pragma unroll
                    for (int iy = 0; iy < ORDER; ++iy)
                        lds_add(a + iy * R, mul_rint(w[1][iy], wxz[iz]));
                }
            }
        }
    }
}

// Writes the region of LDS accumulator `acc` (R^3 cells per grid at plane stride PZ) scaled by
// `inv` into the tile's region `out` (27 pieces, region_index).
template <typename Acc, typename Scale>
__device__ __forceinline__ void flush_region(const Acc *acc, int T, int R, int PZ, int GS, int ncell,
                                             Scale inv, float *out) {
    const int R3 = R * R * R;
    // flush cursor: this thread's first LDS cell and the stride blockDim.x in (a, b, c)
    const int sc = blockDim.x / (R * R), sb = (blockDim.x / R) % R, sa = blockDim.x % R;
    RegionCursor rc;
    rc.init(threadIdx.x, R);
    for (int i = threadIdx.x; i < ncell; i += blockDim.x) {
        const Acc v = acc[rc.g * GS + rc.c * PZ + rc.b * R + rc.a];
        out[rc.g * R3 + region_index(rc.a, rc.b, rc.c, T, R)] = (float)((Scale)v * inv);
        rc.advance(sa, sb, sc, R);
    }
}

template <int ORDER>
__device__ __forceinline__ double deposit_wmax(int single) {
    return ORDER == 4 ? 8.0 / 27.0 : ORDER == 3 ? 27.0 / 64.0 : (ORDER == 2 && single) ? 3.375 : 1.0;
}

// Largest value over the work-group (all threads get it; `red` holds blockDim.x / 64 ints).
__device__ __forceinline__ int block_max(int v, int *red) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
    const int wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if ((threadIdx.x & 63) == 0) red[wave] = v;
    __syncthreads();
    v = red[0];
    for (int w = 1; w < nw; ++w) v = max(v, red[w]);
    return v;
}

// One work-group per tile: particles per home cell, the tile's accumulator width (above), the
// LDS pass(es) over all its particles and the region write.
template <int ORDER, int NG, int RC>
__global__ __launch_bounds__(kDepositThreads) void k_deposit(int T, int R_,
                                                             const uint32_t *offsets,
                                                             const Pk3 *spk,
                                                             float *regions) {
    // 32-bit cells of NG grids, or 64-bit cells of one slab: NG kMaxR (kMaxR^2 + kPadZ) / 2 of
    // them, a whole grid for NG = 2 and ceil(R / 2) planes for NG = 1
    __shared__ __align__(16) int acc[NG * kMaxR * (kMaxR * kMaxR + kPadZ)];
    __shared__ int red[kDepositThreads / 64];
    const int R = RC ? RC : R_;
    const int R3 = R * R * R, ncell = NG * R3;
    const int PZ = R * R + kPadZ, GS = R * PZ;        // padded LDS plane and grid strides
    constexpr int single = NG == 1;
    const int per = deposit_per_thread<ORDER>(single);
    const int64_t tile = blockIdx.x;
    float *out = regions + tile * (int64_t)ncell;
    const uint32_t b = offsets[tile], e = offsets[tile + 1];
    const uint32_t chunk = (uint32_t)blockDim.x * per;      // particles per preload
    const int sbits = deposit_scale_bits<ORDER>(single);
    const float fix = (float)(1 << sbits);
    // 32-bit cells hold B (w_max 2^s + 1/2) < 2^32 (weights >= 0; 2^31 for the single-grid CIC)
    // for B contributions per cell: B = n always fits for small tiles; otherwise bound B by the
    // fullest (ORDER+1)-cell window of home-cell counts along z times (ORDER+1)^2
    constexpr bool sgn = ORDER == 2 && NG == 1;
    constexpr int W = ORDER + 1;
    const double lim = sgn ? 2147483648.0 : 4294967296.0;
    const double per_contrib = deposit_wmax<ORDER>(single) * (double)fix + 0.5;
    bool wide = false;
    if ((double)(e - b) * per_contrib >= lim) {
        const int T3 = T * T * T;                            // <= 4096 counters at acc[0..)
        for (int i = threadIdx.x; i < T3; i += blockDim.x) acc[i] = 0;
        __syncthreads();
        for (uint32_t i = b + threadIdx.x; i < e; i += blockDim.x)
            atomicAdd(&acc[((spk[i].c[2] >> 28) * T + (spk[i].c[1] >> 28)) * T + (spk[i].c[0] >> 28)], 1);
        __syncthreads();
        int best = 0;
        if ((int)threadIdx.x < T * T) {                     // one (x, y) column per thread
            const int col = threadIdx.x;
            int win = 0;
            for (int z = 0; z < T + W - 1; ++z) {
                if (z < T) win += acc[z * T * T + col];
                if (z >= W) win -= acc[(z - W) * T * T + col];
                best = max(best, win);
            }
        }
        best = block_max(best, red);
        const double bound = fmin((double)(e - b), (double)(W * W) * best);
        wide = bound * per_contrib >= lim;
    }
    __syncthreads();                                         // counters read before the clear
    if (!wide) {
        for (int i = threadIdx.x; i < NG * GS; i += blockDim.x) acc[i] = 0;
        for (uint32_t c0 = b; c0 < e; c0 += chunk)
            deposit_chunk<ORDER, NG, RC, false>(acc, R, GS, per, fix, Slab{0, 0, R}, c0,
                                                min(e, c0 + chunk), spk);
        __syncthreads();
        if (sgn) flush_region(acc, T, R, PZ, GS, ncell, 1.0f / fix, out);
        else flush_region(reinterpret_cast<unsigned *>(acc), T, R, PZ, GS, ncell, 1.0f / fix, out);
        return;
    }
    unsigned long long *acc64 = reinterpret_cast<unsigned long long *>(acc);
    const double inv = 1.0 / (double)fix;
    const int half = (R + 1) / 2;
    for (int pass = 0; pass < 2; ++pass) {
        const Slab sl = NG == 2 ? Slab{pass, 0, R} : Slab{0, pass ? half : 0, pass ? R : half};
        const int nsl = (sl.zhi - sl.zlo) * PZ;
        for (int i = threadIdx.x; i < nsl; i += blockDim.x) acc64[i] = 0;
        for (uint32_t c0 = b; c0 < e; c0 += chunk)
            deposit_chunk<ORDER, NG, RC, true>(acc64, R, GS, per, fix, sl, c0, min(e, c0 + chunk),
                                               spk);
        __syncthreads();
        for (int i = threadIdx.x; i < (sl.zhi - sl.zlo) * R * R; i += blockDim.x) {
            const int c = i / (R * R), rem = i - c * (R * R), bb = rem / R, aa = rem - bb * R;
            out[sl.g * R3 + region_index(aa, bb, sl.zlo + c, T, R)] =
                (float)((double)(long long)acc64[c * PZ + bb * R + aa] * inv);
        }
        __syncthreads();
    }
}

// grid[g][z][y][x] = sum of the region entries covering (x,y,z), in a fixed order: for each
// (x, y) option (home first) its home-z entry plus its z-neighbour's entry.  One work-group per
// (tile, grid); thread (x, y) of the T x T face holds its T cells along z in registers and issues
// each contribution's loads as one independent batch.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 4)))
void k_gather(int N, int T, int NT, int R, int ngrids, const float *regions, float *grids) {
    const int tile = blockIdx.x, g = blockIdx.y;
    const int tx = tile % NT, ty = (tile / NT) % NT, tz = tile / (NT * NT);
    const int cx = threadIdx.x % 16, cy = threadIdx.x / 16;
    if (cx >= T || cy >= T) return;
    const int R3 = R * R * R;
    // along each axis: option 0 = home (region coordinate cell + 2), option 1 = the neighbour
    // whose halo covers the cell (previous tile for cells < 2, next tile for cells >= T - 2)
    int nx = 1, ny = 1;
    int ntx[2] = {tx, tx}, nix[2] = {cx + kHalo, 0};
    int nty[2] = {ty, ty}, niy[2] = {cy + kHalo, 0};
    if (cx < kHalo) { ntx[1] = (tx + NT - 1) % NT; nix[1] = cx + T + kHalo; nx = 2; }
    else if (cx >= T - kHalo) { ntx[1] = (tx + 1) % NT; nix[1] = cx - T + kHalo; nx = 2; }
    if (cy < kHalo) { nty[1] = (ty + NT - 1) % NT; niy[1] = cy + T + kHalo; ny = 2; }
    else if (cy >= T - kHalo) { nty[1] = (ty + 1) % NT; niy[1] = cy - T + kHalo; ny = 2; }
    const int tzm = (tz + NT - 1) % NT, tzp = (tz + 1) % NT;
    float s[16];
    // cells beyond T re-read cell T-1 and are never stored: unconditional loads stay batched
    // (a guarded load becomes a branch with a wait after it)
    for (int o = 0; o < nx * ny; ++o) {
        const int ox = o % nx, oy = o / nx;
        // tile numbers as tile_of forms them (24-bit products, below 2^31 tiles)
        const int yx = __mul24(nty[oy], NT) + ntx[ox];
        const int64_t tb = __mul24(__mul24(tz, NT), NT) + yx;
        const int64_t tm = __mul24(__mul24(tzm, NT), NT) + yx;
        const int64_t tp = __mul24(__mul24(tzp, NT), NT) + yx;
        const float *r = regions + (tb * ngrids + g) * R3;
        const float *rm = regions + (tm * ngrids + g) * R3;
        const float *rp = regions + (tp * ngrids + g) * R3;
        float v[16], lo[2], hi[2];
#pragma unroll
        for (int k = 0; k < 16; ++k) v[k] = r[region_index(nix[ox], niy[oy], min(k, T - 1) + kHalo, T, R)];
        // z-neighbours of the same (x, y) option: cells 0, 1 from the previous z tile, cells
        // T-2, T-1 from the next one
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            lo[k] = rm[region_index(nix[ox], niy[oy], k + T + kHalo, T, R)];
            hi[k] = rp[region_index(nix[ox], niy[oy], k, T, R)];
        }
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            float t = v[k];
            if (k < 2) t += lo[k < 2 ? k : 0];
            if (k == T - 2) t += hi[0];
            if (k == T - 1) t += hi[1];
            s[k] = o == 0 ? t : s[k] + t;
        }
    }
    const int64_t V = (int64_t)N * N * N;
    float *out = grids + (int64_t)g * V + ((int64_t)(tz * T) * N + ty * T + cy) * N + tx * T + cx;
#pragma unroll
    for (int k = 0; k < 16; ++k)
        if (k < T) out[(int64_t)k * N * N] = s[k];
}

// The interlaced gather fused with the first axis of the density's FFT: row y of plane z of
// grid1 + i grid2, each cell summed from its region entries in k_gather's order (so the cells
// are bitwise k_gather's), then transformed along x in LDS: out[z][y][k] = sum_x (grid1 +
// i grid2)[z][y][x] e^{-2 pi i x k / N}, the forward DFT of length N = A B taken as the
// inverse four-step transform of the conjugate, conjugated again when written.  In a row,
// x = j + B m sits at m (B + 1) + j and, after the two steps, F[k1 + A k2] at k1 (B + 1) + k2.
// One work-group per (8 rows, plane z; 16 rows on 192, so that each thread has whole quads).
// Thread t sums the quads (4 cells along x) t + 256 i:
// along x a tile's home cells are contiguous in every piece, so a quad's home entries are one
// 16-byte load per grid and option (piece offsets, row lengths T and quad starts are multiples
// of 4 floats), and a quad at a tile's x edge reads its 2 neighbour entries as one 8-byte load
// (any other quad issues the same load on its own tile's entries T + 2, T + 3 and drops them;
// pointing it at the quad's own home entries instead, the line the home load fetches, measured
// 119 against 106 us at 256^3).
// The y options are taken only on rows at a tile's y edge and the z neighbour only on planes at
// a tile's z edge (whole waves or the whole work-group).  The tile edge T (8, 12 or 16) is a
// template parameter.  tw[k] = e^{2 pi i k / N}.
constexpr int xfft_rows(int N) { return N / 4 * 8 % 256 == 0 ? 8 : 16; }

template <int N, int A, int B, int T>
__global__ __launch_bounds__(256) void k_gather_xfft(const float *regions, const float2 *tw,
                                                     float2 *out) {
    static_assert(A * B == N && N % 64 == 0 && N % T == 0 && T % 4 == 0 && T >= 8, "shapes");
    constexpr int NT = N / T, R = T + 2 * kHalo, R3 = R * R * R;
    constexpr int kXfftRows = xfft_rows(N);
    constexpr int PB = B + 1, RS = A * PB + 1, QPR = N / 4, QPT = QPR * kXfftRows / 256;
    constexpr int QB = QPT % 2 ? QPT : 2;        // quads per batch (all loads in flight)
    static_assert(QPR * kXfftRows % 256 == 0 && QPT % QB == 0, "quads per thread");
    static_assert((int64_t)NT * NT * NT * 2 * R3 < (int64_t)1 << 31, "32-bit region offsets");
    __shared__ float2 rows[kXfftRows * RS];
    const int y0 = blockIdx.x * kXfftRows, z = blockIdx.y;
    const int tz = z / T, kz = z - tz * T;
    // the z-neighbour entry of this plane: cells kz < 2 add the previous z tile's entry at
    // kz + T + 2, cells kz >= T - 2 the next tile's at kz - T + 2
    const bool zn = kz < kHalo || kz >= T - kHalo;
    const int tzn = kz < kHalo ? (tz + NT - 1) % NT : (tz + 1) % NT;
    const int czh = kz + kHalo, czn = kz < kHalo ? kz + T + kHalo : kz - T + kHalo;
    const int zb = tz * NT * NT, znb = tzn * NT * NT;
    // region offset of entry (a, b, c) of tile (tzb + ty NT + tx), grid 0 (grid 1 is R3 further)
    auto at = [&](int tzb, int ty, int tx, int a, int b, int c) {
        return (tzb + ty * NT + tx) * 2 * R3 + region_index(a, b, c, T, R);
    };
    auto ld4 = [&](int off, float4 (&v)[2]) {
        v[0] = *reinterpret_cast<const float4 *>(regions + off);
        v[1] = *reinterpret_cast<const float4 *>(regions + off + R3);
    };
    auto ld2 = [&](int off, float2 (&v)[2]) {
        v[0] = *reinterpret_cast<const float2 *>(regions + off);
        v[1] = *reinterpret_cast<const float2 *>(regions + off + R3);
    };
    auto add4 = [](float4 a, float4 b) { return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); };
    auto add2 = [](float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); };
    // s[g][j] of the quad's cells j: s = t00 (+ t10 on the 2 cells at an x edge) (+ t01 (+ t11)),
    // t = home-z entry (+ z-neighbour entry), k_gather's order
    auto sum_batch = [&](auto has_zn, int ib) {
        constexpr bool ZN = decltype(has_zn)::value;
        int tx[QB], a0[QB], ntx[QB], a1[QB], xe[QB];   // xe: 0 none, 1 cells 0-1, 2 cells 2-3
        int ty[QB], b0[QB], nty[QB], b1[QB], ye[QB];
        float4 s[QB][2];
#pragma unroll
        for (int q = 0; q < QB; ++q) {
            const int gq = threadIdx.x + 256 * (ib + q);
            const int x = 4 * (gq % QPR), y = y0 + gq / QPR;
            tx[q] = x / T;
            ty[q] = y / T;
            const int cx = x - tx[q] * T, cy = y - ty[q] * T;
            a0[q] = cx + kHalo;
            b0[q] = b1[q] = cy + kHalo;
            ntx[q] = tx[q];
            nty[q] = ty[q];
            a1[q] = T + kHalo;          // a quad away from the x edges reads (and drops) this
            xe[q] = ye[q] = 0;
            if (cx == 0) { ntx[q] = (tx[q] + NT - 1) % NT; a1[q] = T + kHalo; xe[q] = 1; }
            else if (cx == T - 4) { ntx[q] = (tx[q] + 1) % NT; a1[q] = 0; xe[q] = 2; }
            if (cy < kHalo) { nty[q] = (ty[q] + NT - 1) % NT; b1[q] = cy + T + kHalo; ye[q] = 1; }
            else if (cy >= T - kHalo) { nty[q] = (ty[q] + 1) % NT; b1[q] = cy - T + kHalo; ye[q] = 1; }
        }
        // quad q's options (0, oy) and (1, oy) into s (oy = 0 assigns, oy = 1 adds)
        auto options = [&](int q, int oy, const float4 (&h)[2], const float4 (&hn)[2],
                           const float2 (&e)[2], const float2 (&en)[2]) {
#pragma unroll
            for (int g = 0; g < 2; ++g) {
                const float4 t0 = ZN ? add4(h[g], hn[g]) : h[g];
                const float2 t1 = ZN ? add2(e[g], en[g]) : e[g];
                float4 v = oy ? add4(s[q][g], t0) : t0;
                if (xe[q] == 1) { v.x += t1.x; v.y += t1.y; }
                else if (xe[q] == 2) { v.z += t1.x; v.w += t1.y; }
                s[q][g] = v;
            }
        };
        {
            float4 h[QB][2], hn[QB][2];
            float2 e[QB][2], en[QB][2];
#pragma unroll
            for (int q = 0; q < QB; ++q) {
                ld4(at(zb, ty[q], tx[q], a0[q], b0[q], czh), h[q]);
                ld2(at(zb, ty[q], ntx[q], a1[q], b0[q], czh), e[q]);
                if constexpr (ZN) {
                    ld4(at(znb, ty[q], tx[q], a0[q], b0[q], czn), hn[q]);
                    ld2(at(znb, ty[q], ntx[q], a1[q], b0[q], czn), en[q]);
                }
            }
#pragma unroll
            for (int q = 0; q < QB; ++q) options(q, 0, h[q], hn[q], e[q], en[q]);
        }
#pragma unroll
        for (int q = 0; q < QB; ++q) {
            if (!ye[q]) continue;
            float4 h[2], hn[2];
            float2 e[2], en[2];
            ld4(at(zb, nty[q], tx[q], a0[q], b1[q], czh), h);
            ld2(at(zb, nty[q], ntx[q], a1[q], b1[q], czh), e);
            if constexpr (ZN) {
                ld4(at(znb, nty[q], tx[q], a0[q], b1[q], czn), hn);
                ld2(at(znb, nty[q], ntx[q], a1[q], b1[q], czn), en);
            }
            options(q, 1, h, hn, e, en);
        }
#pragma unroll
        for (int q = 0; q < QB; ++q) {
            const int gq = threadIdx.x + 256 * (ib + q);
            const int x = 4 * (gq % QPR), yr = gq / QPR;
            const float re[4] = {s[q][0].x, s[q][0].y, s[q][0].z, s[q][0].w};
            const float im[4] = {s[q][1].x, s[q][1].y, s[q][1].z, s[q][1].w};
#pragma unroll
            for (int j = 0; j < 4; ++j)
                rows[yr * RS + ((x + j) / B) * PB + (x + j) % B] = make_float2(re[j], -im[j]);
        }
    };
    for (int ib = 0; ib < QPT; ib += QB) {
        if (zn) sum_batch(std::true_type{}, ib);
        else sum_batch(std::false_type{}, ib);
    }
    __syncthreads();
    // step 1: per row and j < B, the A entries x = j + B m over m, times W_N^{j k1}, in place
    for (int t = threadIdx.x; t < kXfftRows * B; t += 256) {
        const int r = t / B, j = t - r * B;
        float2 *rr = rows + r * RS + j;
        float2 a[A];
#pragma unroll
        for (int m = 0; m < A; ++m) a[m] = rr[m * PB];
        rdft<A>(a);
#pragma unroll
        for (int k1 = 0; k1 < A; ++k1) rr[k1 * PB] = k1 ? cmul(a[k1], tw[j * k1]) : a[0];
    }
    __syncthreads();
    // step 2: per row and k1 < A, the B entries k1 (B + 1) + j over j, in place
    for (int t = threadIdx.x; t < kXfftRows * A; t += 256) {
        const int r = t / A, k1 = t - r * A;
        float2 *rr = rows + r * RS + k1 * PB;
        float2 a[B];
#pragma unroll
        for (int j = 0; j < B; ++j) a[j] = rr[j];
        rdft<B>(a);
#pragma unroll
        for (int k2 = 0; k2 < B; ++k2) rr[k2] = a[k2];
    }
    __syncthreads();
    float2 *o = out + ((int64_t)z * N + y0) * N;
#pragma unroll
    for (int i = 0; i < N * kXfftRows / 256; ++i) {
        const int e = threadIdx.x + 256 * i;
        const int k = e % N, yr = e / N;
        const float2 f = rows[yr * RS + (k % A) * PB + k / A];
        o[e] = make_float2(f.x, -f.y);
    }
}

__global__ void k_fill_u32(uint32_t *p, uint32_t v, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace

void launch_fill_u32(uint32_t *p, uint32_t v, int64_t n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s, p, v, n);
}

AssignPlan make_assign_plan(const pbi4_assign_desc &d, int64_t npart) {
    PBI4_CHECK(d.grid >= 4, PBI4_EINVAL, "grid must be >= 4");
    PBI4_CHECK(d.order >= 1 && d.order <= 4, PBI4_EINVAL, "interpolation order must be 1..4");
    PBI4_CHECK(d.interlacing == 0 || d.interlacing == 1, PBI4_EINVAL, "interlacing must be 0/1");
    PBI4_CHECK(d.pos_mode == PBI4_POS_NORMALISED || d.pos_mode == PBI4_POS_BOX, PBI4_EINVAL,
               "bad pos_mode");
    PBI4_CHECK(d.pos_mode == PBI4_POS_NORMALISED || d.box > 0.0f, PBI4_EINVAL,
               "box must be > 0");
    AssignPlan p{};
    p.N = d.grid;
    // tile edge: the largest divisor of N in [4, 16] whose tiles hold at most kTileParticles
    // particles on average (one work-group deposits a tile, so crowded tiles serialise the
    // deposit), but not below 8 cells while a divisor in [8, 16] exists (the region's 2 + 2 halo
    // cells per axis cost (T+4)^3 / T^3 in region traffic)
    p.T = 0;
    int largest = 0, smallest8 = 0;
    for (int t = 16; t >= 4; --t) {
        if (p.N % t) continue;
        if (!largest) largest = t;
        if (t >= 8) smallest8 = t;
        const double per_tile = (double)npart * t * t * t / ((double)p.N * p.N * p.N);
        if (!p.T && per_tile <= kTileParticles && (t >= 8 || !smallest8)) p.T = t;
    }
    PBI4_CHECK(largest != 0, PBI4_EUNSUPPORTED,
               "grid " + std::to_string(p.N) + " has no divisor in [4, 16] for the tile size");
    if (!p.T) p.T = smallest8 ? smallest8 : largest;
    p.NT = p.N / p.T;
    p.R = p.T + 2 * kHalo;
    p.ntiles = (int64_t)p.NT * p.NT * p.NT;
    p.order = d.order;
    p.interlacing = d.interlacing;
    p.ngrids = d.interlacing ? 2 : 1;
    p.pos_mode = d.pos_mode;
    p.box = d.box;
    for (int i = 0; i < 3; ++i) p.rsd[i] = d.rsd[i];
    return p;
}

// Work-groups of the two sort passes (both split the particles the same way).
static int sort_wgs(int64_t npart) {
    return (int)std::max<int64_t>(1, std::min<int64_t>(2048, ceil_div(npart, 8192)));
}
// Per-(work-group, tile) first slots of the direct-counter sort (none for hashed counters).
static size_t sort_base_entries(const AssignPlan &p, int64_t npart) {
    return p.ntiles > kLdsDirectMax ? 0 : (size_t)sort_wgs(npart) * (size_t)p.ntiles;
}

size_t assign_workspace(const AssignPlan &p, int64_t npart) {
    Arena a(nullptr, 0);
    a.take<uint32_t>(p.ntiles + 1);     // counts
    a.take<uint32_t>(p.ntiles + 1);     // offsets
    a.take<uint32_t>(p.ntiles + 1);     // cursor
    a.take<Pk3>(npart);                 // sorted packed coordinates
    a.take<float>((size_t)p.ntiles * p.ngrids * p.R * p.R * p.R);   // regions
    a.take<uint32_t>(sort_base_entries(p, npart) + 1);              // wg_base
    return a.used + 256;
}

const float *assign_run(const AssignPlan &p, const float *pos, const float *vel, int64_t npart,
                       float *grids, Arena &ws, hipStream_t s, KernelTimer *timer) {
    PBI4_CHECK(npart >= 0 && npart < (int64_t)0xFFFFFFFFll, PBI4_EINVAL,
               "npart must be in [0, 2^32)");
    PBI4_CHECK(pos != nullptr || npart == 0, PBI4_EINVAL, "pos is NULL");
    uint32_t *counts = ws.take<uint32_t>(p.ntiles + 1);
    uint32_t *offsets = ws.take<uint32_t>(p.ntiles + 1);
    uint32_t *cursor = ws.take<uint32_t>(p.ntiles + 1);
    Pk3 *spk = ws.take<Pk3>(npart);
    const int R3 = p.R * p.R * p.R;
    float *regions = ws.take<float>((size_t)p.ntiles * p.ngrids * R3);
    uint32_t *wg_base = ws.take<uint32_t>(sort_base_entries(p, npart) + 1);

    PosXform x{};
    x.N = p.N;
    x.pos_mode = p.pos_mode;
    x.box = p.box;
    for (int i = 0; i < 3; ++i) x.rsd[i] = p.rsd[i];
    x.has_vel = vel != nullptr && p.pos_mode == PBI4_POS_BOX;

    const int hashed = p.ntiles > kLdsDirectMax;
    const size_t lds = (hashed ? 2 * (size_t)kHashSlots + 1 : (size_t)p.ntiles) * sizeof(uint32_t);
    const size_t lds_scatter = lds + ((size_t)kSortThreads * kScatterPer + 1) * sizeof(uint32_t);
    const int nwg = sort_wgs(npart);   // the same split in both passes (wg_base relies on it)

    {
        TimedScope ts(timer, "pbi4_sort", s);
        launch_fill_u32(counts, 0u, p.ntiles + 1, s);
        if (npart > 0)
            hipLaunchKernelGGL(k_tile_hist, dim3(nwg), dim3(kSortThreads), lds, s, x, p.T, p.NT,
                               p.ntiles, pos, vel, npart, counts, wg_base, hashed);
        hipLaunchKernelGGL(k_tile_scan, dim3(1), dim3(1024), 0, s, counts, p.ntiles, offsets, cursor);
        if (npart > 0 && !hashed)
            hipLaunchKernelGGL(k_tile_scatter_direct, dim3(nwg), dim3(kSortThreads), lds, s, x,
                               p.T, p.NT, p.ntiles, pos, vel, npart, offsets, wg_base, spk);
        else if (npart > 0)
            hipLaunchKernelGGL(k_tile_scatter, dim3(nwg), dim3(kSortThreads), lds_scatter, s, x, p.T,
                               p.NT, p.ntiles, pos, vel, npart, cursor, spk, hashed);
    }
    {
        TimedScope ts(timer, "pbi4_deposit", s);
        auto launch = [&](auto kern) {
            hipLaunchKernelGGL(kern, dim3((unsigned)p.ntiles), dim3(kDepositThreads), 0, s, p.T,
                               p.R, offsets, spk, regions);
        };
#define PBI4_DEPOSIT(O, G, RC) launch(k_deposit<O, G, RC>)
        const bool two = p.ngrids == 2;
        if (p.R == 12) {             // 8-cell tiles
            switch (p.order) {
                case 1: two ? PBI4_DEPOSIT(1, 2, 12) : PBI4_DEPOSIT(1, 1, 12); break;
                case 2: two ? PBI4_DEPOSIT(2, 2, 12) : PBI4_DEPOSIT(2, 1, 12); break;
                case 3: two ? PBI4_DEPOSIT(3, 2, 12) : PBI4_DEPOSIT(3, 1, 12); break;
                default: two ? PBI4_DEPOSIT(4, 2, 12) : PBI4_DEPOSIT(4, 1, 12); break;
            }
        } else if (p.R == kMaxR) {   // 16-cell tiles
            switch (p.order) {
                case 1: two ? PBI4_DEPOSIT(1, 2, kMaxR) : PBI4_DEPOSIT(1, 1, kMaxR); break;
                case 2: two ? PBI4_DEPOSIT(2, 2, kMaxR) : PBI4_DEPOSIT(2, 1, kMaxR); break;
                case 3: two ? PBI4_DEPOSIT(3, 2, kMaxR) : PBI4_DEPOSIT(3, 1, kMaxR); break;
                default: two ? PBI4_DEPOSIT(4, 2, kMaxR) : PBI4_DEPOSIT(4, 1, kMaxR); break;
            }
        } else {
            switch (p.order) {
                case 1: two ? PBI4_DEPOSIT(1, 2, 0) : PBI4_DEPOSIT(1, 1, 0); break;
                case 2: two ? PBI4_DEPOSIT(2, 2, 0) : PBI4_DEPOSIT(2, 1, 0); break;
                case 3: two ? PBI4_DEPOSIT(3, 2, 0) : PBI4_DEPOSIT(3, 1, 0); break;
                default: two ? PBI4_DEPOSIT(4, 2, 0) : PBI4_DEPOSIT(4, 1, 0); break;
            }
        }
#undef PBI4_DEPOSIT
    }
    if (grids) {
        TimedScope ts(timer, "pbi4_gather", s);
        hipLaunchKernelGGL(k_gather, dim3((unsigned)p.ntiles, p.ngrids), dim3(256), 0, s, p.N, p.T,
                           p.NT, p.R, p.ngrids, regions, grids);
    }
    PBI4_HIP(hipGetLastError());
    return regions;
}

bool gather_xfft_supported(const AssignPlan &p) {
    if (p.ngrids != 2 || p.R != p.T + 2 * kHalo) return false;
    switch (p.N) {
        case 128: case 256: case 512: return p.T == 8 || p.T == 16;
        case 192: case 384: return p.T == 8 || p.T == 12 || p.T == 16;
        default: return false;
    }
}

void launch_gather_xfft(const AssignPlan &p, const float *regions, const float2 *tw, float2 *out,
                        hipStream_t s) {
    PBI4_CHECK(gather_xfft_supported(p), PBI4_EUNSUPPORTED,
               "fused gather + x transform: interlaced grids 128, 192, 256, 384, 512 with tile "
               "edge 8, 12 or 16 only");
    PBI4_CHECK((reinterpret_cast<uintptr_t>(regions) & 15) == 0, PBI4_EINVAL,
               "tile regions must be 16-byte aligned");
    const dim3 grid(p.N / xfft_rows(p.N), p.N);
    auto run = [&](auto kern) { hipLaunchKernelGGL(kern, grid, dim3(256), 0, s, regions, tw, out); };
#define PBI4_GX(N, A, B) \
    (p.T == 8 ? run(k_gather_xfft<N, A, B, 8>) : p.T == 16 ? run(k_gather_xfft<N, A, B, 16>) \
              : run(k_gather_xfft<N, A, B, (N % 12 == 0 ? 12 : 16)>))
    switch (p.N) {
        case 128: PBI4_GX(128, 16, 8); break;
        case 192: PBI4_GX(192, 16, 12); break;
        case 256: PBI4_GX(256, 16, 16); break;
        case 384: PBI4_GX(384, 16, 24); break;
        default: PBI4_GX(512, 16, 32); break;
    }
#undef PBI4_GX
    PBI4_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------------------------
// Catalogue front-end: the redshift-space shift, periodic wrap and normalisation that
// IO.read_catalogues applies to the positions a reader returned (pbi4/pbi4.py:151-154, 162-163,
// 179-180, 213-216 for the shift; :241-243 for the wrap), restated with numpy's exact arithmetic:
//   shift  : RSD_HZ  s = p + ((v * los_d) / Hz) * (1 + z),  RSD_VEL  s = p + v * los_d,
//            both in FP64 (the float64 los promotes the product), then stored in the output type;
//   wrap   : (s + box) % box in the output type, numpy's remainder (fmod, then + box when the
//            sign differs from box's; an exact zero is +0);
//   scale  : / box, correctly rounded, then min(1, x) (1 - nextafter(0, 1) is 1.0 in FP64).
// Input (n, 3) rows as the readers build them, output (3, n) as the reference returns them.
namespace {

template <typename T>
__device__ __forceinline__ T exact_remainder(T a, T b) {
    // fmod is exact; on [0, 2b) it is one exact subtraction (Sterbenz), the common case
    T r = (a >= T(0) && a < b + b) ? (a < b ? a : a - b) : fmod(a, b);
    if (r != T(0)) {
        if ((r < T(0)) != (b < T(0))) r += b;
    } else {
        r = T(0);
    }
    return r;
}

template <typename TP, typename TV, typename TO>
__global__ __launch_bounds__(256) void k_catalogue_positions(const TP *pos, const TV *vel, int64_t n,
                                                             int rsd, double los0, double los1,
                                                             double los2, double hz, double zf,
                                                             double box, TO *out) {
#pragma clang fp contract(off)
    const TO b = (TO)box;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double los = d == 0 ? los0 : d == 1 ? los1 : los2;
            TO s;
            if (rsd == PBI4_RSD_NONE) {
                s = (TO)pos[3 * i + d];
            } else {
                double t = (double)vel[3 * i + d] * los;
                if (rsd == PBI4_RSD_HZ) t = (t / hz) * zf;
                s = (TO)((double)pos[3 * i + d] + t);
            }
            const TO w = exact_remainder<TO>(s + b, b);
            const TO x = w / b;
            out[d * n + i] = x > TO(1) ? TO(1) : x;
        }
    }
}

template <typename TP, typename TV, typename TO>
void launch_catalogue_t(const void *pos, const void *vel, int64_t n, int rsd, const double *los,
                        double hz, double zf, double box, void *out, hipStream_t s) {
    const unsigned blocks = (unsigned)std::min<int64_t>(ceil_div(n, 256), 1 << 16);
    hipLaunchKernelGGL((k_catalogue_positions<TP, TV, TO>), dim3(blocks), dim3(256), 0, s,
                       static_cast<const TP *>(pos), static_cast<const TV *>(vel), n, rsd, los[0],
                       los[1], los[2], hz, zf, box, static_cast<TO *>(out));
}

}  // namespace

void launch_catalogue_positions(const void *pos, int pos_f64, const void *vel, int vel_f64,
                                int64_t n, int rsd, const double *los, double hz, double zf,
                                double box, void *out, int out_f64, hipStream_t s) {
    const int key = pos_f64 * 4 + vel_f64 * 2 + out_f64;
    switch (key) {
    case 0: launch_catalogue_t<float, float, float>(pos, vel, n, rsd, los, hz, zf, box, out, s); break;
    case 1: launch_catalogue_t<float, float, double>(pos, vel, n, rsd, los, hz, zf, box, out, s); break;
    case 2: launch_catalogue_t<float, double, float>(pos, vel, n, rsd, los, hz, zf, box, out, s); break;
    case 3: launch_catalogue_t<float, double, double>(pos, vel, n, rsd, los, hz, zf, box, out, s); break;
    case 4: launch_catalogue_t<double, float, float>(pos, vel, n, rsd, los, hz, zf, box, out, s); break;
    case 5: launch_catalogue_t<double, float, double>(pos, vel, n, rsd, los, hz, zf, box, out, s); break;
    case 6: launch_catalogue_t<double, double, float>(pos, vel, n, rsd, los, hz, zf, box, out, s); break;
    default: launch_catalogue_t<double, double, double>(pos, vel, n, rsd, los, hz, zf, box, out, s); break;
    }
    PBI4_HIP(hipGetLastError());
}

}  // namespace pbi4
