// C ABI of libpbi4hip.so: context, cached tables and rocFFT plans, entry points.
#include <dlfcn.h>
#include <rocfft/rocfft.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "common.h"

namespace pbi4 {

[[noreturn]] void fail(int code, const std::string &msg) { throw Error{code, msg}; }

TimedScope::TimedScope(KernelTimer *t_, const char *k, hipStream_t s_) : t(t_), s(s_) {
    if (t && !t->name.empty() && t->name == k) {
        hipEvent_t a;
        if (hipEventCreate(&a) == hipSuccess && hipEventCreate(&b) == hipSuccess) {
            (void)hipEventRecord(a, s);
            t->events.push_back({a, b});
        }
    }
}
TimedScope::~TimedScope() {
    if (b) (void)hipEventRecord(b, s);
}

}  // namespace pbi4

using namespace pbi4;

namespace {

// ------------------------------------------------------------------ host-built tables
// These follow the reference formulas operation by operation in FP64 (no FMA contraction), so
// mode-to-bin and mode-to-shell assignment is bit-exact, ties included.
#pragma clang fp contract(off)

// P(k) bin of (|kx|, |ky|, |kz|): imk = nint(rk/dk - kcenter/kbin - tiny) + 1, with
// kF = REAL(4) twopi/box and float(ik) (powerI4.f90:527-566).  0 = outside 1..nbins.
// Stored [|kz|][|ky|][kx] so the binning kernel's lanes (consecutive kx) read contiguously.
std::vector<uint16_t> build_pk_table(int N, float box, float kbin, float kcenter, int nbins) {
    const int h = N / 2 + 1;
    const double kF = (double)(kTwoPiF32 / box);
    const double dk = (double)kbin * kF;
    const double kcb = (double)(kcenter / kbin);
    std::vector<double> sq(h);
    for (int a = 0; a < h; ++a) {
        const double r = kF * (double)(float)a;
        sq[a] = r * r;
    }
    std::vector<uint16_t> t((size_t)h * h * h);
    for (int a = 0; a < h; ++a)
        for (int b = 0; b < h; ++b)
            for (int c = 0; c < h; ++c) {
                const double rk = std::sqrt(sq[a] + sq[b] + sq[c]);
                const long imk = std::lround(rk / dk - kcb - DBL_MIN) + 1;
                t[((size_t)c * h + b) * h + a] = (imk >= 1 && imk <= nbins) ? (uint16_t)imk : 0;
            }
    return t;
}

// Bispectrum shell of a grid point with per-axis distances d = min(i, N-i):
// idist = nint(sqrt(d2)/kbin - kcenter/kbin - tiny) + 1 (bispectrumI4.f90:69-73).
int shell_of_d2(long d2, float kbin, float kcenter) {
    const double kcb = (double)(kcenter / kbin);
    return (int)(std::lround(std::sqrt((double)d2) / (double)kbin - kcb - DBL_MIN) + 1);
}

#pragma clang fp contract(on)

struct ShellInfo {
    int N, nb, max_d2, R;
    std::vector<int16_t> table;   // d2 -> shell (0 = none)
    std::vector<double> nmodes;   // full-grid modes per shell
    std::vector<int32_t> d2range;  // per shell: smallest, largest d2 (empty shell: 1, 0)
    bool contiguous = false;       // every d2 inside a shell's range belongs to that shell
    int16_t *table_d = nullptr;
    double *nmodes_d = nullptr;
    int32_t *d2range_d = nullptr;
};

struct PkTable {
    uint16_t *d = nullptr;
    std::vector<double> nmodes;
};

struct FftPlan {
    rocfft_plan plan = nullptr;
    size_t work = 0;
};

struct CountsCache {
    std::vector<unsigned long long> T;   // (nb+1)^3
};

}  // namespace

// ------------------------------------------------------------------ allocation-free scratch
// Split-K partials of pbi4_triangle_sums_noalloc: one static device array per GPU, held until
// the process ends.  Host code never uses the symbol's name as a device address: each device's
// address comes from hipGetSymbolAddress (taken once per device, with that device current, when
// a context is created on it) and only that address reaches kernels.
static __device__ __align__(256) unsigned char scratch[PBI4_NOALLOC_SCRATCH_BYTES];

namespace {

std::mutex g_scratch_mu;
std::map<int, void *> g_scratch_ptr;   // device -> address of its `scratch`

// With `device` current: records the device's scratch address (idempotent).
void scratch_setup(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr.count(device)) return;
    void *p = nullptr;
    PBI4_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    PBI4_CHECK(p != nullptr, PBI4_EHIP, "hipGetSymbolAddress gave no address for the scratch array");
    g_scratch_ptr[device] = p;
}

void *scratch_of(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_ptr.find(device);
    PBI4_CHECK(it != g_scratch_ptr.end(), PBI4_EINVAL, "no scratch address for this device");
    return it->second;
}

}  // namespace

struct pbi4_ctx {
    int device = 0;
    std::string err;
    std::mutex mu;
    KernelTimer timer;
    std::map<std::tuple<int, float, float, float, int>, PkTable> pk_tables;
    std::map<std::tuple<int, int, float, float>, std::unique_ptr<ShellInfo>> shells;
    std::map<std::tuple<int, int>, float *> windows;   // (N, order) -> device sinc^p table
    std::map<int, float2 *> phases;                    // N -> device e^{-i theta s}, s in [-N, 3N/2]
    std::map<int, float2 *> twiddles;                  // N' -> device e^{2 pi i k / N'}, k < N'
    std::map<std::tuple<int, int, int>, FftPlan> plans; // (kind, N, batch/ext)
    std::map<std::tuple<int, int, float, float>, CountsCache> counts;
    std::map<std::tuple<int, int, float, float, int, int>, int32_t *> tri_tables;
    std::vector<void *> owned;   // device allocations owned by the context
};

namespace {

// message of the calling thread's last failed call made without a context
thread_local std::string g_err_noctx;

int set_err(pbi4_ctx *ctx, const Error &e) {
    if (ctx)
        ctx->err = e.msg;
    else
        g_err_noctx = e.msg;
    return e.code;
}

// Workspace size of a call made without a context: 0 with the reason kept for
// pbi4_last_error(NULL) when the descriptor is refused.
template <class F>
size_t sized(F &&f) {
    try {
        return f();
    } catch (const Error &e) {
        g_err_noctx = e.msg;
    } catch (const std::exception &e) {
        g_err_noctx = e.what();
    }
    return 0;
}

template <class F>
int guarded(pbi4_ctx *ctx, F &&f) {
    try {
        if (ctx) {
            PBI4_HIP(hipSetDevice(ctx->device));
        }
        f();
        return PBI4_OK;
    } catch (const Error &e) {
        return set_err(ctx, e);
    } catch (const std::exception &e) {
        return set_err(ctx, Error{PBI4_EINVAL, e.what()});
    }
}

template <class T>
T *dev_upload(pbi4_ctx *ctx, const std::vector<T> &v) {
    T *d = nullptr;
    PBI4_HIP(hipMalloc(&d, std::max<size_t>(1, v.size()) * sizeof(T)));
    ctx->owned.push_back(d);
    if (!v.empty()) PBI4_HIP(hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return d;
}

void fft_check(rocfft_status st, const char *what) {
    if (st != rocfft_status_success)
        fail(PBI4_EFFT, std::string(what) + " failed (rocfft status " + std::to_string((int)st) + ")");
}

// ------------------------------------------------------------------ how rocFFT builds kernels
// rocFFT compiles its kernels at run time through hiprtc, and also takes prebuilt ones from a
// system kernel cache.  Built with default flags they contain packed FP32 VALU operations, which
// on this hardware have given wrong results, with no error, while matrix-core work ran beside
// them: with two serving slots, the last c2r batch of one slot's shell maps changed while the
// other slot's triangle GEMM ran, and never with rocFFT built without them.  So every rocFFT
// kernel this library runs is built with packed FP32 off (as the library's own kernels are),
// into a kernel cache of the library's own, and none comes from the system cache.  rocFFT reads
// these settings when it first plans a transform in the process: they are set when the library
// is loaded (the Python package sets the same ones on import), and the first plan checks that
// rocFFT uses the library's cache, which fails when rocFFT was in use before the library loaded.
constexpr const char *kNoPackedFp32 = "-Xclang -target-feature -Xclang -packed-fp32-ops";
std::string g_rocfft_cache;    // the kernel cache file this library told rocFFT to use

std::string library_dir() {
    Dl_info info{};
    if (dladdr(reinterpret_cast<void *>(&library_dir), &info) && info.dli_fname) {
        const std::string p(info.dli_fname);
        const size_t k = p.rfind('/');
        return k == std::string::npos ? std::string(".") : p.substr(0, k);
    }
    return ".";
}

bool writable_dir(const std::string &d) {
    (void)mkdir(d.c_str(), 0755);
    return access(d.c_str(), W_OK | X_OK) == 0;
}

void rocfft_build_settings() {
    const char *cur = getenv("HIPRTC_COMPILE_OPTIONS_APPEND");
    std::string opts = cur ? cur : "";
    if (opts.find("packed-fp32-ops") == std::string::npos) {
        opts += (opts.empty() ? "" : " ");
        opts += kNoPackedFp32;
        setenv("HIPRTC_COMPILE_OPTIONS_APPEND", opts.c_str(), 1);
    }
    setenv("ROCFFT_RTC_SYS_CACHE_PATH", "/nonexistent/pbi4_amd/no_system_kernel_cache.db", 1);
    const char *given = getenv("PBI4_ROCFFT_CACHE");
    std::string path;
    if (given && *given) {
        path = given;
    } else {
        std::string dir = library_dir() + "/_rocfft_cache";
        if (!writable_dir(dir)) {
            const char *tmp = getenv("TMPDIR");
            dir = std::string(tmp && *tmp ? tmp : "/tmp") + "/pbi4_amd_rocfft_" +
                  std::to_string((long)getuid());
            (void)writable_dir(dir);
        }
        path = dir + "/kernels.db";
    }
    setenv("ROCFFT_RTC_CACHE_PATH", path.c_str(), 1);
    g_rocfft_cache = path;
}

struct RocfftSettingsAtLoad {
    RocfftSettingsAtLoad() { rocfft_build_settings(); }
};
RocfftSettingsAtLoad g_rocfft_settings_at_load;

struct RocfftInit {
    RocfftInit() { rocfft_setup(); }
};

// After the first plan: rocFFT has opened its kernel cache by now; if it is not the one this
// library named, rocFFT was set up before the library loaded and may run packed-FP32 kernels.
void check_rocfft_cache() {
    static bool checked = false;
    if (checked) return;
    checked = true;
    const size_t k = g_rocfft_cache.rfind('/');
    const std::string dir = k == std::string::npos ? "." : g_rocfft_cache.substr(0, k);
    if (access(dir.c_str(), W_OK) == 0 && access(g_rocfft_cache.c_str(), F_OK) != 0)
        fail(PBI4_EUNSUPPORTED,
             "rocFFT is not using this library's kernel cache (" + g_rocfft_cache +
                 "): it was set up in this process before libpbi4hip.so was loaded, so its "
                 "kernels may contain packed FP32 operations; load pbi4_amd (import pbi4_amd, or "
                 "dlopen libpbi4hip.so) before the process's first GPU FFT");
}

// kind 0: r2c of N^3 with `batch` grids (default layouts).
// kind 1: c2r of N^3 with `batch` maps, real output per-map distance `ext` floats.
// kind 2: forward c2c of one N^3 field from planar real / imaginary parts.
// kind 3: 2D c2r of `batch` N^2 planes (the maps after their z transform).
// kind 4: forward c2c over (y, x) of `batch` N^2 planes from planar parts (the interlaced density
//         whose z axis the combine transforms).
FftPlan &get_plan(pbi4_ctx *ctx, int kind, int N, int batch, int64_t ext) {
    static RocfftInit init;
    auto key = std::make_tuple(kind * 1000000 + batch, N, (int)ext);
    auto it = ctx->plans.find(key);
    if (it != ctx->plans.end()) return it->second;
    FftPlan p;
    size_t len[3] = {(size_t)N, (size_t)N, (size_t)N};
    rocfft_plan_description desc = nullptr;
    fft_check(rocfft_plan_description_create(&desc), "rocfft_plan_description_create");
    if (kind == 2) {
        // forward c2c of one N^3 field given as planar parts (real part, imaginary part)
        size_t strides[3] = {1, (size_t)N, (size_t)N * N};
        const size_t dist = (size_t)N * N * N;
        fft_check(rocfft_plan_description_set_data_layout(
                      desc, rocfft_array_type_complex_planar, rocfft_array_type_complex_interleaved,
                      nullptr, nullptr, 3, strides, dist, 3, strides, dist),
                  "rocfft_plan_description_set_data_layout");
    }
    if (kind == 4) {
        // forward c2c over (y, x) of `batch` N x N planes given as planar parts
        size_t strides[2] = {1, (size_t)N};
        const size_t dist = (size_t)N * N;
        fft_check(rocfft_plan_description_set_data_layout(
                      desc, rocfft_array_type_complex_planar, rocfft_array_type_complex_interleaved,
                      nullptr, nullptr, 2, strides, dist, 2, strides, dist),
                  "rocfft_plan_description_set_data_layout");
    }
    if (kind == 3) {
        // c2r over (y, x) of `batch` N x N planes: half spectra [y][x'] in, real [y][x] out
        const size_t h = N / 2 + 1;
        const size_t ld = ext ? (size_t)ext : h;
        size_t in_strides[2] = {1, ld}, out_strides[2] = {1, (size_t)N};
        fft_check(rocfft_plan_description_set_data_layout(
                      desc, rocfft_array_type_hermitian_interleaved, rocfft_array_type_real, nullptr,
                      nullptr, 2, in_strides, ld * N, 2, out_strides, (size_t)N * N),
                  "rocfft_plan_description_set_data_layout");
    }
    if (kind == 1) {
        const size_t h = N / 2 + 1;
        size_t in_strides[3] = {1, h, h * N};
        size_t out_strides[3] = {1, (size_t)N, (size_t)N * N};
        fft_check(rocfft_plan_description_set_data_layout(
                      desc, rocfft_array_type_hermitian_interleaved, rocfft_array_type_real, nullptr,
                      nullptr, 3, in_strides, h * N * N, 3, out_strides, (size_t)ext),
                  "rocfft_plan_description_set_data_layout");
    }
    fft_check(rocfft_plan_create(&p.plan, rocfft_placement_notinplace,
                                 kind == 0                  ? rocfft_transform_type_real_forward
                                 : (kind == 2 || kind == 4) ? rocfft_transform_type_complex_forward
                                                            : rocfft_transform_type_real_inverse,
                                 rocfft_precision_single, (kind == 3 || kind == 4) ? 2 : 3, len,
                                 (size_t)batch,
                                 desc),
              "rocfft_plan_create");
    rocfft_plan_description_destroy(desc);
    fft_check(rocfft_plan_get_work_buffer_size(p.plan, &p.work), "rocfft_plan_get_work_buffer_size");
    ctx->plans[key] = p;
    check_rocfft_cache();
    return ctx->plans[key];
}

void fft_exec(FftPlan &p, void *in, void *out, void *work, hipStream_t s, void *in2 = nullptr) {
    rocfft_execution_info info = nullptr;
    fft_check(rocfft_execution_info_create(&info), "rocfft_execution_info_create");
    fft_check(rocfft_execution_info_set_stream(info, s), "rocfft_execution_info_set_stream");
    if (p.work) fft_check(rocfft_execution_info_set_work_buffer(info, work, p.work),
                          "rocfft_execution_info_set_work_buffer");
    void *ib[2] = {in, in2}, *ob[1] = {out};
    const rocfft_status st = rocfft_execute(p.plan, ib, ob, info);
    rocfft_execution_info_destroy(info);
    fft_check(st, "rocfft_execute");
}

// tWk of fcomb (powerI4.f90:278-292), evaluated in FP64 like the reference (x**nintp as
// repeated squaring), stored as float.
float *get_window(pbi4_ctx *ctx, int N, int order) {
    auto key = std::make_tuple(N, order);
    auto it = ctx->windows.find(key);
    if (it != ctx->windows.end()) return it->second;
    const int h = N / 2 + 1;
    std::vector<float> w(h);
    const double tpiN = (double)kTwoPiF32 / (double)N;
    w[0] = 1.0f;
    for (int i = 1; i < h; ++i) {
        const double rk = tpiN * (double)i;
        double x = std::sin(rk / 2.0) / (rk / 2.0);
        unsigned n = (unsigned)order;
        double y = (n % 2) ? x : 1.0;
        while (n >>= 1) {
            x = x * x;
            if (n % 2) y = y * x;
        }
        w[i] = (float)y;
    }
    return ctx->windows[key] = dev_upload(ctx, w);
}

// The interlacing phase of fcomb (powerI4.f90:251-378), e^{-i theta s} for s = ix +- iy +- iz in
// [-N, 3N/2] at index s + N, theta = twopi / N / 2 from the reference's REAL(4) twopi, evaluated
// in FP64 and rounded once.
float2 *get_phase(pbi4_ctx *ctx, int N) {
    auto it = ctx->phases.find(N);
    if (it != ctx->phases.end()) return it->second;
    const double th = (double)kTwoPiF32 / (double)N / 2.0;
    std::vector<float2> t(5 * (size_t)N / 2 + 1);
    for (int k = 0; k < (int)t.size(); ++k) {
        const double a = th * (double)(k - N);
        t[k] = make_float2((float)std::cos(a), (float)-std::sin(a));
    }
    return ctx->phases[N] = dev_upload(ctx, t);
}

// e^{2 pi i k / N'} for k < N', from FP64, with entry N'-k set to the exact conjugate of k.
float2 *get_twiddle(pbi4_ctx *ctx, int Np) {
    auto it = ctx->twiddles.find(Np);
    if (it != ctx->twiddles.end()) return it->second;
    std::vector<float2> t(Np);
    for (int k = 0; k <= Np / 2; ++k) {
        const double a = 2.0 * M_PI * (double)k / (double)Np;
        t[k] = make_float2((float)std::cos(a), (float)std::sin(a));
        if (k > 0) t[Np - k] = make_float2(t[k].x, -t[k].y);
    }
    return ctx->twiddles[Np] = dev_upload(ctx, t);
}

// The maps' inverse transform, one of four routes:
//  * band-limited maps grid (N' < N, so 3R < N'): the z transform is fused into the shell filter
//    (k_filter_z); the 2D c2r over (y, x) that finishes each plane is k_c2r2d (one pass through
//    LDS) on the maps grids it is compiled for, else rocFFT's batched 2D c2r;
//  * full grid (N' = N, an aliasing request) with 2R < N on 128, 192, 256 or 384: the same fused
//    filter, then the column and row passes of launch_maps_two_pass;
//  * otherwise the shell filter writes half spectra and rocFFT's 3D c2r transforms them.
// On the routes of the library's own the fused filter writes only the columns x' < C =
// min(R + 1, N'/2 + 1) and, per shell, only the lines (x', q_y) inside the shell's disk
// x'^2 + q_y^2 <= largest d2: the others are zero, and the transforms take them as zero without
// reading them.  PBI4_MAPS_C2R=rocfft keeps the rocFFT routes (comparisons and timing).
enum class MapsPath { Rocfft3d, Rocfft2d, LdsPlane, TwoPass };
struct MapsRoute {
    MapsPath path = MapsPath::Rocfft3d;
    int ld = 0;          // row stride of the half spectra (complex entries)
    int cols = 0;        // columns the fused filter writes (0 = all N'/2 + 1)
    int disk_only = 0;   // the fused filter writes only the lines inside each shell's disk
    C2r2dPlan plane{};   // LdsPlane
};

MapsRoute maps_route(int Np, int grid, const ShellInfo &sh, int64_t Kstride) {
    MapsRoute r;
    const int h = Np / 2 + 1, C = std::min(sh.R + 1, h);
    const char *mode = std::getenv("PBI4_MAPS_C2R");
    const bool ours = !(mode && std::strcmp(mode, "rocfft") == 0);
    if (Np < grid && sh.contiguous) {
        // rows padded to 64 entries: 512-byte rows for rocFFT's column pass
        r.ld = (h + 63) / 64 * 64;
        if (ours && make_c2r2d_plan(Np, C, r.ld, Kstride, r.plane)) {
            r.path = MapsPath::LdsPlane;
            r.cols = C;
            r.disk_only = 1;
            r.plane.nb = sh.nb;
            r.plane.shell_d2 = sh.d2range_d;
        } else {
            r.path = MapsPath::Rocfft2d;
        }
        return r;
    }
    if (ours && Np == grid && sh.contiguous && 2 * sh.R < Np && maps_two_pass_supported(Np)) {
        r.path = MapsPath::TwoPass;
        r.cols = C;
        r.disk_only = 1;
        r.ld = (C + 15) / 16 * 16;   // whole 16-column blocks for the column pass
        return r;
    }
    r.ld = h;
    return r;
}

// The shell filter with the z transform fused (every route but rocFFT's 3D c2r).
bool maps_fused(int Np, int grid, const ShellInfo &sh, int64_t Kstride) {
    return maps_route(Np, grid, sh, Kstride).path != MapsPath::Rocfft3d;
}
int maps_ld(int Np, int grid, const ShellInfo &sh, int64_t Kstride) {
    return maps_route(Np, grid, sh, Kstride).ld;
}
int maps_cols(int Np, int grid, const ShellInfo &sh, int64_t Kstride) {
    return maps_route(Np, grid, sh, Kstride).cols;
}
int maps_disk_only(int Np, int grid, const ShellInfo &sh, int64_t Kstride) {
    return maps_route(Np, grid, sh, Kstride).disk_only;
}

// Full-grid maps: one 3D c2r execution transforms at most this many real outputs (all its
// batched maps together), so no execution spans more than 2^31 elements; a 384^3 request with
// 64 shells x 3 multipoles (1.09e10 outputs, 43 GB) runs as six executions of 32 maps.
constexpr int64_t kMaxC2rElems = (int64_t)1 << 31;

int maps_batch(int Np, int nmaps) {
    const int64_t per = (int64_t)Np * Np * Np;
    const int64_t cap = std::max<int64_t>(1, kMaxC2rElems / per);
    const int64_t runs = ceil_div((int64_t)nmaps, cap);
    return (int)ceil_div((int64_t)nmaps, runs);
}

FftPlan &maps_plan(pbi4_ctx *ctx, const MapsRoute &r, int Np, int nmaps, int64_t Kstride) {
    if (r.path == MapsPath::Rocfft3d) return get_plan(ctx, 1, Np, maps_batch(Np, nmaps), Kstride);
    const bool one = Kstride == (int64_t)Np * Np * Np;
    return get_plan(ctx, 3, Np, one ? nmaps * Np : Np, r.ld);
}

// Work buffer bytes the maps' c2r executions need (the full-grid path may run a shorter last
// batch with a plan of its own); none on the routes without rocFFT.
size_t maps_work(pbi4_ctx *ctx, const ShellInfo &sh, int Np, int grid, int nmaps,
                 int64_t Kstride) {
    const MapsRoute r = maps_route(Np, grid, sh, Kstride);
    if (r.path == MapsPath::LdsPlane || r.path == MapsPath::TwoPass) return 0;
    size_t w = maps_plan(ctx, r, Np, nmaps, Kstride).work;
    if (r.path == MapsPath::Rocfft3d) {
        const int b = maps_batch(Np, nmaps);
        if (nmaps % b) w = std::max(w, get_plan(ctx, 1, Np, nmaps % b, Kstride).work);
    }
    return w;
}

void maps_c2r(pbi4_ctx *ctx, const ShellInfo &sh, int Np, int grid, int nmaps, int64_t Kstride,
              float2 *spectra, float *maps, void *work, hipStream_t s) {
    const MapsRoute r = maps_route(Np, grid, sh, Kstride);
    if (r.path == MapsPath::LdsPlane) {
        TimedScope ts(&ctx->timer, "pbi4_c2r", s);
        launch_c2r2d(r.plane, spectra, get_twiddle(ctx, Np), maps, nmaps, s);
        return;
    }
    if (r.path == MapsPath::TwoPass) {
        TimedScope ts(&ctx->timer, "pbi4_c2r", s);
        launch_maps_two_pass(Np, r.cols, r.ld, nmaps, sh.nb, sh.d2range_d, Kstride,
                             get_twiddle(ctx, Np), spectra, maps, s);
        return;
    }
    FftPlan &p = maps_plan(ctx, r, Np, nmaps, Kstride);
    TimedScope ts(&ctx->timer, "pbi4_c2r", s);
    if (r.path == MapsPath::Rocfft3d) {
        const int b = maps_batch(Np, nmaps);
        const size_t spec = (size_t)Np * Np * (Np / 2 + 1);
        for (int m0 = 0; m0 < nmaps; m0 += b) {
            FftPlan &q = m0 + b <= nmaps ? p : get_plan(ctx, 1, Np, nmaps - m0, Kstride);
            fft_exec(q, spectra + (size_t)m0 * spec, maps + (size_t)m0 * Kstride, work, s);
        }
        return;
    }
    if (Kstride == (int64_t)Np * Np * Np) {
        fft_exec(p, spectra, maps, work, s);
        return;
    }
    const size_t spec = (size_t)Np * Np * r.ld;
    for (int m = 0; m < nmaps; ++m) fft_exec(p, spectra + m * spec, maps + m * Kstride, work, s);
}

// kF3 = twopi*twopi*twopi/box/box/box in REAL(4) (powerI4.f90:271)
double kf3_fcomb(float box) {
    volatile float t = kTwoPiF32 * kTwoPiF32;
    t = t * kTwoPiF32;
    t = t / box;
    t = t / box;
    t = t / box;
    return (double)t;
}

PkTable &get_pk_table(pbi4_ctx *ctx, int N, float box, float kbin, float kcenter, int nbins) {
    auto key = std::make_tuple(N, box, kbin, kcenter, nbins);
    auto it = ctx->pk_tables.find(key);
    if (it != ctx->pk_tables.end()) return it->second;
    PBI4_CHECK(nbins >= 1 && nbins < 65535, PBI4_EINVAL, "nbins must be in [1, 65535)");
    PBI4_CHECK(kbin > 0.0f, PBI4_EINVAL, "kbin must be > 0");
    PkTable t;
    t.d = dev_upload(ctx, build_pk_table(N, box, kbin, kcenter, nbins));
    return ctx->pk_tables[key] = t;
}

ShellInfo &get_shells(pbi4_ctx *ctx, int N, int nb, float kbin, float kcenter) {
    auto key = std::make_tuple(N, nb, kbin, kcenter);
    auto it = ctx->shells.find(key);
    if (it != ctx->shells.end()) return *it->second;
    PBI4_CHECK(nb >= 1 && nb <= 32000, PBI4_EINVAL, "nbins must be in [1, 32000]");
    PBI4_CHECK(kbin > 0.0f, PBI4_EINVAL, "kbin must be > 0");
    auto s = std::make_unique<ShellInfo>();
    s->N = N;
    s->nb = nb;
    const int hN = N / 2;
    s->max_d2 = 3 * hN * hN;
    s->table.assign(s->max_d2 + 1, 0);
    int d2max_in = -1;
    for (long d2 = 0; d2 <= s->max_d2; ++d2) {
        const int sh = shell_of_d2(d2, kbin, kcenter);
        if (sh >= 1 && sh <= nb) {
            s->table[d2] = (int16_t)sh;
            d2max_in = (int)d2;
        }
    }
    // largest |component| of any shell mode (bounded by the per-axis distance N/2)
    s->R = d2max_in < 0 ? 0 : std::min(hN, (int)std::floor(std::sqrt((double)d2max_in) + 1e-9));
    // full-grid modes per shell: per-axis multiplicity of each distance v = min(i, N-i)
    std::vector<double> mult(hN + 1, 0.0);
    for (int i = 0; i < N; ++i) mult[std::min(i, N - i)] += 1.0;
    s->nmodes.assign(nb, 0.0);
    for (int a = 0; a <= hN; ++a)
        for (int b = 0; b <= hN; ++b) {
            const long ab = (long)a * a + (long)b * b;
            if (ab > d2max_in) break;
            for (int c = 0; c <= hN; ++c) {
                const long d2 = ab + (long)c * c;
                if (d2 > d2max_in) break;
                const int sh = s->table[d2];
                if (sh) s->nmodes[sh - 1] += mult[a] * mult[b] * mult[c];
            }
        }
    s->d2range.resize(2 * nb);
    for (int i = 0; i < nb; ++i) {
        s->d2range[2 * i] = 1;
        s->d2range[2 * i + 1] = 0;
    }
    for (int d2 = s->max_d2; d2 >= 0; --d2)
        if (const int sh = s->table[d2]) {
            s->d2range[2 * (sh - 1)] = d2;
            if (s->d2range[2 * (sh - 1) + 1] < d2) s->d2range[2 * (sh - 1) + 1] = d2;
        }
    s->contiguous = true;
    for (int i = 0; i < nb; ++i)
        for (int d2 = s->d2range[2 * i]; d2 <= s->d2range[2 * i + 1]; ++d2)
            if (s->table[d2] != i + 1) s->contiguous = false;
    s->table_d = dev_upload(ctx, s->table);
    s->nmodes_d = dev_upload(ctx, s->nmodes);
    s->d2range_d = dev_upload(ctx, s->d2range);
    auto &ref = *s;
    ctx->shells[key] = std::move(s);
    return ref;
}

bool smooth235(int n) {
    for (int p : {2, 3, 5})
        while (n % p == 0) n /= p;
    return n == 1;
}

// Smallest even 2,3,5-smooth size N' with N' > 3R (no aliasing of q1+q2+q3 on N'), or N when
// the reference's own mod-N closure aliases (3R >= N) or no smaller size exists.
int choose_maps_grid(int N, int R) {
    if (3 * R >= N) return N;
    for (int n = 3 * R + 1; n < N; ++n)
        if (n % 2 == 0 && n >= 4 && smooth235(n)) return n;
    return N;
}

int tri_ishift(float kbin, float kcenter, int iopen) { return (int)(kcenter / kbin) + (iopen ? 1 : 0); }

std::vector<int32_t> triangle_list(int nb, float kbin, float kcenter, int iopen, int cross) {
    const int ishift = tri_ishift(kbin, kcenter, iopen);
    std::vector<int32_t> t;
    for (int l = 1; l <= nb; ++l)
        for (int j = 1; j <= l; ++j) {
            const int imax = cross ? std::min(nb, j + l - 1 + ishift) : j;
            for (int i = std::max(1, l - j + 1 - ishift); i <= imax; ++i) {
                t.push_back(l);
                t.push_back(j);
                t.push_back(i);
            }
        }
    return t;
}

int32_t *get_tri_table(pbi4_ctx *ctx, int nb, float kbin, float kcenter, int iopen, int cross,
                       int64_t itot) {
    auto key = std::make_tuple(nb, 0, kbin, kcenter, iopen, cross);
    auto it = ctx->tri_tables.find(key);
    std::vector<int32_t> t = triangle_list(nb, kbin, kcenter, iopen, cross);
    PBI4_CHECK((int64_t)t.size() / 3 == itot, PBI4_EINVAL,
               "itot does not match the triangle loop (" + std::to_string(t.size() / 3) + ")");
    if (it != ctx->tri_tables.end()) return it->second;
    return ctx->tri_tables[key] = dev_upload(ctx, t);
}

// ------------------------------------------------------------------ workspace layouts

struct DensityWs {
    float2 *G;
    void *fftwork;
};
// Interlaced: one complex FFT of grid1 + i grid2 (the array the reference transforms,
// pbi4.py:337), the two grids passed as its planar real and imaginary parts; on the grids 128,
// 192, 256, 384, 512 rocFFT transforms the z planes over (y, x) and the combine kernel does the z axis
// (PBI4_DENSITY_FFT=rocfft3d keeps rocFFT's 3D transform and the separate combine); single
// grid: r2c.
bool density_fused_z(const pbi4_density_desc &d) {
    const char *e = std::getenv("PBI4_DENSITY_FFT");
    return d.interlacing && zfft_combine_supported(d.grid) && !(e && std::strcmp(e, "rocfft3d") == 0);
}
FftPlan &density_plan(pbi4_ctx *ctx, const pbi4_density_desc &d) {
    if (!d.interlacing) return get_plan(ctx, 0, d.grid, 1, 0);
    return density_fused_z(d) ? get_plan(ctx, 4, d.grid, d.grid, 0) : get_plan(ctx, 2, d.grid, 1, 0);
}
DensityWs density_layout(pbi4_ctx *ctx, const pbi4_density_desc &d, Arena &a) {
    const int64_t N = d.grid;
    FftPlan &p = density_plan(ctx, d);
    DensityWs w;
    w.G = a.take<float2>(d.interlacing ? (size_t)(N * N * N) : (size_t)(N * N * (N / 2 + 1)));
    w.fftwork = a.take<char>(std::max<size_t>(p.work, 1));
    return w;
}

void check_density(const pbi4_density_desc &d) {
    PBI4_CHECK(d.grid >= 4 && d.grid % 2 == 0, PBI4_EINVAL, "grid must be even and >= 4");
    PBI4_CHECK(d.order >= 1 && d.order <= 4, PBI4_EINVAL, "interpolation order must be 1..4");
    PBI4_CHECK(d.box > 0.0f, PBI4_EINVAL, "box must be > 0");
    PBI4_CHECK(d.npart > 0.0, PBI4_EINVAL, "npart must be > 0");
}

BinPlan make_bin_plan(int N, int nbins_pk, int nshell, int multipoles, const double *los, float box) {
    BinPlan b{};
    b.N = N;
    b.nbins_pk = nbins_pk;
    b.nshell = nshell;
    b.multipoles = multipoles;
    for (int i = 0; i < 3; ++i) b.los[i] = (float)los[i];
    b.kF = (float)(double)(kTwoPiF32 / box);
    const int64_t rows = (int64_t)N * N;
    // up to 2048 work-groups of 4 waves: the binning is bound by the latency of its dependent
    // loads and DPP scans and wants every wave slot; each work-group adds one partial per bin to
    // the finalize's reads
    b.nwg = (int)std::max<int64_t>(1, std::min<int64_t>(2048, ceil_div(rows, 8)));
    return b;
}

struct BkLayout {
    int Np, nell, nmaps;
    int64_t K, Kstride;
    GemmPlan g;
    float *bin_partials;
    double *shell_sum;
    float *scale_exp;
    float2 *spectra;
    float *maps;
    void *fftwork;
    float *gemm_partials;
    double *C;
};

void check_bk(const pbi4_bk_desc &d) {
    PBI4_CHECK(d.grid >= 4 && d.grid % 2 == 0, PBI4_EINVAL, "grid must be even and >= 4");
    PBI4_CHECK(d.nbins >= 1, PBI4_EINVAL, "nbins must be >= 1");
    PBI4_CHECK(d.kbin > 0.0f, PBI4_EINVAL, "kbin must be > 0");
    PBI4_CHECK(d.npart > 0, PBI4_EINVAL, "npart must be > 0");
    PBI4_CHECK(d.precision == PBI4_PREC_FP32 || d.precision == PBI4_PREC_BF16X3, PBI4_EINVAL,
               "precision must be PBI4_PREC_FP32 or PBI4_PREC_BF16X3");
    PBI4_CHECK(d.maps_grid == 0 || (d.maps_grid >= 4 && d.maps_grid % 2 == 0 && d.maps_grid <= d.grid),
               PBI4_EINVAL, "maps_grid must be 0 or an even size <= grid");
}

BkLayout bk_layout(pbi4_ctx *ctx, const pbi4_bk_desc &d, Arena &a, int nbins_pk = 0) {
    ShellInfo &sh = get_shells(ctx, d.grid, d.nbins, d.kbin, d.kcenter);
    BkLayout L{};
    L.Np = d.maps_grid ? d.maps_grid : choose_maps_grid(d.grid, sh.R);
    PBI4_CHECK(d.maps_grid == 0 || L.Np == d.grid || 3 * sh.R < L.Np, PBI4_EINVAL,
               "maps_grid too small for the shells (needs > 3 * max|q| = " +
                   std::to_string(3 * sh.R) + ")");
    L.nell = d.multipoles ? 3 : 1;
    L.nmaps = d.nbins * L.nell;
    L.K = (int64_t)L.Np * L.Np * L.Np;
    L.Kstride = (int64_t)align_up((size_t)L.K, 64);
    L.g = make_gemm_plan(d.nbins, L.nell, L.K, L.Kstride, d.precision);
    gemm_set_windows(L.g, triangle_list(d.nbins, d.kbin, d.kcenter, d.iopen, 0), false,
                     tri_ishift(d.kbin, d.kcenter, d.iopen));
    BinPlan b = make_bin_plan(d.grid, nbins_pk, d.nbins, 0, d.los, 1.0f);
    const size_t work = maps_work(ctx, sh, L.Np, d.grid, L.nmaps, L.Kstride);
    L.bin_partials = a.take<float>(bin_partials_floats(b));
    L.shell_sum = a.take<double>(d.nbins);
    L.scale_exp = a.take<float>(d.nbins);
    L.spectra = a.take<float2>((size_t)L.nmaps * L.Np * L.Np * maps_ld(L.Np, d.grid, sh, L.Kstride));
    L.maps = a.take<float>((size_t)L.nmaps * L.Kstride);
    L.fftwork = a.take<char>(std::max<size_t>(work, 1));
    L.gemm_partials = a.take<float>(gemm_partials_floats(L.g));
    L.C = a.take<double>((size_t)L.g.npair * L.g.ncol);
    return L;
}

// Shell filter -> c2r maps -> triangle GEMM -> FP64 finalize, once the shell sums are in L.
void bk_from_shells(pbi4_ctx *ctx, const pbi4_bk_desc *d, BkLayout &L, ShellInfo &sh, int32_t *tri,
                    const float *counts, const void *delta, const void *delta2, const void *delta4,
                    double *bout, hipStream_t s) {
    FilterPlan f{};
    f.N = d->grid;
    f.Np = L.Np;
    f.R = sh.R;
    f.nshell = d->nbins;
    f.nell = L.nell;
    for (int i = 0; i < 3; ++i) f.los[i] = (float)d->los[i];
    f.shell_table = sh.table_d;
    f.max_d2 = sh.max_d2;
    f.shell_d2 = sh.d2range_d;
    f.ld = maps_ld(L.Np, d->grid, sh, L.Kstride);
    f.cols = maps_cols(L.Np, d->grid, sh, L.Kstride);
    f.disk_only = maps_disk_only(L.Np, d->grid, sh, L.Kstride);
    f.f2 = static_cast<const float2 *>(delta2);
    f.f4 = static_cast<const float2 *>(delta4);
    PBI4_CHECK(d->multipoles || (!delta2 && !delta4), PBI4_EINVAL,
               "delta2/delta4 are only used with multipoles");
    {
        TimedScope ts(&ctx->timer, "pbi4_filter", s);
        if (maps_fused(L.Np, d->grid, sh, L.Kstride))
            launch_filter_z(f, static_cast<const float2 *>(delta), L.shell_sum,
                            get_twiddle(ctx, L.Np), L.spectra, L.scale_exp, s);
        else
            launch_filter(f, static_cast<const float2 *>(delta), L.shell_sum, sh.nmodes_d,
                          L.spectra, L.scale_exp, s);
    }
    maps_c2r(ctx, sh, L.Np, d->grid, L.nmaps, L.Kstride, L.spectra, L.maps, L.fftwork, s);
    launch_gemm(L.g, L.maps, L.gemm_partials, s, &ctx->timer);
    launch_gemm_reduce(L.g, L.gemm_partials, L.C, s);
    // kf3 = kf**3 in REAL(4) (bispectrumI4.f90:46); powSN = 1/dble(npart)/kf3 (:47)
    volatile float kfv = d->kf;
    volatile float k2 = kfv * kfv;
    volatile float k3 = k2 * kfv;
    const double kf3 = (double)k3;
    const double powSN = 1.0 / (double)d->npart / kf3;
    const double gridD = (double)d->grid * d->grid * d->grid;
    const double ratio = (double)d->grid / (double)L.Np;
    launch_bk_finalize(d->multipoles ? 11 : 9, d->itot, tri, L.C, L.g,
                       L.scale_exp, L.shell_sum, sh.nmodes_d, counts, kf3, powSN, gridD,
                       ratio * ratio * ratio, (double)d->kcenter, (double)d->kbin, bout, s);
}

// Cross bispectrum workspace: the maps of field 1 then field 2 on the maps grid (2 nb maps), the
// GEMM run with nb pair rows from field 1 and 2 nb columns (field 2 in columns nb .. 2nb-1).
struct BkCrossLayout {
    int Np;
    int64_t K, Kstride;
    GemmPlan g;
    float *bin_partials;
    double *shell_sum;   // 2 nb: field 1, field 2
    float *scale_exp;    // 2 nb
    float2 *spectra;
    float *maps;
    void *fftwork;
    float *gemm_partials;
    double *C;
};

BkCrossLayout bk_cross_layout(pbi4_ctx *ctx, const pbi4_bk_desc &d, Arena &a) {
    PBI4_CHECK(!d.multipoles, PBI4_EINVAL, "the cross bispectrum has no multipoles");
    ShellInfo &sh = get_shells(ctx, d.grid, d.nbins, d.kbin, d.kcenter);
    BkCrossLayout L{};
    L.Np = d.maps_grid ? d.maps_grid : choose_maps_grid(d.grid, sh.R);
    PBI4_CHECK(d.maps_grid == 0 || L.Np == d.grid || 3 * sh.R < L.Np, PBI4_EINVAL,
               "maps_grid too small for the shells (needs > 3 * max|q| = " +
                   std::to_string(3 * sh.R) + ")");
    const int nmaps = 2 * d.nbins;
    L.K = (int64_t)L.Np * L.Np * L.Np;
    L.Kstride = (int64_t)align_up((size_t)L.K, 64);
    L.g = make_gemm_plan(d.nbins, 2, L.K, L.Kstride, d.precision);
    gemm_set_windows(L.g, triangle_list(d.nbins, d.kbin, d.kcenter, d.iopen, 1), true,
                     tri_ishift(d.kbin, d.kcenter, d.iopen));
    BinPlan b = make_bin_plan(d.grid, 0, d.nbins, 0, d.los, 1.0f);
    const size_t work = maps_work(ctx, sh, L.Np, d.grid, nmaps, L.Kstride);
    L.bin_partials = a.take<float>(bin_partials_floats(b));
    L.shell_sum = a.take<double>(nmaps);
    L.scale_exp = a.take<float>(nmaps);
    L.spectra = a.take<float2>((size_t)nmaps * L.Np * L.Np * maps_ld(L.Np, d.grid, sh, L.Kstride));
    L.maps = a.take<float>((size_t)nmaps * L.Kstride);
    L.fftwork = a.take<char>(std::max<size_t>(work, 1));
    L.gemm_partials = a.take<float>(gemm_partials_floats(L.g));
    L.C = a.take<double>((size_t)L.g.npair * L.g.ncol);
    return L;
}
}  // namespace

// ==================================================================== C ABI

extern "C" {

int pbi4_version(void) { return 100; }

int pbi4_ctx_create(int device, pbi4_ctx **out) {
    if (!out) return PBI4_EINVAL;
    auto *c = new pbi4_ctx();
    c->device = device;
    int rc = guarded(c, [&] {
        int n = 0;
        PBI4_HIP(hipGetDeviceCount(&n));
        PBI4_CHECK(device >= 0 && device < n, PBI4_EINVAL, "no such device");
        scratch_setup(device);
    });
    if (rc != PBI4_OK) {
        delete c;
        *out = nullptr;
        return rc;
    }
    *out = c;
    return PBI4_OK;
}

int pbi4_ctx_destroy(pbi4_ctx *ctx) {
    if (!ctx) return PBI4_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipDeviceSynchronize();
    for (auto &kv : ctx->plans)
        if (kv.second.plan) rocfft_plan_destroy(kv.second.plan);
    for (void *p : ctx->owned) (void)hipFree(p);
    for (auto &e : ctx->timer.events) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    delete ctx;
    return PBI4_OK;
}

const char *pbi4_last_error(const pbi4_ctx *ctx) { return ctx ? ctx->err.c_str() : g_err_noctx.c_str(); }

size_t pbi4_assign_workspace(const pbi4_assign_desc *d, int64_t npart) {
    return sized([&] {
        PBI4_CHECK(d != nullptr, PBI4_EINVAL, "NULL descriptor");
        return assign_workspace(make_assign_plan(*d, npart), npart);
    });
}

int pbi4_assign(pbi4_ctx *ctx, const pbi4_assign_desc *d, const float *pos, const float *vel,
                int64_t npart, float *grids, void *ws, size_t ws_bytes, void *stream) {
    return guarded(ctx, [&] {
        PBI4_CHECK(d && grids && ws, PBI4_EINVAL, "NULL argument");
        PBI4_CHECK(npart >= 0 && (npart == 0 || pos), PBI4_EINVAL,
                   "positions missing or a negative particle count");
        AssignPlan p = make_assign_plan(*d, npart);
        Arena a(ws, ws_bytes);
        std::lock_guard<std::mutex> g(ctx->mu);
        assign_run(p, pos, vel, npart, grids, a, (hipStream_t)stream, &ctx->timer);
    });
}

size_t pbi4_density_workspace(pbi4_ctx *ctx, const pbi4_density_desc *d) {
    size_t r = 0;
    guarded(ctx, [&] {
        check_density(*d);
        std::lock_guard<std::mutex> g(ctx->mu);
        Arena a(nullptr, 0);
        density_layout(ctx, *d, a);
        r = a.used + 256;
    });
    return r;
}

int pbi4_compute_density(pbi4_ctx *ctx, const pbi4_density_desc *d, const float *grids,
                         void *delta, void *ws, size_t ws_bytes, void *stream) {
    return guarded(ctx, [&] {
        PBI4_CHECK(d && grids && delta && ws, PBI4_EINVAL, "NULL argument");
        check_density(*d);
        std::lock_guard<std::mutex> g(ctx->mu);
        hipStream_t s = (hipStream_t)stream;
        Arena a(ws, ws_bytes);
        DensityWs w = density_layout(ctx, *d, a);
        const int64_t V = (int64_t)d->grid * d->grid * d->grid;
        FftPlan &p = density_plan(ctx, *d);
        {
            TimedScope ts(&ctx->timer, "pbi4_r2c", s);
            fft_exec(p, const_cast<float *>(grids), w.G, w.fftwork, s,
                     d->interlacing ? const_cast<float *>(grids) + V : nullptr);
        }
        const double kF3 = kf3_fcomb(d->box);
        const double cf = (d->interlacing ? 1.0 / (4.0 * kF3) : 1.0 / kF3) / d->npart;
        TimedScope ts(&ctx->timer, "pbi4_combine", s);
        if (density_fused_z(*d))
            launch_zfft_combine(d->grid, get_window(ctx, d->grid, d->order), get_phase(ctx, d->grid),
                                get_twiddle(ctx, d->grid), (float)cf, w.G,
                                static_cast<float2 *>(delta), s);
        else
            launch_combine(d->grid, d->interlacing, get_window(ctx, d->grid, d->order),
                           d->interlacing ? get_phase(ctx, d->grid) : nullptr, (float)cf, w.G,
                           static_cast<float2 *>(delta), s);
    });
}

int pbi4_assign_density(pbi4_ctx *ctx, const pbi4_assign_desc *ad, const pbi4_density_desc *dd,
                        const float *pos, const float *vel, int64_t npart, float *grids,
                        void *delta, void *ws, size_t ws_bytes, void *stream) {
    bool fused = false;
    int rc = guarded(ctx, [&] {
        PBI4_CHECK(ad && dd && grids && delta && ws, PBI4_EINVAL, "NULL argument");
        PBI4_CHECK(npart >= 0 && (npart == 0 || pos), PBI4_EINVAL,
                   "positions missing or a negative particle count");
        check_density(*dd);
        PBI4_CHECK(ad->grid == dd->grid && ad->order == dd->order &&
                       ad->interlacing == dd->interlacing,
                   PBI4_EINVAL, "assignment and density descriptors differ in grid/order/interlacing");
        AssignPlan p = make_assign_plan(*ad, npart);
        // the fused gather reads the regions in 16-byte vectors: a workspace off that alignment
        // takes the two-step route
        fused = gather_xfft_supported(p) && !std::getenv("PBI4_DENSITY_FFT") &&
                (reinterpret_cast<uintptr_t>(ws) & 15) == 0;
        if (!fused) return;
        hipStream_t s = (hipStream_t)stream;
        Arena a(ws, ws_bytes);
        std::lock_guard<std::mutex> g(ctx->mu);
        const float *regions = assign_run(p, pos, vel, npart, nullptr, a, s, &ctx->timer);
        const int N = dd->grid;
        const float2 *tw = get_twiddle(ctx, N);
        float2 *G = reinterpret_cast<float2 *>(grids);
        {
            TimedScope ts(&ctx->timer, "pbi4_gather", s);
            launch_gather_xfft(p, regions, tw, G, s);
        }
        {
            TimedScope ts(&ctx->timer, "pbi4_r2c", s);
            launch_yfft(N, tw, G, s);
        }
        const double cf = 1.0 / (4.0 * kf3_fcomb(dd->box)) / dd->npart;
        TimedScope ts(&ctx->timer, "pbi4_combine", s);
        launch_zfft_combine(N, get_window(ctx, N, dd->order), get_phase(ctx, N), tw, (float)cf, G,
                            static_cast<float2 *>(delta), s);
    });
    if (rc != PBI4_OK || fused) return rc;
    rc = pbi4_assign(ctx, ad, pos, vel, npart, grids, ws, ws_bytes, stream);
    if (rc != PBI4_OK) return rc;
    return pbi4_compute_density(ctx, dd, grids, delta, ws, ws_bytes, stream);
}

int pbi4_fcomb(pbi4_ctx *ctx, int grid, int nintp, int interlacing, float box, const void *fdcl,
               void *dcl, void *stream) {
    return guarded(ctx, [&] {
        PBI4_CHECK(fdcl && dcl && fdcl != dcl, PBI4_EINVAL, "fdcl and dcl must be distinct buffers");
        PBI4_CHECK(grid >= 2 && grid % 2 == 0, PBI4_EINVAL, "grid must be even");
        PBI4_CHECK(nintp >= 1 && nintp <= 4 && box > 0.0f, PBI4_EINVAL, "bad nintp / box");
        std::lock_guard<std::mutex> g(ctx->mu);
        const double kF3 = kf3_fcomb(box);
        const double cf = interlacing ? 1.0 / (4.0 * kF3) : 1.0 / kF3;
        launch_fcomb_full(grid, interlacing, get_window(ctx, grid, nintp), (float)cf,
                          static_cast<const float2 *>(fdcl), static_cast<float2 *>(dcl),
                          (hipStream_t)stream);
    });
}

int pbi4_compute_qk(pbi4_ctx *ctx, int grid, int norder, const int *idx, const void *deltak,
                    void *out, void *stream) {
    return guarded(ctx, [&] {
        PBI4_CHECK(idx && deltak && out, PBI4_EINVAL, "NULL argument");
        PBI4_CHECK(norder == 2 || norder == 4, PBI4_EINVAL, "norder must be 2 or 4");
        for (int i = 0; i < norder; ++i)
            PBI4_CHECK(idx[i] >= 1 && idx[i] <= 3, PBI4_EINVAL, "indices are 1..3");
        launch_qk(grid, norder, idx, static_cast<const float2 *>(deltak), static_cast<float2 *>(out),
                  (hipStream_t)stream);
    });
}

int pbi4_catalogue_positions(pbi4_ctx *ctx, const void *pos, int pos_f64, const void *vel, int vel_f64,
                             int64_t n, int rsd, const double los[3], double hz, double zf, double box,
                             void *out, int out_f64, void *stream) {
    return guarded(ctx, [&] {
        PBI4_CHECK(pos && out && los && n >= 0, PBI4_EINVAL, "NULL argument or negative count");
        PBI4_CHECK(rsd == PBI4_RSD_NONE || rsd == PBI4_RSD_HZ || rsd == PBI4_RSD_VEL, PBI4_EINVAL,
                   "unknown rsd mode");
        PBI4_CHECK(rsd == PBI4_RSD_NONE || vel, PBI4_EINVAL, "a redshift-space shift needs vel");
        PBI4_CHECK(box > 0.0 && (rsd != PBI4_RSD_HZ || hz != 0.0), PBI4_EINVAL, "box must be > 0, hz != 0");
        if (n == 0) return;
        launch_catalogue_positions(pos, pos_f64 != 0, vel, vel_f64 != 0, n, rsd, los, hz, zf, box,
                                   out, out_f64 != 0, (hipStream_t)stream);
    });
}

int pbi4_reduce_array(pbi4_ctx *ctx, int L, int lx, int N, int nx, int in_bytes, int out_bytes,
                      const void *din, void *dout, void *stream) {
    return guarded(ctx, [&] {
        PBI4_CHECK(din && dout && din != dout, PBI4_EINVAL, "din and dout must be distinct buffers");
        PBI4_CHECK(N >= 1 && N <= L, PBI4_EINVAL, "the requested grid must not exceed the stored one");
        PBI4_CHECK(lx >= N / 2 + 1 && lx <= L && nx >= N / 2 + 1 && nx <= N, PBI4_EINVAL,
                   "x extents must cover x <= N/2");
        PBI4_CHECK((in_bytes == 8 || in_bytes == 16) && (out_bytes == 8 || out_bytes == 16),
                   PBI4_EINVAL, "element sizes are 8 (complex float) or 16 (complex double)");
        launch_reduce_array(L, lx, N, nx, in_bytes, out_bytes, din, dout, (hipStream_t)stream);
    });
}

size_t pbi4_pk_workspace(pbi4_ctx *ctx, const pbi4_pk_desc *d) {
    (void)ctx;
    BinPlan b = make_bin_plan(d->grid, d->nbins, 0, d->multipoles, d->los, d->box > 0 ? d->box : 1.0f);
    return align_up(bin_partials_floats(b) * sizeof(float), 256) + 256;
}

int pbi4_measure_pk(pbi4_ctx *ctx, const pbi4_pk_desc *d, const void *delta, const void *delta_b,
                    const void *delta2, const void *delta4, double *pout, void *ws,
                    size_t ws_bytes, void *stream) {
    return guarded(ctx, [&] {
        PBI4_CHECK(d && delta && pout && ws, PBI4_EINVAL, "NULL argument");
        PBI4_CHECK(d->grid >= 4 && d->grid % 2 == 0, PBI4_EINVAL, "grid must be even and >= 4");
        PBI4_CHECK(d->box > 0.0f, PBI4_EINVAL, "box must be > 0");
        std::lock_guard<std::mutex> g(ctx->mu);
        hipStream_t s = (hipStream_t)stream;
        PkTable &t = get_pk_table(ctx, d->grid, d->box, d->kbin, d->kcenter, d->nbins);
        BinPlan b = make_bin_plan(d->grid, d->nbins, 0, d->multipoles, d->los, d->box);
        b.pk_table = t.d;
        b.cross = static_cast<const float2 *>(delta_b);
        b.f2 = static_cast<const float2 *>(delta2);
        b.f4 = static_cast<const float2 *>(delta4);
        PBI4_CHECK(d->multipoles || (!delta2 && !delta4), PBI4_EINVAL,
                   "delta2/delta4 are only used with multipoles");
        PBI4_CHECK(!d->multipoles || !delta_b, PBI4_EINVAL,
                   "a cross spectrum is real-space only (pbi4/pbi4.py:434-436)");
        Arena a(ws, ws_bytes);
        float *part = a.take<float>(bin_partials_floats(b));
        const double kF = (double)(kTwoPiF32 / d->box);
        {
            TimedScope ts(&ctx->timer, "pbi4_bin", s);
            launch_bin(b, static_cast<const float2 *>(delta), part, s);
        }
        launch_bin_finalize(b, part, kF * kF * kF, pout, nullptr, s);
    });
}

int pbi4_count_triangles(pbi4_ctx *ctx, int nbins, int grid, float kbin, float kcenter, int iopen,
                         int cross, int64_t itot, int64_t *ntr) {
    return guarded(ctx, [&] {
        PBI4_CHECK(ntr, PBI4_EINVAL, "NULL argument");
        PBI4_CHECK(grid >= 4 && grid % 2 == 0, PBI4_EINVAL, "grid must be even and >= 4");
        std::lock_guard<std::mutex> g(ctx->mu);
        std::vector<int32_t> tri = triangle_list(nbins, kbin, kcenter, iopen, cross);
        PBI4_CHECK((int64_t)tri.size() / 3 == itot, PBI4_EINVAL,
                   "itot does not match the triangle loop (" + std::to_string(tri.size() / 3) + ")");
        ShellInfo &sh = get_shells(ctx, grid, nbins, kbin, kcenter);
        auto key = std::make_tuple(grid, nbins, kbin, kcenter);
        auto it = ctx->counts.find(key);
        if (it == ctx->counts.end()) {
            const int N = grid, hN = N / 2, S = nbins + 1;
            // residues per axis with distance <= R
            std::vector<int> ax;
            for (int i = 0; i < N; ++i)
                if (std::min(i, N - i) <= sh.R) ax.push_back(i <= hN ? i : i - N);
            std::vector<std::vector<int32_t>> per(nbins);
            for (int x : ax)
                for (int y : ax)
                    for (int z : ax) {
                        const long d2 = (long)x * x + (long)y * y + (long)z * z;
                        const int s = d2 <= sh.max_d2 ? sh.table[d2] : 0;
                        if (s) {
                            auto &v = per[s - 1];
                            v.push_back(x);
                            v.push_back(y);
                            v.push_back(z);
                            v.push_back(s);
                        }
                    }
            std::vector<int32_t> modes;
            std::vector<int64_t> off(nbins + 1, 0);
            for (int s = 0; s < nbins; ++s) {
                off[s] = (int64_t)modes.size() / 4;
                modes.insert(modes.end(), per[s].begin(), per[s].end());
            }
            off[nbins] = (int64_t)modes.size() / 4;
            // canonical representatives a >= b >= c >= 0 and their orbit sizes
            std::vector<int32_t> canon, w;
            for (int a0 = 0; a0 <= std::min(sh.R, hN); ++a0)
                for (int b0 = 0; b0 <= a0; ++b0)
                    for (int c0 = 0; c0 <= b0; ++c0) {
                        const long d2 = (long)a0 * a0 + (long)b0 * b0 + (long)c0 * c0;
                        const int s = d2 <= sh.max_d2 ? sh.table[d2] : 0;
                        if (!s) continue;
                        int perms = (a0 == b0 && b0 == c0) ? 1 : (a0 == b0 || b0 == c0) ? 3 : 6;
                        int signs = 1;
                        for (int v : {a0, b0, c0})
                            if (v != 0 && 2 * v != N) signs *= 2;
                        canon.insert(canon.end(), {a0, b0, c0, s});
                        w.push_back(perms * signs);
                    }
            std::vector<int32_t> modes_v = modes;
            int32_t *modes_d = dev_upload(ctx, modes_v);
            int64_t *off_d = dev_upload(ctx, off);
            int32_t *canon_d = dev_upload(ctx, canon);
            int32_t *w_d = dev_upload(ctx, w);
            unsigned long long *T_d = nullptr;
            const size_t nT = (size_t)S * S * S;
            PBI4_HIP(hipMalloc(&T_d, nT * sizeof(unsigned long long)));
            PBI4_HIP(hipMemset(T_d, 0, nT * sizeof(unsigned long long)));
            count_pairs_gpu(N, nbins, modes_d, off[nbins], off_d, canon_d, w_d, (int64_t)w.size(),
                            sh.table_d, sh.max_d2, T_d, nullptr);
            CountsCache cc;
            cc.T.resize(nT);
            PBI4_HIP(hipMemcpy(cc.T.data(), T_d, nT * sizeof(unsigned long long), hipMemcpyDeviceToHost));
            PBI4_HIP(hipFree(T_d));
            it = ctx->counts.emplace(key, std::move(cc)).first;
        }
        const int S = nbins + 1;
        for (int64_t m = 0; m < itot; ++m) {
            const int l = tri[3 * m], j = tri[3 * m + 1], i = tri[3 * m + 2];
            const int a = std::min(i, j), b = std::max(i, j);
            ntr[m] = (int64_t)it->second.T[((size_t)a * S + b) * S + l];
        }
    });
}

int pbi4_bk_maps_grid(pbi4_ctx *ctx, const pbi4_bk_desc *d) {
    int r = 0;
    int rc = guarded(ctx, [&] {
        check_bk(*d);
        std::lock_guard<std::mutex> g(ctx->mu);
        ShellInfo &sh = get_shells(ctx, d->grid, d->nbins, d->kbin, d->kcenter);
        r = d->maps_grid ? d->maps_grid : choose_maps_grid(d->grid, sh.R);
    });
    return rc == PBI4_OK ? r : rc;
}

size_t pbi4_bk_workspace(pbi4_ctx *ctx, const pbi4_bk_desc *d) {
    size_t r = 0;
    guarded(ctx, [&] {
        check_bk(*d);
        std::lock_guard<std::mutex> g(ctx->mu);
        Arena a(nullptr, 0);
        bk_layout(ctx, *d, a);
        r = a.used + 256;
    });
    return r;
}

int pbi4_measure_bk(pbi4_ctx *ctx, const pbi4_bk_desc *d, const float *counts, const void *delta,
                    const void *delta2, const void *delta4, double *bout, void *ws,
                    size_t ws_bytes, void *stream) {
    return guarded(ctx, [&] {
        PBI4_CHECK(d && counts && delta && bout && ws, PBI4_EINVAL, "NULL argument");
        check_bk(*d);
        std::lock_guard<std::mutex> g(ctx->mu);
        hipStream_t s = (hipStream_t)stream;
        ShellInfo &sh = get_shells(ctx, d->grid, d->nbins, d->kbin, d->kcenter);
        int32_t *tri = get_tri_table(ctx, d->nbins, d->kbin, d->kcenter, d->iopen, 0, d->itot);
        Arena a(ws, ws_bytes);
        BkLayout L = bk_layout(ctx, *d, a);
        // shell powers (Parseval) for pow(s) and the map scales
        BinPlan b = make_bin_plan(d->grid, 0, d->nbins, 0, d->los, 1.0f);
        b.shell_table = sh.table_d;
        b.max_d2 = sh.max_d2;
        launch_bin(b, static_cast<const float2 *>(delta), L.bin_partials, s);
        launch_bin_finalize(b, L.bin_partials, 1.0, nullptr, L.shell_sum, s);
        bk_from_shells(ctx, d, L, sh, tri, counts, delta, delta2, delta4, bout, s);
    });
}

size_t pbi4_bk_cross_workspace(pbi4_ctx *ctx, const pbi4_bk_desc *d) {
    size_t r = 0;
    guarded(ctx, [&] {
        PBI4_CHECK(d, PBI4_EINVAL, "NULL argument");
        check_bk(*d);
        std::lock_guard<std::mutex> g(ctx->mu);
        Arena a(nullptr, 0);
        bk_cross_layout(ctx, *d, a);
        r = a.used + 256;
    });
    return r;
}

int pbi4_measure_bk_cross(pbi4_ctx *ctx, const pbi4_bk_desc *d, const float *counts,
                          const void *delta1, const void *delta2, double *bout, void *ws,
                          size_t ws_bytes, void *stream) {
    return guarded(ctx, [&] {
        PBI4_CHECK(d && counts && delta1 && delta2 && bout && ws, PBI4_EINVAL, "NULL argument");
        check_bk(*d);
        std::lock_guard<std::mutex> g(ctx->mu);
        hipStream_t s = (hipStream_t)stream;
        ShellInfo &sh = get_shells(ctx, d->grid, d->nbins, d->kbin, d->kcenter);
        int32_t *tri = get_tri_table(ctx, d->nbins, d->kbin, d->kcenter, d->iopen, 1, d->itot);
        Arena a(ws, ws_bytes);
        BkCrossLayout L = bk_cross_layout(ctx, *d, a);
        const int nb = d->nbins;
        const size_t spec = (size_t)L.Np * L.Np * maps_ld(L.Np, d->grid, sh, L.Kstride);
        const void *fields[2] = {delta1, delta2};
        for (int f = 0; f < 2; ++f) {
            // shell powers (Parseval) -> map scales; shell filter into maps f*nb .. f*nb+nb-1
            BinPlan b = make_bin_plan(d->grid, 0, nb, 0, d->los, 1.0f);
            b.shell_table = sh.table_d;
            b.max_d2 = sh.max_d2;
            launch_bin(b, static_cast<const float2 *>(fields[f]), L.bin_partials, s);
            launch_bin_finalize(b, L.bin_partials, 1.0, nullptr, L.shell_sum + f * nb, s);
            FilterPlan fp{};
            fp.N = d->grid;
            fp.Np = L.Np;
            fp.R = sh.R;
            fp.nshell = nb;
            fp.nell = 1;
            for (int i = 0; i < 3; ++i) fp.los[i] = (float)d->los[i];
            fp.shell_table = sh.table_d;
            fp.max_d2 = sh.max_d2;
            fp.shell_d2 = sh.d2range_d;
            fp.ld = maps_ld(L.Np, d->grid, sh, L.Kstride);
            fp.cols = maps_cols(L.Np, d->grid, sh, L.Kstride);
            fp.disk_only = maps_disk_only(L.Np, d->grid, sh, L.Kstride);
            TimedScope ts(&ctx->timer, "pbi4_filter", s);
            if (maps_fused(L.Np, d->grid, sh, L.Kstride))
                launch_filter_z(fp, static_cast<const float2 *>(fields[f]), L.shell_sum + f * nb,
                                get_twiddle(ctx, L.Np), L.spectra + f * nb * spec,
                                L.scale_exp + f * nb, s);
            else
                launch_filter(fp, static_cast<const float2 *>(fields[f]), L.shell_sum + f * nb,
                              sh.nmodes_d, L.spectra + f * nb * spec, L.scale_exp + f * nb, s);
        }
        maps_c2r(ctx, sh, L.Np, d->grid, 2 * nb, L.Kstride, L.spectra, L.maps, L.fftwork, s);
        launch_gemm(L.g, L.maps, L.gemm_partials, s, &ctx->timer);
        launch_gemm_reduce(L.g, L.gemm_partials, L.C, s);
        // kf3 = kf**3 in REAL(4) (bispectrumI4.f90:569)
        volatile float kfv = d->kf;
        volatile float k2 = kfv * kfv;
        volatile float k3 = k2 * kfv;
        const double gridD = (double)d->grid * d->grid * d->grid;
        const double ratio = (double)d->grid / (double)L.Np;
        launch_bk_cross_finalize(d->itot, tri, L.C, L.g, L.scale_exp, counts, (double)k3,
                                 gridD, ratio * ratio * ratio, (double)d->kcenter, (double)d->kbin,
                                 bout, s);
    });
}

size_t pbi4_pk_bk_workspace(pbi4_ctx *ctx, const pbi4_pk_desc *pd, const pbi4_bk_desc *d) {
    size_t r = 0;
    guarded(ctx, [&] {
        check_bk(*d);
        std::lock_guard<std::mutex> g(ctx->mu);
        Arena a(nullptr, 0);
        bk_layout(ctx, *d, a, pd->nbins);
        r = a.used + 256;
    });
    return r;
}

int pbi4_measure_pk_bk(pbi4_ctx *ctx, const pbi4_pk_desc *pd, const pbi4_bk_desc *d,
                       const float *counts, const void *delta, double *pout, double *bout,
                       void *ws, size_t ws_bytes, void *stream) {
    return guarded(ctx, [&] {
        PBI4_CHECK(pd && d && counts && delta && pout && bout && ws, PBI4_EINVAL, "NULL argument");
        check_bk(*d);
        PBI4_CHECK(pd->grid == d->grid && pd->multipoles == d->multipoles, PBI4_EINVAL,
                   "P(k) and bispectrum descriptors disagree on grid / multipoles");
        PBI4_CHECK(pd->box > 0.0f, PBI4_EINVAL, "box must be > 0");
        std::lock_guard<std::mutex> g(ctx->mu);
        hipStream_t s = (hipStream_t)stream;
        ShellInfo &sh = get_shells(ctx, d->grid, d->nbins, d->kbin, d->kcenter);
        PkTable &t = get_pk_table(ctx, pd->grid, pd->box, pd->kbin, pd->kcenter, pd->nbins);
        int32_t *tri = get_tri_table(ctx, d->nbins, d->kbin, d->kcenter, d->iopen, 0, d->itot);
        Arena a(ws, ws_bytes);
        BkLayout L = bk_layout(ctx, *d, a, pd->nbins);
        // one pass over delta: P_l bins and shell powers together
        BinPlan b = make_bin_plan(d->grid, pd->nbins, d->nbins, pd->multipoles, pd->los, pd->box);
        b.pk_table = t.d;
        b.shell_table = sh.table_d;
        b.max_d2 = sh.max_d2;
        {
            TimedScope ts(&ctx->timer, "pbi4_bin", s);
            launch_bin(b, static_cast<const float2 *>(delta), L.bin_partials, s);
        }
        const double kF = (double)(kTwoPiF32 / pd->box);
        launch_bin_finalize(b, L.bin_partials, kF * kF * kF, pout, L.shell_sum, s);
        bk_from_shells(ctx, d, L, sh, tri, counts, delta, nullptr, nullptr, bout, s);
    });
}

size_t pbi4_triangle_sums_workspace(int nb, int nell, int64_t K, int64_t Kstride, int precision) {
    return sized([&] {
        GemmPlan g = make_gemm_plan(nb, nell, K, Kstride, precision);
        return align_up(gemm_partials_floats(g) * sizeof(float), 256) + 256;
    });
}

int pbi4_triangle_sums(pbi4_ctx *ctx, int nb, int nell, int64_t K, int64_t Kstride, int precision,
                       const float *maps, double *C, void *ws, size_t ws_bytes, void *stream) {
    return guarded(ctx, [&] {
        PBI4_CHECK(maps && C && ws, PBI4_EINVAL, "NULL argument");
        PBI4_CHECK(nb >= 1 && (nell == 1 || nell == 3), PBI4_EINVAL, "bad nb / nell");
        PBI4_CHECK(Kstride >= K && K > 0, PBI4_EINVAL, "bad K / Kstride");
        PBI4_CHECK(precision == PBI4_PREC_FP32 || precision == PBI4_PREC_BF16X3, PBI4_EINVAL,
                   "bad precision");
        std::lock_guard<std::mutex> g(ctx->mu);
        GemmPlan gp = make_gemm_plan(nb, nell, K, Kstride, precision);
        Arena a(ws, ws_bytes);
        float *part = a.take<float>(gemm_partials_floats(gp));
        launch_gemm(gp, maps, part, (hipStream_t)stream, &ctx->timer);
        launch_gemm_reduce(gp, part, C, (hipStream_t)stream);
    });
}

int pbi4_triangle_sums_noalloc(pbi4_ctx *ctx, int nb, int nell, int64_t K, int64_t Kstride,
                               int precision, const float *maps, double *C, void *stream) {
    return guarded(ctx, [&] {
        PBI4_CHECK(maps && C, PBI4_EINVAL, "NULL argument");
        PBI4_CHECK(nb >= 1 && (nell == 1 || nell == 3), PBI4_EINVAL, "bad nb / nell");
        PBI4_CHECK(Kstride >= K && K > 0, PBI4_EINVAL, "bad K / Kstride");
        PBI4_CHECK(precision == PBI4_PREC_FP32 || precision == PBI4_PREC_BF16X3, PBI4_EINVAL,
                   "bad precision");
        std::lock_guard<std::mutex> g(ctx->mu);
        GemmPlan gp = make_gemm_plan(nb, nell, K, Kstride, precision);
        const size_t need = gemm_partials_floats(gp) * sizeof(float);
        PBI4_CHECK(need <= (size_t)PBI4_NOALLOC_SCRATCH_BYTES, PBI4_ENOMEM,
                   "triangle sums need " + std::to_string(need) +
                       " bytes of partials, more than the allocation-free scratch holds (" +
                       std::to_string((size_t)PBI4_NOALLOC_SCRATCH_BYTES) +
                       "); use pbi4_triangle_sums with a workspace");
        // the same plan and kernels as pbi4_triangle_sums: every partial the reduce reads is
        // written by the GEMM first, so nothing in the scratch needs clearing
        float *part = static_cast<float *>(scratch_of(ctx->device));
        launch_gemm(gp, maps, part, (hipStream_t)stream, nullptr);
        launch_gemm_reduce(gp, part, C, (hipStream_t)stream);
    });
}

int pbi4_timing_enable(pbi4_ctx *ctx, const char *kernel_name) {
    return guarded(ctx, [&] {
        std::lock_guard<std::mutex> g(ctx->mu);
        for (auto &e : ctx->timer.events) {
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
        ctx->timer.events.clear();
        ctx->timer.name = kernel_name ? kernel_name : "";
    });
}

int pbi4_timing_read(pbi4_ctx *ctx, double *total_ms, int64_t *launches) {
    return guarded(ctx, [&] {
        std::lock_guard<std::mutex> g(ctx->mu);
        double t = 0.0;
        for (auto &e : ctx->timer.events) {
            PBI4_HIP(hipEventSynchronize(e.second));
            float ms = 0.0f;
            PBI4_HIP(hipEventElapsedTime(&ms, e.first, e.second));
            t += ms;
        }
        if (total_ms) *total_ms = t;
        if (launches) *launches = (int64_t)ctx->timer.events.size();
    });
}

}  // extern "C"
