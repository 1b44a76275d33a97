// Complex arithmetic and compile-time register DFTs shared by the HIP translation units
// (internal linkage in each).
#pragma once

#include "common.h"

namespace pbi4 {
namespace {

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 conjf2(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 times_i(float2 a) { return make_float2(-a.y, a.x); }

// In-place inverse DFT of R values, y_k = sum_m a_m e^{+2 pi i m k / R}.  Constants are cos and sin
// of 2 pi / 3, 2 pi / 5 and 4 pi / 5 rounded once from their double values.
template <int R>
__device__ __forceinline__ void idft(float2 *a) {
    if constexpr (R == 2) {
        const float2 t = a[0];
        a[0] = cadd(t, a[1]);
        a[1] = csub(t, a[1]);
    } else if constexpr (R == 3) {
        const float2 s = cadd(a[1], a[2]);
        const float2 u = times_i(cscale(csub(a[1], a[2]), 0.8660254037844386f));
        const float2 t = make_float2(a[0].x - 0.5f * s.x, a[0].y - 0.5f * s.y);
        a[0] = cadd(a[0], s);
        a[1] = cadd(t, u);
        a[2] = csub(t, u);
    } else if constexpr (R == 4) {
        const float2 s02 = cadd(a[0], a[2]), d02 = csub(a[0], a[2]);
        const float2 s13 = cadd(a[1], a[3]), d13 = times_i(csub(a[1], a[3]));
        a[0] = cadd(s02, s13);
        a[2] = csub(s02, s13);
        a[1] = cadd(d02, d13);
        a[3] = csub(d02, d13);
    } else {
        const float c1 = 0.30901699437494745f, c2 = -0.8090169943749475f;
        const float s1 = 0.9510565162951535f, s2 = 0.5877852522924731f;
        const float2 b1 = cadd(a[1], a[4]), b2 = cadd(a[2], a[3]);
        const float2 d1 = csub(a[1], a[4]), d2 = csub(a[2], a[3]);
        const float2 t1 = make_float2(a[0].x + c1 * b1.x + c2 * b2.x, a[0].y + c1 * b1.y + c2 * b2.y);
        const float2 t2 = make_float2(a[0].x + c2 * b1.x + c1 * b2.x, a[0].y + c2 * b1.y + c1 * b2.y);
        const float2 u1 = times_i(make_float2(s1 * d1.x + s2 * d2.x, s1 * d1.y + s2 * d2.y));
        const float2 u2 = times_i(make_float2(s2 * d1.x - s1 * d2.x, s2 * d1.y - s1 * d2.y));
        a[0] = cadd(a[0], cadd(b1, b2));
        a[1] = cadd(t1, u1);
        a[4] = csub(t1, u1);
        a[2] = cadd(t2, u2);
        a[3] = csub(t2, u2);
    }
}

// Compile-time inverse DFT of L values in registers, natural order in and out:
// X[k0 + M k1] = sum_m W_L^{m k0} W_r^{m k1} Y_m[k0] with L = r M, Y_m the length-M transform of
// x[m + r i] (decimation in time); radices 5, 4, 3, 2 first to last.  The twiddles W_L^e =
// e^{2 pi i e / L} are compile-time constants from a Taylor series in double, rounded once.
constexpr double kPiD = 3.14159265358979323846;
constexpr double cx_sin(double x) {
    double term = x, sum = x;
    for (int n = 1; n < 30; ++n) {
        term *= -x * x / ((2.0 * n) * (2.0 * n + 1.0));
        sum += term;
    }
    return sum;
}
constexpr double cx_cos(double x) {
    double term = 1.0, sum = 1.0;
    for (int n = 1; n < 30; ++n) {
        term *= -x * x / ((2.0 * n - 1.0) * (2.0 * n));
        sum += term;
    }
    return sum;
}
template <int L>
struct TwTab {
    float re[L], im[L];
    constexpr TwTab() : re(), im() {
        for (int e = 0; e < L; ++e) {
            const int f = 2 * e > L ? e - L : e;            // angle in [-pi, pi]
            const double a = 2.0 * kPiD * (double)f / (double)L;
            re[e] = (float)cx_cos(a);
            im[e] = (float)cx_sin(a);
        }
    }
};
template <int L>
__device__ __forceinline__ float2 tw_const(int e) {
    constexpr TwTab<L> t{};
    return make_float2(t.re[e % L], t.im[e % L]);
}

template <int L>
__device__ __forceinline__ void rdft(float2 *x) {
    if constexpr (L == 1) {
        return;
    } else if constexpr (L == 2 || L == 3 || L == 4 || L == 5) {
        idft<L>(x);
    } else {
        constexpr int r = L % 5 == 0 ? 5 : L % 4 == 0 ? 4 : L % 3 == 0 ? 3 : 2;
        constexpr int M = L / r;
        static_assert(r * M == L, "lengths are 2,3,5-smooth");
        float2 sub[r][M];
#pragma unroll
        for (int m = 0; m < r; ++m) {
#pragma unroll
            for (int i = 0; i < M; ++i) sub[m][i] = x[m + r * i];
            rdft<M>(sub[m]);
        }
#pragma unroll
        for (int k0 = 0; k0 < M; ++k0) {
            float2 a[r];
#pragma unroll
            for (int m = 0; m < r; ++m) a[m] = m * k0 ? cmul(sub[m][k0], tw_const<L>(m * k0)) : sub[m][k0];
            idft<r>(a);
#pragma unroll
            for (int k1 = 0; k1 < r; ++k1) x[k0 + M * k1] = a[k1];
        }
    }
}

}  // namespace
}  // namespace pbi4
