// Shared host/device helpers of librbx (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>

#include "rbx.h"

namespace rbx {

// ---- thread-local error string + status helpers (host) ----
void set_error(const char *fmt, ...);
const std::string &last_error();

#define RBX_CHECK_ARG(cond, ...)            \
    do {                                    \
        if (!(cond)) {                      \
            ::rbx::set_error(__VA_ARGS__);  \
            return RBX_ERR_ARG;             \
        }                                   \
    } while (0)

#define RBX_CHECK_HIP(expr)                                                          \
    do {                                                                             \
        hipError_t e_ = (expr);                                                      \
        if (e_ != hipSuccess) {                                                      \
            ::rbx::set_error("%s failed: %s", #expr, hipGetErrorString(e_));         \
            return RBX_ERR_HIP;                                                      \
        }                                                                            \
    } while (0)

// Launch shape of one call: `in` (NULL = all built in) checked and copied to `out`.
constexpr int kMaxLaunchSplit = 16;
int read_launch(const rbx_launch *in, rbx_launch &out, const char *who);

inline int launch_status(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("launch of %s failed: %s", what, hipGetErrorString(e));
        return RBX_ERR_HIP;
    }
    return RBX_OK;
}

// ---- wavefront-level helpers ----
// LDS written by some lanes of a wavefront and then read by other lanes of the
// same wavefront: keeps the compiler from moving accesses across this point
// (the LDS serves one wavefront's requests in order).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// ---- device math ----
constexpr double kPi = 3.14159265358979323846;
constexpr double kLog2e = 1.4426950408889634074;
constexpr double kTcmb = 2.725;  // utils.py:77

// e^{-tau} for tau >= 0 (or NaN): FP64 range reduction t = -tau*log2(e) = n + r,
// |r| <= 1/2, then 2^r with the hardware FP32 exp2 and an exact FP64 scaling.
// Relative error ~1e-7, independent of tau.
// Branch-free: n is clamped at -1100 (the result underflows to 0 there), and a
// NaN tau gives a NaN r, so NaN propagates as with exp().
__device__ __forceinline__ double exp_neg_mixed(double tau) {
    const double t = -tau * kLog2e;
    double n = __builtin_rint(t);
    n = n < -1100.0 ? -1100.0 : n;
    const float r = static_cast<float>(t - n);
    const float e = __builtin_amdgcn_exp2f(r);
    return __builtin_ldexp(static_cast<double>(e), static_cast<int>(n));
}

// Exponential integral E2(x), x >= 0, FP64.  Numerical Recipes (3rd ed., 6.3)
// / Abramowitz & Stegun 5.1.11 power series for x <= 1 and the modified-Lentz
// continued fraction (A&S 5.1.22) above, with n = 2.
__device__ inline double expint_e2(double x) {
    const double euler = 0.5772156649015328606;
    const double eps = 1.0e-16;
    if (x != x) return x;
    if (x <= 0.0) return 1.0;
    if (x > 740.0) return 0.0;
    if (x > 1.0) {
        double b = x + 2.0;
        double c = 1.0e300;
        double d = 1.0 / b;
        double h = d;
        for (int i = 1; i < 200; ++i) {
            double an = -static_cast<double>(i) * (1.0 + i);
            b += 2.0;
            d = 1.0 / (an * d + b);
            c = b + an / c;
            double del = c * d;
            h *= del;
            if (fabs(del - 1.0) <= 2.3e-16) break;
        }
        return h * exp(-x);
    }
    // nm1 = 1: ans = 1/nm1, series terms; the i == nm1 term uses psi(2) = 1 - euler
    double ans = 1.0;
    double fact = 1.0;
    for (int i = 1; i < 200; ++i) {
        fact *= -x / i;
        double del;
        if (i != 1) {
            del = -fact / (i - 1);
        } else {
            double psi = 1.0 - euler;
            del = fact * (-log(x) + psi);
        }
        ans += del;
        if (fabs(del) < fabs(ans) * eps) break;
    }
    return ans;
}

}  // namespace rbx
