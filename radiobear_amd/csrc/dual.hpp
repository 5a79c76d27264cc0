// Forward-mode dual numbers for the absorption tangents: value plus the
// derivatives with respect to (T, q_NH3, q_H2O) of one layer.
#pragma once

#include <hip/hip_runtime.h>

namespace rbx {

using ::exp;
using ::log;
using ::pow;

constexpr int kNd = 3;   // d/dT, d/dq_NH3, d/dq_H2O

struct Dual {
    double v;
    double d[kNd];
    __device__ Dual() : v(0.0), d{0.0, 0.0, 0.0} {}
    __device__ Dual(double x) : v(x), d{0.0, 0.0, 0.0} {}  // NOLINT: constants promote
    __device__ Dual(double x, double a, double b, double c) : v(x), d{a, b, c} {}
};

__device__ inline double val(double x) { return x; }
__device__ inline double val(const Dual &x) { return x.v; }

__device__ inline Dual operator+(const Dual &a, const Dual &b) {
    return {a.v + b.v, a.d[0] + b.d[0], a.d[1] + b.d[1], a.d[2] + b.d[2]};
}
__device__ inline Dual operator-(const Dual &a, const Dual &b) {
    return {a.v - b.v, a.d[0] - b.d[0], a.d[1] - b.d[1], a.d[2] - b.d[2]};
}
__device__ inline Dual operator-(const Dual &a) { return {-a.v, -a.d[0], -a.d[1], -a.d[2]}; }
__device__ inline Dual operator*(const Dual &a, const Dual &b) {
    return {a.v * b.v, a.d[0] * b.v + a.v * b.d[0], a.d[1] * b.v + a.v * b.d[1], a.d[2] * b.v + a.v * b.d[2]};
}
__device__ inline Dual operator/(const Dual &a, const Dual &b) {
    const double r = 1.0 / b.v;   // one division; the quotient and tangents multiply by it
    const double q = a.v * r;
    return {q, (a.d[0] - q * b.d[0]) * r, (a.d[1] - q * b.d[1]) * r, (a.d[2] - q * b.d[2]) * r};
}
__device__ inline Dual operator+(const Dual &a, double b) { return {a.v + b, a.d[0], a.d[1], a.d[2]}; }
__device__ inline Dual operator+(double b, const Dual &a) { return a + b; }
__device__ inline Dual operator-(const Dual &a, double b) { return {a.v - b, a.d[0], a.d[1], a.d[2]}; }
__device__ inline Dual operator-(double b, const Dual &a) { return {b - a.v, -a.d[0], -a.d[1], -a.d[2]}; }
__device__ inline Dual operator*(const Dual &a, double b) { return {a.v * b, a.d[0] * b, a.d[1] * b, a.d[2] * b}; }
__device__ inline Dual operator*(double b, const Dual &a) { return a * b; }
__device__ inline Dual operator/(const Dual &a, double b) {
    const double r = 1.0 / b;
    return {a.v * r, a.d[0] * r, a.d[1] * r, a.d[2] * r};
}
__device__ inline Dual operator/(double a, const Dual &b) {
    const double q = a / b.v;
    const double s = -q / b.v;
    return {q, s * b.d[0], s * b.d[1], s * b.d[2]};
}
__device__ inline Dual chain(const Dual &a, double f, double df) {
    return {f, df * a.d[0], df * a.d[1], df * a.d[2]};
}
__device__ inline Dual exp(const Dual &a) {
    const double e = ::exp(a.v);
    return chain(a, e, e);
}
__device__ inline Dual log(const Dual &a) { return chain(a, ::log(a.v), 1.0 / a.v); }
__device__ inline Dual pow(const Dual &a, double p) {
    const double y = ::pow(a.v, p);
    return chain(a, y, p * ::pow(a.v, p - 1.0));
}
__device__ inline Dual pow(const Dual &a, const Dual &p) {  // a^p, a > 0 (NaN otherwise, as pow)
    const double y = ::pow(a.v, p.v);
    const double la = ::log(a.v);
    const double da = p.v * ::pow(a.v, p.v - 1.0);
    return {y, da * a.d[0] + y * la * p.d[0], da * a.d[1] + y * la * p.d[1], da * a.d[2] + y * la * p.d[2]};
}
// min(a, 1) as the plug-ins clamp: derivative 0 where the clamp is active
__device__ inline Dual fmin1(const Dual &a) { return a.v > 1.0 ? Dual(1.0) : a; }
__device__ inline double fmin1(double a) { return a > 1.0 ? 1.0 : a; }

}  // namespace rbx
