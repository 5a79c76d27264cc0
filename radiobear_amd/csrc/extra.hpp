// The optional absorbers, evaluated per (layer, frequency) in FP64: co_ddb (the
// Neptune default), clouds_idp and the tabulated H2 CIA h2_orton.  None of them
// fits the canonical line form of the main line sums (co_ddb's Voigt is a
// complex rational function of the detuning, the other two are closed forms or
// table interpolations), and each costs at most a few thousand FP64 operations
// per entry, so they run in their own small kernel only when selected.
#pragma once

#include "catalog.hpp"

namespace rbx {

struct Cplx {
    double re, im;
};

__device__ inline Cplx cmul(Cplx a, Cplx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }

// a / b by Smith's scaling (no overflow of |b|^2), as Python's and numpy's complex division
__device__ inline Cplx cdiv(Cplx a, Cplx b) {
    if (fabs(b.re) >= fabs(b.im)) {
        const double r = b.im / b.re, d = b.re + b.im * r;
        return {(a.re + a.im * r) / d, (a.im - a.re * r) / d};
    }
    const double r = b.re / b.im, d = b.re * r + b.im;
    return {(a.re * r + a.im) / d, (a.im * r - a.re) / d};
}

// ---- co_ddb (constituents/co/co_ddb.py:22-97) ----
constexpr double kCoT0 = 296.0;          // co_ddb.py:7
constexpr double kCoHck = 1.438396;      // co_ddb.py:8
constexpr double kCoGHz = 29.9792458;    // co_ddb.py:9
constexpr double kCoCoef = 7.244E+21;    // co_ddb.py:6
// Voigt rational approximation (Janssen p67, co_ddb.py:24-29); numerator degree 6
__constant__ double kVoigtA[8] = {122.60793178, 214.38238869, 181.92853309, 93.15558046,
                                  30.18014220,  5.91262621,   0.56418958,   0.0};
__constant__ double kVoigtB[8] = {122.60793178, 352.73062511, 457.33447878, 348.70391772,
                                  170.35400182, 53.99290691,  10.47985711,  1.0};

// shape: 0 'voigt', 1 'vvw', 2 'diff' (sums of line shapes, no strengths, co_ddb.py:86-91);
// 3 the pressure blend weighted by the line strengths (co_ddb.py:92-93).
// pco = P q_CO (bar).
__device__ inline double co_ddb(const DevBlock &B, double P, double T, double ph2, double phe, double pco, double f,
                                int shape) {
    const double gamma = pow(kCoT0 / T, 0.7) * (1.960 * ph2 + 1.200 * phe + 6.000 * pco);   // co_ddb.py:45-52
    const double g2 = gamma * gamma;
    const double w = fmin(fmax((P - 0.001) / (0.1 - 0.001), 0.0), 1.0);                     // co_ddb.py:58-62
    const bool voigt = P <= 0.1 || shape == 0 || shape == 2;                                  // co_ddb.py:68
    const bool vvw = P >= 0.001 || shape == 1 || shape == 2;                                  // co_ddb.py:80
    const double betaD = 4.3e-7 * sqrt(T / 28.0) * f;                                         // co_ddb.py:70
    const double vpre = kCoGHz * (1.0 / (sqrt(kPi) * betaD));
    const double f2 = f * f;
    const double eT = -((1.0 / T) - (1.0 / kCoT0)) * kCoHck;
    double sum = 0.0;
    for (int i = 0; i < B.n; ++i) {
        const double f0 = B.col[0][i];
        double sv = 0.0, sw = 0.0;
        if (voigt) {
            const Cplx xi = {gamma / betaD, (f - f0) / betaD};
            Cplx num = {kVoigtA[7], 0.0}, den = {kVoigtB[7], 0.0};
#pragma unroll
            for (int j = 6; j >= 0; --j) {
                num = cmul(num, xi);
                num.re += kVoigtA[j];
                den = cmul(den, xi);
                den.re += kVoigtB[j];
            }
            sv = vpre * cdiv(num, den).re;
        }
        if (vvw) {
            const double num = gamma * f2 + gamma * (f0 * f0 + g2);
            const double d = f2 - f0 * f0 - g2;
            const double r = f / f0;
            sw = kCoGHz * 2.0 * (r * r) * num / (kPi * (d * d + 4.0 * f2 * g2));
        }
        if (shape == 0)
            sum += sv;
        else if (shape == 1)
            sum += sw;
        else if (shape == 2)
            sum += sv - sw;
        else
            sum += (w * sw + (1.0 - w) * sv) * (B.col[1][i] * exp(eT * B.col[2][i]));
    }
    return kCoCoef * (pco / kCoT0) * pow(kCoT0 / T, 3.5) * sum;                             // co_ddb.py:95
}

// ---- clouds_idp (constituents/clouds/clouds_idp.py:6-101) ----
__constant__ double kCloudFr[9] = {1.0E8, 3.0E8, 1.0E9, 2.0E9, 3.0E9, 5.0E9, 1.0E10, 3.0E10, 1.0E11};
__constant__ double kCloudEimag[9] = {8.0E-3, 1.5E-3, 8.0E-4, 1.0E-3, 1.2E-3, 1.5E-3, 3.0E-3, 8.0E-3, 2.0E-2};
// bulk densities and refractive indices n = re - i im of the RBX_CL_* rows (clouds_idp.py:17-52);
// the first two are water, whose permittivity depends on f and T
__constant__ double kCloudBulk[RBX_NUM_CLOUDS] = {0.9, 1.0, 1.2, 1.6, 1.5, 1.0};
__constant__ double kCloudN[RBX_NUM_CLOUDS][2] = {{0, 0}, {0, 0}, {1.7, 0.005}, {1.3, 0.0001},
                                                  {1.15, 0.0001}, {1.3, 0.00001}};

// permittivity of water (clouds_idp.py:68-101): Debye above 0 C, the tabulated loss below
__device__ inline Cplx water_permittivity(double f, double T) {
    const double Tc = T - 273.0;
    const double fHz = f * 1.0E9;
    double E1, E2;
    if (Tc >= 0.0) {
        const double RelT = 1.1109E-10 - Tc * 3.824E-12 + (Tc * Tc) * 6.938E-14 - (Tc * Tc * Tc) * 5.096E-16;
        double E0 = 88.045 - 0.4147 * Tc + (Tc * Tc) * 6.295E-4 + (Tc * Tc * Tc) * 1.075E-5;
        if (E0 < 0.0) E0 = 0.0;
        const double EINF = 4.9;
        const double fr = fHz * RelT;
        E1 = EINF + (E0 - EINF) / (1.0 + fr * fr);
        E2 = fr * (E0 - EINF) / (1.0 + fr * fr);
        if (E2 < 0.0) E2 = 0.0;
    } else {
        E1 = 3.17;
        const double LF = log10(fHz);
        int j = 0;
        while (j < 7 && kCloudFr[j + 1] < fHz) ++j;   // first j with FR[j+1] >= f, else the last interval
        const double LF0 = log10(kCloudFr[j]), LF1 = log10(kCloudFr[j + 1]);
        const double DLF = (LF - LF0) / (LF1 - LF0);
        const double X0 = log10(kCloudEimag[j]), X1 = log10(kCloudEimag[j + 1]);
        E2 = pow(10.0, X0 + DLF * (X1 - X0));
    }
    return {E1, -E2};
}

// sum over the selected cloud types (bit RBX_CL_x of `types`) of 3 k (rho / rho_bulk) (-Im K),
// K = (e - 1)/(e + 2), k = 2 pi f / c in cm^-1 (acloud, clouds_idp.py:58-62)
__device__ inline double clouds_idp(const double *cloud, int nl, int layer, double T, double f, int types) {
    const double k = 2.0 * kPi * f / kCoGHz;
    double a = 0.0;
    for (int t = 0; t < RBX_NUM_CLOUDS; ++t) {
        if (!((types >> t) & 1)) continue;
        const double fraction = cloud[static_cast<size_t>(t) * nl + layer] / kCloudBulk[t];
        Cplx e;
        if (t <= RBX_CL_SOLN) {
            e = water_permittivity(f, T);
        } else {
            const Cplx n = {kCloudN[t][0], -kCloudN[t][1]};
            e = cmul(n, n);
        }
        const Cplx K = cdiv({e.re - 1.0, e.im}, {e.re + 2.0, e.im});
        a += 3.0 * k * fraction * (-K.im);
    }
    return a < 0.0 ? 0.0 : a;   // clouds_idp.py:47-49
}

// ---- h2_orton (constituents/h2/h2_orton.py:17-222) ----
constexpr int kOrtonN = RBX_ORTON_NTEMP;

// h2_jj (constituents/h2/h2_jj.py:7-22) in dB/km, the units h2_orton calls it in
__device__ inline double h2_jj_db(double f, double T, double ph2, double phe, double pch4) {
    const double th = 273.0 / T;
    const double cf = 3.9522E-14 * (f * f) * ph2;
    return cf * (ph2 * pow(th, 3.12) + 1.382 * phe * pow(th, 2.24) + 9.322 * pch4 * pow(th, 3.34)) * kDbPerInvCm;
}

// value at t of the not-a-knot cubic spline through (x[i], y[i]), the interpolant of
// scipy's interp1d(kind='cubic') that h2_orton.py:191-207 uses: the second derivatives
// M solve the tridiagonal system of the interior rows with the two not-a-knot rows
// (equal third derivatives across x[1] and x[N-2]) folded into the first and last.
__device__ inline double spline_not_a_knot(const double *x, const double *y, double t) {
    constexpr int N = kOrtonN;
    double h[N - 1];
#pragma unroll
    for (int i = 0; i < N - 1; ++i) h[i] = x[i + 1] - x[i];
    // rows 1..N-2: lo M[i-1] + di M[i] + up M[i+1] = r
    double lo[N], di[N], up[N], r[N];
#pragma unroll
    for (int i = 1; i < N - 1; ++i) {
        lo[i] = h[i - 1];
        di[i] = 2.0 * (h[i - 1] + h[i]);
        up[i] = h[i];
        r[i] = 6.0 * ((y[i + 1] - y[i]) / h[i] - (y[i] - y[i - 1]) / h[i - 1]);
    }
    // M[0] = ((h0 + h1) M[1] - h0 M[2]) / h1; M[N-1] likewise from M[N-2], M[N-3]
    di[1] += h[0] * (h[0] + h[1]) / h[1];
    up[1] -= h[0] * h[0] / h[1];
    di[N - 2] += h[N - 2] * (h[N - 3] + h[N - 2]) / h[N - 3];
    lo[N - 2] -= h[N - 2] * h[N - 2] / h[N - 3];
    // Thomas elimination on rows 1..N-2
#pragma unroll
    for (int i = 2; i < N - 1; ++i) {
        const double m = lo[i] / di[i - 1];
        di[i] -= m * up[i - 1];
        r[i] -= m * r[i - 1];
    }
    double M[N];
    M[N - 2] = r[N - 2] / di[N - 2];
#pragma unroll
    for (int i = N - 3; i >= 1; --i) M[i] = (r[i] - up[i] * M[i + 1]) / di[i];
    M[0] = ((h[0] + h[1]) * M[1] - h[0] * M[2]) / h[1];
    M[N - 1] = ((h[N - 3] + h[N - 2]) * M[N - 2] - h[N - 2] * M[N - 3]) / h[N - 3];
    // the interval [x[k], x[k+1]] holding t (t == x[N-1] uses the last one)
    double v = 0.0;
#pragma unroll
    for (int k = 0; k < N - 1; ++k) {
        const bool here = (t >= x[k] && t <= x[k + 1]) && (k == 0 || t > x[k]);
        if (here) {
            const double a = x[k + 1] - t, b = t - x[k], hk = h[k];
            v = M[k] * a * a * a / (6.0 * hk) + M[k + 1] * b * b * b / (6.0 * hk) +
                (y[k] - M[k] * hk * hk / 6.0) * (a / hk) + (y[k + 1] - M[k + 1] * hk * hk / 6.0) * (b / hk);
        }
    }
    return v;
}

// the table values of one table at frequency f for every tabulated temperature: the
// quadratic in f through the rows of the first tabulated frequency above f (at least
// the second) and its two neighbours, in linear units (h2_orton.py:74-97)
__device__ inline void orton_column(const DevBlock &B, int table, int ifreq, double f, double *y) {
    const double *ft = B.col[0];
    const double *lv = B.col[1] + (static_cast<size_t>(table) * B.n + ifreq - 1) * kOrtonN;
    const double X1 = ft[ifreq - 1];
    const double X21 = ft[ifreq] - ft[ifreq - 1];
    const double X32 = ft[ifreq + 1] - ft[ifreq];
    const double X212 = ft[ifreq] * ft[ifreq] - ft[ifreq - 1] * ft[ifreq - 1];
    const double X322 = ft[ifreq + 1] * ft[ifreq + 1] - ft[ifreq] * ft[ifreq];
    const double DQ = X212 * X32 - X322 * X21;
#pragma unroll
    for (int l = 0; l < kOrtonN; ++l) {
        const double Y1 = exp(lv[l]), Y2 = exp(lv[kOrtonN + l]), Y3 = exp(lv[2 * kOrtonN + l]);
        const double Y21 = Y2 - Y1, Y32 = Y3 - Y2;
        const double AQ = (X32 * Y21 - X21 * Y32) / DQ;
        const double BQ = (X212 * Y32 - X322 * Y21) / DQ;
        const double CQ = Y1 - AQ * (X1 * X1) - BQ * X1;
        y[l] = AQ * (f * f) + BQ * f + CQ;
    }
}

// state 0 'e', 1 'n'; ph2, phe, pch4 partial pressures in bar
__device__ inline double h2_orton(const DevBlock &B, double T, double ph2, double phe, double pch4, double f,
                                  int state) {
    constexpr double atm = 1.01325, T0 = 273.0;   // h2_orton.py:9-10
    const double *ft = B.col[0];
    const double *Tt = B.col[2];
    // first tabulated frequency above f (h2_orton.py:77-81), kept where its neighbours exist
    int lo = 0, hi = B.n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (ft[mid] > f)
            hi = mid;
        else
            lo = mid + 1;
    }
    const int ifreq = min(max(lo, 1), B.n - 2);
    double yh2[kOrtonN], yhe[kOrtonN], ych4[kOrtonN];
    orton_column(B, 0 + state, ifreq, f, yh2);    // eh2h2 / nh2h2
    orton_column(B, 2 + state, ifreq, f, yhe);    // eh2he / nh2he
    orton_column(B, 4 + state, ifreq, f, ych4);   // eh2ch4 / nh2ch4
    const double a2 = ph2 / atm, ae = phe / atm, am = pch4 / atm;
    auto total = [&](double ah2, double ahe, double ach4, double Tj) {
        const double r = T0 / Tj;
        return a2 * (ah2 * a2 + ahe * ae + ach4 * am) * (r * r);
    };
    if (T < Tt[0]) {   // quartic extrapolation down (h2_orton.py:145-170)
        const double v0 = total(yh2[0], yhe[0], ych4[0], Tt[0]);
        const double v1 = total(yh2[1], yhe[1], ych4[1], Tt[1]);
        const double X21n = pow(Tt[1], 4.0) - pow(Tt[0], 4.0);
        const double AQ = -1.0 * (v1 - v0) / X21n;
        const double CQ = v0 + AQ * pow(Tt[0], 4.0);
        return CQ - AQ * pow(T, 4.0);
    }
    if (T > Tt[kOrtonN - 1]) {   // h2_jj matched at the hottest table temperature (h2_orton.py:171-190)
        const double Tn = Tt[kOrtonN - 1];
        const double anear = total(yh2[kOrtonN - 1], yhe[kOrtonN - 1], ych4[kOrtonN - 1], Tn);
        return h2_jj_db(f, T, ph2, phe, pch4) * (anear / h2_jj_db(f, Tn, ph2, phe, pch4));
    }
    double x[kOrtonN];
#pragma unroll
    for (int l = 0; l < kOrtonN; ++l) x[l] = Tt[l];
    return total(spline_not_a_knot(x, yh2, T), spline_not_a_knot(x, yhe, T), spline_not_a_knot(x, ych4, T), T);
}

}  // namespace rbx
