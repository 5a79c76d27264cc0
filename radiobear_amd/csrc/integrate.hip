// Brightness integration along ray paths, and its Jacobian.
//
// Lane = frequency, wavefront = one ray x 64 consecutive frequencies.  The
// step loop follows Brightness.single (brightness.py:53-117):
//   dtau = (a0 + a1) ds/2, tau_{i+1} = tau_i + dtau          (FP64 prefix sum; in the
//   FP32 mode the increments, the weights and the trapezoid sums are FP32)
//   W_{i+1} = a1 e^{-tau} (point) or 2 a1 E2(tau) (disc), W_0 = 0
//   intW += (W_{i+1} + W_i) ds/2, Tb_lyr += (T1 W_{i+1} + T0 W_i) ds/2
//   Tb = T_cmb if Tb_lyr < T_cmb else Tb_lyr / intW
// Path point k is layer k (the reference's paths are ingress-only), so alpha
// for a step is one coalesced row read alpha[k][f0..f0+63]; the ray's ds and
// T[k] are wave-uniform scalar loads.  Blocks map frequency groups to XCDs
// (block b -> group b % ngroups) so each XCD's L2 holds a slice of the table, or,
// ray-major (launch.int_map 2), every group of a ray block to one XCD so the
// ray's ds row is fetched once and the whole table stays in that XCD's L2.
#include <algorithm>

#include "common.hpp"

namespace rbx {

constexpr int kWaves = 4;
constexpr int kLanes = 64;

// One path point's entry in the per-ray factor table: the half step h_k times log2(e) (the FP32
// integration's tau scan carries tau log2(e) in FP64) and point k + 1's trapezoid factors
// (h_k + h_{k+1}, T_{k+1} (h_k + h_{k+1})) in FP32.
struct alignas(16) PathFactor {
    double hl;
    float hh, thh;
};

struct IntArgs {
    const void *alpha;
    int f32;
    const double *T;
    int nl, nf;
    const double *hds;
    int ld;
    const int *nsteps;
    int nray;
    int mode, fast_exp;
    double tau_cut;
    double *Tb, *intW;
    int ngroups;
    int ray_major;   // block order: 0 frequency group major, 1 ray major (XCD-aware)
    // several atmospheres laid end to end (null: one atmosphere of nl layers):
    // ray r uses rows [layer_off[a], layer_off[a + 1]) of alpha and T, a = ray_atm[r]
    const int *ray_atm, *layer_off;
    int natm;
    const int *order;   // wavefront w integrates ray order[w] (null: ray w)
    // column j of the table is request frequency cols[j]: Tb and intW go to column cols[j] of
    // their rows (null: column j; rbx_integrate_cols)
    const int *cols;
    // FP32 point rays: each ray's path factors [nray][ld], written by integrate_factors_kernel
    // (null: each wavefront stages the trapezoid factors in LDS)
    const PathFactor *factors;
};


constexpr int kXcds = 8;   // MI355X: blocks are dealt round-robin to 8 XCDs (b % 8)

// The frequency group and ray block (kWaves rays) of this block.  Ray major: the
// blocks of XCD x (b = x, x + 8, ...) run ngroups groups of ray block x, then of x + 8, ...
__device__ inline void block_tile(const struct IntArgs &A, int &group, int &rblock) {
    const int b = blockIdx.x;
    if (A.ray_major) {
        const int q = b / kXcds;
        group = q % A.ngroups;
        rblock = (q / A.ngroups) * kXcds + b % kXcds;
    } else {
        group = b % A.ngroups;
        rblock = b / A.ngroups;
    }
}

// The ray a wavefront integrates (-1: none) and its first layer row and layer
// count; false when its atmosphere index or offsets are out of range.
__device__ inline int ray_of(const IntArgs &A, int w) {
    w = __builtin_amdgcn_readfirstlane(w);   // wavefront-uniform (threadIdx.x / 64 is)
    if (w >= A.nray) return -1;
    const int r = __builtin_amdgcn_readfirstlane(A.order ? A.order[w] : w);
    return (r >= 0 && r < A.nray) ? r : -1;
}

__device__ inline bool ray_span(const IntArgs &A, int ray, int &base, int &nl) {
    base = 0;
    nl = A.nl;
    if (!A.ray_atm) return true;
    // the ray is the wavefront's: keep its atmosphere's offsets in scalar registers
    const int a = __builtin_amdgcn_readfirstlane(A.ray_atm[ray]);
    if (a < 0 || a >= A.natm) return false;
    base = __builtin_amdgcn_readfirstlane(A.layer_off[a]);
    const int end = __builtin_amdgcn_readfirstlane(A.layer_off[a + 1]);
    nl = end - base;
    return base >= 0 && end <= A.nl && nl >= 2 && nl - 1 <= A.ld;
}

// The reference carries a NaN half step or temperature through tau and the sums to a NaN Tb
// (a ray that misses a shell gets a NaN ds, raypath.py:190-209): an optical-depth cut must not
// skip one, so the cut is off for a ray whose path (steps [0, n)) holds one.  Frequency
// independent; one pass of the wavefront over the ray's rows.
__device__ inline bool path_has_nan(const double *h, const double *Tl, int n) {
    bool bad = false;
    for (int k = static_cast<int>(threadIdx.x % kLanes); k < n; k += kLanes)
        bad = bad || __builtin_isnan(h[k]) || __builtin_isnan(Tl[k + 1]);
    return __any(bad);
}

__device__ inline double load_alpha(const void *p, int f32, size_t i) {
    return f32 ? static_cast<double>(static_cast<const float *>(p)[i]) : static_cast<const double *>(p)[i];
}

// e^-tau in FP64 for tau >= 0 (point rays of the FP64 integration): with n = floor(64 tau
// log2(e)), e^-tau = 2^-(n >> 6) 2^-((n & 63) / 64) e^-w, w = tau - n ln2/64 in about
// [0, 0.0109) (ln2/64 split in two parts, Cody and Waite, so w carries an ulp); the middle
// factor from a 64-entry table, e^-w as its Taylor polynomial to w^6 (remainder < 4e-18).
// Relative error about an ulp (2.2e-16 against a 60-digit reference over tau in [0, 700]) in 15
// instructions where the library exp takes about 22.  Past tau = 1100 the result is 0 (as
// exp's underflow); a NaN tau gives NaN.
constexpr int kExpTab = 64;
__constant__ double kExp2Tab[kExpTab] = {   // 2^(-j/64), correctly rounded
    1.0, 0.9892280131939755, 0.9785720620877001, 0.9680308967461472,
    0.9576032806985737, 0.9472879907934828, 0.93708381705515, 0.9269895625416927,
    0.9170040432046712, 0.9071260877501994, 0.8973545375015536, 0.8876882462632606,
    0.8781260801866497, 0.8686669176368531, 0.859309649061239, 0.8500531768592617,
    0.8408964152537145, 0.8318382901633682, 0.8228777390769825, 0.8140137109286739,
    0.8052451659746271, 0.7965710756711335, 0.7879904225539432, 0.7795022001189185,
    0.7711054127039704, 0.7627990753722692, 0.7545822137967114, 0.7464538641456324,
    0.7384130729697497, 0.7304588970903235, 0.7225904034885233, 0.714806669195985,
    0.7071067811865476, 0.6994898362691556, 0.691954940981916, 0.6845012114872953,
    0.6771277734684463, 0.6698337620266515, 0.6626183215798707, 0.6554806057623822,
    0.6484197773255048, 0.6414350080393891, 0.6345254785958666, 0.6276903785123455,
    0.620928906036742, 0.614240268053435, 0.6076236799902345, 0.6010783657263515,
    0.5946035575013605, 0.5881984958251406, 0.5818624293887887, 0.5755946149764913,
    0.5693943173783458, 0.5632608093041209, 0.5571933712979462, 0.5511912916539204,
    0.5452538663326288, 0.5393803988785599, 0.5335702003384118, 0.5278225891802786,
    0.5221368912137069, 0.5165124395106142, 0.5109485743270583, 0.5054446430258502,
};
constexpr double kLn2By64Hi = 0.010830424696223417;   // 36 significant bits: n kLn2By64Hi is exact
constexpr double kLn2By64Lo = 2.572804622327669e-14;
__device__ __forceinline__ double exp_neg_tab(double tau, const double *tab) {
    const double fl = floor((tau > 1100.0 ? 1100.0 : tau) * (64.0 * 1.4426950408889634));
    const int n = static_cast<int>(fl);
    const double v = fma(fl, kLn2By64Lo, fma(fl, kLn2By64Hi, -tau));   // -w
    double p = fma(1.0 / 720.0, v, 1.0 / 120.0);
    p = fma(p, v, 1.0 / 24.0);
    p = fma(p, v, 1.0 / 6.0);
    p = fma(p, v, 0.5);
    p = fma(p, v, 1.0);
    p = fma(p, v, 1.0);
    return ldexp(tab[n & (kExpTab - 1)] * p, -(n >> 6));
}

// One trapezoid step of brightness.py:93-101 (DISC: W = 2 a1 E2(tau)).  `tab`: the 64-entry
// table of exp_neg_tab (FP64 point rays).
template <int MODE, bool FAST>
struct Step {
    double tau = 0.0, W0 = 0.0, TW0 = 0.0, intW = 0.0, num = 0.0, a0 = 0.0;
    const double *tab = nullptr;
    __device__ __forceinline__ void operator()(double a1, double hh, double T1) {
        tau = tau + (a0 + a1) * hh;
        double W1;
        if constexpr (MODE == 1)
            W1 = 2.0 * a1 * expint_e2(tau);
        else
            W1 = a1 * (FAST ? exp_neg_mixed(tau) : exp_neg_tab(tau, tab));
        intW = intW + (W1 + W0) * hh;
        const double TW1 = T1 * W1;
        num = num + (TW1 + TW0) * hh;
        a0 = a1;
        W0 = W1;
        TW0 = TW1;
    }
};

// Path steps are taken U at a time: the U rows of alpha, ds and T of a batch are
// all requested before the first is used, so one memory latency covers U steps.
template <class AT, int MODE, bool FAST, int U>
__global__ __launch_bounds__(kWaves * kLanes) void integrate_kernel(IntArgs A) {
    __shared__ double etab[kExpTab];
    if constexpr (MODE == 0 && !FAST) {   // (before any thread returns)
        for (int k = threadIdx.x; k < kExpTab; k += blockDim.x) etab[k] = kExp2Tab[k];
        __syncthreads();
    }
    const int wave = threadIdx.x / kLanes;
    const int lane = threadIdx.x % kLanes;
    int group, rblock;
    block_tile(A, group, rblock);
    const int ray = ray_of(A, rblock * kWaves + wave);
    if (ray < 0) return;
    const int j = group * kLanes + lane;
    const bool ok = j < A.nf;
    const int jj = ok ? j : A.nf - 1;
    int base, nlr;
    const bool valid = ray_span(A, ray, base, nlr);
    const int S = valid ? min(A.nsteps[ray], nlr - 1) : -1;   // never read past the ray's ds row or table
    const double *h = A.hds + static_cast<size_t>(ray) * A.ld;
    const size_t nf = A.nf;
    const AT *al = static_cast<const AT *>(A.alpha) + static_cast<size_t>(base) * nf + jj;
    const double *Tl = A.T + base;
    Step<MODE, FAST> st;
    st.tab = etab;
    if (S > 1) {
        const int n = S - 1;
        const double cut = (A.tau_cut > 0.0 && !path_has_nan(h, Tl, n)) ? A.tau_cut : 0.0;
        st.a0 = static_cast<double>(al[0]);
        int i = 0;
        for (; i + U <= n; i += U) {
            AT ab[U];
            double hb[U], Tb[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                ab[u] = al[(i + u + 1) * nf];
                hb[u] = h[i + u];
                Tb[u] = Tl[i + u + 1];
            }
#pragma unroll
            for (int u = 0; u < U; ++u) st(static_cast<double>(ab[u]), hb[u], Tb[u]);
            if (cut > 0.0 && __all(!ok || st.tau > cut)) {
                i = n;
                break;
            }
        }
        for (; i < n; ++i) st(static_cast<double>(al[(i + 1) * nf]), h[i], Tl[i + 1]);
    }
    const double intW = st.intW, num = st.num;
    if (!ok) return;
    const size_t o = static_cast<size_t>(ray) * A.nf + (A.cols ? A.cols[j] : j);
    double Tb;
    if (!valid)
        Tb = __builtin_nan("");
    else if (S < 0 || num < kTcmb)
        Tb = kTcmb;
    else
        Tb = num / intW;
    A.Tb[o] = Tb;
    if (A.intW) A.intW[o] = (S < 0) ? 0.0 : intW;
}

// point k + 1's trapezoid factors of a path of n steps: (h_k + h_{k+1}, T_{k+1} (h_k + h_{k+1}))
// rounded to FP32, with h_n = 0 (brightness.py:99-101 regrouped by path point)
__device__ __forceinline__ float2 trapezoid_factor(const double *h, const double *Tl, int k, int n) {
    const double hh = h[k] + (k + 1 < n ? h[k + 1] : 0.0);
    return make_float2(static_cast<float>(hh), static_cast<float>(Tl[k + 1] * hh));
}

// Every frequency group of a ray stages the same factors; this writes them once per ray,
// fac[ray][k] for k < n, and at k = n (a slot every path leaves free) hh = NaN when the path
// holds a NaN half step or temperature (the integration then takes no cut) -- one wavefront
// per ray.  They depend on the paths and T only, not on alpha, so a caller can write them
// beside the absorption.
__global__ __launch_bounds__(kWaves * kLanes) void integrate_factors_kernel(IntArgs A, PathFactor *fac) {
    const int wave = threadIdx.x / kLanes, lane = threadIdx.x % kLanes;
    const int ray = ray_of(A, static_cast<int>(blockIdx.x) * kWaves + wave);
    if (ray < 0) return;
    int base, nlr;
    const bool valid = ray_span(A, ray, base, nlr);
    const int S = valid ? min(A.nsteps[ray], nlr - 1) : -1;
    if (S <= 1) return;
    const int n = S - 1;
    const double *h = A.hds + static_cast<size_t>(ray) * A.ld, *Tl = A.T + base;
    PathFactor *out = fac + static_cast<size_t>(ray) * A.ld;
    bool bad = false;
    for (int k = lane; k < n; k += kLanes) {
        const float2 f = trapezoid_factor(h, Tl, k, n);
        out[k] = PathFactor{h[k] * kLog2e, f.x, f.y};
        bad = bad || __builtin_isnan(h[k]) || __builtin_isnan(Tl[k + 1]);
    }
    bad = __any(bad);
    if (lane == 0) out[n] = PathFactor{0.0, bad ? __builtin_nanf("") : 0.0f, 0.0f};
}

// ---- FP32 mode, point rays: FP32 weights and trapezoid sums, FP64 path and tau ----
// The half steps and the tau prefix sum stay FP64.  The scan carries t = tau log2(e):
// t_{i+1} = t_i + (a0 + a1) (h_i log2 e) in an FP64 FMA at every step, with a0 + a1 (two FP32
// table entries) added in FP32 and h_i log2 e the FP64 half step times log2(e) (a wave-uniform
// scalar load, U steps per s_load batch, from the factor table or formed from the ray's half
// step; the same FP64 product either way), and the cut is taken at t > tau_cut log2(e).  W, T W
// and the two trapezoid sums are FP32, added into FP64 totals after every batch of U steps
// (blocked summation).  e^-tau is the hardware exp2 of -t with t rounded to FP32: the rounding
// moves the exponent by at most t 2^-24, so a weight moves by at most tau e^-tau 2^-24 ~ 2e-8
// of its alpha (largest at tau = 1), below the FP32 sums' own error; past t = 150 the result
// is 0.  The trapezoid sums are taken by path point:
// sum_i (W_i + W_{i+1}) h_i = sum_p W_p (h_{p-1} + h_p) (W_0 = 0, h_n = 0), and the same with
// T_p W_p, so a step adds W_p times two staged factors (two FMAs, where the pairwise form took
// two adds, two FMAs and a multiply).  Each wavefront first copies its ray's factors to LDS as
// FP32 pairs (h_{p-1} + h_p, T_p (h_{p-1} + h_p)), so a step reads them with one broadcast LDS
// load; with FAC they come instead from the per-ray table integrate_factors_kernel wrote, read
// with scalar loads (no LDS).  A wavefront stopped by the cut at point p takes W_p h_p and
// T_p W_p h_p back out (its sums end at p, as the pairwise form's do).
// BUF: the table rows are read as raw buffer loads whose row offset is a scalar register
// (the row's byte offset computed on the scalar unit), so a step's load costs no vector
// address arithmetic; for tables under 2 GiB.
constexpr int kBufWord3 = 0x00020000;   // buffer descriptor word 3: 32-bit data format, no swizzle

template <int U, bool BUF, bool FAC>
__global__ __launch_bounds__(kWaves * kLanes) void integrate_f32_kernel(IntArgs A) {
    extern __shared__ float2 hT_all[];   // [kWaves][ld] (none with FAC)
    const int wave = threadIdx.x / kLanes;
    const int lane = threadIdx.x % kLanes;
    int group, rblock;
    block_tile(A, group, rblock);
    const int ray = ray_of(A, rblock * kWaves + wave);
    if (ray < 0) return;
    const int j = group * kLanes + lane;
    const bool ok = j < A.nf;
    const int jj = ok ? j : A.nf - 1;
    int base, nlr;
    const bool valid = ray_span(A, ray, base, nlr);
    const int S = valid ? min(A.nsteps[ray], nlr - 1) : -1;   // never read past the ray's ds row or table
    const double *h = A.hds + static_cast<size_t>(ray) * A.ld;
    const size_t nf = A.nf;
    const float *al = static_cast<const float *>(A.alpha) + static_cast<size_t>(base) * nf + jj;
    const double *Tl = A.T + base;
    // the ray's factors: staged in this wavefront's LDS, or (FAC) read from the factor table
    float2 *stage = hT_all + static_cast<size_t>(wave) * A.ld;
    [[maybe_unused]] const PathFactor *fac = FAC ? A.factors + static_cast<size_t>(ray) * A.ld : nullptr;
    auto factor = [&](int k) -> float2 {   // point k + 1's trapezoid factors
        if constexpr (FAC)
            return make_float2(fac[k].hh, fac[k].thh);
        else
            return stage[k];
    };
    auto hlog = [&](int k) -> double {   // h_k log2(e)
        if constexpr (FAC)
            return fac[k].hl;
        else
            return h[k] * kLog2e;
    };
    [[maybe_unused]] __amdgpu_buffer_rsrc_t rows;
    [[maybe_unused]] const int nfb = static_cast<int>(nf) * 4, lane_b = jj * 4;
    if constexpr (BUF)
        rows = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<float *>(static_cast<const float *>(A.alpha) + static_cast<size_t>(base) * nf), 0,
            max(nlr, 1) * nfb, kBufWord3);
    auto row = [&](int r) -> float {   // table row r of the ray's atmosphere, this lane's column
        if constexpr (BUF)
            return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rows, lane_b, r * nfb, 0));
        else
            return al[static_cast<size_t>(r) * nf];
    };
    double t = 0.0, intW = 0.0, num = 0.0;   // t = tau log2(e)
    if (S > 1) {
        const int n = S - 1;
        bool bad = false;   // a NaN half step or temperature: no cut (path_has_nan)
        if constexpr (FAC) {
            bad = __builtin_isnan(fac[n].hh);   // the table marks such a path past its last factor
        } else {
            for (int k = lane; k < n; k += kLanes) {
                stage[k] = trapezoid_factor(h, Tl, k, n);
                bad = bad || __builtin_isnan(h[k]) || __builtin_isnan(Tl[k + 1]);
            }
            bad = __any(bad);
            wave_sync();
        }
        const double cut = (A.tau_cut > 0.0 && !bad) ? A.tau_cut * kLog2e : 0.0;   // on t
        const unsigned long long idle = __builtin_amdgcn_ballot_w64(!ok);   // lanes past the last column
        float a0 = row(0), W1 = 0.0f;
        int i = 0;
        for (; i + U <= n; i += U) {
            float ab[U];
            double hb[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                ab[u] = row(i + u + 1);
                hb[u] = hlog(i + u);
            }
            float iW = 0.0f, nm = 0.0f;
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const float2 ht = factor(i + u);
                t = fma(static_cast<double>(a0 + ab[u]), hb[u], t);
                W1 = ab[u] * __builtin_amdgcn_exp2f(-static_cast<float>(t));
                iW = fmaf(W1, ht.x, iW);
                nm = fmaf(W1, ht.y, nm);
                a0 = ab[u];
            }
            intW += static_cast<double>(iW);
            num += static_cast<double>(nm);
            if (cut > 0.0 && (__builtin_amdgcn_ballot_w64(t > cut) | idle) == __builtin_amdgcn_read_exec()) {
                const int p = i + U;   // the sums end at point p
                if (p < n) {
                    const double hp = h[p] * static_cast<double>(W1);
                    intW -= hp;
                    num -= hp * Tl[p];
                }
                i = n;
                break;
            }
        }
        for (; i < n; ++i) {
            const float a1 = row(i + 1);
            const float2 ht = factor(i);
            t = fma(static_cast<double>(a0 + a1), hlog(i), t);
            W1 = a1 * __builtin_amdgcn_exp2f(-static_cast<float>(t));
            intW += static_cast<double>(W1 * ht.x);
            num += static_cast<double>(W1 * ht.y);
            a0 = a1;
        }
    }
    if (!ok) return;
    const size_t o = static_cast<size_t>(ray) * A.nf + (A.cols ? A.cols[j] : j);
    double Tb;
    if (!valid)
        Tb = __builtin_nan("");
    else if (S < 0 || num < kTcmb)
        Tb = kTcmb;
    else
        Tb = num / intW;
    A.Tb[o] = Tb;
    if (A.intW) A.intW[o] = (S < 0) ? 0.0 : intW;
}

// The forward sums of one ray (brightness.py:93-101) for the Jacobian and the adjoint:
// the integrated weight D = intW, the T-weighted sum num and the last path point reached.
// The optical-depth cut is checked after every 8 steps; the steps are taken 8 at a time,
// their loads issued together, so one memory latency covers 8 steps.
template <bool FAST>
__device__ __forceinline__ double exp_neg(double tau) {
    if constexpr (FAST)
        return exp_neg_mixed(tau);
    else
        return exp(-tau);
}

template <bool FAST>
__device__ inline void forward_sums(const IntArgs &A, const double *h, const double *Tl, size_t row0, int S, int jj,
                                    bool ok, double &intW, double &num, int &last) {
    constexpr int U = 8;
    double tau = 0.0, W0 = 0.0, TW0 = 0.0;
    intW = 0.0;
    num = 0.0;
    last = 0;
    auto step = [&](double a0, double a1, double hh, double T1) {
        tau = tau + (a0 + a1) * hh;
        const double W1 = a1 * exp_neg<FAST>(tau);
        intW = intW + (W1 + W0) * hh;
        const double TW1 = T1 * W1;
        num = num + (TW1 + TW0) * hh;
        W0 = W1;
        TW0 = TW1;
    };
    if (S <= 1) return;
    const int n = S - 1;
    const double cut = (A.tau_cut > 0.0 && !path_has_nan(h, Tl, n)) ? A.tau_cut : 0.0;
    const size_t nf = A.nf;
    double a0 = load_alpha(A.alpha, A.f32, row0 + jj);
    int i = 0;
    for (; i + U <= n; i += U) {
        double ab[U], hb[U], tb[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            ab[u] = load_alpha(A.alpha, A.f32, row0 + static_cast<size_t>(i + u + 1) * nf + jj);
            hb[u] = h[i + u];
            tb[u] = Tl[i + u + 1];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            step(a0, ab[u], hb[u], tb[u]);
            a0 = ab[u];
        }
        last = i + U;
        if (cut > 0.0 && __all(!ok || tau > cut)) return;
    }
    for (; i < n; ++i) {
        const double a1 = load_alpha(A.alpha, A.f32, row0 + static_cast<size_t>(i + 1) * nf + jj);
        step(a0, a1, h[i], Tl[i + 1]);
        a0 = a1;
        last = i + 1;
    }
}

// ---- Jacobian: two forward passes per (ray, freq) ----
// With c_k = h_{k-1} + h_k (h_{-1} = h_{S-1} = 0), D = intW, N = Tb_lyr:
//   lambda_k = dTb/dW_k = c_k (T_k - Tb) / D,   sum_k lambda_k W_k = (N - Tb D)/D = 0,
//   dTb/dT_k|direct = c_k W_k / D,
//   dTb/dalpha_k = lambda_k e^{-tau_k} + h_{k-1} P_{k-1} + h_k P_k,  P_k = sum_{j<=k} lambda_j W_j,
// the last form using Lambda_k = sum_{j>k} lambda_j W_j = -P_k.  e^{-tau} is
// recomputed in the second pass, never stored.
struct JacArgs {
    IntArgs I;
    const double *dalpha;  // [3][nl][nf]
    float *J;              // [nray][3][ldj][nf]
    int ldj;               // layer rows per ray in J (>= the layers of any ray's atmosphere)
};

__global__ __launch_bounds__(kWaves * kLanes) void jacobian_kernel(JacArgs Aj) {
    const IntArgs &A = Aj.I;
    const int wave = threadIdx.x / kLanes;
    const int lane = threadIdx.x % kLanes;
    const int group = blockIdx.x % A.ngroups;
    const int ray = ray_of(A, (blockIdx.x / A.ngroups) * kWaves + wave);
    if (ray < 0) return;
    const int j = group * kLanes + lane;
    const bool ok = j < A.nf;
    const int jj = ok ? j : A.nf - 1;
    int base, nlr;
    const bool valid = ray_span(A, ray, base, nlr) && nlr <= Aj.ldj;
    const int S = valid ? min(A.nsteps[ray], nlr - 1) : -1;   // never read past the ray's ds row or table
    const double *h = A.hds + static_cast<size_t>(ray) * A.ld;
    const size_t nlnf = static_cast<size_t>(A.nl) * A.nf;        // dalpha plane
    const size_t jrow = static_cast<size_t>(Aj.ldj) * A.nf;      // J plane of one ray
    float *JT = Aj.J + static_cast<size_t>(ray) * 3 * jrow;
    float *Jq = JT + jrow;
    float *Jw = Jq + jrow;
    const size_t row0 = static_cast<size_t>(base) * A.nf;
    const void *alpha = A.f32 ? static_cast<const void *>(static_cast<const float *>(A.alpha) + row0)
                              : static_cast<const void *>(static_cast<const double *>(A.alpha) + row0);
    const double *Tl = A.T + base;
    const double *dal = Aj.dalpha + row0;

    // pass 1: Tb and intW (the recurrence of integrate_kernel, the cut checked every 8 steps)
    double intW, num;
    int last;   // last path point reached
    if (A.fast_exp)
        forward_sums<true>(A, h, Tl, row0, S, jj, ok, intW, num, last);
    else
        forward_sums<false>(A, h, Tl, row0, S, jj, ok, intW, num, last);
    const bool live = S > 1 && !(num < kTcmb);
    const double Tb = !valid ? __builtin_nan("") : ((S < 0 || num < kTcmb) ? kTcmb : num / intW);
    if (ok && A.Tb) A.Tb[static_cast<size_t>(ray) * A.nf + j] = Tb;

    // pass 2: gradients along the path; zero elsewhere
    int k = 0;
    if (live) {
        const double invD = 1.0 / intW;
        double P = 0.0;          // P_{k-1}
        double hprev = 0.0;      // h_{k-1}
        double a_k = load_alpha(alpha, A.f32, jj);
        double tau = 0.0;
        for (k = 0; k <= last; ++k) {
            const double hk = (k < last) ? h[k] : 0.0;   // the cut (or path end) closes the sums
            double e = 1.0, W = 0.0;
            if (k > 0) {
                e = A.fast_exp ? exp_neg_mixed(tau) : exp(-tau);
                W = a_k * e;
            }
            const double ck = hprev + hk;
            const double lam = ck * (Tl[k] - Tb) * invD;
            const double Pk = P + lam * W;
            const double ga = (k > 0 ? lam * e : 0.0) + hprev * P + hk * Pk;
            const size_t o = static_cast<size_t>(k) * A.nf + jj;
            const double dT = ck * W * invD + ga * dal[o];
            const double dq = ga * dal[nlnf + o];
            const double dw = ga * dal[2 * nlnf + o];
            if (ok) {
                JT[static_cast<size_t>(k) * A.nf + j] = static_cast<float>(dT);
                Jq[static_cast<size_t>(k) * A.nf + j] = static_cast<float>(dq);
                Jw[static_cast<size_t>(k) * A.nf + j] = static_cast<float>(dw);
            }
            if (k < last) {
                const double a_next = load_alpha(alpha, A.f32, static_cast<size_t>(k + 1) * A.nf + jj);
                tau = tau + (a_k + a_next) * hk;
                a_k = a_next;
            }
            P = Pk;
            hprev = hk;
        }
    }
    if (ok) {
        for (; k < Aj.ldj; ++k) {
            JT[static_cast<size_t>(k) * A.nf + j] = 0.0f;
            Jq[static_cast<size_t>(k) * A.nf + j] = 0.0f;
            Jw[static_cast<size_t>(k) * A.nf + j] = 0.0f;
        }
    }
}

// ---- adjoint of the integration (vector-Jacobian product) ----
// For upstream gradients gTb[ray][f] this accumulates, over every ray,
//   galpha[l][f] = sum_ray gTb dTb/dalpha_l   and   gT[l] = sum_ray,f gTb dTb/dT_l (direct part)
// where the direct part is the dependence through T_l W_l; the dependence of alpha on T is
// galpha's, which rbx_alpha_vjp contracts with d alpha / d(T, q).  Per ray and frequency the
// derivative is the one jacobian_kernel writes (a forward pass for Tb and the integrated
// weight D, then a second pass along the path with e^-tau recomputed, never stored).  Here the
// second pass of a wavefront's NR rays runs interleaved, layer by layer, so the rays' terms
// add up in registers and reach HBM as one FP64 atomic add per (wavefront, layer, frequency)
// instead of one per ray; the direct T part is summed over the wavefront's 64 frequencies
// first (one atomic add per wavefront and layer).  Lanes whose gTb is 0 contribute nothing
// (not 0 x NaN), so rays and frequencies a loss masks out are skipped even where Tb is NaN; a
// NaN Tb under a non-zero gTb propagates.
struct AdjArgs {
    IntArgs I;
    const double *gTb;   // [nray][nf]
    double *galpha;      // [nl][nf] result
    double *gT;          // [nl] result
    // the atomic adds go to `copies` copies of galpha, [copies][nl][nf] (ray block b adds to
    // copy b % copies), and of gT per frequency group, [copies][ngroups][nl], zeroed first and
    // summed into galpha and gT at the end: thousands of wavefronts of one atmosphere add to
    // the same rows at the same time, and spreading them over copies keeps them from queueing
    // at the same addresses.  With one copy these are galpha and gT themselves (gT then holds
    // [ngroups][nl] partial sums only if ngroups is 1; see launch_adjoint).
    double *acc_alpha, *acc_T;
    int copies;
};

constexpr int kAdjChunk = 8;   // layers summed in registers between atomic flushes

// One ray of the wavefront: its path and atmosphere rows.
struct AdjRay {
    const double *h;     // half steps (cm)
    int steps;           // path steps: S - 1 (0 or less: no path)
    double g;            // gTb of this lane (0: the lane contributes nothing)
};

// Both passes for NR rays of one atmosphere (first alpha row row0 = base x nf, temperatures Tl),
// interleaved layer by layer so the rays share each alpha row and temperature load, and the
// next layer's loads are issued one step ahead.
// Pass 1, per ray: tau, the integrated weight D and the T-weighted sum (the recurrence of
// integrate_kernel; the optical-depth cut checked every 8 steps per ray, off for a path with a
// NaN half step or temperature), and the last path point reached.
// Pass 2, per layer k and ray, along the same points with e^-tau recomputed:
//   c_k = h_{k-1} + h_k, gl = gTb lambda_k = c_k (T_k - Tb) u with u = gTb / D,
//   W_k = alpha_k e^-tau_k,  P_k = P_{k-1} + gl W_k (gTb times the prefix of lambda W),
//   gTb dTb/dalpha_k = gl e^-tau_k + h_{k-1} P_{k-1} + h_k P_k,   gTb dTb/dT_k = c_k W_k u,
// with h_last = 0 closing each ray's sums (the path end or the cut); the rays' terms add up in
// registers and go to the accumulators with one atomic add per layer (gT summed over the
// wavefront's frequencies first).
template <int NR, bool FAST>
__device__ inline void adj_rays(const AdjArgs &Aa, const AdjRay (&R)[NR], const double *Tl, size_t row0, int base,
                                int j, int jj, bool ok) {
    const IntArgs &A = Aa.I;
    const size_t nf = A.nf;
    int kmax = 0;                          // steps of the longest path
    double cut[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        kmax = max(kmax, R[r].steps);
        cut[r] = (A.tau_cut > 0.0 && R[r].steps > 0 && !path_has_nan(R[r].h, Tl, R[r].steps)) ? A.tau_cut : 0.0;
    }
    if (kmax <= 0) return;
    // ---- pass 1 ----
    double tau[NR], W0[NR], TW0[NR], iw[NR], nm[NR];
    int last[NR];
    bool run[NR];                          // wavefront-uniform
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        tau[r] = W0[r] = TW0[r] = iw[r] = nm[r] = 0.0;
        last[r] = 0;
        run[r] = R[r].steps > 0;
    }
    double a0 = load_alpha(A.alpha, A.f32, row0 + jj);
    double a_n = load_alpha(A.alpha, A.f32, row0 + nf + jj), T_n = Tl[1];
    double h_n[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) h_n[r] = R[r].steps > 0 ? R[r].h[0] : 0.0;
    for (int i = 0; i < kmax; ++i) {
        const double a1 = a_n, T1 = T_n;
        double hh[NR];
#pragma unroll
        for (int r = 0; r < NR; ++r) hh[r] = h_n[r];
        if (i + 1 < kmax) {   // the next step's loads, one step ahead
            a_n = load_alpha(A.alpha, A.f32, row0 + static_cast<size_t>(i + 2) * nf + jj);
            T_n = Tl[i + 2];
#pragma unroll
            for (int r = 0; r < NR; ++r) h_n[r] = i + 1 < R[r].steps ? R[r].h[i + 1] : 0.0;
        }
        bool any = false;
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            if (!run[r] || i >= R[r].steps) continue;
            tau[r] = tau[r] + (a0 + a1) * hh[r];
            const double W1 = a1 * exp_neg<FAST>(tau[r]);
            iw[r] = iw[r] + (W1 + W0[r]) * hh[r];
            const double TW1 = T1 * W1;
            nm[r] = nm[r] + (TW1 + TW0[r]) * hh[r];
            W0[r] = W1;
            TW0[r] = TW1;
            last[r] = i + 1;
            if ((i & 7) == 7 && cut[r] > 0.0 && __all(!ok || tau[r] > cut[r])) run[r] = false;
            any = true;
        }
        a0 = a1;
        if (!any) break;
    }
    // per ray: Tb, u = gTb / D, and whether it has a gradient (on the planet, more than one
    // point, Tb above the T_cmb floor, some lane with gTb != 0)
    double Tb[NR], u[NR];
    int kend = -1;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        const bool live = R[r].steps > 0 && !(nm[r] < kTcmb) && ok && R[r].g != 0.0;
        Tb[r] = nm[r] / iw[r];
        u[r] = live ? R[r].g / iw[r] : 0.0;
        if (!__any(live)) last[r] = -1;
        kend = max(kend, last[r]);
    }
    if (kend < 0) return;
    // ---- pass 2 ----
    const int group = blockIdx.x % A.ngroups;
    const int copy = static_cast<int>((blockIdx.x / A.ngroups) % Aa.copies);
    double *acc_a = Aa.acc_alpha + static_cast<size_t>(copy) * A.nl * nf;
    double *acc_t = Aa.acc_T + (static_cast<size_t>(copy) * A.ngroups + group) * A.nl;
    double P[NR], hp[NR];   // per ray: P_{k-1}, h_{k-1}
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        tau[r] = 0.0;
        P[r] = 0.0;
        hp[r] = 0.0;
        h_n[r] = last[r] > 0 ? R[r].h[0] : 0.0;
    }
    double a_prev = 0.0;
    a_n = load_alpha(A.alpha, A.f32, row0 + jj);
    T_n = Tl[0];
    for (int k = 0; k <= kend; ++k) {
        const double a_k = a_n, T_k = T_n;
        double hk[NR];
#pragma unroll
        for (int r = 0; r < NR; ++r) hk[r] = k < last[r] ? h_n[r] : 0.0;   // the path end closes the sums
        if (k < kend) {
            a_n = load_alpha(A.alpha, A.f32, row0 + static_cast<size_t>(k + 1) * nf + jj);
            T_n = Tl[k + 1];
#pragma unroll
            for (int r = 0; r < NR; ++r) h_n[r] = k + 1 < last[r] ? R[r].h[k + 1] : 0.0;
        }
        double sa = 0.0, st = 0.0;
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            if (k > last[r]) continue;   // wavefront-uniform: this ray has ended
            double e = 1.0, W = 0.0;
            if (k > 0) {
                tau[r] = tau[r] + (a_prev + a_k) * hp[r];
                e = exp_neg<FAST>(tau[r]);
                W = a_k * e;
            }
            const double ck = hp[r] + hk[r];
            const double gl = ck * (T_k - Tb[r]) * u[r];
            const double Pn = P[r] + gl * W;
            const double ga = (k > 0 ? gl * e : 0.0) + hp[r] * P[r] + hk[r] * Pn;
            const bool on = u[r] != 0.0;
            sa += on ? ga : 0.0;
            st += on ? ck * W * u[r] : 0.0;
            P[r] = Pn;
            hp[r] = hk[r];
        }
        a_prev = a_k;
        if (ok && sa != 0.0) unsafeAtomicAdd(&acc_a[row0 + static_cast<size_t>(k) * nf + j], sa);
        const double t = wave_sum(ok ? st : 0.0);
        if (t != 0.0 && threadIdx.x % kLanes == 0) unsafeAtomicAdd(&acc_t[base + k], t);
    }
}

// ---- kernel: wavefront w takes schedule slots [w NR, (w + 1) NR) ----
template <int NR, bool FAST>
__global__ __launch_bounds__(kWaves * kLanes) void adjoint_kernel(AdjArgs Aa) {
    const IntArgs &A = Aa.I;
    const int wave = threadIdx.x / kLanes;
    const int lane = threadIdx.x % kLanes;
    const int group = blockIdx.x % A.ngroups;
    const int slot0 = ((blockIdx.x / A.ngroups) * kWaves + wave) * NR;
    const int j = group * kLanes + lane;
    const bool ok = j < A.nf;
    const int jj = ok ? j : A.nf - 1;
    AdjRay R[NR];
    int base0 = -1, bases[NR];
    bool same = true;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        R[r].h = A.hds;
        R[r].steps = 0;
        R[r].g = 0.0;
        bases[r] = -1;
        const int ray = ray_of(A, slot0 + r);
        if (ray < 0) continue;
        int base, nlr;
        if (!ray_span(A, ray, base, nlr)) continue;    // bad atmosphere index: Tb NaN, no gradient
        R[r].h = A.hds + static_cast<size_t>(ray) * A.ld;
        R[r].steps = min(A.nsteps[ray], nlr - 1) - 1;
        R[r].g = ok ? Aa.gTb[static_cast<size_t>(ray) * A.nf + j] : 0.0;
        bases[r] = base;
        if (R[r].steps <= 0) continue;
        if (base0 < 0)
            base0 = base;
        else if (base != base0)
            same = false;
    }
    if (base0 < 0) return;
    if (same) {
        adj_rays<NR, FAST>(Aa, R, A.T + base0, static_cast<size_t>(base0) * A.nf, base0, j, jj, ok);
    } else {   // rays of several atmospheres in one wavefront: one at a time (one copy of the code)
#pragma unroll 1
        for (int r = 0; r < NR; ++r) {
            AdjRay one[1] = {R[0]};
            int b = bases[0];
#pragma unroll
            for (int q = 1; q < NR; ++q) {
                if (q == r) {
                    one[0] = R[q];
                    b = bases[q];
                }
            }
            if (one[0].steps <= 0) continue;
            adj_rays<1, FAST>(Aa, one, A.T + b, static_cast<size_t>(b) * A.nf, b, j, jj, ok);
        }
    }
}

__global__ __launch_bounds__(256) void adjoint_clear_kernel(double *acc_alpha, size_t na, double *acc_T, size_t nt) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < na) acc_alpha[i] = 0.0;
    if (i < nt) acc_T[i] = 0.0;
}

// galpha = the sum of the copies (in copy order); gT[l] = the sum over copies and frequency groups
__global__ __launch_bounds__(256) void adjoint_reduce_kernel(const double *acc_alpha, const double *acc_T, int copies,
                                                             int parts_T, size_t n, int nl, double *galpha,
                                                             double *gT) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n && galpha != acc_alpha) {
        double s = 0.0;
        for (int c = 0; c < copies; ++c) s += acc_alpha[static_cast<size_t>(c) * n + i];
        galpha[i] = s;
    }
    if (i < static_cast<size_t>(nl)) {
        double s = 0.0;
        for (int c = 0; c < parts_T; ++c) s += acc_T[static_cast<size_t>(c) * nl + i];
        gT[i] = s;
    }
}

// grad[m][l] = sum_f galpha[l][f] dalpha[m][l][f] (+ gT[l] for m = 0): one wavefront per layer
__global__ __launch_bounds__(256) void alpha_vjp_kernel(const double *dalpha, const double *galpha, const double *gT,
                                                        int nl, int nf, double *grad) {
    const int l = blockIdx.x * (blockDim.x / kLanes) + threadIdx.x / kLanes;
    if (l >= nl) return;
    const int lane = threadIdx.x % kLanes;
    const size_t plane = static_cast<size_t>(nl) * nf;
    const size_t row = static_cast<size_t>(l) * nf;
    double s[3] = {0.0, 0.0, 0.0};
    for (int f = lane; f < nf; f += kLanes) {
        const double g = galpha[row + f];
#pragma unroll
        for (int m = 0; m < 3; ++m) s[m] = fma(g, dalpha[m * plane + row + f], s[m]);
    }
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        const double t = wave_sum(s[m]);
        if (lane == 0) grad[static_cast<size_t>(m) * nl + l] = t + (m == 0 && gT ? gT[l] : 0.0);
    }
}

// ---- per-point profiles of one ray set: tau, W and Tb_lyr at every path point ----
// (Brightness.single keeps these as .tau / .W / .Tb_lyr, brightness.py:55-57, 94-104, 118-120)
struct ProfArgs {
    IntArgs I;
    double *tau, *W, *Tbl;   // [nray][nl][nf]
};

__global__ __launch_bounds__(kWaves * kLanes) void profile_kernel(ProfArgs Ap) {
    const IntArgs &A = Ap.I;
    const int wave = threadIdx.x / kLanes;
    const int lane = threadIdx.x % kLanes;
    const int group = blockIdx.x % A.ngroups;
    const int ray = __builtin_amdgcn_readfirstlane((blockIdx.x / A.ngroups) * kWaves + wave);
    if (ray >= A.nray) return;
    const int j = group * kLanes + lane;
    if (j >= A.nf) return;
    const int S = min(A.nsteps[ray], A.nl - 1);
    const double *h = A.hds + static_cast<size_t>(ray) * A.ld;
    const size_t base = static_cast<size_t>(ray) * A.nl * A.nf + j;
    double tau = 0.0, W0 = 0.0, TW0 = 0.0, num = 0.0;
    int k = 0;
    if (S >= 1) {
        Ap.tau[base] = 0.0;
        Ap.W[base] = 0.0;
        Ap.Tbl[base] = 0.0;
        k = 1;
        double a0 = load_alpha(A.alpha, A.f32, j);
        for (int i = 0; i < S - 1; ++i, ++k) {
            const double hh = h[i];
            const double a1 = load_alpha(A.alpha, A.f32, static_cast<size_t>(i + 1) * A.nf + j);
            tau = tau + (a0 + a1) * hh;
            double W1;
            if (A.mode == 1)
                W1 = 2.0 * a1 * expint_e2(tau);
            else
                W1 = a1 * (A.fast_exp ? exp_neg_mixed(tau) : exp(-tau));
            const double TW1 = A.T[i + 1] * W1;
            num = num + (TW1 + TW0) * hh;
            const size_t o = base + static_cast<size_t>(i + 1) * A.nf;
            Ap.tau[o] = tau;
            Ap.W[o] = W1;
            Ap.Tbl[o] = num;
            a0 = a1;
            W0 = W1;
            TW0 = TW1;
        }
    }
    for (; k < A.nl; ++k) {   // points the ray does not reach
        const size_t o = base + static_cast<size_t>(k) * A.nf;
        Ap.tau[o] = 0.0;
        Ap.W[o] = 0.0;
        Ap.Tbl[o] = 0.0;
    }
    (void)W0;
}

static int check_int_args(const void *alpha, int alpha_precision, const double *T, int nl, int nf,
                          const double *hds, int ld, const int *nsteps, int nray, int precision) {
    RBX_CHECK_ARG(alpha && T && hds && nsteps, "integration: null device pointer");
    RBX_CHECK_ARG(nl >= 2 && nf > 0 && nray >= 0, "integration: bad sizes nl=%d nf=%d nray=%d", nl, nf, nray);
    RBX_CHECK_ARG(alpha_precision == RBX_FP64 || alpha_precision == RBX_FP32, "integration: bad alpha precision");
    RBX_CHECK_ARG(precision == RBX_FP64 || precision == RBX_FP32, "integration: bad precision");
    return RBX_OK;
}

static void fill_int_args(IntArgs &A, const void *alpha, int alpha_precision, const double *T, int nl, int nf,
                          const double *hds, int ld, const int *nsteps, int nray, int precision, double tau_cut) {
    A.alpha = alpha;
    A.f32 = alpha_precision == RBX_FP32;
    A.T = T;
    A.nl = nl;
    A.nf = nf;
    A.hds = hds;
    A.ld = ld;
    A.nsteps = nsteps;
    A.nray = nray;
    A.mode = 0;
    A.fast_exp = precision == RBX_FP32;
    A.tau_cut = tau_cut;
    A.Tb = nullptr;
    A.intW = nullptr;
    A.ngroups = (nf + kLanes - 1) / kLanes;
    A.ray_major = 0;
    A.ray_atm = nullptr;
    A.layer_off = nullptr;
    A.natm = 0;
    A.order = nullptr;
    A.cols = nullptr;
    A.factors = nullptr;
}

// ---- frequency order for the optical-depth cut ----
// A wavefront stops integrating once every one of its 64 frequencies is past the cut, so
// frequencies of very different opacity in one wavefront make it run as long as its most
// transparent lane.  Ordering the columns by the step at which a reference ray (the on-planet
// ray with the longest path) passes the cut groups frequencies that stop together: on 800
// layers x 1024 frequencies x 16384 rays the wavefronts ran 1.45 times the steps their lanes
// needed in request order and 1.07 times in this order.  One block: the reference ray (most
// steps, lowest index), each column's cut step (the path's end when it is never passed), and
// the columns sorted by (cut step, column) -- a stable sort by cut step -- with a bitonic
// network over (step << kOrderColBits | column) in LDS.
constexpr int kOrderThreads = 1024;
constexpr int kOrderColBits = 14;
constexpr int kOrderMaxCols = 1 << kOrderColBits;   // sort keys kept in LDS
constexpr int kOrderMaxSteps = 1 << 17;             // cut steps fit the key's upper 17 bits
constexpr int kOrderBatch = 32;                     // path steps loaded together

template <class T>
__global__ __launch_bounds__(kOrderThreads) void integrate_order_kernel(const T *alpha, int nl, int nf,
                                                                        const double *hds, int ld,
                                                                        const int *nsteps, int nray, double cut,
                                                                        int *cols) {
    __shared__ int key[kOrderMaxCols];
    __shared__ int best[kOrderThreads / 64][2];
    const int tid = threadIdx.x, lane = tid % 64, wave = tid / 64;
    int bs = -1, br = 0x7fffffff;   // (steps, ray) of this thread's best ray
    for (int r = tid; r < nray; r += kOrderThreads) {
        const int s = nsteps[r];
        if (s > bs) {
            bs = s;
            br = r;
        }
    }
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) {
        const int os = __shfl_xor(bs, m), orr = __shfl_xor(br, m);
        if (os > bs || (os == bs && orr < br)) {
            bs = os;
            br = orr;
        }
    }
    if (lane == 0) {
        best[wave][0] = bs;
        best[wave][1] = br;
    }
    __syncthreads();
    bs = best[0][0];
    br = best[0][1];
    for (int w = 1; w < kOrderThreads / 64; ++w)
        if (best[w][0] > bs || (best[w][0] == bs && best[w][1] < br)) {
            bs = best[w][0];
            br = best[w][1];
        }
    const int S = bs > 1 ? min(bs, nl) - 1 : 0;   // path steps of the reference ray
    const double *h = hds + static_cast<size_t>(br < nray ? br : 0) * ld;
    // the steps in batches of kOrderBatch whose loads are all issued before the first sum
    // (rows past the path re-read its last row and are not summed), so a column costs
    // S / kOrderBatch memory latencies rather than S
    for (int j = tid; j < nf; j += kOrderThreads) {
        double tau = 0.0, a0 = static_cast<double>(alpha[j]);
        int k = S;
        for (int i0 = 0; i0 < S && k == S; i0 += kOrderBatch) {
            double a[kOrderBatch], hb[kOrderBatch];
#pragma unroll
            for (int t = 0; t < kOrderBatch; ++t) {
                const int i = min(i0 + t, S - 1);
                a[t] = static_cast<double>(alpha[static_cast<size_t>(i + 1) * nf + j]);
                hb[t] = h[i];
            }
#pragma unroll
            for (int t = 0; t < kOrderBatch; ++t) {
                if (k == S && i0 + t < S) {
                    tau += (a0 + a[t]) * hb[t];
                    if (tau > cut) k = i0 + t;
                }
                a0 = a[t];
            }
        }
        key[j] = (k << kOrderColBits) | j;
    }
    int n2 = 1;
    while (n2 < nf) n2 <<= 1;
    for (int j = nf + tid; j < n2; j += kOrderThreads) key[j] = 0x7fffffff;   // pads sort last
    __syncthreads();
    for (int size = 2; size <= n2; size <<= 1)
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int t = tid; t < n2 / 2; t += kOrderThreads) {
                const int lo = 2 * t - (t & (stride - 1)), hi = lo + stride;
                const int a = key[lo], b = key[hi];
                if ((a > b) == ((lo & size) == 0)) {
                    key[lo] = b;
                    key[hi] = a;
                }
            }
            __syncthreads();
        }
    for (int r = tid; r < nf; r += kOrderThreads) cols[r] = key[r] & (kOrderMaxCols - 1);
}

// out[l][j] = in[l][cols[j]] (the table in the order of cols)
template <class T>
__global__ __launch_bounds__(256) void table_cols_kernel(const T *in, int nl, int nf, const int *cols, T *out) {
    const size_t n = static_cast<size_t>(nl) * nf;
    for (size_t e = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < n;
         e += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const size_t l = e / nf;
        const int j = static_cast<int>(e - l * nf);
        out[e] = in[l * nf + cols[j]];
    }
}

}  // namespace rbx

static int launch_integrate(rbx::IntArgs A, const rbx_launch &L, void *stream) {
    const long long rblocks = (A.nray + rbx::kWaves - 1) / rbx::kWaves;
    // ray major pads the ray blocks to a multiple of the XCD count (padding blocks find no ray).
    // Built-in choice: ray major while the whole table fits one XCD's L2 (4 MiB) beside the
    // streamed ds rows (measured on the 800 x 1024 table: 2.7 ms ray major in FP32, 3.3 ms
    // group major; the 6.5 MB FP64 table is faster group major).
    const int forced = L.int_map;
    const double table = static_cast<double>(A.nl) * A.nf * (A.f32 ? 4.0 : 8.0);
    A.ray_major = forced ? forced == 2 : table <= 3.5 * 1024 * 1024;
    const long long blocks = static_cast<long long>(A.ngroups) *
                             (A.ray_major ? (rblocks + rbx::kXcds - 1) / rbx::kXcds * rbx::kXcds : rblocks);
    RBX_CHECK_ARG(blocks < (1LL << 31), "rbx_integrate: too many blocks");
    const dim3 grid(static_cast<unsigned>(blocks)), block(rbx::kWaves * rbx::kLanes);
    hipStream_t st = static_cast<hipStream_t>(stream);
    const bool u4 = L.int_unroll == 4;
#define RBX_INT_LAUNCH(AT, MODE, FAST, U) hipLaunchKernelGGL((rbx::integrate_kernel<AT, MODE, FAST, U>), grid, block, 0, st, A)
    if (A.mode == 1) {
        if (A.f32) RBX_INT_LAUNCH(float, 1, false, 4); else RBX_INT_LAUNCH(double, 1, false, 4);
    } else if (A.f32) {
        const size_t lds = sizeof(float2) * rbx::kWaves * static_cast<size_t>(A.ld);
        if (A.fast_exp && (A.factors || lds <= 64 * 1024)) {
            const bool buf = static_cast<size_t>(A.nl) * static_cast<size_t>(A.nf) * 4 < (size_t(1) << 31);
            const size_t L = A.factors ? 0 : lds;
#define RBX_F32_LAUNCH(U, B)                                                                                 \
    do {                                                                                                    \
        if (A.factors) hipLaunchKernelGGL((rbx::integrate_f32_kernel<U, B, true>), grid, block, L, st, A);  \
        else hipLaunchKernelGGL((rbx::integrate_f32_kernel<U, B, false>), grid, block, L, st, A);           \
    } while (0)
            if (u4 && buf) RBX_F32_LAUNCH(4, true);
            else if (u4) RBX_F32_LAUNCH(4, false);
            else if (buf) RBX_F32_LAUNCH(8, true);
            else RBX_F32_LAUNCH(8, false);
#undef RBX_F32_LAUNCH
        } else if (A.fast_exp) { if (u4) RBX_INT_LAUNCH(float, 0, true, 4); else RBX_INT_LAUNCH(float, 0, true, 8); }
        else { if (u4) RBX_INT_LAUNCH(float, 0, false, 4); else RBX_INT_LAUNCH(float, 0, false, 8); }
    } else {
        if (A.fast_exp) { if (u4) RBX_INT_LAUNCH(double, 0, true, 4); else RBX_INT_LAUNCH(double, 0, true, 8); }
        else { if (u4) RBX_INT_LAUNCH(double, 0, false, 4); else RBX_INT_LAUNCH(double, 0, false, 8); }
    }
#undef RBX_INT_LAUNCH
    return rbx::launch_status("integrate_kernel");
}

extern "C" int rbx_integrate(const void *alpha, int alpha_precision, const double *T, int nl, int nf,
                             const double *hds, int ld, const int *nsteps, int nray, int mode, int precision,
                             double tau_cut, double *Tb, double *intW, const rbx_launch *launch, void *stream) {
    int rc = rbx::check_int_args(alpha, alpha_precision, T, nl, nf, hds, ld, nsteps, nray, precision);
    if (rc) return rc;
    rbx_launch L;
    rc = rbx::read_launch(launch, L, "rbx_integrate");
    if (rc) return rc;
    RBX_CHECK_ARG(ld >= nl - 1, "rbx_integrate: ld=%d < nl-1=%d", ld, nl - 1);
    RBX_CHECK_ARG(Tb != nullptr, "rbx_integrate: null Tb");
    RBX_CHECK_ARG(mode == 0 || mode == 1, "rbx_integrate: bad mode %d", mode);
    if (nray == 0) return RBX_OK;
    rbx::IntArgs A;
    rbx::fill_int_args(A, alpha, alpha_precision, T, nl, nf, hds, ld, nsteps, nray, precision, tau_cut);
    A.mode = mode;
    A.Tb = Tb;
    A.intW = intW;
    return launch_integrate(A, L, stream);
}

extern "C" int rbx_integrate_order(const void *alpha, int alpha_precision, int nl, int nf, const double *hds,
                                   int ld, const int *nsteps, int nray, double tau_cut, int *cols, void *alpha_cols,
                                   void *stream) {
    RBX_CHECK_ARG(alpha && hds && nsteps && cols && alpha_cols, "rbx_integrate_order: null device pointer");
    RBX_CHECK_ARG(alpha_precision == RBX_FP64 || alpha_precision == RBX_FP32,
                  "rbx_integrate_order: bad alpha precision %d", alpha_precision);
    RBX_CHECK_ARG(nl >= 1 && nf >= 1 && nray >= 1, "rbx_integrate_order: nl=%d nf=%d nray=%d", nl, nf, nray);
    RBX_CHECK_ARG(nf <= rbx::kOrderMaxCols, "rbx_integrate_order: %d frequencies, at most %d", nf,
                  rbx::kOrderMaxCols);
    RBX_CHECK_ARG(nl <= rbx::kOrderMaxSteps, "rbx_integrate_order: %d layers, at most %d", nl,
                  rbx::kOrderMaxSteps);
    RBX_CHECK_ARG(ld >= nl - 1, "rbx_integrate_order: ld=%d < nl-1=%d", ld, nl - 1);
    RBX_CHECK_ARG(tau_cut > 0.0, "rbx_integrate_order: the order is for a positive tau cut, got %g", tau_cut);
    const hipStream_t st = static_cast<hipStream_t>(stream);
    if (alpha_precision == RBX_FP32)
        hipLaunchKernelGGL(rbx::integrate_order_kernel<float>, dim3(1), dim3(rbx::kOrderThreads), 0, st,
                           static_cast<const float *>(alpha), nl, nf, hds, ld, nsteps, nray, tau_cut, cols);
    else
        hipLaunchKernelGGL(rbx::integrate_order_kernel<double>, dim3(1), dim3(rbx::kOrderThreads), 0, st,
                           static_cast<const double *>(alpha), nl, nf, hds, ld, nsteps, nray, tau_cut, cols);
    int rc = rbx::launch_status("integrate_order_kernel");
    if (rc) return rc;
    const size_t n = static_cast<size_t>(nl) * nf;
    const unsigned blocks = static_cast<unsigned>(std::min<size_t>((n + 255) / 256, 4096));
    if (alpha_precision == RBX_FP32)
        hipLaunchKernelGGL(rbx::table_cols_kernel<float>, dim3(blocks), dim3(256), 0, st,
                           static_cast<const float *>(alpha), nl, nf, cols, static_cast<float *>(alpha_cols));
    else
        hipLaunchKernelGGL(rbx::table_cols_kernel<double>, dim3(blocks), dim3(256), 0, st,
                           static_cast<const double *>(alpha), nl, nf, cols, static_cast<double *>(alpha_cols));
    return rbx::launch_status("table_cols_kernel");
}

extern "C" int rbx_integrate_cols(const void *alpha_cols, int alpha_precision, const int *cols, const double *T,
                                  int nl, int nf, const double *hds, int ld, const int *nsteps, int nray, int mode,
                                  int precision, double tau_cut, double *Tb, double *intW, const void *factors,
                                  const rbx_launch *launch, void *stream) {
    int rc = rbx::check_int_args(alpha_cols, alpha_precision, T, nl, nf, hds, ld, nsteps, nray, precision);
    if (rc) return rc;
    rbx_launch L;
    rc = rbx::read_launch(launch, L, "rbx_integrate_cols");
    if (rc) return rc;
    RBX_CHECK_ARG(cols != nullptr, "rbx_integrate_cols: null column order");
    RBX_CHECK_ARG(ld >= nl - 1, "rbx_integrate_cols: ld=%d < nl-1=%d", ld, nl - 1);
    RBX_CHECK_ARG(Tb != nullptr, "rbx_integrate_cols: null Tb");
    RBX_CHECK_ARG(mode == 0 || mode == 1, "rbx_integrate_cols: bad mode %d", mode);
    if (nray == 0) return RBX_OK;
    rbx::IntArgs A;
    rbx::fill_int_args(A, alpha_cols, alpha_precision, T, nl, nf, hds, ld, nsteps, nray, precision, tau_cut);
    A.mode = mode;
    A.Tb = Tb;
    A.intW = intW;
    A.cols = cols;
    // FP32 point rays read each ray's path factors from rbx_integrate_factors' table
    RBX_CHECK_ARG(reinterpret_cast<uintptr_t>(factors) % 16 == 0, "rbx_integrate_cols: factors not 16-byte aligned");
    if (factors && mode == 0 && A.f32 && A.fast_exp) A.factors = static_cast<const rbx::PathFactor *>(factors);
    return launch_integrate(A, L, stream);
}

extern "C" int rbx_integrate_factors(const double *T, int nl, const double *hds, int ld, const int *nsteps,
                                     int nray, void *factors, void *stream) {
    RBX_CHECK_ARG(T && hds && nsteps && factors, "rbx_integrate_factors: null device pointer");
    RBX_CHECK_ARG(reinterpret_cast<uintptr_t>(factors) % 16 == 0, "rbx_integrate_factors: factors not 16-byte aligned");
    RBX_CHECK_ARG(nl >= 1 && nray >= 0, "rbx_integrate_factors: nl=%d nray=%d", nl, nray);
    RBX_CHECK_ARG(ld >= nl - 1 && ld >= 1, "rbx_integrate_factors: ld=%d < nl-1=%d", ld, nl - 1);
    if (nray == 0) return RBX_OK;
    rbx::IntArgs A;
    rbx::fill_int_args(A, nullptr, RBX_FP32, T, nl, 1, hds, ld, nsteps, nray, RBX_FP32, 0.0);
    const unsigned blocks = static_cast<unsigned>((nray + rbx::kWaves - 1) / rbx::kWaves);
    hipLaunchKernelGGL(rbx::integrate_factors_kernel, dim3(blocks), dim3(rbx::kWaves * rbx::kLanes), 0,
                       static_cast<hipStream_t>(stream), A, static_cast<rbx::PathFactor *>(factors));
    return rbx::launch_status("integrate_factors_kernel");
}

extern "C" int rbx_integrate_batch(const void *alpha, int alpha_precision, const double *T, int nl_total, int nf,
                                   const int *layer_off, int natm, const double *hds, int ld, const int *nsteps,
                                   const int *ray_atm, const int *order, int nray, int mode, int precision,
                                   double tau_cut, double *Tb, double *intW, const rbx_launch *launch,
                                   void *stream) {
    int rc = rbx::check_int_args(alpha, alpha_precision, T, nl_total, nf, hds, ld, nsteps, nray, precision);
    if (rc) return rc;
    rbx_launch L;
    rc = rbx::read_launch(launch, L, "rbx_integrate_batch");
    if (rc) return rc;
    RBX_CHECK_ARG(layer_off && ray_atm && natm > 0, "rbx_integrate_batch: null offsets / ray map or natm=%d", natm);
    RBX_CHECK_ARG(ld >= 1, "rbx_integrate_batch: ld=%d", ld);
    RBX_CHECK_ARG(Tb != nullptr, "rbx_integrate_batch: null Tb");
    RBX_CHECK_ARG(mode == 0 || mode == 1, "rbx_integrate_batch: bad mode %d", mode);
    if (nray == 0) return RBX_OK;
    rbx::IntArgs A;
    rbx::fill_int_args(A, alpha, alpha_precision, T, nl_total, nf, hds, ld, nsteps, nray, precision, tau_cut);
    A.mode = mode;
    A.Tb = Tb;
    A.intW = intW;
    A.ray_atm = ray_atm;
    A.layer_off = layer_off;
    A.natm = natm;
    A.order = order;
    return launch_integrate(A, L, stream);
}

static int launch_jacobian(const rbx::JacArgs &Aj, void *stream) {
    const long long blocks = static_cast<long long>(Aj.I.ngroups) * ((Aj.I.nray + rbx::kWaves - 1) / rbx::kWaves);
    RBX_CHECK_ARG(blocks < (1LL << 31), "rbx_jacobian: too many blocks");
    hipLaunchKernelGGL(rbx::jacobian_kernel, dim3(static_cast<unsigned>(blocks)), dim3(rbx::kWaves * rbx::kLanes),
                       0, static_cast<hipStream_t>(stream), Aj);
    return rbx::launch_status("jacobian_kernel");
}

extern "C" int rbx_jacobian(const void *alpha, int alpha_precision, const double *dalpha, const double *T, int nl,
                            int nf, const double *hds, int ld, const int *nsteps, int nray, double tau_cut, float *J,
                            double *Tb, void *stream) {
    int rc = rbx::check_int_args(alpha, alpha_precision, T, nl, nf, hds, ld, nsteps, nray, alpha_precision);
    if (rc) return rc;
    RBX_CHECK_ARG(ld >= nl - 1, "rbx_jacobian: ld=%d < nl-1=%d", ld, nl - 1);
    RBX_CHECK_ARG(dalpha != nullptr && J != nullptr, "rbx_jacobian: null dalpha or J");
    if (nray == 0) return RBX_OK;
    rbx::JacArgs Aj;
    rbx::fill_int_args(Aj.I, alpha, alpha_precision, T, nl, nf, hds, ld, nsteps, nray, alpha_precision, tau_cut);
    Aj.I.Tb = Tb;
    Aj.dalpha = dalpha;
    Aj.J = J;
    Aj.ldj = nl;
    return launch_jacobian(Aj, stream);
}

extern "C" int rbx_jacobian_batch(const void *alpha, int alpha_precision, const double *dalpha, const double *T,
                                  int nl_total, int nf, const int *layer_off, int natm, const double *hds, int ld,
                                  const int *nsteps, const int *ray_atm, const int *order, int nray, int precision,
                                  double tau_cut, float *J, int ldj, double *Tb, void *stream) {
    int rc = rbx::check_int_args(alpha, alpha_precision, T, nl_total, nf, hds, ld, nsteps, nray, precision);
    if (rc) return rc;
    RBX_CHECK_ARG(layer_off && ray_atm && natm > 0, "rbx_jacobian_batch: null offsets / ray map or natm=%d", natm);
    RBX_CHECK_ARG(ld >= 1 && ldj >= 2, "rbx_jacobian_batch: ld=%d ldj=%d", ld, ldj);
    RBX_CHECK_ARG(dalpha != nullptr && J != nullptr, "rbx_jacobian_batch: null dalpha or J");
    if (nray == 0) return RBX_OK;
    rbx::JacArgs Aj;
    rbx::fill_int_args(Aj.I, alpha, alpha_precision, T, nl_total, nf, hds, ld, nsteps, nray, precision, tau_cut);
    Aj.I.Tb = Tb;
    Aj.I.ray_atm = ray_atm;
    Aj.I.layer_off = layer_off;
    Aj.I.natm = natm;
    Aj.I.order = order;
    Aj.dalpha = dalpha;
    Aj.J = J;
    Aj.ldj = ldj;
    return launch_jacobian(Aj, stream);
}

extern "C" int rbx_integrate_profiles(const void *alpha, int alpha_precision, const double *T, int nl, int nf,
                                      const double *hds, int ld, const int *nsteps, int nray, int mode,
                                      int precision, double *tau, double *W, double *Tb_lyr, void *stream) {
    int rc = rbx::check_int_args(alpha, alpha_precision, T, nl, nf, hds, ld, nsteps, nray, precision);
    if (rc) return rc;
    RBX_CHECK_ARG(ld >= nl - 1, "rbx_integrate_profiles: ld=%d < nl-1=%d", ld, nl - 1);
    RBX_CHECK_ARG(tau && W && Tb_lyr, "rbx_integrate_profiles: null output");
    RBX_CHECK_ARG(mode == 0 || mode == 1, "rbx_integrate_profiles: bad mode %d", mode);
    if (nray == 0) return RBX_OK;
    rbx::ProfArgs Ap;
    rbx::fill_int_args(Ap.I, alpha, alpha_precision, T, nl, nf, hds, ld, nsteps, nray, precision, 0.0);
    Ap.I.mode = mode;
    Ap.tau = tau;
    Ap.W = W;
    Ap.Tbl = Tb_lyr;
    const long long blocks = static_cast<long long>(Ap.I.ngroups) * ((nray + rbx::kWaves - 1) / rbx::kWaves);
    RBX_CHECK_ARG(blocks < (1LL << 31), "rbx_integrate_profiles: too many blocks");
    hipLaunchKernelGGL(rbx::profile_kernel, dim3(static_cast<unsigned>(blocks)), dim3(rbx::kWaves * rbx::kLanes), 0,
                       static_cast<hipStream_t>(stream), Ap);
    return rbx::launch_status("profile_kernel");
}

namespace rbx {
// Rays per wavefront and copies of the accumulators of one adjoint call: launch.adj_rays
// (built in 4) and launch.adj_copies, else about one copy per 4 ray blocks of an atmosphere,
// at most 32 (a 16384-ray image of one atmosphere: 32; retrieval batches of 16 rays per
// atmosphere: 1, their atmospheres' rows already differ).
static void adjoint_plan(int nl, int nf, int nray, int natm, const rbx_launch &L, int &nr, int &copies,
                         size_t &bytes) {
    nr = L.adj_rays ? L.adj_rays : 4;
    const long long per_atm = (static_cast<long long>(nray) + natm - 1) / std::max(natm, 1);
    const long long blocks_per_atm = (per_atm + static_cast<long long>(kWaves) * nr - 1) / (kWaves * nr);
    copies = L.adj_copies ? L.adj_copies : static_cast<int>(std::min<long long>(32, std::max(1LL, blocks_per_atm / 4)));
    const int ngroups = (nf + kLanes - 1) / kLanes;
    const size_t n = static_cast<size_t>(nl) * nf;
    bytes = 0;
    if (copies > 1 || ngroups > 1)
        bytes = ((copies > 1 ? static_cast<size_t>(copies) * n : 0) + static_cast<size_t>(copies) * ngroups * nl) *
                sizeof(double);
}
}  // namespace rbx

static int launch_adjoint(rbx::AdjArgs &Aa, const rbx_launch &L, int natm, void *workspace, size_t workspace_bytes,
                          void *stream) {
    int nr, copies;
    size_t need;
    rbx::adjoint_plan(Aa.I.nl, Aa.I.nf, Aa.I.nray, natm, L, nr, copies, need);
    RBX_CHECK_ARG(need == 0 || (workspace && workspace_bytes >= need),
                  "rbx_integrate_adjoint: workspace of %zu bytes, this call needs %zu "
                  "(rbx_integrate_adjoint_workspace_size)",
                  workspace_bytes, need);
    RBX_CHECK_ARG(need == 0 || reinterpret_cast<uintptr_t>(workspace) % 8 == 0,
                  "rbx_integrate_adjoint: workspace not 8-byte aligned");
    const long long slots = (static_cast<long long>(Aa.I.nray) + static_cast<long long>(rbx::kWaves) * nr - 1) /
                            (static_cast<long long>(rbx::kWaves) * nr);
    const long long blocks = static_cast<long long>(Aa.I.ngroups) * slots;
    RBX_CHECK_ARG(blocks < (1LL << 31), "rbx_integrate_adjoint: too many blocks");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const size_t n = static_cast<size_t>(Aa.I.nl) * Aa.I.nf;
    const int parts_T = copies * Aa.I.ngroups;
    Aa.copies = copies;
    double *ws = static_cast<double *>(workspace);
    Aa.acc_alpha = copies > 1 ? ws : Aa.galpha;
    Aa.acc_T = need > 0 ? ws + (copies > 1 ? static_cast<size_t>(copies) * n : 0) : Aa.gT;
    const size_t na = (copies > 1 ? static_cast<size_t>(copies) : 1) * n;
    const size_t nt = static_cast<size_t>(parts_T) * Aa.I.nl;
    const size_t nclear = std::max(na, nt);
    hipLaunchKernelGGL(rbx::adjoint_clear_kernel, dim3(static_cast<unsigned>((nclear + 255) / 256)), dim3(256), 0, st,
                       Aa.acc_alpha, na, Aa.acc_T, nt);
    int rc = rbx::launch_status("adjoint_clear_kernel");
    if (rc) return rc;
    if (blocks > 0) {
        const dim3 grid(static_cast<unsigned>(blocks)), block(rbx::kWaves * rbx::kLanes);
#define RBX_ADJ_LAUNCH(NR)                                                                   \
    do {                                                                                     \
        if (Aa.I.fast_exp)                                                                   \
            hipLaunchKernelGGL((rbx::adjoint_kernel<NR, true>), grid, block, 0, st, Aa);     \
        else                                                                                 \
            hipLaunchKernelGGL((rbx::adjoint_kernel<NR, false>), grid, block, 0, st, Aa);    \
    } while (0)
        switch (nr) {
            case 1: RBX_ADJ_LAUNCH(1); break;
            case 2: RBX_ADJ_LAUNCH(2); break;
            case 4: RBX_ADJ_LAUNCH(4); break;
            default: RBX_ADJ_LAUNCH(8); break;
        }
#undef RBX_ADJ_LAUNCH
        rc = rbx::launch_status("adjoint_kernel");
        if (rc) return rc;
    }
    if (need == 0) return RBX_OK;   // one copy, one frequency group: the sums are galpha and gT
    const size_t nred = std::max(n, static_cast<size_t>(Aa.I.nl));
    hipLaunchKernelGGL(rbx::adjoint_reduce_kernel, dim3(static_cast<unsigned>((nred + 255) / 256)), dim3(256), 0, st,
                       Aa.acc_alpha, Aa.acc_T, copies, parts_T, n, Aa.I.nl, Aa.galpha, Aa.gT);
    return rbx::launch_status("adjoint_reduce_kernel");
}

extern "C" int rbx_integrate_adjoint_workspace_size(int nl, int nf, int nray, int natm, const rbx_launch *launch,
                                                    size_t *bytes) {
    RBX_CHECK_ARG(bytes, "rbx_integrate_adjoint_workspace_size: null output");
    RBX_CHECK_ARG(nl >= 2 && nf > 0 && nray >= 0 && natm >= 1, "rbx_integrate_adjoint_workspace_size: nl=%d nf=%d "
                  "nray=%d natm=%d", nl, nf, nray, natm);
    rbx_launch L;
    int rc = rbx::read_launch(launch, L, "rbx_integrate_adjoint_workspace_size");
    if (rc) return rc;
    int nr, copies;
    rbx::adjoint_plan(nl, nf, nray, natm, L, nr, copies, *bytes);
    return RBX_OK;
}

extern "C" int rbx_integrate_adjoint(const void *alpha, int alpha_precision, const double *T, int nl, int nf,
                                     const double *hds, int ld, const int *nsteps, int nray, const int *order,
                                     int precision, double tau_cut, const double *gTb, double *galpha, double *gT,
                                     const rbx_launch *launch, void *workspace, size_t workspace_bytes,
                                     void *stream) {
    int rc = rbx::check_int_args(alpha, alpha_precision, T, nl, nf, hds, ld, nsteps, nray, precision);
    if (rc) return rc;
    rbx_launch L;
    rc = rbx::read_launch(launch, L, "rbx_integrate_adjoint");
    if (rc) return rc;
    RBX_CHECK_ARG(ld >= nl - 1, "rbx_integrate_adjoint: ld=%d < nl-1=%d", ld, nl - 1);
    RBX_CHECK_ARG(gTb && galpha && gT, "rbx_integrate_adjoint: null gTb, galpha or gT");
    rbx::AdjArgs Aa;
    rbx::fill_int_args(Aa.I, alpha, alpha_precision, T, nl, nf, hds, ld, nsteps, nray, precision, tau_cut);
    Aa.I.order = order;
    Aa.gTb = gTb;
    Aa.galpha = galpha;
    Aa.gT = gT;
    return launch_adjoint(Aa, L, 1, workspace, workspace_bytes, stream);
}

extern "C" int rbx_integrate_adjoint_batch(const void *alpha, int alpha_precision, const double *T, int nl_total,
                                           int nf, const int *layer_off, int natm, const double *hds, int ld,
                                           const int *nsteps, const int *ray_atm, const int *order, int nray,
                                           int precision, double tau_cut, const double *gTb, double *galpha,
                                           double *gT, const rbx_launch *launch, void *workspace,
                                           size_t workspace_bytes, void *stream) {
    int rc = rbx::check_int_args(alpha, alpha_precision, T, nl_total, nf, hds, ld, nsteps, nray, precision);
    if (rc) return rc;
    rbx_launch L;
    rc = rbx::read_launch(launch, L, "rbx_integrate_adjoint_batch");
    if (rc) return rc;
    RBX_CHECK_ARG(layer_off && ray_atm && natm > 0, "rbx_integrate_adjoint_batch: null offsets / ray map or natm=%d",
                  natm);
    RBX_CHECK_ARG(ld >= 1, "rbx_integrate_adjoint_batch: ld=%d", ld);
    RBX_CHECK_ARG(gTb && galpha && gT, "rbx_integrate_adjoint_batch: null gTb, galpha or gT");
    rbx::AdjArgs Aa;
    rbx::fill_int_args(Aa.I, alpha, alpha_precision, T, nl_total, nf, hds, ld, nsteps, nray, precision, tau_cut);
    Aa.I.ray_atm = ray_atm;
    Aa.I.layer_off = layer_off;
    Aa.I.natm = natm;
    Aa.I.order = order;
    Aa.gTb = gTb;
    Aa.galpha = galpha;
    Aa.gT = gT;
    return launch_adjoint(Aa, L, natm, workspace, workspace_bytes, stream);
}

extern "C" int rbx_alpha_vjp(const double *dalpha, const double *galpha, const double *gT, int nl, int nf,
                             double *grad, void *stream) {
    RBX_CHECK_ARG(dalpha && galpha && grad, "rbx_alpha_vjp: null dalpha, galpha or grad");
    RBX_CHECK_ARG(nl > 0 && nf > 0, "rbx_alpha_vjp: nl=%d nf=%d", nl, nf);
    const int per = 256 / rbx::kLanes;
    hipLaunchKernelGGL(rbx::alpha_vjp_kernel, dim3(static_cast<unsigned>((nl + per - 1) / per)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), dalpha, galpha, gT, nl, nf, grad);
    return rbx::launch_status("alpha_vjp_kernel");
}
