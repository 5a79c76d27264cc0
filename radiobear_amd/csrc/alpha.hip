// Per-layer absorption alpha[layer][freq] (cm^-1) and its tangents: line sums on VALU.
//
// Every line shape of the Jupiter absorbers is written in one canonical form
// in x = f^2 (GHz^2):
//
//     term(x) = (a x + c) / ((x - D)^2 + g x + G),   alpha_c(f) = x * (sum term + K)
//
//   Ben-Reuven (nh3_hs inversion, h2s_ddb, ph3_jh), nh3_hs.py:198-255,
//     h2s_ddb.py:76-79, ph3_jh.py:80-87:
//       D = (f0+delta)^2 + gamma^2 - zeta^2, a = W (gamma-zeta), c = W (gamma+zeta) D,
//       g = 4 gamma^2, G = 0, W = line weight / f0^2 ((f/f0)^2 of the shape).
//   Gross (nh3_hs rotational and v2), nh3_hs.py:285-291, 322-328:
//       D = f0^2, a = 0, c = W gamma, g = 4 gamma^2, G = 0.
//   Van Vleck-Weisskopf with base (h2o_bk), h2o_bk.py:174-189, after putting the
//   two resonances over one denominator:
//       D = s^2 - df^2, G = 4 s^2 df^2, g = 0, a = 2 df W, c = a (s^2 + df^2),
//       and the per-line base subtraction -2 W base x becomes the layer term K.
//   Collision-induced H2 (h2_jj_ddb.py:7-38) and the H2O continuum
//   (h2o_bk.py:102-119) are alpha = K x.
//   nh3_sjs (nh3_sjs.py:93-122) has frequency-dependent widths, so its 200
//   lines are summed with per-(line, frequency) coefficients; it is only used
//   below 400 bar (nh3_hs_sjs.py:14-25).
//
// Forward path: the line-sum kernel's grid is (layer x group of four 64K-column
// wavefront tiles, part of the line list); a tile may straddle the 30 GHz
// switch of nh3_hs (its columns then read two coefficient sets).  Each 256-thread
// block stages 256 lines at a time in LDS:
// thread t derives line t's canonical coefficients for this layer (FP64, with
// exp/pow), then every thread sums all staged lines for its K frequencies,
// reading the coefficients as LDS broadcasts.  A finish kernel adds the parts
// and the per-layer closed forms.  The optional absorbers co_ddb, clouds_idp and
// h2_orton (extra.hpp) run in a small kernel of their own before the finish.
//
// FP64 mode sums in FP64 with IEEE division.  FP32 mode: the detuning x - D
// is formed in FP64 (cancellation near line centre) and rounded, the rest of
// the term is FP32 with the hardware reciprocal (two lines per reciprocal),
// partial sums of kSub lines are FP32 and are accumulated in FP64 (blocked
// summation).  Line weights are scaled per layer and block by a power of two
// so FP32 never sees the 1e-39..1e-19 raw intensities.
//
// The tangent kernel runs the same coefficient code on dual numbers
// (value, d/dT, d/dq_NH3, d/dq_H2O) and differentiates the canonical term in
// the line sum: d term = (dN - term dQ) / Q.
#include <algorithm>
#include <atomic>
#include <cstdint>
#include <mutex>
#include <type_traits>

#include "catalog.hpp"
#include "dual.hpp"

namespace rbx {

constexpr int kThreads = 256;
constexpr int kSub = 8;    // FP32 partial-sum length (lines unrolled per step)
constexpr int kMaxSplit = 16;  // most parts a layer's line sums are split into

// physical constants as the plug-ins spell them
constexpr double kHsHcKb = 19.858252418E-24 / 1.38E-23;    // nh3_hs.py:75-76
constexpr double kHsCoef = 1.0E6 * 6.02297E23 / 8.31432E7; // nh3_hs.py:81
constexpr double kGHzToInvCm = 1.0 / 29.9792458;           // nh3_hs.py:70
constexpr double kHck = 1.438396;                          // h2s_ddb.py:8
constexpr double kGeisa = 7.244E+21;                       // h2s_ddb.py:6
constexpr double kGHz = 29.9792458;                        // h2s_ddb.py:9
constexpr double kDbPerInvCm = 434294.5;
}  // namespace rbx

#include "extra.hpp"

namespace rbx {

// inversion-line coefficient sets: 0 nh3_hs <= 30 GHz, 1 nh3_hs and nh3_dbs > 30 GHz
// (nh3_hs.py:165-194), 2 nh3_dbs <= 30 GHz (the Bellotti/Steffes set, nh3_dbs.py:135-150):
// gnu_H2, gnu_He, gnu_NH3, GAMMA_H2, GAMMA_He, GAMMA_NH3,
// zeta_H2, zeta_He, zeta_NH3, Z_H2, Z_He, Z_NH3, d, Con
__constant__ double kHsInv[3][14] = {
    {1.640, 0.75, 0.852, 0.7756, 0.666, 1.0, 1.262, 0.3, 0.5296, 0.7964, 0.667, 1.554, -0.0498, 0.9301},
    {1.7465, 0.9779, 0.7298, 0.8202, 1.0, 1.0, 1.2163, 0.0291, 0.5152, 0.8873, 0.8994, 2.0 / 3.0,
     -0.0627, 0.9862},
    {1.6937, 0.6997, 0.7523, 0.8085, 1.0, 1.0, 1.3263, 0.1607, 0.6162, 0.8199, 0.0, 1.3832, -0.0139, 0.9619}};

// rbx_absorbers.nh3: 1 nh3_hs_sjs, 2 nh3_hs, 3 nh3_sjs, 4 nh3_dbs_sjs, 5 nh3_dbs.  nh3_dbs is
// nh3_hs with coefficient set 2 below 30 GHz (nh3_dbs.py:92-313); nh3_dbs_sjs blends it with
// nh3_sjs exactly as nh3_hs_sjs does (nh3_dbs_sjs.py:6-26).
__host__ __device__ inline bool nh3_has_hs(int sel) { return sel == 1 || sel == 2 || sel == 4 || sel == 5; }
__host__ __device__ inline bool nh3_has_sjs(int sel) { return sel == 1 || sel == 3 || sel == 4; }
__host__ __device__ inline bool nh3_blended(int sel) { return sel == 1 || sel == 4; }
__host__ __device__ inline int nh3_low_set(int sel) { return sel >= 4 ? 2 : 0; }

// one layer: T and partial pressures (bar); S = double or Dual
template <class S>
struct LayerT {
    double P;
    S T, h2, he, ch4, nh3, h2o, h2s, ph3;
};

template <class S>
struct CoefT {
    S D, a, c, g, G;
};

constexpr int kFarStreams = 7;   // near-mask streams: h2o, h2s, inversion set 0, set 1, rot, v2, ph3

struct AlphaArgs {
    DevCatalog cat;
    rbx_absorbers abs;
    const double *P, *T, *mix;
    const double *cloud;   // [RBX_NUM_CLOUDS][nl] or null (clouds_idp)
    int nl;
    const double *freq, *freq_nh3;
    int nf, nf_low;
    const double *scale;
    void *alpha;       // forward: alpha [nl][nf]; tangent: dalpha [3][nl][nf] (double)
    double *by_c;
    int tiles_total;   // forward kernel: wavefront tiles of tile_width columns
    int tile_width;    // columns per tile
    double *value;     // tangent: alpha [nl][nf] (float64) or null
    double *ws;        // forward: partial line sums [kSlots][nsplit][nl][nf]
    double *ws_base;   // forward: partial h2o base terms [nsplit][nl]
    double *ws_extra;  // forward: optional absorbers [kNumExtra][nl][nf] (when any is selected)
    // far-line series (far.hpp): near-line masks [nl][tiles_total][far_words], moments
    // [nl][far_split][tiles_total][kFarMom], h2o base terms [far_split][nl]
    unsigned int *near;
    double *far_mom, *far_base;
    double *far_prep;   // the segments' coefficient derivers per layer [nl][kSegs][kPrepWords]
    int far_words, far_split;
    int far_off[kFarStreams];   // first mask word of each stream
    double *tan_prep;   // tangent pass: the segments' dual-number derivers per layer
    // tangent pass split into tan_parts parts of a layer's tan_rounds 64-line rounds;
    // tan_seg0[s]: the first round of segment s (h2o, h2s, nh3 inversion, rot, v2, ph3)
    int tan_parts, tan_rounds;
    int tan_seg0[6];
};

// One line's catalogue columns held in registers, read by the coefficient sets below the
// way they read a DevBlock (B.col[c][i] gives column c of this line whatever i is).
template <int NC>
struct LineRow {
    struct Col {
        double v;
        __device__ double operator[](int) const { return v; }
    } col[NC];
    __device__ void load(const DevBlock &B, int i) {
#pragma unroll
        for (int c = 0; c < NC; ++c) col[c].v = B.col[c][i];
    }
};

// ---- per-layer coefficient sets: each derives (D, a, c, g, G) of one line ----
template <class S>
struct InvPrep {   // nh3_hs inversion lines (nh3_hs.py:198-255)
    S Ag, Bg, Az, Bz, eT, pref;
    double d;
    __device__ InvPrep(const LayerT<S> &L, int set) {
        const double *k = kHsInv[set];
        S Tdiv = 300.0 / L.T, t295 = 295.0 / L.T;
        Ag = (k[0] * L.h2) * pow(Tdiv, k[3]) + (k[1] * L.he) * pow(Tdiv, k[4]);
        Bg = k[2] * L.nh3 * pow(t295, k[5]);
        Az = (k[6] * L.h2) * pow(Tdiv, k[9]) + (k[7] * L.he) * pow(Tdiv, k[10]);
        Bz = k[8] * L.nh3 * pow(t295, k[11]);
        d = k[12];
        eT = -(1.0 / L.T - 1.0 / 300.0) * kHsHcKb;
        pref = k[13] * kHsCoef * (L.nh3 / 300.0) * pow(300.0 / L.T, 3.5) * ((2.0 / kPi) * (1.0 / kGHzToInvCm));
    }
    // the two coefficient sets share eT, so a line's weight before pref is common to both
    static constexpr bool kShared = true;
    template <class Blk>
    __device__ S weight(const Blk &B, int i) const {
        const double fo = B.col[0][i];
        return exp(eT * B.col[2][i]) * (B.col[1][i] / (fo * fo));
    }
    template <class Blk>
    __device__ CoefT<S> coef(const Blk &B, int i, const S &w) const {
        const double fo = B.col[0][i], go = B.col[3][i];
        S gam = Ag + Bg * go;
        S zet = Az + Bz * go;
        S fd = fo + d * gam;
        S W = pref * w;
        S Dd = fd * fd + gam * gam - zet * zet;
        return {Dd, W * (gam - zet), W * (gam + zet) * Dd, 4.0 * gam * gam, S(0.0)};
    }
    template <class Blk>
    __device__ CoefT<S> operator()(const Blk &B, int i, S &) const { return coef(B, i, weight(B, i)); }
};

template <class S>
struct RotPrep {   // nh3_hs rotational lines, Gross (nh3_hs.py:257-291)
    static constexpr bool kShared = false;
    S c1, c2, c3, eT, pref;
    __device__ explicit RotPrep(const LayerT<S> &L) {
        S Tdiv = 300.0 / L.T;
        c1 = 0.2984 * L.h2 * pow(Tdiv, 0.8730);
        c2 = 0.75 * L.he * pow(Tdiv, 2.0 / 3.0);
        c3 = 3.1789 * L.nh3 * Tdiv;
        eT = (1.0 / 300.0 - 1.0 / L.T) * kHsHcKb;
        pref = 2.4268 * kHsCoef * (L.nh3 / 300.0) * pow(300.0 / L.T, 3.5) * ((1.0 / kGHzToInvCm) * (4.0 / kPi));
    }
    template <class Blk>
    __device__ CoefT<S> operator()(const Blk &B, int i, S &) const {
        const double fo = B.col[0][i];
        S gam = c1 * B.col[4][i] + c2 * B.col[5][i] + c3 * B.col[3][i];
        S W = pref * B.col[1][i] * exp(eT * B.col[2][i]);
        return {S(fo * fo), S(0.0), W * gam, 4.0 * gam * gam, S(0.0)};
    }
};

template <class S>
struct V2Prep {    // nh3_hs v2 roto-vibrational lines, Gross (nh3_hs.py:293-328)
    static constexpr bool kShared = false;
    S gam, eT, pref;
    __device__ explicit V2Prep(const LayerT<S> &L) {
        S Tdiv = 300.0 / L.T;
        gam = L.h2 * 1.4 * pow(Tdiv, 0.73) + L.he * 0.68 * pow(Tdiv, 0.5716) + L.nh3 * 9.5 * Tdiv;
        eT = (1.0 / 300.0 - 1.0 / L.T) * kHsHcKb;
        pref = 1.1206 * kHsCoef * (L.nh3 / 300.0) * pow(300.0 / L.T, 3.5) * ((1.0 / kGHzToInvCm) * (4.0 / kPi));
    }
    template <class Blk>
    __device__ CoefT<S> operator()(const Blk &B, int i, S &) const {
        const double fo = B.col[0][i];
        S W = pref * B.col[1][i] * exp(eT * B.col[2][i]);
        return {S(fo * fo), S(0.0), W * gam, 4.0 * gam * gam, S(0.0)};
    }
};

template <class S>
struct H2sPrep {   // h2s_ddb, Ben-Reuven with zeta = gamma (h2s_ddb.py:42-87)
    static constexpr bool kShared = false;
    S Ag, Bg, del, eT, pref;
    __device__ explicit H2sPrep(const LayerT<S> &L) {
        const double T0 = 296.0;
        S tg = pow(T0 / L.T, 0.7);
        Ag = tg * (1.960 * L.h2 + 1.200 * L.he);
        Bg = tg * L.h2s;
        del = 1.28 * L.h2s;
        eT = -((1.0 / L.T) - (1.0 / T0)) * kHck;
        pref = kGeisa * (L.h2s / T0) * pow(T0 / L.T, 3.5) * (kGHz * 2.0 / kPi);
    }
    template <class Blk>
    __device__ CoefT<S> operator()(const Blk &B, int i, S &) const {
        const double f0 = B.col[0][i];
        S gam = Ag + Bg * B.col[3][i];
        S W = pref * B.col[1][i] * exp(eT * B.col[2][i]) / (f0 * f0);
        S fd = f0 + del;
        S D = fd * fd;
        return {D, S(0.0), W * 2.0 * gam * D, 4.0 * gam * gam, S(0.0)};
    }
};

template <class S>
struct Ph3Prep {   // ph3_jh, Ben-Reuven with zeta = delta = 0 (ph3_jh.py:64-108)
    static constexpr bool kShared = false;
    S A, Bs, eT, pref;
    __device__ explicit Ph3Prep(const LayerT<S> &L) {
        const double T0 = 300.0;
        A = pow(T0 / L.T, 2.0 / 3.0) * (3.2930 * L.h2 + 1.6803 * L.he);
        Bs = (T0 / L.T) * 4.2157 * L.ph3;
        eT = -((1.0 / L.T) - (1.0 / T0)) * kHck;
        pref = kGeisa * (L.ph3 / T0) * pow(T0 / L.T, 3.5) * (kGHz * 2.0 / kPi);
    }
    template <class Blk>
    __device__ CoefT<S> operator()(const Blk &B, int i, S &) const {
        const double f0 = B.col[0][i];
        S gam = A * B.col[4][i] + Bs * B.col[5][i];
        S W = pref * (B.col[1][i] * B.col[3][i]) * exp(eT * B.col[2][i]) / (f0 * f0);
        S D = f0 * f0 + gam * gam;
        return {D, W * gam, W * gam * D, 4.0 * gam * gam, S(0.0)};
    }
};

template <class S>
struct H2oPrep {   // h2o_bk lines, VVW with base (h2o_bk.py:83-99, 130-189)
    static constexpr bool kShared = false;
    S lnTh, Th25, Th, h2o, h2, he, pref;
    __device__ explicit H2oPrep(const LayerT<S> &L) {
        const double M = 8.314472 / 0.46151805;
        Th = 300.0 / L.T;
        lnTh = log(Th);
        Th25 = pow(Th, 2.5);
        S density = (M * L.h2o) / (8.314472e-5 * L.T);
        density = 0.997317 * (density / M) * 6.0221415e23 * (1.0 / 1e6);
        pref = density * (4.342945 * 1e-4 / kDbPerInvCm / kPi);
        h2o = L.h2o;
        h2 = L.h2;
        he = L.he;
    }
    // also adds the line's base contribution (-2 W base) to `base`
    template <class Blk>
    __device__ CoefT<S> operator()(const Blk &B, int i, S &base) const {
        const double fo = B.col[0][i];
        S Sl = B.col[1][i] * Th25 * exp(B.col[2][i] * (1.0 - Th));
        S df = B.col[3][i] * h2o * exp(B.col[4][i] * lnTh) + B.col[5][i] * h2 * exp(B.col[6][i] * lnTh) +
               B.col[7][i] * he * exp(B.col[8][i] * lnTh);
        const double s = fo + B.col[9][i];
        S W = pref * Sl / (fo * fo);
        S a = 2.0 * df * W;
        const double s2 = s * s;
        S df2 = df * df;
        base = base + (-2.0) * W * (df / (562500.0 + df2));
        return {s2 - df2, a, a * (s2 + df2), S(0.0), 4.0 * s2 * df2};
    }
};

// ---- LDS staging: up to two coefficient sets per line (the nh3_hs <= / > 30 GHz sets) ----
struct Staged64 {
    double D[2][kThreads], a[2][kThreads], c[2][kThreads], g[2][kThreads], G[2][kThreads];
};
// FP32 mode stages the denominator as (x - Ds)^2 + h with Ds = D - g/2 and
// h = g (D - g/4) + G, the same polynomial in x as (x - D)^2 + g x + G; it is
// evaluated in FP64 (h < 0 for lines wider than their centre, where the two
// terms cancel) and rounded once, so each line costs one FMA less.
// (set 1 starts 4 entries later than a multiple of the bank count, so a wavefront
// whose columns read both sets does not conflict)
struct Staged32 {
    alignas(16) double D[2][kThreads + 4];   // Ds
    alignas(16) double h[2][kThreads + 4];
    alignas(16) float2 ac[2][kThreads + 4];  // a, c (scaled)
    alignas(16) float2 dh[2][kThreads + 4];  // Ds, h rounded to FP32 (the FP32 fast path)
    // per set and column wavefront: bit j = staged lines [j kSub, (j + 1) kSub) are all
    // FP32-safe for that wavefront's columns (byte w written by staging wavefront w)
    unsigned int fast[2][kThreads / 64];
};

// Column range [lo, hi] of x = f^2 (var 0) and of x for NH3 (var 1) per wavefront of a block.
struct WaveRanges {
    double lo[2][kThreads / 64], hi[2][kThreads / 64];
};

// FP32 detuning is safe for a line against columns x in [lo, hi] when its shifted centre
// Ds lies below 0.6 lo or above hi / 0.6: then |x - Ds| >= (x + |Ds|) / 4 at every column,
// so (float)x - (float)Ds carries at most about five FP32 roundings relative to x - Ds.
// A line with h < 0 also needs (x - Ds)^2 >= 4 |h| there, so that (x - Ds)^2 + h does not
// cancel.  An empty range (lo = +inf) is safe; NaN coefficients are not.
__device__ inline bool fp32_safe(double Ds, double h, double lo, double hi) {
    double dmin;
    if (Ds <= 0.6 * lo)
        dmin = lo - Ds;
    else if (0.6 * Ds >= hi)
        dmin = Ds - hi;
    else
        return false;
    return h >= 0.0 || dmin * dmin >= -4.0 * h;
}

// bit j set when bits [8 j, 8 j + 8) of m are all set (kSub = 8 lines per chunk)
__device__ inline unsigned int full_chunks(unsigned long long m) {
    unsigned int b = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) b |= ((m >> (8 * j)) & 0xFFull) == 0xFFull ? 1u << j : 0u;
    return b;
}

template <class S>
__device__ inline void shifted_denominator(const CoefT<S> &q, S &Ds, S &h) {
    Ds = q.D - 0.5 * q.g;
    h = q.g * (q.D - 0.25 * q.g) + q.G;
}
template <int N>
struct SjsLines {
    double f0[N], w[N], g0[N];
    double dw[kNd][N];
};
using StagedSjs = SjsLines<kThreads>;

union SharedFwd {
    Staged64 s64;
    Staged32 s32;
    StagedSjs sjs;
    double red[kThreads];
};

// coefficient sets `need` of line i (the inversion sets share the line weight)
template <class Prep, class S>
__device__ __forceinline__ void line_coefs(const Prep *prep, int need, const DevBlock &B, int i, S &base,
                                           CoefT<S> (&q)[2]) {
    if constexpr (Prep::kShared) {
        const S w = prep[0].weight(B, i);
        if (need & 1) q[0] = prep[0].coef(B, i, w);
        if (need & 2) q[1] = prep[1].coef(B, i, w);
    } else {
        if (need & 1) q[0] = prep[0](B, i, base);
    }
}

// ---- forward line sums of the staged lines for K frequencies per lane ----
// set[k]: the coefficient set column k reads (the nh3_hs inversion lines have a
// <= 30 GHz and a > 30 GHz set; every other block has one).  UNI: every column of
// the wavefront reads the same set, so one LDS read serves all K columns.
template <int K, bool UNI, class ST = Staged32>
__device__ __forceinline__ void sum_staged_f32(const ST &st, const int (&set)[K], int jb, int je,
                                               const double (&x)[K], const float (&xf)[K], double (&acc)[K]) {
    // chunks of kSub lines that are FP32-safe for this wavefront's columns take the all-FP32
    // path (the same sums with (float)x - (float)Ds and d^2 + h in FP32); the mask is
    // wavefront-uniform, so the choice is a scalar branch per chunk
    const unsigned int fast =
        UNI ? __builtin_amdgcn_readfirstlane(st.fast[set[0]][threadIdx.x / 64]) : 0u;
    for (int j0 = jb; j0 < je; j0 += kSub) {
        float p[K];
#pragma unroll
        for (int k = 0; k < K; ++k) p[k] = 0.0f;
        if (UNI && ((fast >> (j0 / kSub)) & 1u)) {
#pragma unroll
            for (int u = 0; u < kSub; u += 2) {
                const float4 dp = reinterpret_cast<const float4 *>(st.dh[set[0]])[(j0 + u) / 2];
                const float4 qp = reinterpret_cast<const float4 *>(st.ac[set[0]])[(j0 + u) / 2];
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    const float d1 = xf[k] - dp.x, d2 = xf[k] - dp.z;
                    const float e1 = fmaf(d1, d1, dp.y), e2 = fmaf(d2, d2, dp.w);
                    const float n1 = fmaf(qp.x, xf[k], qp.y);
                    const float n2 = fmaf(qp.z, xf[k], qp.w);
                    const float num = fmaf(n2, e1, n1 * e2);
                    p[k] = fmaf(num, __builtin_amdgcn_rcpf(e1 * e2), p[k]);
                }
            }
        } else {
            // two lines over one denominator: n1/e1 + n2/e2 = (n1 e2 + n2 e1)/(e1 e2),
            // one reciprocal per pair (denominators stay within 1e-17..1e27 here).  The
            // mixed-set path (one wavefront per layer at most) reads K times as much
            // and is not unrolled, to stay within the uniform path's registers.
            constexpr int kPairUnroll = UNI ? kSub / 2 : 1;
#pragma unroll kPairUnroll
            for (int u = 0; u < kSub; u += 2) {
                double2 Dp[K], hp[K];
                float4 qp[K];
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    if (UNI && k > 0) {
                        Dp[k] = Dp[0];
                        hp[k] = hp[0];
                        qp[k] = qp[0];
                    } else {   // pairs of lines as 16-byte LDS reads
                        Dp[k] = reinterpret_cast<const double2 *>(st.D[set[k]])[(j0 + u) / 2];
                        hp[k] = reinterpret_cast<const double2 *>(st.h[set[k]])[(j0 + u) / 2];
                        qp[k] = reinterpret_cast<const float4 *>(st.ac[set[k]])[(j0 + u) / 2];
                    }
                }
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    // (x - Ds)^2 + h in FP64: h < 0 when a Ben-Reuven line's coupling exceeds its
                    // centre frequency (zeta > f0 + delta), and the sum then cancels
                    const double d1 = x[k] - Dp[k].x, d2 = x[k] - Dp[k].y;
                    const float e1 = static_cast<float>(fma(d1, d1, hp[k].x));
                    const float e2 = static_cast<float>(fma(d2, d2, hp[k].y));
                    const float n1 = fmaf(qp[k].x, xf[k], qp[k].y);
                    const float n2 = fmaf(qp[k].z, xf[k], qp[k].w);
                    const float num = fmaf(n2, e1, n1 * e2);
                    p[k] = fmaf(num, __builtin_amdgcn_rcpf(e1 * e2), p[k]);
                }
            }
        }
#pragma unroll
        for (int k = 0; k < K; ++k) acc[k] += static_cast<double>(p[k]);
    }
}

// 1 / x to within an ulp or two: the hardware estimate refined by two Newton steps
__device__ inline double recip(double x) {
    double r = __builtin_amdgcn_rcp(x);
    r = fma(r, fma(-x, r, 1.0), r);
    return fma(r, fma(-x, r, 1.0), r);
}

// FP64 sums: two lines over one denominator, n1/e1 + n2/e2 = (n1 e2 + n2 e1) / (e1 e2), and
// the reciprocal as the hardware estimate refined twice (a few ulps, where the IEEE division
// took twice the instructions); the denominators stay within 1e-17..1e27, so their products
// are far from FP64's range limits.  An odd last line pairs with a zero term.
template <int K, bool UNI, class ST = Staged64>
__device__ __forceinline__ void sum_staged_f64(const ST &st, const int (&set)[K], int jb, int je,
                                               const double (&x)[K], double (&acc)[K]) {
    for (int j = jb; j < je; j += 2) {
        const bool two = j + 1 < je;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int m = UNI ? set[0] : set[k];
            const double d1 = x[k] - st.D[m][j];
            const double n1 = fma(st.a[m][j], x[k], st.c[m][j]);
            const double e1 = fma(d1, d1, fma(st.g[m][j], x[k], st.G[m][j]));
            double n2 = 0.0, e2 = 1.0;
            if (two) {
                const double d2 = x[k] - st.D[m][j + 1];
                n2 = fma(st.a[m][j + 1], x[k], st.c[m][j + 1]);
                e2 = fma(d2, d2, fma(st.g[m][j + 1], x[k], st.G[m][j + 1]));
            }
            acc[k] += fma(n1, e2, n2 * e1) * recip(e1 * e2);
        }
    }
}

// ---- forward line sum over lines [lb, le) of one block for K frequencies per lane ----
// prep[0..1]: coefficient sets (two only for the nh3_hs inversion lines, whose
// <= / > 30 GHz constants differ); `need` is the block's bitmask of sets to
// stage; `active`: the wavefront has columns (the others only help staging).
// Staged lines are padded to a multiple of kSub with zero terms so the sum
// loop has a fixed trip count and unrolls.
template <int K, bool F32, class Prep>
__device__ void line_sum(const DevBlock &B, const Prep *prep, int need, bool active, const int (&set)[K],
                         const double *scale, const double (&x)[K], double (&acc)[K], SharedFwd &sh, double &base,
                         int lb, int le, const WaveRanges &rng, int var) {
    float xf[K];
#pragma unroll
    for (int k = 0; k < K; ++k) xf[k] = static_cast<float>(x[k]);
    bool uni = true;
#pragma unroll
    for (int k = 0; k < K; ++k) uni = uni && set[k] == set[0];
    uni = __all(uni && set[0] == __builtin_amdgcn_readfirstlane(set[0]));
    for (int l0 = lb; l0 < le; l0 += kThreads) {
        const int i = l0 + threadIdx.x;
        const int cnt = min(kThreads, le - l0);
        const int cntp = (cnt + kSub - 1) / kSub * kSub;
        unsigned int safe[2] = {0u, 0u};   // FP32 mode: bit w = safe for wavefront w's columns
        if (static_cast<int>(threadIdx.x) < cntp) {
            CoefT<double> qs[2] = {{0.0, 0.0, 0.0, 0.0, 1.0}, {0.0, 0.0, 0.0, 0.0, 1.0}};   // padding: 0
            if (i < le) line_coefs(prep, need, B, i, base, qs);
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                if (!(need & (1 << s))) continue;
                const CoefT<double> &q = qs[s];
                if constexpr (F32) {
                    double Ds, h;
                    shifted_denominator(q, Ds, h);
                    sh.s32.D[s][threadIdx.x] = Ds;
                    sh.s32.h[s][threadIdx.x] = h;
                    sh.s32.ac[s][threadIdx.x] =
                        make_float2(static_cast<float>(q.a * scale[s]), static_cast<float>(q.c * scale[s]));
                    sh.s32.dh[s][threadIdx.x] = make_float2(static_cast<float>(Ds), static_cast<float>(h));
#pragma unroll
                    for (int w = 0; w < kThreads / 64; ++w)
                        safe[s] |= fp32_safe(Ds, h, rng.lo[var][w], rng.hi[var][w]) ? 1u << w : 0u;
                } else {
                    sh.s64.D[s][threadIdx.x] = q.D;
                    sh.s64.a[s][threadIdx.x] = q.a;
                    sh.s64.c[s][threadIdx.x] = q.c;
                    sh.s64.g[s][threadIdx.x] = q.g;
                    sh.s64.G[s][threadIdx.x] = q.G;
                }
            }
        }
        if constexpr (F32) {
            // chunk masks: staging wavefront sw holds lines [64 sw, 64 sw + 64) = chunks
            // [8 sw, 8 sw + 8); lanes past cntp report unsafe (those chunks are not read)
            const int sw = threadIdx.x / 64;
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                if (!(need & (1 << s))) continue;
#pragma unroll
                for (int w = 0; w < kThreads / 64; ++w) {
                    const unsigned int b = full_chunks(__ballot((safe[s] >> w) & 1u));
                    if (threadIdx.x % 64 == 0) reinterpret_cast<unsigned char *>(&sh.s32.fast[s][w])[sw] = b;
                }
            }
        }
        __syncthreads();
        if (active) {
            if constexpr (F32) {
                if (uni)
                    sum_staged_f32<K, true>(sh.s32, set, 0, cntp, x, xf, acc);
                else
                    sum_staged_f32<K, false>(sh.s32, set, 0, cntp, x, xf, acc);
            } else {
                if (uni)
                    sum_staged_f64<K, true>(sh.s64, set, 0, cntp, x, acc);
                else
                    sum_staged_f64<K, false>(sh.s64, set, 0, cntp, x, acc);
            }
        }
        __syncthreads();
    }
}

// power of two s with |w s| ~ 1 (FP32 range for the line weights)
__device__ inline double pow2_scale(double w) {
    if (!(w > 0.0) || !isfinite(w)) return 1.0;
    return ldexp(1.0, -ilogb(w));
}

__device__ inline double block_reduce_sum(double v, double *red) {
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}


// nh3_sjs Spilker coefficient set (nh3_sjs.py:52-77): GH2, GHe, GNH3, ZH2, ZHe, ZNH3, D
template <class S>
__device__ inline void sjs_spilker(const S &T, const S &Ph2, S *c) {
    const double EPS = 1E-12;
    S rexp = 8.79 * exp(-T / 83.0);
    S GH2a = exp(9.024 - T / 20.3) - 0.9918 + Ph2;
    GH2a = pow(GH2a, rexp);  // NaN for a negative base, as numpy
    if (val(GH2a) < EPS) {
        c[0] = 1.690; c[1] = 0.750; c[2] = 0.60; c[3] = 1.35; c[4] = 0.30; c[5] = 0.20; c[6] = -0.45;
    } else {
        GH2a = 2.122 * exp(-T / 116.8) / GH2a;
        GH2a = 2.34 * (1.0 - GH2a);
        c[0] = GH2a;
        c[1] = 0.46 + T / 3000.0;
        c[2] = 0.74;
        c[3] = 5.7465 - 7.7644 * GH2a + 9.1931 * GH2a * GH2a - 5.6816 * GH2a * GH2a * GH2a +
               1.2307 * GH2a * GH2a * GH2a * GH2a;
        c[4] = 0.28 - T / 1750.0;
        c[5] = 0.50;
        c[6] = -0.45;
    }
}

// nh3_sjs for this thread's K frequencies (FP64; S = double or Dual).  WAVE:
// each wavefront stages and sums its own 64 lines at a time in its own staging
// buffer `sg` (wavefronts of a block may be on different layers); otherwise the
// whole block stages kThreads lines per round in `sg`.
template <int K, bool WAVE = false, class S, class Stage>
__device__ void nh3_sjs(const DevBlock &B, const LayerT<S> &L, const double (&f)[K], S (&out)[K], Stage &sg) {
    const double T0 = 296.0;
    const double J[7] = {1.690, 0.750, 0.6, 1.350, 0.300, 0.200, -0.45};
    S Sp[7];
    sjs_spilker(L.T, L.h2, Sp);
    const S tf = pow(T0 / L.T, 2.0 / 3.0);
    const S eT = -((1.0 / L.T) - (1.0 / T0)) * kHck;
    S Ag[K], Bg[K], Az[K], Bz[K], del[K], sum[K];
    double x[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        S c[7];
        if (f[k] <= 26.0) {
            for (int m = 0; m < 7; ++m) c[m] = Sp[m];
        } else if (f[k] >= 34.0) {
            for (int m = 0; m < 7; ++m) c[m] = J[m];
        } else {
            const double flfh = (26.0 - 34.0) / (f[k] - 26.0);
            for (int m = 0; m < 7; ++m) c[m] = Sp[m] + (Sp[m] - J[m]) / flfh;
        }
        Ag[k] = tf * (c[0] * L.h2 + c[1] * L.he);
        Bg[k] = tf * c[2] * L.nh3;
        Az[k] = tf * (c[3] * L.h2 + c[4] * L.he);
        Bz[k] = tf * c[5] * L.nh3;
        del[k] = c[6] * L.nh3;
        x[k] = f[k] * f[k];
        sum[k] = 0.0;
    }
    constexpr int chunk = WAVE ? 64 : kThreads;
    const int me = WAVE ? threadIdx.x % 64 : threadIdx.x;
    for (int l0 = 0; l0 < B.n; l0 += chunk) {
        const int i = l0 + me;
        if (i < B.n) {
            const double f0 = B.col[0][i];
            const S w = B.col[1][i] * exp(eT * B.col[2][i]) / (f0 * f0);
            sg.f0[me] = f0;
            sg.w[me] = val(w);
            if constexpr (std::is_same<S, Dual>::value) {
                for (int m = 0; m < kNd; ++m) sg.dw[m][me] = w.d[m];
            }
            sg.g0[me] = B.col[3][i];
        }
        if (WAVE)
            wave_sync();
        else
            __syncthreads();
        const int cnt = min(chunk, B.n - l0);
        for (int j = 0; j < cnt; ++j) {
            const double f0 = sg.f0[j], g0 = sg.g0[j];
            S w;
            if constexpr (std::is_same<S, Dual>::value)
                w = Dual(sg.w[j], sg.dw[0][j], sg.dw[1][j], sg.dw[2][j]);
            else
                w = sg.w[j];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                S gam = Ag[k] + Bg[k] * g0;
                S zet = Az[k] + Bz[k] * g0;
                S fd = f0 + del[k];
                S Dd = fd * fd + gam * gam - zet * zet;
                S num = (gam - zet) * x[k] + (gam + zet) * Dd;
                S d = x[k] - Dd;
                S den = d * d + 4.0 * x[k] * gam * gam;
                sum[k] = sum[k] + w * num / den;
            }
        }
        if (WAVE)
            wave_sync();
        else
            __syncthreads();
    }
    const S pref = kGeisa * (L.nh3 / T0) * pow(T0 / L.T, 3.5) * (kGHz * 2.0 / kPi) * (1.0 + L.P / 1.0E5);
#pragma unroll
    for (int k = 0; k < K; ++k) out[k] = pref * x[k] * sum[k];
}

// h2_jj_ddb layer constant: alpha_h2 = K x (h2_jj_ddb.py:7-38)
template <class S>
__device__ inline S h2_cia(const LayerT<S> &L, int state) {
    const S th = 273.0 / L.T;
    S pre;
    if (state == 1) {
        pre = fmin1(pow(L.T / 55.0, 2.7));
        pre = fmin1(pre * pow(L.T / 120.0, 0.55));
    } else {
        pre = fmin1(pow(L.T / 40.0, 2.5));
    }
    return 3.9522E-14 * L.h2 * pre *
           (L.h2 * pow(th, 3.12) + 1.382 * L.he * pow(th, 2.24) + 9.322 * L.ch4 * pow(th, 3.34));
}

// h2o_bk continuum: alpha = K x (h2o_bk.py:102-125)
template <class S>
__device__ inline S h2o_continuum(const LayerT<S> &L) {
    const S Th = 300.0 / L.T;
    const double Cf_he = ((1.0 / 0.001) * (1.0 / 0.001)) * 1.03562010226e-10;
    const double Cf_h2 = ((1.0 / 0.001) * (1.0 / 0.001)) * 5.07722009423e-11;
    const S Cs1 = 3.1e-07 * pow(Th, 12.0);
    const S Th3 = Th * Th * Th;
    const S w = L.h2o / 0.001;
    return (4.342945 * (Cf_he * L.he * L.h2o * Th3 + Cf_h2 * L.h2 * L.h2o * Th3) + 4.342945 * Cs1 * (w * w)) /
           kDbPerInvCm;
}

__device__ inline LayerT<double> layer_of(const AlphaArgs &A, int layer) {
    LayerT<double> L;
    L.P = A.P[layer];
    L.T = A.T[layer];
    const double *mx = A.mix;
    const int nl = A.nl;
    L.h2 = L.P * mx[RBX_X_H2 * nl + layer];
    L.he = L.P * mx[RBX_X_HE * nl + layer];
    L.ch4 = L.P * mx[RBX_X_CH4 * nl + layer];
    L.nh3 = L.P * mx[RBX_X_NH3 * nl + layer];
    L.h2o = L.P * mx[RBX_X_H2O * nl + layer];
    L.h2s = L.P * mx[RBX_X_H2S * nl + layer];
    L.ph3 = L.P * mx[RBX_X_PH3 * nl + layer];
    return L;
}

// nh3_hs_sjs / nh3_dbs_sjs pressure blend (nh3_hs_sjs.py:6-26, nh3_dbs_sjs.py:6-26)
template <class S>
__device__ inline S nh3_select(int sel, double P, const S &hs, const S &sj) {
    if (!nh3_blended(sel)) return nh3_has_hs(sel) ? hs : sj;
    if (P < 400.0) return hs;
    if (P > 2000.0) return sj;
    const double W = (P - 400.0) / (2000.0 - 400.0);
    return W * sj + (1.0 - W) * hs;
}

// ---- forward path: split line sums, then a per-column finish ----
//
// A layer's lines are visited as one sequence of staging rounds (kThreads lines
// each) over the segments h2o, h2s, nh3 inversion, nh3 rotational, nh3 v2, ph3.
// alpha_lines_kernel block (tile group, layer, part) sums the rounds
// [part R / nsplit, (part + 1) R / nsplit) of that sequence into four partial
// sums per column (h2o, h2s, nh3_hs, ph3 slots, each already divided by its
// FP32 scale) and writes them to the workspace; splitting the lines lets the
// grid be cut into enough equal blocks to keep every CU busy to the end.
// alpha_finish_kernel then adds the parts in order, applies the closed forms
// (H2 CIA, H2O continuum), the nh3_hs negative clamp, nh3_sjs and the
// pressure blend, and scale-sums the constituents (alpha.py:151-192).
constexpr int kSlots = 4;
enum { kSlotH2o = 0, kSlotH2s = 1, kSlotNh3 = 2, kSlotPh3 = 3 };

__device__ __host__ inline int rounds_of(int n) { return (n + kThreads - 1) / kThreads; }

// which NH3 parts a layer needs (nh3_hs_sjs.py:14-25)
__device__ inline bool nh3_uses_hs(int sel, double P) { return nh3_has_hs(sel) && (!nh3_blended(sel) || P <= 2000.0); }
__device__ inline bool nh3_uses_sjs(int sel, double P) { return nh3_has_sjs(sel) && (!nh3_blended(sel) || P >= 400.0); }

// A layer's staging rounds per segment (h2o, h2s, nh3 inversion, rotational, v2,
// ph3), the slot each segment sums into, and the round range of each part.
constexpr int kSegs = 6;
constexpr int kSegSlot[kSegs] = {kSlotH2o, kSlotH2s, kSlotNh3, kSlotNh3, kSlotNh3, kSlotPh3};
constexpr int kSegKind[kSegs] = {RBX_LINES_H2O, RBX_LINES_H2S, RBX_LINES_NH3_INV,
                                 RBX_LINES_NH3_ROT, RBX_LINES_NH3_V2, RBX_LINES_PH3};

struct LinePlan {
    int nr[kSegs];
    int R;
    __device__ LinePlan(const AlphaArgs &A, double P) {
        const bool hs = A.abs.nh3 && nh3_uses_hs(A.abs.nh3, P);
        const bool on[kSegs] = {A.abs.h2o != 0, A.abs.h2s != 0, hs, hs, hs, A.abs.ph3 != 0};
        R = 0;
        for (int s = 0; s < kSegs; ++s) {
            nr[s] = on[s] ? rounds_of(A.cat.blk[kSegKind[s]].n) : 0;
            R += nr[s];
        }
    }
    __device__ int begin(int part, int nsplit) const {
        return static_cast<int>(static_cast<long long>(part) * R / nsplit);
    }
    // bitmask of the slots whose segments part `part` has rounds of
    __device__ int slots(int part, int nsplit) const {
        const int r0 = begin(part, nsplit), r1 = begin(part + 1, nsplit);
        int m = 0, off = 0;
        for (int s = 0; s < kSegs; ++s) {
            if (max(r0, off) < min(r1, off + nr[s])) m |= 1 << kSegSlot[s];
            off += nr[s];
        }
        return m;
    }
};

template <int K, bool F32>
__global__ __launch_bounds__(kThreads, K >= 4 ? 3 : 4) void alpha_lines_kernel(AlphaArgs A) {
    __shared__ SharedFwd sh;
    // wavefront tile t = columns [t 64K, (t + 1) 64K); a tile may straddle the nh3_hs
    // 30 GHz switch (column nf_low): each column reads its own coefficient set
    const int bpl = (A.tiles_total + kThreads / 64 - 1) / (kThreads / 64);   // blocks per layer
    const int layer = blockIdx.x / bpl, bx = blockIdx.x % bpl;
    const int part = blockIdx.y, nsplit = gridDim.y;
    const LayerT<double> L = layer_of(A, layer);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / 64), lane = threadIdx.x % 64;
    const int t = bx * (kThreads / 64) + wave;
    const bool active = t < A.tiles_total;                // the block's last wavefronts may have no columns
    const int c0 = bx * kThreads * K, c1 = min(c0 + kThreads * K, A.nf);   // the block's columns
    const int need_inv = (c0 < A.nf_low ? 1 : 0) | (c1 > A.nf_low ? 2 : 0);
    const int one[K] = {};                                // every other block: set 0

    int col[K], set[K];
    bool ok[K];
    double x[K], xn[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        col[k] = t * A.tile_width + lane + k * 64;
        ok[k] = active && col[k] < A.nf;
        const int cc = min(col[k], A.nf - 1);
        set[k] = cc < A.nf_low ? 0 : 1;
        const double f = A.freq[cc];
        const double fn = A.freq_nh3 ? A.freq_nh3[cc] : f;
        x[k] = f * f;
        xn[k] = fn * fn;
    }

    // each wavefront's column range of x and of x for NH3 (the FP32 fast path's test)
    __shared__ WaveRanges rng;
    if constexpr (F32) {
        double lo[2] = {INFINITY, INFINITY}, hi[2] = {-INFINITY, -INFINITY};
#pragma unroll
        for (int k = 0; k < K; ++k) {
            if (!ok[k]) continue;
            lo[0] = fmin(lo[0], x[k]);
            hi[0] = fmax(hi[0], x[k]);
            lo[1] = fmin(lo[1], xn[k]);
            hi[1] = fmax(hi[1], xn[k]);
        }
#pragma unroll
        for (int m = 32; m > 0; m >>= 1) {
#pragma unroll
            for (int v = 0; v < 2; ++v) {
                lo[v] = fmin(lo[v], __shfl_xor(lo[v], m));
                hi[v] = fmax(hi[v], __shfl_xor(hi[v], m));
            }
        }
        if (lane == 0) {
#pragma unroll
            for (int v = 0; v < 2; ++v) {
                rng.lo[v][wave] = lo[v];
                rng.hi[v][wave] = hi[v];
            }
        }
        __syncthreads();
    }

    const DevBlock *blk = A.cat.blk;
    const LinePlan plan(A, L.P);
    const int *nr = plan.nr;
    const int r0 = plan.begin(part, nsplit), r1 = plan.begin(part + 1, nsplit);
    int seg = 0, off = 0, lb = 0, le = 0;
    // lines [lb, le) of the next segment that fall in this part's rounds
    auto next = [&](int n) {
        const int a = max(r0, off), b = min(r1, off + nr[seg]);
        lb = (a - off) * kThreads;
        le = min(n, (b - off) * kThreads);
        off += nr[seg++];
        return a < b;
    };

    // each slot's segments are consecutive: a slot's partial is written as soon as
    // its last segment is summed, so only one accumulator row is live at a time
    // (the finish kernel reads exactly the slots a part touches)
    const int touched = plan.slots(part, nsplit);
    const size_t plane = static_cast<size_t>(A.nl) * A.nf;
    double acc[K] = {};
    auto store = [&](int q) {
        if (touched & (1 << q)) {
            double *w = A.ws + (static_cast<size_t>(q) * nsplit + part) * plane + static_cast<size_t>(layer) * A.nf;
#pragma unroll
            for (int k = 0; k < K; ++k)
                if (ok[k]) w[col[k]] = acc[k];
        }
#pragma unroll
        for (int k = 0; k < K; ++k) acc[k] = 0.0;
    };
    double base = 0.0, unused = 0.0;
    bool did_h2o = false;
    if (next(blk[RBX_LINES_H2O].n)) {
        H2oPrep<double> prep(L);
        const double s[2] = {F32 ? pow2_scale(prep.pref * blk[RBX_LINES_H2O].wscale) : 1.0, 1.0};
        double a2[K] = {};
        line_sum<K, F32>(blk[RBX_LINES_H2O], &prep, 1, active, one, s, x, a2, sh, base, lb, le, rng, 0);
#pragma unroll
        for (int k = 0; k < K; ++k) acc[k] = a2[k] / s[0];
        did_h2o = true;
    }
    store(kSlotH2o);
    if (next(blk[RBX_LINES_H2S].n)) {
        H2sPrep<double> prep(L);
        const double s[2] = {F32 ? pow2_scale(prep.pref * blk[RBX_LINES_H2S].wscale) : 1.0, 1.0};
        double a2[K] = {};
        line_sum<K, F32>(blk[RBX_LINES_H2S], &prep, 1, active, one, s, x, a2, sh, unused, lb, le, rng, 0);
#pragma unroll
        for (int k = 0; k < K; ++k) acc[k] = a2[k] / s[0];
    }
    store(kSlotH2s);
    if (next(blk[RBX_LINES_NH3_INV].n)) {
        const InvPrep<double> prep[2] = {InvPrep<double>(L, nh3_low_set(A.abs.nh3)), InvPrep<double>(L, 1)};
        const double s[2] = {F32 ? pow2_scale(prep[0].pref * blk[RBX_LINES_NH3_INV].wscale) : 1.0,
                             F32 ? pow2_scale(prep[1].pref * blk[RBX_LINES_NH3_INV].wscale) : 1.0};
        double a2[K] = {};
        line_sum<K, F32>(blk[RBX_LINES_NH3_INV], prep, need_inv, active, set, s, xn, a2, sh, unused, lb, le, rng, 1);
#pragma unroll
        for (int k = 0; k < K; ++k) acc[k] += a2[k] / s[set[k]];
    }
    if (next(blk[RBX_LINES_NH3_ROT].n)) {
        RotPrep<double> prep(L);
        const double s[2] = {F32 ? pow2_scale(prep.pref * blk[RBX_LINES_NH3_ROT].wscale) : 1.0, 1.0};
        double a2[K] = {};
        line_sum<K, F32>(blk[RBX_LINES_NH3_ROT], &prep, 1, active, one, s, xn, a2, sh, unused, lb, le, rng, 1);
#pragma unroll
        for (int k = 0; k < K; ++k) acc[k] += a2[k] / s[0];
    }
    if (next(blk[RBX_LINES_NH3_V2].n)) {
        V2Prep<double> prep(L);
        const double s[2] = {F32 ? pow2_scale(prep.pref * blk[RBX_LINES_NH3_V2].wscale) : 1.0, 1.0};
        double a2[K] = {};
        line_sum<K, F32>(blk[RBX_LINES_NH3_V2], &prep, 1, active, one, s, xn, a2, sh, unused, lb, le, rng, 1);
#pragma unroll
        for (int k = 0; k < K; ++k) acc[k] += a2[k] / s[0];
    }
    store(kSlotNh3);
    if (next(blk[RBX_LINES_PH3].n)) {
        Ph3Prep<double> prep(L);
        const double s[2] = {F32 ? pow2_scale(prep.pref * blk[RBX_LINES_PH3].wscale) : 1.0, 1.0};
        double a2[K] = {};
        line_sum<K, F32>(blk[RBX_LINES_PH3], &prep, 1, active, one, s, x, a2, sh, unused, lb, le, rng, 0);
#pragma unroll
        for (int k = 0; k < K; ++k) acc[k] = a2[k] / s[0];
    }
    store(kSlotPh3);
    if (did_h2o) base = block_reduce_sum(base, sh.red);   // did_h2o is the same for the whole block

    if ((touched & (1 << kSlotH2o)) && bx == 0 && threadIdx.x == 0)
        A.ws_base[static_cast<size_t>(part) * A.nl + layer] = base;
}

// optional absorbers, one thread per (layer, column), into ws_extra (extra.hpp)
enum { kExtraClouds = 0, kExtraCo = 1, kExtraH2 = 2, kNumExtra = 3 };

__host__ __device__ inline bool has_extra(const rbx_absorbers &ab) { return ab.clouds || ab.co || ab.h2 >= 3; }

__global__ __launch_bounds__(kThreads) void alpha_extra_kernel(AlphaArgs A) {
    const int bpl = (A.nf + kThreads - 1) / kThreads;
    const int layer = blockIdx.x / bpl;
    const int col = (blockIdx.x % bpl) * kThreads + threadIdx.x;
    if (col >= A.nf) return;
    const int nl = A.nl;
    const double P = A.P[layer], T = A.T[layer], f = A.freq[col];
    const double ph2 = P * A.mix[RBX_X_H2 * nl + layer];
    const double phe = P * A.mix[RBX_X_HE * nl + layer];
    const size_t plane = static_cast<size_t>(nl) * A.nf;
    double *out = A.ws_extra + static_cast<size_t>(layer) * A.nf + col;
    if (A.abs.clouds) out[kExtraClouds * plane] = clouds_idp(A.cloud, nl, layer, T, f, A.abs.cloud_types);
    if (A.abs.co)
        out[kExtraCo * plane] = co_ddb(A.cat.blk[RBX_LINES_CO], P, T, ph2, phe, P * A.mix[RBX_X_CO * nl + layer], f,
                                       A.abs.co_shape);
    if (A.abs.h2 >= 3)
        out[kExtraH2 * plane] = h2_orton(A.cat.blk[RBX_TABLE_H2_ORTON], T, ph2, phe,
                                         P * A.mix[RBX_X_CH4 * nl + layer], f, A.abs.h2 - 3);
}

template <bool F32>
__global__ __launch_bounds__(kThreads) void alpha_finish_kernel(AlphaArgs A, int nsplit) {
    __shared__ SharedFwd sh;
    const int bpl = (A.nf + kThreads - 1) / kThreads;
    const int layer = blockIdx.x / bpl;
    const int nl = A.nl;
    const LayerT<double> L = layer_of(A, layer);
    const int col = (blockIdx.x % bpl) * kThreads + threadIdx.x;
    const bool ok = col < A.nf;
    const int cc = ok ? col : A.nf - 1;
    const double f[1] = {A.freq[cc]};
    const double fn = A.freq_nh3 ? A.freq_nh3[cc] : f[0];
    const double x = f[0] * f[0], xn = fn * fn;
    const size_t plane = static_cast<size_t>(nl) * A.nf;
    const size_t o = static_cast<size_t>(layer) * A.nf + cc;
    const LinePlan plan(A, L.P);
    auto parts = [&](int q) {
        const double *w = A.ws + static_cast<size_t>(q) * nsplit * plane + o;
        double s = 0.0;
        for (int p = 0; p < nsplit; ++p)
            if (plan.slots(p, nsplit) & (1 << q)) s += w[static_cast<size_t>(p) * plane];
        return s;
    };

    double tot = 0.0;
    auto emit = [&](int c, double v) {
        const double sc = A.scale ? A.scale[c * nl + layer] : 1.0;
        tot += v * sc;
        if (A.by_c && ok) A.by_c[(static_cast<size_t>(c) * nl + layer) * A.nf + col] = v;
    };
    // the total sums the constituents in name order: clouds, co, h2, h2o, h2s, nh3, ph3
    const double *extra = A.ws_extra + o;
    if (A.abs.clouds) emit(RBX_C_CLOUDS, extra[kExtraClouds * plane]);
    if (A.abs.co) emit(RBX_C_CO, extra[kExtraCo * plane]);
    if (A.abs.h2) emit(RBX_C_H2, A.abs.h2 <= 2 ? h2_cia(L, A.abs.h2) * x : extra[kExtraH2 * plane]);
    if (A.abs.h2o) {
        double base = 0.0;
        for (int p = 0; p < nsplit; ++p)
            if (plan.slots(p, nsplit) & (1 << kSlotH2o)) base += A.ws_base[static_cast<size_t>(p) * nl + layer];
        emit(RBX_C_H2O, x * (parts(kSlotH2o) + base + h2o_continuum(L)));
    }
    if (A.abs.h2s) emit(RBX_C_H2S, x * parts(kSlotH2s));
    if (A.abs.nh3) {
        const int sel = A.abs.nh3;
        double hs = 0.0, sj[1] = {0.0};
        if (nh3_uses_hs(sel, L.P)) {
            hs = xn * parts(kSlotNh3);
            if (hs < 0.0) hs = 1.0E-8;   // nh3_hs.py:337-338
        }
        if (nh3_uses_sjs(sel, L.P)) nh3_sjs<1>(A.cat.blk[RBX_LINES_NH3_SJS], L, f, sj, sh.sjs);   // block-uniform
        emit(RBX_C_NH3, nh3_select(sel, L.P, hs, sj[0]));
    }
    if (A.abs.ph3) emit(RBX_C_PH3, x * parts(kSlotPh3));
    if (!ok) return;
    if (F32)
        static_cast<float *>(A.alpha)[o] = static_cast<float>(tot);
    else
        static_cast<double *>(A.alpha)[o] = tot;
}

#include "far.hpp"

// ---- tangent kernel: alpha and dalpha[3][nl][nf] = d alpha / d(T, q_NH3, q_H2O) ----
//
// Work item = (layer, tile of 64 consecutive columns), one per wavefront; the
// wavefronts of a block are independent (they may be on different layers), so
// each stages its own lines: lane l derives line l0 + l's coefficients and
// their (T, q_own) tangents on dual numbers, writes them to the wavefront's LDS
// buffer, and all 64 lanes sum the 64 staged lines.  A tile may straddle the
// nh3_hs 30 GHz switch: the inversion lines are then staged with both
// coefficient sets and each lane reads its own.  Per line and column:
//   t = N/Q, N = a x + c, Q = (x - D)^2 + g x + G,
//   dt = (dN - t dQ)/Q, dN = da x + dc, dQ = -2 (x - D) dD + dg x + dG.
// FP32 mode stores the staged values in FP32 (FP64 D), scales a, c and their
// tangents by the layer's power of two, forms x - D in FP64 and sums blocks of kSubT
// lines in FP32 before accumulating in FP64; FP64 mode is plain FP64.  The denominator
// Q = (x - Ds)^2 + h is formed from the FP32 rounding of x - Ds in one FP32 FMA where
// that stays accurate, else in FP64 and rounded: the FP32 form's relative error is
// about (A + 1) FP32 roundings, A = (2 (x - Ds)^2 + |h|) / Q.  A <= 2 when h >= 0; when
// h < 0 and Ds < 0, A is largest at x = 0 over x = f^2 >= 0, A0 = (2 Ds^2 - h) / (Ds^2 + h),
// and FP32 is taken when A0 <= kTanAmp (2e-6); any other line (Ben-Reuven lines far
// wider than their centre, h2s_ddb's deep in the atmosphere) takes FP64, and so does
// every line of a staged buffer holding one.
constexpr int kWaveLines = 64;
constexpr int kSetPad = 4;   // moves set 1 to other LDS banks than set 0
constexpr int kSubT = 4;
constexpr double kTanAmp = 32.0;

template <bool F32>
struct TanLines;
template <>
struct TanLines<true> {   // shifted denominator (x - Ds)^2 + h, as the forward FP32 staging
    alignas(16) double D[2][kWaveLines + kSetPad];       // Ds
    alignas(16) double hd[2][kWaveLines + kSetPad];      // h (the FP64 denominator's)
    alignas(16) float h[2][kWaveLines + kSetPad];
    alignas(16) float2 ac[2][kWaveLines + kSetPad];      // a s, c s
    alignas(16) float4 dq[2][2][kWaveLines + kSetPad];   // [set][slot] d(a s), d(c s), dh, -2 dDs
    unsigned long long exact[2];                          // [set] lines whose denominator needs FP64
};
template <>
struct TanLines<false> {
    double v[2][5][kWaveLines + kSetPad];      // [set] D, a, c, g, G
    double d[2][2][5][kWaveLines + kSetPad];   // [set][slot]
};

// each wavefront's staging buffer: its share of the lines the block's wavefronts sum
template <bool F32>
union WaveTan {
    TanLines<F32> lines;
    SjsLines<64> sjs;
};

// W wavefronts: the frequency tiles of one layer that share the lines' derivation
template <bool F32, int W>
struct SharedTan {
    WaveTan<F32> w[W];
    // each wavefront's coefficient sets of the block being staged (uniform across the
    // wavefront: kept in LDS, not registers, while the staged lines are summed)
    alignas(16) double prep[W][64];
    double base[W][4];   // the wavefronts' shares of the h2o base term (value, 3 tangents)
};

// value + (T, q_own) tangent sums of one line block for this lane's column.
// prep[0..1]: coefficient sets (two only for the nh3_hs inversion lines);
// `need` = sets any lane of the block uses, `mine` = this lane's set;
// scale[s] = FP32 power-of-two weight scale of set s (1 in FP64 mode).
// The block's nw wavefronts (frequency tiles of one layer) take the lines in rounds of
// nw x 64: wavefront wi derives and stages the round's lines [wi 64, wi 64 + 64) into its
// buffer bufs[wi], and after a block barrier every wavefront sums all nw buffers, in line
// order, for its own columns -- each line is derived once per block rather than once per
// tile.  Every wavefront of the block must make the same calls (same lb, le).
template <bool F32, class Prep>
__device__ void line_sum_tan(const DevBlock &B, const Prep *prep, int need, int mine, const double *scale,
                             int qslot, double x, double &acc, double (&dacc)[2], WaveTan<F32> *bufs, int nw,
                             int wi, Dual &base, int lb, int le) {
    const int lane = threadIdx.x % 64;
    const float xf = static_cast<float>(x);
    for (int l0 = lb; l0 < le; l0 += nw * kWaveLines) {
        TanLines<F32> &st = bufs[wi].lines;
        const int b0 = l0 + wi * kWaveLines;
        const int i = b0 + lane;
        const int cnt = max(0, min(kWaveLines, le - b0));
        const int cntp = (cnt + kSubT - 1) / kSubT * kSubT;
        [[maybe_unused]] bool exact[2] = {false, false};
        if (lane < cntp) {
            CoefT<Dual> qs[2] = {{Dual(0.0), Dual(0.0), Dual(0.0), Dual(0.0), Dual(1.0)},
                                 {Dual(0.0), Dual(0.0), Dual(0.0), Dual(0.0), Dual(1.0)}};   // padding: t = 0
            if (i < le) line_coefs(prep, need, B, i, base, qs);
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                if (!(need & (1 << s))) continue;
                const CoefT<Dual> &q = qs[s];
                const int sl[2] = {0, qslot};
                if constexpr (F32) {
                    const double sc = scale[s];
                    Dual Ds, h;
                    shifted_denominator(q, Ds, h);
                    st.D[s][lane] = Ds.v;
                    st.hd[s][lane] = h.v;
                    st.h[s][lane] = static_cast<float>(h.v);
                    exact[s] = !(h.v >= 0.0 || (Ds.v < 0.0 && 2.0 * Ds.v * Ds.v - h.v <=
                                                                   kTanAmp * (Ds.v * Ds.v + h.v)));
                    st.ac[s][lane] = make_float2(static_cast<float>(q.a.v * sc), static_cast<float>(q.c.v * sc));
#pragma unroll
                    for (int m = 0; m < 2; ++m) {
                        const int d = sl[m];
                        st.dq[s][m][lane] =
                            make_float4(static_cast<float>(q.a.d[d] * sc), static_cast<float>(q.c.d[d] * sc),
                                        static_cast<float>(h.d[d]), static_cast<float>(-2.0 * Ds.d[d]));
                    }
                } else {
                    const Dual *c5[5] = {&q.D, &q.a, &q.c, &q.g, &q.G};
#pragma unroll
                    for (int c = 0; c < 5; ++c) {
                        st.v[s][c][lane] = c5[c]->v;
                        st.d[s][0][c][lane] = c5[c]->d[0];
                        st.d[s][1][c][lane] = c5[c]->d[qslot];
                    }
                }
            }
        }
        if constexpr (F32) {
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                const unsigned long long m = __builtin_amdgcn_ballot_w64(exact[s]);
                if (lane == 0) st.exact[s] = m;
            }
        }
        __syncthreads();
        for (int w = 0; w < nw; ++w) {
        const TanLines<F32> &sw = bufs[w].lines;
        const int cw = max(0, min(kWaveLines, le - (l0 + w * kWaveLines)));
        const int cwp = (cw + kSubT - 1) / kSubT * kSubT;
        if constexpr (F32) {
            const double2 *sD = reinterpret_cast<const double2 *>(sw.D[mine]);
            const double2 *sHd = reinterpret_cast<const double2 *>(sw.hd[mine]);
            const float4 *sH = reinterpret_cast<const float4 *>(sw.h[mine]);
            const float4 *sAC = reinterpret_cast<const float4 *>(sw.ac[mine]);
            // the buffer holds a line of a set this block reads whose denominator needs FP64: then
            // all of its lines take the FP64 denominator (a choice per buffer keeps the FP32 loop
            // free of branches)
            const bool exact_lines =
                uniform(((need & 1) ? sw.exact[0] : 0ull) | ((need & 2) ? sw.exact[1] : 0ull)) != 0ull;
            auto block = [&](int j0, auto fp64_denominator) {
                float p = 0.0f, dp[2] = {0.0f, 0.0f};
#pragma unroll
                for (int u = 0; u < kSubT; ++u) {
                    const int j = j0 + u;
                    // lines j, j + 1 come as one 16-byte read of D and of ac, lines j0 .. j0 + 3
                    // as one of h
                    const double2 D2 = sD[j / 2];
                    const float4 AC2 = sAC[j / 2];
                    const double dd = x - ((u & 1) ? D2.y : D2.x);
                    const float ddf = static_cast<float>(dd);
                    float Q;
                    if constexpr (decltype(fp64_denominator)::value) {
                        const double2 H2 = sHd[j / 2];
                        Q = static_cast<float>(fma(dd, dd, (u & 1) ? H2.y : H2.x));
                    } else {
                        const float4 H4 = sH[j / 4];
                        Q = fmaf(ddf, ddf, (u & 3) == 0 ? H4.x : (u & 3) == 1 ? H4.y : (u & 3) == 2 ? H4.z : H4.w);
                    }
                    const float2 ac = (u & 1) ? make_float2(AC2.z, AC2.w) : make_float2(AC2.x, AC2.y);
                    const float N = fmaf(ac.x, xf, ac.y);
                    const float r = __builtin_amdgcn_rcpf(Q);
                    const float t = N * r;
                    p += t;
#pragma unroll
                    for (int m = 0; m < 2; ++m) {
                        const float4 dq = sw.dq[mine][m][j];
                        const float dN = fmaf(dq.x, xf, dq.y);
                        const float dQ = fmaf(ddf, dq.w, dq.z);   // d[(x - Ds)^2 + h], dq.w = -2 dDs
                        dp[m] = fmaf(fmaf(-t, dQ, dN), r, dp[m]);
                    }
                }
                acc += static_cast<double>(p);
                dacc[0] += static_cast<double>(dp[0]);
                dacc[1] += static_cast<double>(dp[1]);
            };
            if (exact_lines)
                for (int j0 = 0; j0 < cwp; j0 += kSubT) block(j0, std::true_type{});
            else
                for (int j0 = 0; j0 < cwp; j0 += kSubT) block(j0, std::false_type{});
        } else {
            for (int j = 0; j < cwp; ++j) {
                const double D = sw.v[mine][0][j], a = sw.v[mine][1][j], c = sw.v[mine][2][j];
                const double g = sw.v[mine][3][j], G = sw.v[mine][4][j];
                const double d = x - D;
                const double N = fma(a, x, c);
                const double r = 1.0 / fma(d, d, fma(g, x, G));
                const double t = N * r;
                acc += t;
#pragma unroll
                for (int m = 0; m < 2; ++m) {
                    const double *dd = &sw.d[mine][m][0][0];
                    const int S = kWaveLines + kSetPad;
                    const double dN = fma(dd[1 * S + j], x, dd[2 * S + j]);
                    const double dQ = fma(-2.0 * d, dd[j], fma(dd[3 * S + j], x, dd[4 * S + j]));
                    dacc[m] += (dN - t * dQ) * r;
                }
            }
        }
        }
        __syncthreads();   // (the next round overwrites the buffers)
    }
}

__device__ inline Dual wave_sum(const Dual &v) {
    Dual r;
    r.v = wave_sum(v.v);
    for (int m = 0; m < kNd; ++m) r.d[m] = wave_sum(v.d[m]);
    return r;
}

// The tangent pass in one launch (TAN_ALL) or split (TAN_PART, then TAN_FINISH): with
// several parts per layer, part p of a (layer, tile) sums the staging rounds
// [p R / parts, (p + 1) R / parts) of the layer's sequence of 64-line rounds over the
// segments h2o, h2s, nh3 inversion, rotational, v2, ph3 and writes its per-constituent
// partial sums (value, d/dT, d/dq of the constituent's own species) to the workspace; the
// finish adds the parts in order and applies what needs whole sums (the H2O base term, the
// nh3_hs clamp, nh3_sjs, the closed forms, the constituent scales).
constexpr int kTanAll = 0, kTanPart = 1, kTanFinish = 2;
constexpr int kTanSegs = 6;
constexpr int kTanSlots = 4;   // h2o, h2s, nh3_hs, ph3 (the forward kSlot* order)
constexpr int kTanComps = 3;   // value, d/dT, d/dq_own

__device__ __host__ inline int tan_rounds(int n) { return (n + kWaveLines - 1) / kWaveLines; }

// The layer's coefficient sets on dual numbers (value, d/dT, d/dq_NH3, d/dq_H2O), as the
// tangent pass builds them from the layer state.
__device__ inline LayerT<Dual> dual_layer(const AlphaArgs &A, int layer) {
    const LayerT<double> L0 = layer_of(A, layer);
    LayerT<Dual> L;
    L.P = L0.P;
    L.T = Dual(L0.T, 1.0, 0.0, 0.0);
    L.h2 = L0.h2;
    L.he = L0.he;
    L.ch4 = L0.ch4;
    L.nh3 = Dual(L0.nh3, 0.0, L0.P, 0.0);   // d(P q)/dq = P
    L.h2o = Dual(L0.h2o, 0.0, 0.0, L0.P);
    L.h2s = L0.h2s;
    L.ph3 = L0.ph3;
    return L;
}

// Every (layer, tile) wavefront of the tangent pass, and every part of one when it is split,
// would build the layer's dual-number derivers again, so tangent_prep_kernel writes them once
// per layer, [nl][kTanPrepWords] doubles, segment s at kTanPrepOff[s] (the nh3_hs inversion
// lines' two sets one after the other), and the wavefronts read them there.
template <class P>
constexpr int prep_words() {
    static_assert(sizeof(P) % sizeof(double) == 0, "a deriver is whole doubles");
    return static_cast<int>(sizeof(P) / sizeof(double));
}
constexpr int kTanPrepOff[7] = {
    0,
    prep_words<H2oPrep<Dual>>(),
    prep_words<H2oPrep<Dual>>() + prep_words<H2sPrep<Dual>>(),
    prep_words<H2oPrep<Dual>>() + prep_words<H2sPrep<Dual>>() + 2 * prep_words<InvPrep<Dual>>(),
    prep_words<H2oPrep<Dual>>() + prep_words<H2sPrep<Dual>>() + 2 * prep_words<InvPrep<Dual>>() +
        prep_words<RotPrep<Dual>>(),
    prep_words<H2oPrep<Dual>>() + prep_words<H2sPrep<Dual>>() + 2 * prep_words<InvPrep<Dual>>() +
        prep_words<RotPrep<Dual>>() + prep_words<V2Prep<Dual>>(),
    prep_words<H2oPrep<Dual>>() + prep_words<H2sPrep<Dual>>() + 2 * prep_words<InvPrep<Dual>>() +
        prep_words<RotPrep<Dual>>() + prep_words<V2Prep<Dual>>() + prep_words<Ph3Prep<Dual>>()};
constexpr int kTanPrepWords = kTanPrepOff[6];
static_assert(sizeof(Dual) % sizeof(double) == 0, "dual numbers are whole doubles");

template <class P>
__device__ inline void put_prep(double *row, const P &p) {
    const double *w = reinterpret_cast<const double *>(&p);
    for (int k = 0; k < prep_words<P>(); ++k) row[k] = w[k];
}

__global__ __launch_bounds__(256) void tangent_prep_kernel(AlphaArgs A) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= A.nl * kTanSegs) return;
    const int layer = e / kTanSegs, seg = e % kTanSegs;
    const LayerT<Dual> L = dual_layer(A, layer);
    double *row = A.tan_prep + static_cast<size_t>(layer) * kTanPrepWords + kTanPrepOff[seg];
    switch (seg) {
        case 0: put_prep(row, H2oPrep<Dual>(L)); break;
        case 1: put_prep(row, H2sPrep<Dual>(L)); break;
        case 2:
            put_prep(row, InvPrep<Dual>(L, nh3_low_set(A.abs.nh3)));
            put_prep(row + prep_words<InvPrep<Dual>>(), InvPrep<Dual>(L, 1));
            break;
        case 3: put_prep(row, RotPrep<Dual>(L)); break;
        case 4: put_prep(row, V2Prep<Dual>(L)); break;
        default: put_prep(row, Ph3Prep<Dual>(L)); break;
    }
}

// Blocks: the line-summing modes (TAN_ALL, TAN_PART) take W consecutive 64-column tiles of one
// layer per block (block = layer x group of W tiles, W = blockDim.x / 64; a wavefront past the
// last tile derives its share of the lines and sums none), so that the block derives each line
// once (line_sum_tan); block barriers follow, so the whole block returns together or not at
// all.  TAN_FINISH takes any W (layer, tile) items per block, independent of each other.
template <bool F32, int MODE, int W>
__device__ __forceinline__ void alpha_tangent_body(const AlphaArgs &A, SharedTan<F32, W> &sh) {
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / 64), lane = threadIdx.x % 64;
    const int tiles = (A.nf + 63) / 64;
    int layer, tile, g = 0;
    if constexpr (MODE == kTanFinish) {
        const long long item = static_cast<long long>(blockIdx.x) * W + wave;
        if (item >= static_cast<long long>(A.nl) * tiles) return;   // the whole wavefront: no block barriers
        layer = static_cast<int>(item / tiles);
        tile = static_cast<int>(item % tiles);
    } else {
        const int G = (tiles + W - 1) / W;                           // groups of W tiles per layer
        layer = static_cast<int>(blockIdx.x / G);
        g = static_cast<int>(blockIdx.x % G);
        if (layer >= A.nl) return;                                   // the whole block
        tile = g * W + wave;
    }
    const int col = tile * 64 + lane;
    const bool ok = col < A.nf;
    const int cc = ok ? col : A.nf - 1;
    const int mine = cc < A.nf_low ? 0 : 1;                      // nh3_hs set of this lane's column
    // coefficient sets any column of the block's tiles uses (the block stages them for all)
    const int gc0 = g * W * 64, gc1 = min((g + 1) * W * 64, A.nf);
    const int need = MODE == kTanFinish ? 0 : (gc0 < A.nf_low ? 1 : 0) | (gc1 > A.nf_low ? 2 : 0);
    const int parts = A.tan_parts;
    const int part = MODE == kTanPart ? static_cast<int>(blockIdx.y) : 0;
    const int r0 = MODE == kTanPart ? part * A.tan_rounds / parts : 0;
    const int r1 = MODE == kTanPart ? (part + 1) * A.tan_rounds / parts : 0;

    const LayerT<double> L0 = layer_of(A, layer);
    const int nl = A.nl;
    const LayerT<Dual> L = dual_layer(A, layer);
    // the layer's derivers from their table (typed views of its doubles)
    const double *trow = MODE != kTanFinish ? A.tan_prep + static_cast<size_t>(layer) * kTanPrepWords : nullptr;
    const double f[1] = {A.freq[cc]};
    const double fn = A.freq_nh3 ? A.freq_nh3[cc] : f[0];
    const double x = f[0] * f[0], xn = fn * fn;

    // workspace of the split pass: slot sums [parts][kTanSlots][kTanComps][nl][nf], then
    // the H2O base terms [parts][kTanComps][nl]
    const size_t plane = static_cast<size_t>(nl) * A.nf;
    const size_t o = static_cast<size_t>(layer) * A.nf + col;
    double *ws_base = A.ws + static_cast<size_t>(parts) * kTanSlots * kTanComps * plane;
    auto slot_out = [&](int slot, int own, const Dual &v) {
        if (!ok) return;
        double *w = A.ws + (static_cast<size_t>(part) * kTanSlots + slot) * kTanComps * plane + o;
        w[0] = v.v;
        w[plane] = v.d[0];
        w[2 * plane] = v.d[own];
    };
    auto slot_in = [&](int slot, int own) {
        Dual r;
        for (int p = 0; p < parts; ++p) {
            const double *w = A.ws + (static_cast<size_t>(p) * kTanSlots + slot) * kTanComps * plane + o;
            r.v += w[0];
            r.d[0] += w[plane];
            r.d[own] += w[2 * plane];
        }
        return r;
    };
    // the lines of one segment this pass sums (all of them unless split)
    auto range = [&](int seg, int n, int &lb, int &le) {
        lb = 0;
        le = n;
        if (MODE != kTanPart) return;
        const int g0 = A.tan_seg0[seg];
        const int ra = max(r0, g0), rb = min(r1, g0 + tan_rounds(n));
        lb = min(n, max(0, ra - g0) * kWaveLines);
        le = rb > ra ? min(n, (rb - g0) * kWaveLines) : lb;
    };

    // total_layer_alpha (alpha.py:151-192): constituents scale-summed as each completes
    Dual tot, unused;
    auto emit = [&](int c, const Dual &v) { tot = tot + v * (A.scale ? A.scale[c * nl + layer] : 1.0); };
    // one line block: x (sum / scale) with the value, d/dT and d/dq_own parts, added to out
    auto lines = [&](int seg, const DevBlock &B, const auto *prep, int nset, int qslot, double xx, Dual &out,
                     Dual &base) {
        double s[2] = {1.0, 1.0};
        if (F32)
            for (int k = 0; k < nset; ++k) s[k] = pow2_scale(prep[k].pref.v * B.wscale);
        double acc = 0.0, dacc[2] = {0.0, 0.0};
        const int nd = nset == 2 ? need : 1, me = nset == 2 ? mine : 0;
        int lb, le;
        range(seg, B.n, lb, le);
        using Prep = std::remove_const_t<std::remove_reference_t<decltype(prep[0])>>;
        static_assert(2 * sizeof(Prep) <= sizeof(sh.prep[0]), "coefficient sets exceed their LDS slot");
        Prep *lp = reinterpret_cast<Prep *>(sh.prep[wave]);
        if (lane == 0)
            for (int k = 0; k < nset; ++k) lp[k] = prep[k];
        wave_sync();
        line_sum_tan<F32>(B, lp, nd, me, s, qslot, xx, acc, dacc, sh.w, W, wave, base, lb, le);
        wave_sync();   // the next block's sets overwrite these
        Dual r(xx * acc / s[me]);
        r.d[0] = xx * dacc[0] / s[me];
        r.d[qslot] = xx * dacc[1] / s[me];
        out = out + r;
    };

    if (MODE != kTanPart && A.abs.h2) emit(RBX_C_H2, h2_cia(L, A.abs.h2) * x);
    if (A.abs.h2o) {
        Dual v, base;
        if (MODE == kTanFinish) {
            v = slot_in(kSlotH2o, 2);
            for (int p = 0; p < parts; ++p) {
                base.v += ws_base[(static_cast<size_t>(p) * kTanComps) * nl + layer];
                base.d[0] += ws_base[(static_cast<size_t>(p) * kTanComps + 1) * nl + layer];
                base.d[2] += ws_base[(static_cast<size_t>(p) * kTanComps + 2) * nl + layer];
            }
        } else {
            lines(0, A.cat.blk[RBX_LINES_H2O], reinterpret_cast<const H2oPrep<Dual> *>(trow + kTanPrepOff[0]), 1, 2, x,
                  v, base);
            // each wavefront derived its share of the lines: the block's base term, added in order
            base = wave_sum(base);
            if (lane == 0) {
                sh.base[wave][0] = base.v;
                for (int m = 0; m < kNd; ++m) sh.base[wave][1 + m] = base.d[m];
            }
            __syncthreads();
            base = Dual();
            for (int w = 0; w < W; ++w) {
                base.v += sh.base[w][0];
                for (int m = 0; m < kNd; ++m) base.d[m] += sh.base[w][1 + m];
            }
            __syncthreads();
        }
        if (MODE == kTanPart) {
            slot_out(kSlotH2o, 2, v);
            if (tile == 0 && lane == 0) {
                ws_base[(static_cast<size_t>(part) * kTanComps) * nl + layer] = base.v;
                ws_base[(static_cast<size_t>(part) * kTanComps + 1) * nl + layer] = base.d[0];
                ws_base[(static_cast<size_t>(part) * kTanComps + 2) * nl + layer] = base.d[2];
            }
        } else {
            emit(RBX_C_H2O, v + (base + h2o_continuum(L)) * x);
        }
    }
    if (A.abs.h2s) {
        Dual v;
        if (MODE == kTanFinish) {
            v = slot_in(kSlotH2s, 1);
        } else {
            lines(1, A.cat.blk[RBX_LINES_H2S], reinterpret_cast<const H2sPrep<Dual> *>(trow + kTanPrepOff[1]), 1, 1, x,
                  v, unused);
        }
        if (MODE == kTanPart)
            slot_out(kSlotH2s, 1, v);
        else
            emit(RBX_C_H2S, v);
    }
    if (A.abs.nh3) {
        const int sel = A.abs.nh3;
        Dual hs, sj[1];
        if (nh3_uses_hs(sel, L0.P)) {
            if (MODE == kTanFinish) {
                hs = slot_in(kSlotNh3, 1);
            } else {
                lines(2, A.cat.blk[RBX_LINES_NH3_INV], reinterpret_cast<const InvPrep<Dual> *>(trow + kTanPrepOff[2]), 2,
                      1, xn, hs, unused);
                lines(3, A.cat.blk[RBX_LINES_NH3_ROT], reinterpret_cast<const RotPrep<Dual> *>(trow + kTanPrepOff[3]), 1,
                      1, xn, hs, unused);
                lines(4, A.cat.blk[RBX_LINES_NH3_V2], reinterpret_cast<const V2Prep<Dual> *>(trow + kTanPrepOff[4]), 1,
                      1, xn, hs, unused);
            }
            if (MODE == kTanPart) slot_out(kSlotNh3, 1, hs);
            if (hs.v < 0.0) hs = Dual(1.0E-8);   // nh3_hs.py:337-338
        }
        if (MODE != kTanPart) {
            if (nh3_uses_sjs(sel, L0.P)) nh3_sjs<1, true>(A.cat.blk[RBX_LINES_NH3_SJS], L, f, sj, sh.w[wave].sjs);
            emit(RBX_C_NH3, nh3_select(sel, L0.P, hs, sj[0]));
        }
    }
    if (A.abs.ph3) {
        // ph3 does not depend on q_NH3: its slot-1 tangent is identically zero
        Dual v;
        if (MODE == kTanFinish) {
            v = slot_in(kSlotPh3, 1);
        } else {
            lines(5, A.cat.blk[RBX_LINES_PH3], reinterpret_cast<const Ph3Prep<Dual> *>(trow + kTanPrepOff[5]), 1, 1, x,
                  v, unused);
        }
        if (MODE == kTanPart)
            slot_out(kSlotPh3, 1, v);
        else
            emit(RBX_C_PH3, v);
    }

    if (MODE == kTanPart || !ok) return;
    double *out = static_cast<double *>(A.alpha);
#pragma unroll
    for (int m = 0; m < kNd; ++m) out[m * plane + o] = tot.d[m];
    if (A.value) A.value[o] = tot.v;
}

// FP32: ask for 3 wavefronts per SIMD (the staged-line buffers take 35 KB per
// block); FP64's 65 KB of staging allows 2 blocks per CU whatever the registers.
template <int MODE, int W>
__global__ __launch_bounds__(kThreads, 3) void alpha_tangent_kernel_f32(AlphaArgs A) {
    __shared__ SharedTan<true, W> sh;
    alpha_tangent_body<true, MODE, W>(A, sh);
}

template <int MODE, int W>
__global__ __launch_bounds__(kThreads) void alpha_tangent_kernel_f64(AlphaArgs A) {
    __shared__ SharedTan<false, W> sh;
    alpha_tangent_body<false, MODE, W>(A, sh);
}

// check that the catalogue holds what the selected absorbers read
static int check_catalog(const rbx_catalog *cat, const rbx_absorbers &ab, const char *who) {
    RBX_CHECK_ARG(ab.h2 >= 0 && ab.h2 <= 4, "%s: bad h2 selector %d", who, ab.h2);
    RBX_CHECK_ARG(ab.co >= 0 && ab.co <= 1, "%s: bad co selector %d", who, ab.co);
    RBX_CHECK_ARG(ab.co_shape >= 0 && ab.co_shape <= 3, "%s: bad co_shape %d", who, ab.co_shape);
    RBX_CHECK_ARG(ab.clouds >= 0 && ab.clouds <= 1, "%s: bad clouds selector %d", who, ab.clouds);
    RBX_CHECK_ARG(ab.cloud_types >= 0 && ab.cloud_types < (1 << RBX_NUM_CLOUDS), "%s: bad cloud_types %d", who,
                  ab.cloud_types);
    RBX_CHECK_ARG(ab.h2o >= 0 && ab.h2o <= 1, "%s: bad h2o selector %d", who, ab.h2o);
    RBX_CHECK_ARG(ab.h2s >= 0 && ab.h2s <= 1, "%s: bad h2s selector %d", who, ab.h2s);
    RBX_CHECK_ARG(ab.nh3 >= 0 && ab.nh3 <= 5, "%s: bad nh3 selector %d", who, ab.nh3);
    RBX_CHECK_ARG(ab.ph3 >= 0 && ab.ph3 <= 1, "%s: bad ph3 selector %d", who, ab.ph3);
    struct Need {
        bool on;
        int kind;
        const char *what;
    } needs[] = {{ab.h2o != 0, RBX_LINES_H2O, "h2o_bk"},
                 {ab.h2s != 0, RBX_LINES_H2S, "h2s_ddb"},
                 {ab.ph3 != 0, RBX_LINES_PH3, "ph3_jh"},
                 {nh3_has_hs(ab.nh3), RBX_LINES_NH3_INV, "nh3_hs / nh3_dbs"},
                 {nh3_has_sjs(ab.nh3), RBX_LINES_NH3_SJS, "nh3_sjs"},
                 {ab.co != 0, RBX_LINES_CO, "co_ddb"},
                 {ab.h2 >= 3, RBX_TABLE_H2_ORTON, "h2_orton"}};
    for (const Need &n : needs) {
        if (n.on && cat->dev.blk[n.kind].n <= 0) {
            set_error("%s: absorber %s selected but its line block %d is empty", who, n.what, n.kind);
            return RBX_ERR_STATE;
        }
    }
    return RBX_OK;
}

static int fill_args(AlphaArgs &A, const rbx_catalog *cat, const rbx_absorbers *absorbers, const double *P,
                     const double *T, const double *mix, const double *cloud, int nl, const double *freq,
                     const double *freq_nh3, int nf, int nf_nh3_low, const double *scale, void *out, double *by_c,
                     const char *who) {
    RBX_CHECK_ARG(cat && absorbers, "%s: null catalogue or absorber selection", who);
    RBX_CHECK_ARG(P && T && mix && freq && out, "%s: null device pointer", who);
    RBX_CHECK_ARG(!absorbers->clouds || cloud, "%s: clouds_idp selected but the cloud table is null", who);
    RBX_CHECK_ARG(nl > 0 && nf > 0, "%s: nl=%d nf=%d must be positive", who, nl, nf);
    RBX_CHECK_ARG(static_cast<long long>(nl) * ((nf + 63) / 64) < (1LL << 31), "%s: nl=%d x nf=%d too large", who,
                  nl, nf);
    RBX_CHECK_ARG(nf_nh3_low >= 0 && nf_nh3_low <= nf, "%s: nf_nh3_low=%d outside [0, %d]", who, nf_nh3_low, nf);
    int rc = check_catalog(cat, *absorbers, who);
    if (rc) return rc;
    A.cat = cat->dev;
    A.abs = *absorbers;
    A.P = P;
    A.T = T;
    A.mix = mix;
    A.cloud = cloud;
    A.nl = nl;
    A.freq = freq;
    A.freq_nh3 = freq_nh3;
    A.nf = nf;
    A.nf_low = nf_nh3_low;
    A.scale = scale;
    A.alpha = out;
    A.by_c = by_c;
    return RBX_OK;
}


// frequencies per lane of the forward kernel: launch.alpha_k, else 2 for wide requests (each
// staged line is read once for two columns; on 800 layers x 1024 frequencies 4.2 ms against
// 4.8 ms with one) and 1 below 256 frequencies (fewer idle lanes)
static int alpha_tile_k(int nf, const rbx_launch &L) {
    const int k = L.alpha_k;
    if (k == 1 || k == 2 || k == 4) return k;
    return nf >= 256 ? 2 : 1;
}

static void set_wave_tiles(AlphaArgs &A, int K) {
    A.tile_width = 64 * K;   // one wavefront tile
    A.tiles_total = (A.nf + A.tile_width - 1) / A.tile_width;
}

template <int K>
static const void *lines_kernel_ptr(int precision) {
    return precision == RBX_FP32 ? reinterpret_cast<const void *>(&alpha_lines_kernel<K, true>)
                                 : reinterpret_cast<const void *>(&alpha_lines_kernel<K, false>);
}

// resident blocks of the line-sum kernel on the current device (cached per device and variant)
static int resident_blocks(int K, int precision, int &out) {
    constexpr int kMaxDev = 64;
    static std::atomic<int> cache[kMaxDev][3][2];
    int dev = 0;
    RBX_CHECK_HIP(hipGetDevice(&dev));
    const int ki = K == 4 ? 2 : K - 1, pi = precision == RBX_FP32 ? 1 : 0;
    if (dev < kMaxDev && cache[dev][ki][pi].load() > 0) {
        out = cache[dev][ki][pi].load();
        return RBX_OK;
    }
    const void *fn = K == 4 ? lines_kernel_ptr<4>(precision)
                            : (K == 2 ? lines_kernel_ptr<2>(precision) : lines_kernel_ptr<1>(precision));
    int per_cu = 0, cus = 0;
    RBX_CHECK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kThreads, 0));
    RBX_CHECK_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    out = std::max(1, per_cu * cus);
    if (dev < kMaxDev) cache[dev][ki][pi].store(out);
    return RBX_OK;
}

// Parts per layer of the line sums: launch.alpha_split, else the split that minimises
// (dispatch rounds of equal blocks) x (staging rounds per block + 0.2, the block's own
// set-up), i.e. the grid's tail.
static int alpha_split(const AlphaArgs &A, const rbx_launch &L, int K, int precision, int &nsplit) {
    const DevBlock *blk = A.cat.blk;
    const bool hs = nh3_has_hs(A.abs.nh3);
    const int R = (A.abs.h2o ? rounds_of(blk[RBX_LINES_H2O].n) : 0) +
                  (A.abs.h2s ? rounds_of(blk[RBX_LINES_H2S].n) : 0) +
                  (hs ? rounds_of(blk[RBX_LINES_NH3_INV].n) + rounds_of(blk[RBX_LINES_NH3_ROT].n) +
                            rounds_of(blk[RBX_LINES_NH3_V2].n)
                      : 0) +
                  (A.abs.ph3 ? rounds_of(blk[RBX_LINES_PH3].n) : 0);
    if (L.alpha_split > 0) {
        nsplit = std::min(L.alpha_split, kMaxSplit);
        return RBX_OK;
    }
    nsplit = 1;
    if (R <= 1) return RBX_OK;
    int slots = 0;
    int rc = resident_blocks(K, precision, slots);
    if (rc) return rc;
    const long long items = static_cast<long long>((A.tiles_total + 3) / 4) * A.nl;
    double best = 0.0;
    for (int n = 1; n <= std::min(kMaxSplit, R); ++n) {
        const long long waves = (items * n + slots - 1) / slots;
        const double cost = static_cast<double>(waves) * ((R + n - 1) / n + 0.2);
        if (n == 1 || cost < best * 0.999) {
            best = cost;
            nsplit = n;
        }
    }
    return RBX_OK;
}

// far-line series (far.hpp): launch.alpha_series 1 off, 2 on; built in on for
// requests of at least kSeriesMinTiles wavefront tiles while the workspace stays under
// kSeriesMaxWorkspace.  (With fewer tiles few (line, tile) pairs are far and the extra pass
// costs more than it saves: one 20-frequency tile of 150 layers took 0.37 ms per step against
// 0.21 ms with the direct sums.)
constexpr size_t kSeriesMaxWorkspace = size_t(4) << 30;
constexpr int kSeriesMinTiles = 4;
static bool series_on(const AlphaArgs &A, const rbx_launch &L) {
    const int t = L.alpha_series;
    return t == 2 || (t != 1 && A.tiles_total >= kSeriesMinTiles);
}

// near-mask streams: one bit per line of every staging round of each selected segment
static void far_layout(AlphaArgs &A) {
    const DevBlock *blk = A.cat.blk;
    const bool hs = nh3_has_hs(A.abs.nh3);
    const int n[kFarStreams] = {A.abs.h2o ? blk[RBX_LINES_H2O].n : 0,     A.abs.h2s ? blk[RBX_LINES_H2S].n : 0,
                                hs ? blk[RBX_LINES_NH3_INV].n : 0,        hs ? blk[RBX_LINES_NH3_INV].n : 0,
                                hs ? blk[RBX_LINES_NH3_ROT].n : 0,        hs ? blk[RBX_LINES_NH3_V2].n : 0,
                                A.abs.ph3 ? blk[RBX_LINES_PH3].n : 0};
    int at = 0;
    for (int s = 0; s < kFarStreams; ++s) {
        A.far_off[s] = at;
        at += rounds_of(n[s]) * (kThreads / 32);
    }
    A.far_words = at;
}

// parts per layer of the classification pass: launch.far_split, else enough blocks
// (layers x tile groups x parts) to fill the device a few times over
static int far_split(const AlphaArgs &A, const rbx_launch &L) {
    const DevBlock *blk = A.cat.blk;
    const bool hs = nh3_has_hs(A.abs.nh3);
    const int R = (A.abs.h2o ? rounds_of(blk[RBX_LINES_H2O].n) : 0) +
                  (A.abs.h2s ? rounds_of(blk[RBX_LINES_H2S].n) : 0) +
                  (hs ? rounds_of(blk[RBX_LINES_NH3_INV].n) + rounds_of(blk[RBX_LINES_NH3_ROT].n) +
                            rounds_of(blk[RBX_LINES_NH3_V2].n)
                      : 0) +
                  (A.abs.ph3 ? rounds_of(blk[RBX_LINES_PH3].n) : 0);
    if (L.far_split > 0) return std::min(L.far_split, kMaxSplit);
    const long long items = static_cast<long long>(A.nl) * ((A.tiles_total + kFarTiles - 1) / kFarTiles);
    const long long want = (4096 + items - 1) / items;
    return static_cast<int>(std::max(1LL, std::min<long long>(want, std::min(kMaxSplit, std::max(R, 1)))));
}

static size_t alpha_ws_bytes(const AlphaArgs &A, int precision, int nsplit, bool series) {
    const size_t plane = static_cast<size_t>(A.nl) * A.nf;
    const size_t extra = has_extra(A.abs) ? static_cast<size_t>(kNumExtra) * plane : 0;
    size_t b = (static_cast<size_t>(kSlots) * nsplit * plane + static_cast<size_t>(nsplit) * A.nl + extra) *
               sizeof(double);
    if (series)
        b += (static_cast<size_t>(A.nl) * A.far_split * A.tiles_total * kFarMom + static_cast<size_t>(A.far_split) * A.nl +
              static_cast<size_t>(A.nl) * kSegs * kPrepWords) *
                 sizeof(double) +
             static_cast<size_t>(A.nl) * A.tiles_total * A.far_words * sizeof(unsigned int);
    return b;
}

template <int K>
static int launch_alpha(AlphaArgs &A, int precision, int nsplit, hipStream_t st) {
    const int waves = kThreads / 64;
    dim3 grid(static_cast<unsigned>((A.tiles_total + waves - 1) / waves * static_cast<long long>(A.nl)), nsplit);
    if (precision == RBX_FP32)
        hipLaunchKernelGGL((alpha_lines_kernel<K, true>), grid, dim3(kThreads), 0, st, A);
    else
        hipLaunchKernelGGL((alpha_lines_kernel<K, false>), grid, dim3(kThreads), 0, st, A);
    int rc = launch_status("alpha_lines_kernel");
    if (rc) return rc;
    dim3 g2(static_cast<unsigned>((A.nf + kThreads - 1) / kThreads * static_cast<long long>(A.nl)));
    if (has_extra(A.abs)) {
        hipLaunchKernelGGL(alpha_extra_kernel, g2, dim3(kThreads), 0, st, A);
        rc = launch_status("alpha_extra_kernel");
        if (rc) return rc;
    }
    if (precision == RBX_FP32)
        hipLaunchKernelGGL((alpha_finish_kernel<true>), g2, dim3(kThreads), 0, st, A, nsplit);
    else
        hipLaunchKernelGGL((alpha_finish_kernel<false>), g2, dim3(kThreads), 0, st, A, nsplit);
    return launch_status("alpha_finish_kernel");
}

// the series path: classification pass, near sums per tile, then the same finish (one part)
template <int K, bool F32>
static int launch_series(AlphaArgs &A, hipStream_t st) {
    const long long groups = (A.tiles_total + kFarTiles - 1) / kFarTiles;
    hipLaunchKernelGGL(alpha_prep_kernel<F32>, dim3(static_cast<unsigned>((static_cast<long long>(A.nl) * kSegs + 255) / 256)),
                       dim3(256), 0, st, A);
    int rc = launch_status("alpha_prep_kernel");
    if (rc) return rc;
    dim3 g1(static_cast<unsigned>(groups * A.nl), static_cast<unsigned>(A.far_split));
    hipLaunchKernelGGL(alpha_far_kernel<F32>, g1, dim3(kThreads), 0, st, A);
    rc = launch_status("alpha_far_kernel");
    if (rc) return rc;
    dim3 g2(static_cast<unsigned>(static_cast<long long>(A.tiles_total) * A.nl));
    hipLaunchKernelGGL((alpha_near_kernel<K, F32>), g2, dim3(kThreads), 0, st, A);
    rc = launch_status("alpha_near_kernel");
    if (rc) return rc;
    dim3 g3(static_cast<unsigned>((A.nf + kThreads - 1) / kThreads * static_cast<long long>(A.nl)));
    if (has_extra(A.abs)) {
        hipLaunchKernelGGL(alpha_extra_kernel, g3, dim3(kThreads), 0, st, A);
        rc = launch_status("alpha_extra_kernel");
        if (rc) return rc;
    }
    hipLaunchKernelGGL((alpha_finish_kernel<F32>), g3, dim3(kThreads), 0, st, A, 1);
    return launch_status("alpha_finish_kernel");
}

template <int K>
static int launch_series(AlphaArgs &A, int precision, hipStream_t st) {
    return precision == RBX_FP32 ? launch_series<K, true>(A, st) : launch_series<K, false>(A, st);
}

// workspace of one rbx_alpha call: the tile width, split and path it will use
static int plan_alpha(AlphaArgs &A, const rbx_launch &L, int precision, int &K, int &nsplit, bool &series,
                      size_t &bytes) {
    K = alpha_tile_k(A.nf, L);
    set_wave_tiles(A, K);
    series = series_on(A, L);
    if (series) {
        far_layout(A);
        series = A.far_words > 0;   // no line lists selected: the plain path (one part)
    }
    if (series) {
        A.far_split = far_split(A, L);
        nsplit = 1;
        // the moments and near masks grow with layers x tiles x lines: past
        // kSeriesMaxWorkspace the built-in choice is the direct sums (launch.alpha_series = 2
        // forces the series)
        if (L.alpha_series != 2 && alpha_ws_bytes(A, precision, 1, true) > kSeriesMaxWorkspace)
            series = false;
    }
    if (!series) {
        int rc = alpha_split(A, L, K, precision, nsplit);
        if (rc) return rc;
    }
    bytes = alpha_ws_bytes(A, precision, nsplit, series);
    return RBX_OK;
}

// Lay the planned workspace out and launch (rbx_alpha and rbx_alpha_noalloc).
static int run_planned(AlphaArgs &A, int precision, int K, int nsplit, bool series, void *workspace, hipStream_t st) {
    const int nl = A.nl, nf = A.nf;
    A.ws = static_cast<double *>(workspace);
    A.ws_base = A.ws + static_cast<size_t>(kSlots) * nsplit * nl * nf;
    A.ws_extra = A.ws_base + static_cast<size_t>(nsplit) * nl;
    if (series) {
        const size_t plane = static_cast<size_t>(nl) * nf;
        A.far_mom = A.ws_extra + (has_extra(A.abs) ? static_cast<size_t>(kNumExtra) * plane : 0);
        A.far_base = A.far_mom + static_cast<size_t>(nl) * A.far_split * A.tiles_total * kFarMom;
        A.far_prep = A.far_base + static_cast<size_t>(A.far_split) * nl;
        A.near = reinterpret_cast<unsigned int *>(A.far_prep + static_cast<size_t>(nl) * kSegs * kPrepWords);
        switch (K) {
            case 4: return launch_series<4>(A, precision, st);
            case 2: return launch_series<2>(A, precision, st);
            default: return launch_series<1>(A, precision, st);
        }
    }
    switch (K) {
        case 4: return launch_alpha<4>(A, precision, nsplit, st);
        case 2: return launch_alpha<2>(A, precision, nsplit, st);
        default: return launch_alpha<1>(A, precision, nsplit, st);
    }
}

}  // namespace rbx

extern "C" int rbx_alpha_workspace_size(const rbx_catalog *cat, const rbx_absorbers *absorbers, int nl, int nf,
                                        int nf_nh3_low, int precision, const rbx_launch *launch, size_t *bytes) {
    RBX_CHECK_ARG(bytes, "rbx_alpha_workspace_size: null output");
    RBX_CHECK_ARG(precision == RBX_FP64 || precision == RBX_FP32, "rbx_alpha_workspace_size: bad precision %d",
                  precision);
    RBX_CHECK_ARG(cat && absorbers, "rbx_alpha_workspace_size: null catalogue or absorber selection");
    RBX_CHECK_ARG(nl > 0 && nf > 0 && static_cast<long long>(nl) * ((nf + 63) / 64) < (1LL << 31),
                  "rbx_alpha_workspace_size: nl=%d nf=%d", nl, nf);
    RBX_CHECK_ARG(nf_nh3_low >= 0 && nf_nh3_low <= nf, "rbx_alpha_workspace_size: nf_nh3_low=%d", nf_nh3_low);
    int rc = rbx::check_catalog(cat, *absorbers, "rbx_alpha_workspace_size");
    if (rc) return rc;
    rbx_launch L;
    rc = rbx::read_launch(launch, L, "rbx_alpha_workspace_size");
    if (rc) return rc;
    rbx::AlphaArgs A{};
    A.cat = cat->dev;
    A.abs = *absorbers;
    A.nl = nl;
    A.nf = nf;
    A.nf_low = nf_nh3_low;
    int K = 1, nsplit = 1;
    bool series = false;
    return rbx::plan_alpha(A, L, precision, K, nsplit, series, *bytes);
}

extern "C" int rbx_alpha(const rbx_catalog *cat, const rbx_absorbers *absorbers, const double *P, const double *T,
                         const double *mix, const double *cloud, int nl, const double *freq, const double *freq_nh3,
                         int nf,
                         int nf_nh3_low, const double *scale, int precision, void *alpha, double *by_constituent,
                         const rbx_launch *launch, void *workspace, size_t workspace_bytes, void *stream) {
    RBX_CHECK_ARG(precision == RBX_FP64 || precision == RBX_FP32, "rbx_alpha: bad precision %d", precision);
    rbx::AlphaArgs A{};
    int rc = rbx::fill_args(A, cat, absorbers, P, T, mix, cloud, nl, freq, freq_nh3, nf, nf_nh3_low, scale, alpha,
                            by_constituent, "rbx_alpha");
    if (rc) return rc;
    rbx_launch L;
    rc = rbx::read_launch(launch, L, "rbx_alpha");
    if (rc) return rc;
    int K = 1, nsplit = 1;
    bool series = false;
    size_t need = 0;
    rc = rbx::plan_alpha(A, L, precision, K, nsplit, series, need);
    if (rc) return rc;
    RBX_CHECK_ARG(workspace, "rbx_alpha: null workspace (rbx_alpha_workspace_size gives its size)");
    RBX_CHECK_ARG(workspace_bytes >= need, "rbx_alpha: workspace of %zu bytes, this call needs %zu", workspace_bytes,
                  need);
    RBX_CHECK_ARG(reinterpret_cast<uintptr_t>(workspace) % 8 == 0, "rbx_alpha: workspace not 8-byte aligned");
    return rbx::run_planned(A, precision, K, nsplit, series, workspace, static_cast<hipStream_t>(stream));
}

// ---- rbx_alpha without a workspace argument: the partial sums go to a static device array ----
// For callers that may not allocate (code under HIP graph capture, serving loops).  Every
// device that runs the library holds its own copy of `scratch` (RBX_NOALLOC_SCRATCH_BYTES)
// until the process ends; its device address is looked up once per device with
// hipGetSymbolAddress and used for every launch (the array's name is a host-side stand-in
// in host code).  Calls on one device share the array: two that overlap race.  No kernel
// of rbx_alpha reads a workspace entry before writing it, so nothing is cleared per call.
namespace rbx {

static __device__ __align__(256) unsigned char scratch[RBX_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxScratchDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_at[kMaxScratchDevices];

static int scratch_address(void **out) {
    int dev = 0;
    RBX_CHECK_HIP(hipGetDevice(&dev));
    RBX_CHECK_ARG(dev >= 0 && dev < kMaxScratchDevices, "rbx_alpha_noalloc: device %d beyond %d", dev,
                  kMaxScratchDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_at[dev] == nullptr) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr) {
            set_error("rbx_alpha_noalloc: hipGetSymbolAddress(scratch) on device %d failed: %s", dev,
                      hipGetErrorString(e));
            return RBX_ERR_HIP;
        }
        g_scratch_at[dev] = p;
    }
    *out = g_scratch_at[dev];
    return RBX_OK;
}

}  // namespace rbx

extern "C" int rbx_alpha_noalloc(const rbx_catalog *cat, const rbx_absorbers *absorbers, const double *P,
                                 const double *T, const double *mix, const double *cloud, int nl, const double *freq,
                                 const double *freq_nh3, int nf, int nf_nh3_low, const double *scale, int precision,
                                 void *alpha, double *by_constituent, const rbx_launch *launch, void *stream) {
    RBX_CHECK_ARG(precision == RBX_FP64 || precision == RBX_FP32, "rbx_alpha_noalloc: bad precision %d", precision);
    rbx::AlphaArgs A{};
    int rc = rbx::fill_args(A, cat, absorbers, P, T, mix, cloud, nl, freq, freq_nh3, nf, nf_nh3_low, scale, alpha,
                            by_constituent, "rbx_alpha_noalloc");
    if (rc) return rc;
    rbx_launch L;
    rc = rbx::read_launch(launch, L, "rbx_alpha_noalloc");
    if (rc) return rc;
    const size_t cap = RBX_NOALLOC_SCRATCH_BYTES;
    int K = 1, nsplit = 1;
    bool series = false;
    size_t need = 0;
    rc = rbx::plan_alpha(A, L, precision, K, nsplit, series, need);
    if (rc) return rc;
    if (need > cap && L.alpha_series != 2) {
        // the direct sums need far less: take them, with the most parts (up to the plan's) that fit
        rbx_launch D = L;
        D.alpha_series = 1;
        rc = rbx::plan_alpha(A, D, precision, K, nsplit, series, need);
        if (rc) return rc;
        for (int n = nsplit - 1; n >= 1 && need > cap && L.alpha_split == 0; --n) {
            D.alpha_split = n;
            rc = rbx::plan_alpha(A, D, precision, K, nsplit, series, need);
            if (rc) return rc;
        }
    }
    RBX_CHECK_ARG(need <= cap,
                  "rbx_alpha_noalloc: this call needs %zu bytes of scratch, more than its %zu-byte maximum "
                  "(rbx_alpha with a workspace has no such limit)",
                  need, cap);
    void *ws = nullptr;
    rc = rbx::scratch_address(&ws);
    if (rc) return rc;
    return rbx::run_planned(A, precision, K, nsplit, series, ws, static_cast<hipStream_t>(stream));
}

namespace rbx {

// Parts per layer of the tangent pass: launch.tan_split, else 1 when the (layer, tile)
// wavefronts alone give the device about 4 per SIMD, else enough parts for that (a
// 500-layer, 64-frequency request is 500 wavefronts on 1024 SIMDs: one part leaves half of
// them idle and the rest with no latency hiding).
static void plan_tangent(AlphaArgs &A, const rbx_launch &L) {
    const DevBlock *blk = A.cat.blk;
    const bool hs = nh3_has_hs(A.abs.nh3);
    const int n[kTanSegs] = {A.abs.h2o ? blk[RBX_LINES_H2O].n : 0, A.abs.h2s ? blk[RBX_LINES_H2S].n : 0,
                             hs ? blk[RBX_LINES_NH3_INV].n : 0,    hs ? blk[RBX_LINES_NH3_ROT].n : 0,
                             hs ? blk[RBX_LINES_NH3_V2].n : 0,     A.abs.ph3 ? blk[RBX_LINES_PH3].n : 0};
    int r = 0;
    for (int s = 0; s < kTanSegs; ++s) {
        A.tan_seg0[s] = r;
        r += tan_rounds(n[s]);
    }
    A.tan_rounds = r;
    int parts = L.tan_split;
    if (parts == 0) {
        const long long items = static_cast<long long>(A.nl) * ((A.nf + 63) / 64);
        const long long want = 4 * 1024;   // wavefronts: 4 per SIMD of a 256-CU device
        parts = static_cast<int>(std::min<long long>(kMaxLaunchSplit, (want + items - 1) / items));
    }
    A.tan_parts = std::max(1, std::min(parts, std::max(r, 1)));
}

static size_t tangent_ws_doubles_before_prep(const AlphaArgs &A) {
    const size_t plane = static_cast<size_t>(A.nl) * A.nf;
    return static_cast<size_t>(A.tan_parts) * kTanSlots * kTanComps * plane +
           static_cast<size_t>(A.tan_parts) * kTanComps * A.nl;
}
// partial sums of a split pass (none unsplit), then the derivers' table
static size_t tangent_ws_bytes(const AlphaArgs &A) {
    const size_t parts = A.tan_parts > 1 ? tangent_ws_doubles_before_prep(A) : 0;
    return (parts + static_cast<size_t>(A.nl) * kTanPrepWords) * sizeof(double);
}

static int tangent_args(AlphaArgs &A, rbx_launch &L, const rbx_catalog *cat, const rbx_absorbers *absorbers,
                        int nl, int nf, int nf_nh3_low, int precision, const rbx_launch *launch, const char *who) {
    RBX_CHECK_ARG(precision == RBX_FP64 || precision == RBX_FP32, "%s: bad precision %d", who, precision);
    RBX_CHECK_ARG(cat && absorbers, "%s: null catalogue or absorber selection", who);
    RBX_CHECK_ARG(!has_extra(*absorbers), "%s: co_ddb, clouds_idp and h2_orton have no tangent; select them off",
                  who);
    RBX_CHECK_ARG(nl > 0 && nf > 0 && static_cast<long long>(nl) * ((nf + 63) / 64) < (1LL << 31),
                  "%s: nl=%d nf=%d", who, nl, nf);
    RBX_CHECK_ARG(nf_nh3_low >= 0 && nf_nh3_low <= nf, "%s: nf_nh3_low=%d", who, nf_nh3_low);
    int rc = check_catalog(cat, *absorbers, who);
    if (rc) return rc;
    rc = read_launch(launch, L, who);
    if (rc) return rc;
    A.cat = cat->dev;
    A.abs = *absorbers;
    A.nl = nl;
    A.nf = nf;
    A.nf_low = nf_nh3_low;
    plan_tangent(A, L);
    return RBX_OK;
}

}  // namespace rbx

extern "C" int rbx_alpha_tangent_workspace_size(const rbx_catalog *cat, const rbx_absorbers *absorbers, int nl,
                                                int nf, int nf_nh3_low, int precision, const rbx_launch *launch,
                                                size_t *bytes) {
    RBX_CHECK_ARG(bytes, "rbx_alpha_tangent_workspace_size: null output");
    rbx::AlphaArgs A{};
    rbx_launch L;
    int rc = rbx::tangent_args(A, L, cat, absorbers, nl, nf, nf_nh3_low, precision, launch,
                               "rbx_alpha_tangent_workspace_size");
    if (rc) return rc;
    *bytes = rbx::tangent_ws_bytes(A);
    return RBX_OK;
}

extern "C" int rbx_alpha_tangent(const rbx_catalog *cat, const rbx_absorbers *absorbers, const double *P,
                                 const double *T, const double *mix, int nl, const double *freq,
                                 const double *freq_nh3, int nf, int nf_nh3_low, const double *scale, int precision,
                                 double *alpha, double *dalpha, const rbx_launch *launch, void *workspace,
                                 size_t workspace_bytes, void *stream) {
    rbx::AlphaArgs A{};
    rbx_launch L;
    int rc = rbx::tangent_args(A, L, cat, absorbers, nl, nf, nf_nh3_low, precision, launch, "rbx_alpha_tangent");
    if (rc) return rc;
    // (fill_args sets the call's arrays and leaves the tangent plan as it is)
    rc = rbx::fill_args(A, cat, absorbers, P, T, mix, nullptr, nl, freq, freq_nh3, nf, nf_nh3_low, scale, dalpha,
                        nullptr, "rbx_alpha_tangent");
    if (rc) return rc;
    const int parts = A.tan_parts;
    A.value = alpha;
    const size_t need = rbx::tangent_ws_bytes(A);
    RBX_CHECK_ARG(workspace && workspace_bytes >= need,
                  "rbx_alpha_tangent: workspace of %zu bytes, this call needs %zu "
                  "(rbx_alpha_tangent_workspace_size)",
                  workspace_bytes, need);
    RBX_CHECK_ARG(reinterpret_cast<uintptr_t>(workspace) % 8 == 0, "rbx_alpha_tangent: workspace not 8-byte aligned");
    A.ws = static_cast<double *>(workspace);
    A.tan_prep = A.ws + (parts > 1 ? rbx::tangent_ws_doubles_before_prep(A) : 0);
    // line-summing passes: a block per layer and group of W tiles (W = 1, 2 or 4), so the
    // block derives each line once for its W tiles; the finish: 4 independent items a block
    const int tiles = (nf + 63) / 64;
    const int W = tiles >= 3 ? 4 : tiles;
    const long long groups = static_cast<long long>(nl) * ((tiles + W - 1) / W);
    const long long items = static_cast<long long>(nl) * tiles;
    const long long fblocks = (items + rbx::kThreads / 64 - 1) / (rbx::kThreads / 64);
    RBX_CHECK_ARG(groups < (1LL << 31) && fblocks < (1LL << 31), "rbx_alpha_tangent: too many layers x frequencies");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const dim3 gs(static_cast<unsigned>(groups), static_cast<unsigned>(std::max(parts, 1))), bs(64 * W);
    const bool f32 = precision == RBX_FP32;
    hipLaunchKernelGGL(rbx::tangent_prep_kernel, dim3(static_cast<unsigned>((static_cast<long long>(nl) * rbx::kTanSegs + 255) / 256)),
                       dim3(256), 0, st, A);
    rc = rbx::launch_status("tangent_prep_kernel");
    if (rc) return rc;
#define RBX_TAN_LAUNCH(MODE)                                                                                    \
    do {                                                                                                       \
        if (f32) {                                                                                             \
            if (W == 4) hipLaunchKernelGGL((rbx::alpha_tangent_kernel_f32<MODE, 4>), gs, bs, 0, st, A);        \
            else if (W == 2) hipLaunchKernelGGL((rbx::alpha_tangent_kernel_f32<MODE, 2>), gs, bs, 0, st, A);   \
            else hipLaunchKernelGGL((rbx::alpha_tangent_kernel_f32<MODE, 1>), gs, bs, 0, st, A);               \
        } else {                                                                                               \
            if (W == 4) hipLaunchKernelGGL((rbx::alpha_tangent_kernel_f64<MODE, 4>), gs, bs, 0, st, A);        \
            else if (W == 2) hipLaunchKernelGGL((rbx::alpha_tangent_kernel_f64<MODE, 2>), gs, bs, 0, st, A);   \
            else hipLaunchKernelGGL((rbx::alpha_tangent_kernel_f64<MODE, 1>), gs, bs, 0, st, A);               \
        }                                                                                                      \
    } while (0)
    if (parts <= 1) {
        RBX_TAN_LAUNCH(rbx::kTanAll);
        return rbx::launch_status("alpha_tangent_kernel");
    }
    RBX_TAN_LAUNCH(rbx::kTanPart);
#undef RBX_TAN_LAUNCH
    rc = rbx::launch_status("alpha_tangent_kernel (parts)");
    if (rc) return rc;
    const dim3 g1(static_cast<unsigned>(fblocks));
    if (f32)
        hipLaunchKernelGGL((rbx::alpha_tangent_kernel_f32<rbx::kTanFinish, 4>), g1, dim3(rbx::kThreads), 0, st, A);
    else
        hipLaunchKernelGGL((rbx::alpha_tangent_kernel_f64<rbx::kTanFinish, 4>), g1, dim3(rbx::kThreads), 0, st, A);
    return rbx::launch_status("alpha_tangent_kernel (finish)");
}

// ---- reuse path: a saved per-constituent absorption table, rescaled ----
// alpha[l][f] = sum_c scale[c][l] absorb[l][f][c] in the saved column order (the
// reference's Alpha.alpha_data layout, alpha.py:110-192): one thread per
// (layer, frequency), HBM-bound.
namespace rbx {
template <bool F32>
__global__ __launch_bounds__(256) void alpha_rescale_kernel(const double *absorb, int nl, int nf, int nc,
                                                            const double *scale, void *alpha) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= static_cast<long long>(nl) * nf) return;
    const int l = static_cast<int>(i / nf);
    const double *a = absorb + i * nc;
    double t = 0.0;
    for (int c = 0; c < nc; ++c) t += a[c] * (scale ? scale[static_cast<size_t>(c) * nl + l] : 1.0);
    if (F32)
        static_cast<float *>(alpha)[i] = static_cast<float>(t);
    else
        static_cast<double *>(alpha)[i] = t;
}
}  // namespace rbx

extern "C" int rbx_alpha_rescale(const double *absorb, int nl, int nf, int nc, const double *scale, int precision,
                                 void *alpha, void *stream) {
    RBX_CHECK_ARG(absorb && alpha, "rbx_alpha_rescale: null device pointer");
    RBX_CHECK_ARG(nl > 0 && nf > 0 && nc > 0 && nc <= 64, "rbx_alpha_rescale: nl=%d nf=%d nc=%d", nl, nf, nc);
    RBX_CHECK_ARG(precision == RBX_FP64 || precision == RBX_FP32, "rbx_alpha_rescale: bad precision %d", precision);
    const long long n = static_cast<long long>(nl) * nf;
    const long long blocks = (n + 255) / 256;
    RBX_CHECK_ARG(blocks < (1LL << 31), "rbx_alpha_rescale: table too large");
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (precision == RBX_FP32)
        hipLaunchKernelGGL(rbx::alpha_rescale_kernel<true>, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, st,
                           absorb, nl, nf, nc, scale, alpha);
    else
        hipLaunchKernelGGL(rbx::alpha_rescale_kernel<false>, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, st,
                           absorb, nl, nf, nc, scale, alpha);
    return rbx::launch_status("alpha_rescale_kernel");
}
