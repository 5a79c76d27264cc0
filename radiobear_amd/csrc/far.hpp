// Forward line sums with far-line series, FP32 and FP64 modes (included by alpha.hip, after
// the line-sum kernel it replaces when the series are on).
//
// A line's term is a rational function of x = f^2 with two poles p (the roots of its
// denominator x^2 - 2 Ds x + P0, P0 = Ds^2 + h):
//   term(x) = (a x + c) / (x^2 - 2 Ds x + P0).
// Over a wavefront tile whose columns span x in [lo, hi]:
//   * every |p| <= kFarRatio lo ("below"): term = sum_k m_k x^-(k+1), with
//     m_0 = a, m_1 = c + 2 Ds a, m_k = 2 Ds m_{k-1} - P0 m_{k-2};
//   * every |p| >= hi / kFarRatio ("above"): term = sum_k n_k x^k, with
//     n_0 = c / P0, n_1 = (a + 2 Ds n_0) / P0, n_k = (2 Ds n_{k-1} - n_{k-2}) / P0
// (the recurrences match powers of x in term * denominator = a x + c).  kFarTerms = 16
// terms at kFarRatio = 1/3 leave ~4e-8 of a far line's value for two well-separated
// poles and at most 8e-7 at the ratio limit for a near-double pole of a Gross line
// (a = 0: 15 terms carry it, with coefficients growing like (k + 1) p^k); far lines are
// small next to the near ones, and the tables stay within 3e-7 of FP64 like the direct
// FP32 sums.  On 800 layers x 1024 frequencies (0.3-300 GHz, 128-column tiles) about
// 30 % of the (line, tile) pairs are neither: only those are summed per column.  (12
// terms at 1/4 leave more near pairs, 33 %, and were 2 % slower overall.)  FP64 mode uses
// the same 16 terms at ratio 0.15 (at most ~8e-12 of a far line's value) and sums its near
// lines in FP64 from their (D, a, c, g, G), as the direct FP64 sums do.
//
// alpha_prep_kernel (one thread per layer and segment) first writes every layer's
// coefficient derivers to the workspace; both passes below read them from there.
// alpha_far_kernel (block = layer x group of kFarTiles tiles x part of the layer's
// staging rounds; each wavefront a contiguous run of 64-line blocks): each lane derives
// one line's coefficients per coefficient set (once per line and layer, its catalogue
// columns fetched a block ahead), classifies it against every tile of the group (one
// ballot per tile gives the tiles' near-mask words) and adds its series to running sums
// keyed by the first tile it is below for and the last it is above for (prefix and
// suffix sums over the tiles give each tile's moments).  The h2o base term is summed
// here.
// alpha_near_kernel (block = tile x layer, all four wavefronts on the tile's
// columns): lists the tile's near lines from its masks, derives each listed line's
// coefficients again from its catalogue row (the catalogue stays in L2; no per-layer copy
// of the lines goes through HBM), sums them as the line-sum kernel does with the lines
// split between the wavefronts, adds the wavefronts in order and the series of the summed
// moments, and writes the same per-slot partials the finish kernel reads (one part).

constexpr int kFarTerms = 16;
// pole ratio of a far line: 1/3 in FP32 mode, 0.15 in FP64 mode (at most ~8e-12 of a far
// line's value there, against the 1e-9 FP64 parity bar)
template <bool F32>
__host__ __device__ constexpr double far_ratio() { return F32 ? 1.0 / 3.0 : 0.15; }
constexpr int kFarTiles = 8;                          // tiles per alpha_far_kernel block
constexpr int kFarSets = 6;                           // h2o, h2s, inversion set 0, set 1, rot + v2, ph3
constexpr int kFarMom = kFarSets * 2 * kFarTerms;     // moments per (layer, part, tile)
constexpr int kSegStream[kSegs] = {0, 1, 2, 4, 5, 6};
constexpr int kSegFarSet[kSegs] = {0, 1, 2, 4, 4, 5};
constexpr int kSegVar[kSegs] = {0, 0, 1, 1, 1, 0};    // columns' x (0) or x for NH3 (1)
constexpr int kListWords = 128;                       // mask words listed per window
constexpr int kFarRanges = 4;                         // x; x NH3; x NH3 of set 0 columns; of set 1

__device__ inline void tile_cols(const AlphaArgs &A, int tile, int &c0, int &c1) {
    c0 = tile * A.tile_width;
    c1 = min(c0 + A.tile_width, A.nf);
}

// inversion coefficient sets a tile's columns use: bit 0 set 0 (columns < nf_low), bit 1 set 1
__device__ inline int tile_sets(const AlphaArgs &A, int c0, int c1) {
    return (c0 < min(c1, A.nf_low) ? 1 : 0) | (max(c0, A.nf_low) < c1 ? 2 : 0);
}

// [lo, hi] of x, x for NH3, and x for NH3 over the set 0 / set 1 columns of a tile
// (called by a whole wavefront; the results are wavefront-uniform, +inf/-inf when empty)
__device__ inline void tile_ranges(const AlphaArgs &A, int tile, double (&lo)[kFarRanges],
                                   double (&hi)[kFarRanges]) {
#pragma unroll
    for (int v = 0; v < kFarRanges; ++v) {
        lo[v] = INFINITY;
        hi[v] = -INFINITY;
    }
    int c0, c1;
    tile_cols(A, tile, c0, c1);
    for (int c = c0 + static_cast<int>(threadIdx.x % 64); c < c1; c += 64) {
        const double f = A.freq[c];
        const double fn = A.freq_nh3 ? A.freq_nh3[c] : f;
        const double xs[kFarRanges] = {f * f, fn * fn, c < A.nf_low ? fn * fn : NAN, c < A.nf_low ? NAN : fn * fn};
#pragma unroll
        for (int v = 0; v < kFarRanges; ++v) {
            lo[v] = fmin(lo[v], xs[v]);   // fmin / fmax skip the NaN placeholders
            hi[v] = fmax(hi[v], xs[v]);
        }
    }
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) {
#pragma unroll
        for (int v = 0; v < kFarRanges; ++v) {
            lo[v] = fmin(lo[v], __shfl_xor(lo[v], m));
            hi[v] = fmax(hi[v], __shfl_xor(hi[v], m));
        }
    }
}

// moduli (smaller rs, larger rb) of the two roots of x^2 - 2 Ds x + P0, P0 = Ds^2 + h:
// a complex pair of modulus sqrt(P0) when h >= 0, else |Ds| -+ sqrt(-h); in FP32 from the
// FP64 coefficients (the class tests compare FP32 values; NaN coefficients give NaN
// moduli, near everywhere)
__device__ inline void pole_moduli_f32(double Ds, double h, double P0, float &rs, float &rb) {
    const float m = fabsf(static_cast<float>(Ds)), q = __builtin_amdgcn_sqrtf(fabsf(static_cast<float>(h)));
    const float r = __builtin_amdgcn_sqrtf(static_cast<float>(P0));
    const bool complex_pair = h >= 0.0;
    rb = complex_pair ? r : m + q;
    rs = complex_pair ? r : fabsf(m - q);
    if (!(h == h)) rs = rb = NAN;
}

// a copy of a wavefront-uniform value held in scalar registers (each 32-bit word taken from
// the first active lane)
template <class T>
__device__ __forceinline__ T uniform(const T &v) {
    static_assert(sizeof(T) % 4 == 0, "whole 32-bit words");
    T out = v;
    const int *src = reinterpret_cast<const int *>(&v);
    int *dst = reinterpret_cast<int *>(&out);
#pragma unroll
    for (int w = 0; w < static_cast<int>(sizeof(T) / 4); ++w) dst[w] = __builtin_amdgcn_readfirstlane(src[w]);
    return out;
}

// b when second, else a, chosen word by word: choosing between the two objects' addresses
// instead makes the compiler keep both in scratch and read the chosen one from there
template <class T>
__device__ __forceinline__ T pick(const T &a, const T &b, bool second) {
    static_assert(sizeof(T) % 4 == 0, "whole 32-bit words");
    T out = a;
    const int *wa = reinterpret_cast<const int *>(&a), *wb = reinterpret_cast<const int *>(&b);
    int *wo = reinterpret_cast<int *>(&out);
#pragma unroll
    for (int w = 0; w < static_cast<int>(sizeof(T) / 4); ++w) wo[w] = second ? wb[w] : wa[w];
    return out;
}

// The segments' coefficient derivers per layer (the layer's powers of T and partial pressures,
// computed once by alpha_prep_kernel): [nl][kSegs][kPrepWords] doubles, set 0's deriver in
// words 0-7, set 1's (the same one outside the nh3_hs inversion lines) in 8-15, and their FP32
// weight scales (1 in FP64 mode) in 16 and 17.  The passes read a layer's row with uniform loads.
constexpr int kPrepWords = 18;
template <int S>
using SegC = std::integral_constant<int, S>;

template <bool F32>
__global__ __launch_bounds__(256) void alpha_prep_kernel(AlphaArgs A) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= A.nl * kSegs) return;
    const int layer = e / kSegs, seg = e % kSegs;
    const LayerT<double> L = layer_of(A, layer);
    const DevBlock *blk = A.cat.blk;
    double *row = A.far_prep + static_cast<size_t>(e) * kPrepWords;
    auto put = [&](const auto &p0, const auto &p1, double s0, double s1) {
        using Prep = std::decay_t<decltype(p0)>;
        static_assert(sizeof(Prep) <= 8 * sizeof(double) && sizeof(Prep) % sizeof(double) == 0,
                      "a deriver is at most 8 doubles");
        const double *w0 = reinterpret_cast<const double *>(&p0), *w1 = reinterpret_cast<const double *>(&p1);
        for (int k = 0; k < 8; ++k) {
            row[k] = k < static_cast<int>(sizeof(Prep) / sizeof(double)) ? w0[k] : 0.0;
            row[8 + k] = k < static_cast<int>(sizeof(Prep) / sizeof(double)) ? w1[k] : 0.0;
        }
        row[16] = F32 ? s0 : 1.0;
        row[17] = F32 ? s1 : 1.0;
    };
    switch (seg) {
        case 0: {
            const H2oPrep<double> p(L);
            put(p, p, pow2_scale(p.pref * blk[RBX_LINES_H2O].wscale), 1.0);
            break;
        }
        case 1: {
            const H2sPrep<double> p(L);
            put(p, p, pow2_scale(p.pref * blk[RBX_LINES_H2S].wscale), 1.0);
            break;
        }
        case 2: {
            const InvPrep<double> p0(L, nh3_low_set(A.abs.nh3)), p1(L, 1);
            const double w = blk[RBX_LINES_NH3_INV].wscale;
            put(p0, p1, pow2_scale(p0.pref * w), pow2_scale(p1.pref * w));
            break;
        }
        case 3: {
            const RotPrep<double> p(L);
            put(p, p, pow2_scale(p.pref * blk[RBX_LINES_NH3_ROT].wscale), 1.0);
            break;
        }
        case 4: {
            const V2Prep<double> p(L);
            put(p, p, pow2_scale(p.pref * blk[RBX_LINES_NH3_V2].wscale), 1.0);
            break;
        }
        default: {
            const Ph3Prep<double> p(L);
            put(p, p, pow2_scale(p.pref * blk[RBX_LINES_PH3].wscale), 1.0);
            break;
        }
    }
}

// a deriver from its table words (in scalar registers)
template <class Prep>
__device__ __forceinline__ Prep load_prep(const double *w) {
    struct Raw {
        double v[sizeof(Prep) / sizeof(double)];
    } r;
#pragma unroll
    for (int k = 0; k < static_cast<int>(sizeof(Prep) / sizeof(double)); ++k) r.v[k] = w[k];
    return uniform(__builtin_bit_cast(Prep, r));
}

// f(segment as a compile-time constant, set 0 deriver, set 1 deriver, scale 0, scale 1) with
// the layer's derivers of the segment from the table
template <class F>
__device__ __forceinline__ void with_prep(const AlphaArgs &A, int layer, int seg, F &&f) {
    const double *row = A.far_prep + (static_cast<size_t>(layer) * kSegs + seg) * kPrepWords;
    const double s0 = __builtin_amdgcn_readfirstlane(row[16]), s1 = __builtin_amdgcn_readfirstlane(row[17]);
    switch (seg) {
        case 0: {
            const auto p = load_prep<H2oPrep<double>>(row);
            f(SegC<0>{}, p, p, s0, s1);
            break;
        }
        case 1: {
            const auto p = load_prep<H2sPrep<double>>(row);
            f(SegC<1>{}, p, p, s0, s1);
            break;
        }
        case 2: {
            const auto p0 = load_prep<InvPrep<double>>(row), p1 = load_prep<InvPrep<double>>(row + 8);
            f(SegC<2>{}, p0, p1, s0, s1);
            break;
        }
        case 3: {
            const auto p = load_prep<RotPrep<double>>(row);
            f(SegC<3>{}, p, p, s0, s1);
            break;
        }
        case 4: {
            const auto p = load_prep<V2Prep<double>>(row);
            f(SegC<4>{}, p, p, s0, s1);
            break;
        }
        default: {
            const auto p = load_prep<Ph3Prep<double>>(row);
            f(SegC<5>{}, p, p, s0, s1);
            break;
        }
    }
}

// coefficient set v of line i of a segment from its catalogue row (both passes derive a
// line's coefficients this way; `base` collects the h2o base term)
template <class Prep, class Row>
__device__ __forceinline__ CoefT<double> seg_coef(const Prep &p0, const Prep &p1, int v, const Row &row, int i,
                                                  double &base) {
    if constexpr (Prep::kShared) {
        const double w = p0.weight(row, i);
        return v ? p1.coef(row, i, w) : p0.coef(row, i, w);
    } else {
        return p0(row, i, base);
    }
}

// ---- pass 1: far-line series and near masks ----
//
// A line's series coefficients (m_k below, n_k above) do not depend on the tile, only its
// class does: the line is below for the tiles with q lo >= its larger pole modulus and
// above for those with hi / q <= its smaller one.  Ranked by q lo, the tiles a line is
// below for are the last ones (all ranks >= kb, kb = the number of tiles with q lo under
// the modulus); ranked by hi / q, the tiles it is above for are the first ka.  So each
// thread derives its own line's series once per coefficient set and adds them into a
// difference array over the ranks (entry kb of the below array, entry ka of the above
// one); the moments of a tile are the prefix (below) or suffix (above) sums at its rank.
// The line lists are sorted by centre, so a thread's successive lines mostly share kb
// and ka: it keeps running sums in registers for its current kb and ka and adds them to
// the block's difference arrays in LDS (FP64 atomics) only when they change.  Near bits
// come from one wavefront ballot per tile.

constexpr int kFarWaves = kThreads / 64;

struct FarShared {
    // per wavefront: difference arrays, below (entry kb) and above (entry ka)
    double dB[kFarWaves][kFarTiles + 1][kFarTerms];
    double dA[kFarWaves][kFarTiles + 1][kFarTerms];
    double mom[kFarTiles][kFarMom];
    double lo[kFarTiles][kFarRanges], hi[kFarTiles][kFarRanges];
    double bsum[kThreads / 64];
    alignas(16) float thr[2][kFarTiles];   // current coefficient set: q lo, hi / q (+inf: unused)
};

// a used tile's class thresholds for range rv: q lo (below) and hi / q (above), as FP32
template <bool F32>
__device__ inline float far_threshold(const FarShared &sh, int j, int rv, bool above) {
    const double q = far_ratio<F32>();
    return above ? static_cast<float>(sh.hi[j][rv] / q) : static_cast<float>(q * sh.lo[j][rv]);
}

// rank of used tile j among the used tiles by its threshold (ties by index)
template <bool F32>
__device__ inline int tile_rank(const FarShared &sh, unsigned used, int j, int rv, bool above) {
    const float vj = far_threshold<F32>(sh, j, rv, above);
    int r = 0;
#pragma unroll
    for (int i = 0; i < kFarTiles; ++i) {
        const float vi = far_threshold<F32>(sh, i, rv, above);
        r += ((used >> i) & 1u) && (vi < vj || (vi == vj && i < j)) ? 1 : 0;
    }
    return r;
}


// The lanes with `pred` add their running sums acc into entry e (their own) of the
// wavefront's difference array d: for each entry among them, the 16 terms are summed over
// its lanes by a fixed reduce-scatter (halving the terms each lane carries at xor 32, 16, 8
// and 4, then adding over xor 2 and 1: term 8 b5 + 4 b4 + 2 b3 + b2 ends in the lanes with
// those lane bits), so the result does not depend on timing.  Called by the whole wavefront.
__device__ __forceinline__ void far_flush(double (&d)[kFarTiles + 1][kFarTerms], bool pred, int e,
                                          const double (&acc)[kFarTerms]) {
    static_assert(kFarTerms == 16, "the reduce-scatter below is written for 16 terms");
    unsigned long long rem = __ballot(pred);
    const int lane = threadIdx.x % 64;
    const bool b5 = lane & 32, b4 = lane & 16, b3 = lane & 8, b2 = lane & 4;
    while (rem != 0ull) {
        const int b = __builtin_amdgcn_readlane(e, __builtin_ctzll(rem));
        const bool sel = pred && e == b;
        rem &= ~__ballot(sel);
        double w8[8], w4[4], w2[2];
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const double lo = sel ? acc[t] : 0.0, hi = sel ? acc[t + 8] : 0.0;
            w8[t] = (b5 ? hi : lo) + __shfl_xor(b5 ? lo : hi, 32);
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) w4[t] = (b4 ? w8[t + 4] : w8[t]) + __shfl_xor(b4 ? w8[t] : w8[t + 4], 16);
#pragma unroll
        for (int t = 0; t < 2; ++t) w2[t] = (b3 ? w4[t + 2] : w4[t]) + __shfl_xor(b3 ? w4[t] : w4[t + 2], 8);
        double w1 = (b2 ? w2[1] : w2[0]) + __shfl_xor(b2 ? w2[0] : w2[1], 4);
        w1 += __shfl_xor(w1, 2);
        w1 += __shfl_xor(w1, 1);
        if ((lane & 3) == 0) d[b][(b5 ? 8 : 0) + (b4 ? 4 : 0) + (b3 ? 2 : 0) + (b2 ? 1 : 0)] += w1;
    }
}

template <bool F32>
__global__ __launch_bounds__(kThreads, 3) void alpha_far_kernel(AlphaArgs args) {
    // the arguments read where the launch put them (explicit kernel arguments open the
    // kernarg segment): a private copy of the struct would be spilled to scratch per lane
#if defined(__HIP_DEVICE_COMPILE__)
    const AlphaArgs &A = *reinterpret_cast<const AlphaArgs *>(__builtin_amdgcn_kernarg_segment_ptr());
#else
    const AlphaArgs &A = args;   // (host pass: never run)
#endif
    __shared__ FarShared sh;
    const int ngroups = (A.tiles_total + kFarTiles - 1) / kFarTiles;
    const int layer = blockIdx.x / ngroups, g = blockIdx.x % ngroups;
    const int part = blockIdx.y, nsplit = gridDim.y;
    const int tid = threadIdx.x, lane = tid % 64, wave = tid / 64;
    const LayerT<double> L = layer_of(A, layer);

    for (int j = wave; j < kFarTiles; j += kThreads / 64) {
        double lo[kFarRanges], hi[kFarRanges];
        tile_ranges(A, g * kFarTiles + j, lo, hi);   // empty for tiles past the last
        if (lane == 0) {
#pragma unroll
            for (int v = 0; v < kFarRanges; ++v) {
                sh.lo[j][v] = lo[v];
                sh.hi[j][v] = hi[v];
            }
        }
    }
    for (int e = tid; e < kFarTiles * kFarMom; e += kThreads) (&sh.mom[0][0])[e] = 0.0;
    for (int e = tid; e < kFarWaves * (kFarTiles + 1) * kFarTerms; e += kThreads) {
        (&sh.dB[0][0][0])[e] = 0.0;
        (&sh.dA[0][0][0])[e] = 0.0;
    }
    int gc0, gc1;
    tile_cols(A, g * kFarTiles, gc0, gc1);
    gc1 = min(A.nf, (g + 1) * kFarTiles * A.tile_width);
    const int gsets = tile_sets(A, gc0, gc1);   // inversion sets any tile of the group needs
    const int ntiles = min(kFarTiles, A.tiles_total - g * kFarTiles);
    unsigned tile_set_bits[2] = {0u, 0u};       // tiles of the group whose columns use set 0 / 1
    for (int jj = 0; jj < ntiles; ++jj) {
        int c0, c1;
        tile_cols(A, g * kFarTiles + jj, c0, c1);
        const int ts = tile_sets(A, c0, c1);
        tile_set_bits[0] |= (ts & 1) ? 1u << jj : 0u;
        tile_set_bits[1] |= (ts & 2) ? 1u << jj : 0u;
    }
    __syncthreads();

    const LinePlan plan(A, L.P);
    const int r0 = plan.begin(part, nsplit), r1 = plan.begin(part + 1, nsplit);
    double base = 0.0;
    int off = 0;
#pragma unroll 1
    for (int seg = 0; seg < kSegs; ++seg) {
        const int a0 = max(r0, off), b0 = min(r1, off + plan.nr[seg]);
        const int seg_off = off;
        off += plan.nr[seg];
        if (a0 >= b0) continue;
        with_prep(A, layer, seg, [&](auto segc, const auto &p0, const auto &p1, double s0, double s1) {
            constexpr int SEG = decltype(segc)::value;
            constexpr bool inv = SEG == 2;
            const DevBlock &B = A.cat.blk[kSegKind[SEG]];
            const int need = inv ? gsets : 1;
            (void)s0;
            (void)s1;
#pragma unroll 1
            for (int v = 0; v < 2; ++v) {
                if (!((need >> v) & 1)) continue;
                const unsigned used = inv ? tile_set_bits[v] : (1u << ntiles) - 1u;
                const int nv = __popc(used);
                const int rv = kSegVar[SEG] == 0 ? 0 : (inv ? 2 + v : 1);
                // thresholds of the used tiles; unused ones never count (+inf: a finite
                // modulus is below it and never at or above it)
                if (tid < 2 * kFarTiles) {
                    const int j = tid % kFarTiles;
                    sh.thr[tid / kFarTiles][j] = (used >> j) & 1u ? far_threshold<F32>(sh, j, rv, tid >= kFarTiles) : INFINITY;
                }
                __syncthreads();
                const int soff = A.far_off[kSegStream[SEG] + ((inv && v) ? 1 : 0)];
                unsigned int *mrow = A.near + (static_cast<size_t>(layer) * A.tiles_total + g * kFarTiles) * A.far_words + soff;
                double accB[kFarTerms], accA[kFarTerms];
#pragma unroll
                for (int k = 0; k < kFarTerms; ++k) accB[k] = accA[k] = 0.0;
                int cbB = -1, cbA = -1;   // entries the running sums belong to (-1: none yet)
                // wavefront w takes a contiguous run of the part's 64-line blocks (its lanes'
                // successive lines are 64 apart, so their entries change seldom)
                const int nblk = b0 - a0;   // blocks per wavefront (kThreads / 64 per round)
                const int blk0 = (a0 - seg_off) * kFarWaves + wave * nblk;
                // the set-v deriver, chosen once for the part's lines
                const auto pv = pick(p0, p1, v != 0);
                // each line's catalogue columns are fetched one block ahead
                LineRow<kNumCols[kSegKind[SEG]]> next;
                next.load(B, min(blk0 * 64 + lane, B.n - 1));
#pragma unroll 1
                for (int blk = blk0; blk < blk0 + nblk; ++blk) {
                    const int i = blk * 64 + lane;
                    const bool valid = i < B.n;
                    const auto row = next;
                    if (blk + 1 < blk0 + nblk) next.load(B, min(i + 64, B.n - 1));
                    CoefT<double> c = {0.0, 0.0, 0.0, 0.0, 1.0};
                    if (valid) c = seg_coef(p0, pv, 1, row, i, base);
                    double Ds, h;
                    shifted_denominator(c, Ds, h);
                    const double P0 = fma(Ds, Ds, h);
                    // class against every used tile: below when rb <= q lo, else above when
                    // rs >= hi / q, else near (NaN moduli: near everywhere)
                    float frs, frb;
                    pole_moduli_f32(Ds, h, P0, frs, frb);
                    float loq[kFarTiles], hiq[kFarTiles];
#pragma unroll
                    for (int j = 0; j < kFarTiles; j += 4) {
                        const float4 l4 = *reinterpret_cast<const float4 *>(&sh.thr[0][j]);
                        const float4 h4 = *reinterpret_cast<const float4 *>(&sh.thr[1][j]);
                        loq[j] = l4.x; loq[j + 1] = l4.y; loq[j + 2] = l4.z; loq[j + 3] = l4.w;
                        hiq[j] = h4.x; hiq[j + 1] = h4.y; hiq[j + 2] = h4.z; hiq[j + 3] = h4.w;
                    }
                    // (unused tiles: both thresholds +inf, never counted for finite moduli; their
                    // near bits are not stored)
                    int kb = 0, ka = 0;
                    unsigned long long near[kFarTiles];
#pragma unroll
                    for (int j = 0; j < kFarTiles; ++j) {
                        const bool below = frb <= loq[j], above = frs >= hiq[j];
                        kb += below ? 0 : 1;
                        ka += above ? 1 : 0;
                        near[j] = __builtin_amdgcn_ballot_w64(valid && !below && !above);
                    }
                    if (lane == 0) {   // the block's two mask words of every used tile
#pragma unroll
                        for (int j = 0; j < kFarTiles; ++j)
                            if ((used >> j) & 1u)
                                *reinterpret_cast<uint2 *>(mrow + static_cast<size_t>(j) * A.far_words + 2 * blk) =
                                    make_uint2(static_cast<unsigned int>(near[j]), static_cast<unsigned int>(near[j] >> 32));
                    }
                    // below: entry kb (< nv), series m_k; above: entry ka (> 0), series n_k.
                    // A thread whose entry changes first adds its running sums to the block's
                    // array; lanes with no share carry zero coefficients.
                    const bool hasB = valid && kb < nv, hasA = valid && ka > 0 && ka <= nv;
                    const bool swB = hasB && kb != cbB, swA = hasA && ka != cbA;
                    if (__builtin_amdgcn_ballot_w64(swB)) {
                        far_flush(sh.dB[wave], swB && cbB >= 0, cbB, accB);
#pragma unroll
                        for (int k = 0; k < kFarTerms; ++k) accB[k] = swB ? 0.0 : accB[k];
                        cbB = swB ? kb : cbB;
                    }
                    if (__builtin_amdgcn_ballot_w64(swA)) {
                        far_flush(sh.dA[wave], swA && cbA >= 0, cbA, accA);
#pragma unroll
                        for (int k = 0; k < kFarTerms; ++k) accA[k] = swA ? 0.0 : accA[k];
                        cbA = swA ? ka : cbA;
                    }
                    {
                        double a = hasB ? c.a : 0.0, cc = hasB ? c.c : 0.0;
                        double Ds2 = hasB ? 2.0 * Ds : 0.0, p0 = hasB ? P0 : 0.0;
                        // (the series start after the flushes: their terms are not held
                        // live across them)
                        asm volatile("" : "+v"(a), "+v"(cc), "+v"(Ds2), "+v"(p0));
                        double m0 = a, m1 = fma(Ds2, a, cc);
                        accB[0] += m0;
                        accB[1] += m1;
#pragma unroll
                        for (int k = 2; k < kFarTerms; ++k) {
                            const double m2 = fma(Ds2, m1, -p0 * m0);
                            accB[k] += m2;
                            m0 = m1;
                            m1 = m2;
                        }
                    }
                    {
                        double a = hasA ? c.a : 0.0, cc = hasA ? c.c : 0.0;
                        double Ds2 = hasA ? 2.0 * Ds : 0.0, iP = hasA ? recip(P0) : 0.0;
                        asm volatile("" : "+v"(a), "+v"(cc), "+v"(Ds2), "+v"(iP));
                        double n0 = cc * iP, n1 = fma(Ds2, n0, a) * iP;
                        accA[0] += n0;
                        accA[1] += n1;
#pragma unroll
                        for (int k = 2; k < kFarTerms; ++k) {
                            const double n2 = fma(Ds2, n1, -n0) * iP;
                            accA[k] += n2;
                            n0 = n1;
                            n1 = n2;
                        }
                    }
                }
                far_flush(sh.dB[wave], cbB >= 0, cbB, accB);
                far_flush(sh.dA[wave], cbA >= 0, cbA, accA);
                __syncthreads();
                // moments of each used tile: prefix sums of the below array up to its rank
                // by q lo, suffix sums of the above array past its rank by hi / q
                const int ss = kSegFarSet[SEG] + (inv ? v : 0);
                for (int e = tid; e < kFarTiles * kFarTerms; e += kThreads) {
                    const int j = e / kFarTerms, k = e % kFarTerms;
                    if (!((used >> j) & 1u)) continue;
                    const int pb = tile_rank<F32>(sh, used, j, rv, false), pa = tile_rank<F32>(sh, used, j, rv, true);
                    double sb = 0.0, sa = 0.0;
                    for (int w = 0; w < kFarWaves; ++w) {
                        for (int p = 0; p <= pb; ++p) sb += sh.dB[w][p][k];
                        for (int p = pa + 1; p <= nv; ++p) sa += sh.dA[w][p][k];
                    }
                    sh.mom[j][ss * 2 * kFarTerms + k] += sb;
                    sh.mom[j][ss * 2 * kFarTerms + kFarTerms + k] += sa;
                }
                __syncthreads();
                for (int e = tid; e < kFarWaves * (kFarTiles + 1) * kFarTerms; e += kThreads) {
                    (&sh.dB[0][0][0])[e] = 0.0;
                    (&sh.dA[0][0][0])[e] = 0.0;
                }
                __syncthreads();
            }
        });
    }
    for (int e = tid; e < kFarTiles * kFarMom; e += kThreads) {
        const int jj = e / kFarMom, r = e % kFarMom;
        const int t = g * kFarTiles + jj;
        if (t < A.tiles_total)
            A.far_mom[((static_cast<size_t>(layer) * nsplit + part) * A.tiles_total + t) * kFarMom + r] = sh.mom[jj][r];
    }
    // the h2o base term of this part's lines (every group derives them; group 0 writes)
    for (int m = 32; m > 0; m >>= 1) base += __shfl_xor(base, m);
    if (lane == 0) sh.bsum[wave] = base;
    __syncthreads();
    if (tid == 0 && g == 0) {
        double b = 0.0;
        for (int w = 0; w < kThreads / 64; ++w) b += sh.bsum[w];
        A.far_base[static_cast<size_t>(part) * A.nl + layer] = b;
    }
}

// ---- pass 2: near lines summed per tile, plus the series ----

// one wavefront's staged batch of kWaveBatch near lines (the layout sum_staged_f32 reads)
constexpr int kWaveBatch = 64;
struct WaveStage {
    alignas(16) double D[2][kWaveBatch];
    alignas(16) double h[2][kWaveBatch];
    alignas(16) float2 ac[2][kWaveBatch];
    alignas(16) float2 dh[2][kWaveBatch];
    unsigned int fast[2][kThreads / 64];
};

// FP64 mode: the (D, a, c, g, G) the direct FP64 sums read (sum_staged_f64)
struct WaveStage64 {
    double D[2][kWaveBatch], a[2][kWaveBatch], c[2][kWaveBatch], g[2][kWaveBatch], G[2][kWaveBatch];
};

template <int K, bool F32>
struct NearShared {
    std::conditional_t<F32, WaveStage, WaveStage64> st[kThreads / 64];
    union {
        unsigned short list[kListWords * 32];   // a window's near lines
        double red[kThreads / 64][K][64];       // the wavefronts' partial sums at a slot's end
    };
    double far[kFarMom];
    WaveRanges rng;
    int wsum[kThreads / 64];
    alignas(16) double prep[2][8];   // the segment's coefficient sets (read from LDS while deriving)
    double scale[2];                 // and their FP32 weight scales
};

// sum of the far-line series of one moment set at x
__device__ inline double far_eval(const double *mom, double x) {
    double below = 0.0;
    if (x > 0.0) {
        const double u = 1.0 / x;
        double s = mom[kFarTerms - 1];
#pragma unroll
        for (int k = kFarTerms - 2; k >= 0; --k) s = fma(s, u, mom[k]);
        below = s * u;
    }
    const double *n = mom + kFarTerms;
    double s = n[kFarTerms - 1];
#pragma unroll
    for (int k = kFarTerms - 2; k >= 0; --k) s = fma(s, x, n[k]);
    return below + s;
}

// A listed line's coefficients for the near sums, derived from its catalogue row as pass 1
// derives them, with the segment's coefficient sets read from LDS: FP32 (Ds, h, a s, c s),
// FP64 (D, a, c, g, G)
template <class Prep, int NC>
__device__ __forceinline__ CoefT<double> near_coef(const DevBlock &B, int i, int v, const double (&prep)[2][8]) {
    LineRow<NC> row;
    row.load(B, i);
    double unused = 0.0;
    return seg_coef(*reinterpret_cast<const Prep *>(prep[0]), *reinterpret_cast<const Prep *>(prep[1]), v, row, i,
                    unused);
}

template <bool F32>
__device__ __forceinline__ void derive_near(int seg, int v, const DevBlock &B, int i, const double (&prep)[2][8],
                                            double sv, double (&c5)[5], float2 &ac) {
    CoefT<double> c;
    switch (seg) {
        case 0: c = near_coef<H2oPrep<double>, kNumCols[RBX_LINES_H2O]>(B, i, v, prep); break;
        case 1: c = near_coef<H2sPrep<double>, kNumCols[RBX_LINES_H2S]>(B, i, v, prep); break;
        case 2: c = near_coef<InvPrep<double>, kNumCols[RBX_LINES_NH3_INV]>(B, i, v, prep); break;
        case 3: c = near_coef<RotPrep<double>, kNumCols[RBX_LINES_NH3_ROT]>(B, i, v, prep); break;
        case 4: c = near_coef<V2Prep<double>, kNumCols[RBX_LINES_NH3_V2]>(B, i, v, prep); break;
        default: c = near_coef<Ph3Prep<double>, kNumCols[RBX_LINES_PH3]>(B, i, v, prep); break;
    }
    if constexpr (F32) {
        shifted_denominator(c, c5[0], c5[1]);
        ac = make_float2(static_cast<float>(c.a * sv), static_cast<float>(c.c * sv));
    } else {
        c5[0] = c.D;
        c5[1] = c.a;
        c5[2] = c.c;
        c5[3] = c.g;
        c5[4] = c.G;
    }
}

// near lines of one segment and coefficient set v, listed from the tile's mask words; each
// wavefront sums its own batches of kWaveBatch listed lines (batches w, w + 4, ...), staged
// in its own LDS and fetched one batch ahead into registers.  `so` = the stream's first line
// in the staged-line arrays.
template <int K, bool F32, class Derive>
__device__ void near_sum(const AlphaArgs &A, int n, const Derive &derive, int v, bool mixed, bool uni,
                         const unsigned int *mask,
                         int nwords, int var, const int (&set)[K], const double (&x)[K], const float (&xf)[K],
                         double (&acc)[K], NearShared<K, F32> &sh) {
    const int tid = threadIdx.x, lane = tid % 64, wave = tid / 64;
    auto &st = sh.st[wave];
    const double lo = sh.rng.lo[var][0], hi = sh.rng.hi[var][0];
    for (int w0 = 0; w0 < nwords; w0 += kListWords) {
        const int wi = w0 + tid;
        const unsigned int bits = (tid < kListWords && wi < nwords) ? mask[wi] : 0u;
        const int cnt = __popc(bits);
        int incl = cnt;
#pragma unroll
        for (int m = 1; m < 64; m <<= 1) {
            const int y = __shfl_up(incl, m);
            if (lane >= m) incl += y;
        }
        if (lane == 63) sh.wsum[wave] = incl;
        __syncthreads();
        int pos = incl - cnt, total = 0;
#pragma unroll
        for (int w = 0; w < kThreads / 64; ++w) {
            if (w < wave) pos += sh.wsum[w];
            total += sh.wsum[w];
        }
        for (unsigned int b = bits; b != 0u; b &= b - 1u)
            sh.list[pos++] = static_cast<unsigned short>(tid * 32 + __builtin_ctz(b));
        __syncthreads();
        // listed line e's coefficients, derived from its catalogue row when its batch is
        // staged (past the list: a zero term): FP32 (Ds, h, a s, c s), FP64 (D, a, c, g, G)
        const int nb = (total + kWaveBatch - 1) / kWaveBatch;
        for (int b = wave; b < nb; b += kThreads / 64) {
            const int cntp = (min(kWaveBatch, total - b * kWaveBatch) + kSub - 1) / kSub * kSub;
            const int e = b * kWaveBatch + lane;
            const int i = e < total ? w0 * 32 + sh.list[e] : n;
            double c5[5] = {0.0, F32 ? 1.0 : 0.0, 0.0, 0.0, 1.0};
            float2 ac = make_float2(0.0f, 0.0f);
            if (i < n) derive(i, c5, ac);
            if constexpr (F32) {
                st.D[v][lane] = c5[0];
                st.h[v][lane] = c5[1];
                st.ac[v][lane] = ac;
                st.dh[v][lane] = make_float2(static_cast<float>(c5[0]), static_cast<float>(c5[1]));
                if (mixed) {   // the other set's columns add nothing
                    st.D[1 - v][lane] = 0.0;
                    st.h[1 - v][lane] = 1.0;
                    st.ac[1 - v][lane] = make_float2(0.0f, 0.0f);
                    st.dh[1 - v][lane] = make_float2(0.0f, 1.0f);
                }
                const unsigned int fb = full_chunks(__ballot(fp32_safe(c5[0], c5[1], lo, hi)));
                if (lane == 0) st.fast[v][wave] = fb;
            } else {
                st.D[v][lane] = c5[0];
                st.a[v][lane] = c5[1];
                st.c[v][lane] = c5[2];
                st.g[v][lane] = c5[3];
                st.G[v][lane] = c5[4];
                if (mixed) {
                    st.D[1 - v][lane] = 0.0;
                    st.a[1 - v][lane] = 0.0;
                    st.c[1 - v][lane] = 0.0;
                    st.g[1 - v][lane] = 0.0;
                    st.G[1 - v][lane] = 1.0;
                }
            }
            wave_sync();
            if constexpr (F32) {
                if (uni)
                    sum_staged_f32<K, true>(st, set, 0, cntp, x, xf, acc);
                else
                    sum_staged_f32<K, false>(st, set, 0, cntp, x, xf, acc);
            } else {
                if (uni)
                    sum_staged_f64<K, true>(st, set, 0, cntp, x, acc);
                else
                    sum_staged_f64<K, false>(st, set, 0, cntp, x, acc);
            }
            wave_sync();
        }
        __syncthreads();   // every wavefront is done with this window's list
    }
}

template <int K, bool F32>
__global__ __launch_bounds__(kThreads, K >= 4 ? 2 : 4) void alpha_near_kernel(AlphaArgs args) {
#if defined(__HIP_DEVICE_COMPILE__)
    const AlphaArgs &A = *reinterpret_cast<const AlphaArgs *>(__builtin_amdgcn_kernarg_segment_ptr());
#else
    const AlphaArgs &A = args;   // (host pass: never run)
#endif
    __shared__ NearShared<K, F32> sh;
    // layer fastest: blocks are dealt to the XCDs in turn, so each XCD gets every tile (the
    // tiles' near fractions differ several-fold) rather than one tile of every layer.  Last
    // tile first: requests run low to high frequency and line lists grow denser with frequency,
    // so the heavy tiles start first and the light ones fill the tail of the grid (800 layers
    // x 1024 frequencies: step 3.92 -> 3.73 ms).
    const int layer = blockIdx.x % A.nl, tile = A.tiles_total - 1 - static_cast<int>(blockIdx.x / A.nl);
    const int tid = threadIdx.x, lane = tid % 64, wave = __builtin_amdgcn_readfirstlane(tid / 64);
    const int nsplit = A.far_split;
    const LayerT<double> L = layer_of(A, layer);
    int c0, c1;
    tile_cols(A, tile, c0, c1);
    const int tsets = tile_sets(A, c0, c1);
    const bool mixed = tsets == 3;

    int col[K], set[K];
    bool ok[K];
    double x[K], xn[K];
    float xf[K], xnf[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        col[k] = c0 + lane + k * 64;
        ok[k] = col[k] < A.nf;
        const int cc = min(col[k], A.nf - 1);
        set[k] = cc < A.nf_low ? 0 : 1;
        const double f = A.freq[cc];
        const double fn = A.freq_nh3 ? A.freq_nh3[cc] : f;
        x[k] = f * f;
        xn[k] = fn * fn;
        xf[k] = static_cast<float>(x[k]);
        xnf[k] = static_cast<float>(xn[k]);
    }
    const bool uni_set = !mixed;   // every column of the tile reads one inversion set

    if (wave == 0) {
        double lo[kFarRanges], hi[kFarRanges];
        tile_ranges(A, tile, lo, hi);
        if (lane < kThreads / 64) {
#pragma unroll
            for (int v = 0; v < 2; ++v) {
                sh.rng.lo[v][lane] = lo[v];
                sh.rng.hi[v][lane] = hi[v];
            }
        }
    }
    for (int e = tid; e < kFarMom; e += kThreads) {
        double s = 0.0;
        for (int p = 0; p < nsplit; ++p)
            s += A.far_mom[((static_cast<size_t>(layer) * nsplit + p) * A.tiles_total + tile) * kFarMom + e];
        sh.far[e] = s;
    }
    if (tile == 0 && tid == 0) {
        double b = 0.0;
        for (int p = 0; p < nsplit; ++p) b += A.far_base[static_cast<size_t>(p) * A.nl + layer];
        A.ws_base[layer] = b;
    }
    __syncthreads();

    const LinePlan plan(A, L.P);
    const int touched = plan.slots(0, 1);
    const size_t plane = static_cast<size_t>(A.nl) * A.nf;
    const unsigned int *masks = A.near + (static_cast<size_t>(layer) * A.tiles_total + tile) * A.far_words;
    double acc[K] = {};
    auto series = [&](int q, int k) {
        switch (q) {
            case kSlotH2o: return far_eval(sh.far + 0 * 2 * kFarTerms, x[k]);
            case kSlotH2s: return far_eval(sh.far + 1 * 2 * kFarTerms, x[k]);
            case kSlotNh3:
                return far_eval(sh.far + (2 + set[k]) * 2 * kFarTerms, xn[k]) +
                       far_eval(sh.far + 4 * 2 * kFarTerms, xn[k]);
            default: return far_eval(sh.far + 5 * 2 * kFarTerms, x[k]);
        }
    };
    // the wavefronts' partial sums of the same columns, added in wavefront order, plus the series
    auto store = [&](int q) {
        if (touched & (1 << q)) {
#pragma unroll
            for (int k = 0; k < K; ++k) sh.red[wave][k][lane] = acc[k];
            __syncthreads();
            if (wave == 0) {
                double *w = A.ws + static_cast<size_t>(q) * plane + static_cast<size_t>(layer) * A.nf;
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    double t = sh.red[0][k][lane];
#pragma unroll
                    for (int ww = 1; ww < kThreads / 64; ++ww) t += sh.red[ww][k][lane];
                    if (ok[k]) w[col[k]] = t + series(q, k);
                }
            }
            __syncthreads();
        }
#pragma unroll
        for (int k = 0; k < K; ++k) acc[k] = 0.0;
    };
    for (int seg = 0; seg < kSegs; ++seg) {
        if (plan.nr[seg] > 0) {
            // the segment's coefficient sets and weight scales go to LDS and are read there
            // while deriving (held in registers across the line sums they would crowd out the
            // sums' own)
            __syncthreads();   // (the previous segment's derivations are done)
            if (tid < kPrepWords) {
                const double w = A.far_prep[(static_cast<size_t>(layer) * kSegs + seg) * kPrepWords + tid];
                if (tid < 16)
                    sh.prep[tid / 8][tid % 8] = w;
                else
                    sh.scale[tid - 16] = w;
            }
            __syncthreads();
            const bool inv = seg == 2;
            const int var = kSegVar[seg];
            const DevBlock &B = A.cat.blk[kSegKind[seg]];
            const int nwords = plan.nr[seg] * (kThreads / 32);
            double a2[K] = {}, xv[K];
            float xfv[K];
            int sets[K];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                xv[k] = var ? xn[k] : x[k];
                xfv[k] = var ? xnf[k] : xf[k];
                sets[k] = inv ? set[k] : 0;
            }
#pragma unroll 1
            for (int v = 0; v < 2; ++v) {
                if (inv ? !((tsets >> v) & 1) : v > 0) continue;
                const unsigned int *mk = masks + A.far_off[kSegStream[seg] + (inv ? v : 0)];
                const double sv = sh.scale[v];
                auto derive = [&](int i, double (&c5)[5], float2 &ac) { derive_near<F32>(seg, v, B, i, sh.prep, sv, c5, ac); };
                near_sum<K, F32>(A, B.n, derive, inv ? v : 0, inv && mixed, inv ? uni_set : true, mk, nwords, var, sets,
                                 xv, xfv, a2, sh);
            }
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const double sc = sh.scale[inv ? set[k] : 0];
                if (seg == 0 || seg == 1 || seg == 5)
                    acc[k] = a2[k] / sc;
                else
                    acc[k] += a2[k] / sc;
            }
        }
        if (seg == 0) store(kSlotH2o);
        if (seg == 1) store(kSlotH2s);
        if (seg == 4) store(kSlotNh3);
        if (seg == 5) store(kSlotPh3);
    }
}
