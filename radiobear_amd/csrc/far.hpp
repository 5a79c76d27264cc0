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
// alpha_far_kernel (block = layer x group of kFarTiles tiles x part of the layer's
// staging rounds): each thread derives one line's coefficients per round (once per
// line and layer), then thread groups of 16 classify the round's lines against one
// tile each, add the far lines' series coefficients into per-tile moments (FP64,
// in registers, reduced per segment) and write one bit per near line into the tile's
// near mask.  The h2o base term is summed here.
// It also stores every line's staged form per layer: FP32 mode (Ds, h; scaled a, c),
// FP64 mode (D, a, c, g, G).
// alpha_near_kernel (block = tile x layer, all four wavefronts on the tile's
// columns): lists the tile's near lines from its masks, gathers and sums them as the
// line-sum kernel does with the staged lines split between the wavefronts, adds the
// wavefronts in order and the series of the summed moments, and writes the same
// per-slot partials the finish kernel reads (one part).

constexpr int kFarTerms = 16;
// pole ratio of a far line: 1/3 in FP32 mode, 0.15 in FP64 mode (at most ~8e-12 of a far
// line's value there, against the 1e-9 FP64 parity bar)
template <bool F32>
__host__ __device__ constexpr double far_ratio() { return F32 ? 1.0 / 3.0 : 0.15; }
constexpr int kFarTiles = 8;                          // tiles per alpha_far_kernel block
constexpr int kFarSets = 6;                           // h2o, h2s, inversion set 0, set 1, rot + v2, ph3
constexpr int kFarMom = kFarSets * 2 * kFarTerms;     // moments per (layer, part, tile)
constexpr int kRedChunk = 8;                          // accumulator terms reduced per LDS pass
static_assert((2 * kFarTerms) % kRedChunk == 0, "series terms must fill whole reduction passes");
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

// moduli of the two roots of x^2 - 2 Ds x + (Ds^2 + h): complex pair (h >= 0) or real
__device__ inline void pole_moduli(double Ds, double h, double &rs, double &rb) {
    if (h >= 0.0) {
        rs = rb = sqrt(fma(Ds, Ds, h));
    } else {
        const double q = sqrt(-h), m = fabs(Ds);
        rb = m + q;
        rs = fabs(m - q);
    }
}

// the segment's coefficient deriver(s) for a layer and their FP32 weight scales (FP64 mode
// uses none)
template <class F>
__device__ __forceinline__ void with_prep(const AlphaArgs &A, const LayerT<double> &L, int seg, F &&f) {
    const DevBlock *blk = A.cat.blk;
    switch (seg) {
        case 0: {
            const H2oPrep<double> p(L);
            f(&p, pow2_scale(p.pref * blk[RBX_LINES_H2O].wscale), 1.0);
            break;
        }
        case 1: {
            const H2sPrep<double> p(L);
            f(&p, pow2_scale(p.pref * blk[RBX_LINES_H2S].wscale), 1.0);
            break;
        }
        case 2: {
            const InvPrep<double> p[2] = {InvPrep<double>(L, nh3_low_set(A.abs.nh3)), InvPrep<double>(L, 1)};
            const double w = blk[RBX_LINES_NH3_INV].wscale;
            f(p, pow2_scale(p[0].pref * w), pow2_scale(p[1].pref * w));
            break;
        }
        case 3: {
            const RotPrep<double> p(L);
            f(&p, pow2_scale(p.pref * blk[RBX_LINES_NH3_ROT].wscale), 1.0);
            break;
        }
        case 4: {
            const V2Prep<double> p(L);
            f(&p, pow2_scale(p.pref * blk[RBX_LINES_NH3_V2].wscale), 1.0);
            break;
        }
        default: {
            const Ph3Prep<double> p(L);
            f(&p, pow2_scale(p.pref * blk[RBX_LINES_PH3].wscale), 1.0);
            break;
        }
    }
}

// ---- pass 1: classification, moments, near masks ----

constexpr int kFarPairs = kFarTiles + 1;   // (tile, coefficient set) pairs: one tile may need both sets

struct FarShared {
    union {
        struct {   // one staging round, per coefficient set and line
            double a[2][kThreads], c[2][kThreads], Ds[2][kThreads], P0[2][kThreads];
            float2 r[2][kThreads];   // pole moduli (smaller, larger)
        } st;
        double red[kThreads][kRedChunk];   // accumulator reduction, kRedChunk terms at a time
    };
    double mom[kFarTiles][kFarMom];
    double lo[kFarTiles][kFarRanges], hi[kFarTiles][kFarRanges];
    unsigned int maskw[kFarPairs][kThreads / 32];
    double bsum[kThreads / 64];
};

// the (tile, set) pairs of a segment in a group of tiles, in order: every tile of the
// group, with each inversion set its columns use (one set for the other segments)
struct FarPairs {
    int n;
    int j[kFarPairs], set[kFarPairs];
    __device__ FarPairs(const AlphaArgs &A, int g, int ntiles, bool inv) {
        n = 0;
        for (int jj = 0; jj < ntiles; ++jj) {
            int c0, c1;
            tile_cols(A, g * kFarTiles + jj, c0, c1);
            const int ts = inv ? tile_sets(A, c0, c1) : 1;
            for (int v = 0; v < 2; ++v)
                if ((ts >> v) & 1) {
                    j[n] = jj;
                    set[n] = v;
                    ++n;
                }
        }
    }
};

// three blocks per CU (168 VGPRs, a few spills) run the pass in 1.00 ms against 1.18 ms with two
// blocks and no spills; four (128 VGPRs) gain 3 % more for twice the spill traffic
template <bool F32>
__global__ __launch_bounds__(kThreads, 3) void alpha_far_kernel(AlphaArgs A) {
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
    int gc0, gc1;
    tile_cols(A, g * kFarTiles, gc0, gc1);
    gc1 = min(A.nf, (g + 1) * kFarTiles * A.tile_width);
    const int gsets = tile_sets(A, gc0, gc1);   // inversion sets any tile of the group needs
    const int ntiles = min(kFarTiles, A.tiles_total - g * kFarTiles);
    const double q = far_ratio<F32>();
    __syncthreads();

    const LinePlan plan(A, L.P);
    const int r0 = plan.begin(part, nsplit), r1 = plan.begin(part + 1, nsplit);
    double base = 0.0;
    int off = 0;
    for (int seg = 0; seg < kSegs; ++seg) {
        const int a0 = max(r0, off), b0 = min(r1, off + plan.nr[seg]);
        const int seg_off = off;
        off += plan.nr[seg];
        if (a0 >= b0) continue;
        const DevBlock &B = A.cat.blk[kSegKind[seg]];
        const bool inv = seg == 2;
        const int need = inv ? gsets : 1;
        // thread groups of tpp consecutive threads, one per (tile, set) pair; each takes
        // every tpp-th line of a round
        const FarPairs pairs(A, g, ntiles, inv);
        const int tpp = kThreads / pairs.n;
        const int myp = tid / tpp, sidx = tid % tpp;
        const bool active = myp < pairs.n;
        const int myj = active ? pairs.j[myp] : 0, myset = active ? pairs.set[myp] : 0;
        const int rv = kSegVar[seg] == 0 ? 0 : (inv ? 2 + myset : 1);
        const float lo_q = static_cast<float>(q * sh.lo[myj][rv]), hi_q = static_cast<float>(sh.hi[myj][rv] / q);
        double accB[kFarTerms], accA[kFarTerms];
#pragma unroll
        for (int k = 0; k < kFarTerms; ++k) accB[k] = accA[k] = 0.0;

        with_prep(A, L, seg, [&](const auto *prep, double s0, double s1) {
            const double scale[2] = {F32 ? s0 : 1.0, F32 ? s1 : 1.0};
            if (g == 0 && tid < 2)   // (every part that has the segment stores the same value)
                A.far_scale[(static_cast<size_t>(layer) * kSegs + seg) * 2 + tid] = scale[tid];
            for (int r = a0; r < b0; ++r) {
                const int l0 = (r - seg_off) * kThreads;
                const int cnt = min(kThreads, B.n - l0);
                {   // phase 1: thread tid derives line l0 + tid and stores its staged form
                    CoefT<double> qs[2] = {{0.0, 0.0, 0.0, 0.0, 1.0}, {0.0, 0.0, 0.0, 0.0, 1.0}};
                    if (tid < cnt) line_coefs(prep, need, B, l0 + tid, base, qs);
#pragma unroll
                    for (int s = 0; s < 2; ++s) {
                        if (!(need & (1 << s))) continue;
                        double Ds, h, rs, rb;
                        shifted_denominator(qs[s], Ds, h);
                        pole_moduli(Ds, h, rs, rb);
                        // one group stores each stream: the first whose tiles use the set
                        const int writer = (inv && s == 1) ? A.nf_low / A.tile_width / kFarTiles : 0;
                        if (g == writer && tid < cnt) {
                            const size_t o = static_cast<size_t>(layer) * A.far_words * 32 +
                                             static_cast<size_t>(A.far_off[kSegStream[seg] + (inv ? s : 0)]) * 32 +
                                             l0 + tid;
                            if constexpr (F32) {
                                A.near_dh[o] = make_double2(Ds, h);
                                A.near_ac[o] = make_float2(static_cast<float>(qs[s].a * scale[s]),
                                                           static_cast<float>(qs[s].c * scale[s]));
                            } else {
                                const size_t pl = static_cast<size_t>(A.nl) * A.far_words * 32;
                                A.near_c64[o] = qs[s].D;
                                A.near_c64[pl + o] = qs[s].a;
                                A.near_c64[2 * pl + o] = qs[s].c;
                                A.near_c64[3 * pl + o] = qs[s].g;
                                A.near_c64[4 * pl + o] = qs[s].G;
                            }
                        }
                        sh.st.a[s][tid] = qs[s].a;
                        sh.st.c[s][tid] = qs[s].c;
                        sh.st.Ds[s][tid] = Ds;
                        sh.st.P0[s][tid] = fma(Ds, Ds, h);
                        sh.st.r[s][tid] = make_float2(static_cast<float>(rs), static_cast<float>(rb));
                    }
                    if (tid < kFarPairs * (kThreads / 32)) (&sh.maskw[0][0])[tid] = 0u;
                }
                __syncthreads();
                // phase 2: classify against the pair's tile; far lines add their series
                // coefficients, near lines set their bit
                if (active) {
                    for (int i = sidx; i < cnt; i += tpp) {
                        const float2 rr = sh.st.r[myset][i];
                        if (rr.y <= lo_q) {
                            const double a = sh.st.a[myset][i], c = sh.st.c[myset][i];
                            const double Ds2 = 2.0 * sh.st.Ds[myset][i], P0 = sh.st.P0[myset][i];
                            double m0 = a, m1 = fma(Ds2, a, c);
                            accB[0] += m0;
                            accB[1] += m1;
#pragma unroll
                            for (int k = 2; k < kFarTerms; ++k) {
                                const double m2 = fma(Ds2, m1, -P0 * m0);
                                accB[k] += m2;
                                m0 = m1;
                                m1 = m2;
                            }
                        } else if (rr.x >= hi_q) {
                            const double a = sh.st.a[myset][i], c = sh.st.c[myset][i];
                            const double Ds2 = 2.0 * sh.st.Ds[myset][i], iP = 1.0 / sh.st.P0[myset][i];
                            double n0 = c * iP, n1 = fma(Ds2, n0, a) * iP;
                            accA[0] += n0;
                            accA[1] += n1;
#pragma unroll
                            for (int k = 2; k < kFarTerms; ++k) {
                                const double n2 = fma(Ds2, n1, -n0) * iP;
                                accA[k] += n2;
                                n0 = n1;
                                n1 = n2;
                            }
                        } else {   // near (also NaN coefficients: summed directly)
                            atomicOr(&sh.maskw[myp][i >> 5], 1u << (i & 31));
                        }
                    }
                }
                __syncthreads();
                if (tid < pairs.n * (kThreads / 32)) {   // the round's 8 mask words of each pair
                    const int p = tid / (kThreads / 32), w = tid % (kThreads / 32);
                    const int stream = kSegStream[seg] + (inv ? pairs.set[p] : 0);
                    A.near[(static_cast<size_t>(layer) * A.tiles_total + g * kFarTiles + pairs.j[p]) * A.far_words +
                           A.far_off[stream] + l0 / 32 + w] = sh.maskw[p][w];
                }
            }
        });
        // the segment's moments: each pair's threads' accumulators summed through LDS
        // (the staging area), kRedChunk terms at a time, into the pair's tile and set
#pragma unroll
        for (int part = 0; part < 2 * kFarTerms / kRedChunk; ++part) {
#pragma unroll
            for (int e = 0; e < kRedChunk; ++e) {
                const int t = part * kRedChunk + e;
                sh.red[tid][e] = t < kFarTerms ? accB[t] : accA[t - kFarTerms];
            }
            __syncthreads();
            if (tid < pairs.n * kRedChunk) {
                const int p = tid / kRedChunk, e = tid % kRedChunk;
                double sum = 0.0;
                for (int k = 0; k < tpp; ++k) sum += sh.red[p * tpp + k][e];
                const int ss = kSegFarSet[seg] + (inv ? pairs.set[p] : 0);
                sh.mom[pairs.j[p]][ss * 2 * kFarTerms + part * kRedChunk + e] += sum;
            }
            __syncthreads();
        }
    }
    __syncthreads();
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
    double scale[kSegs][2];
    double far[kFarMom];
    WaveRanges rng;
    int wsum[kThreads / 64];
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

// near lines of one segment and coefficient set v, listed from the tile's mask words; each
// wavefront sums its own batches of kWaveBatch listed lines (batches w, w + 4, ...), staged
// in its own LDS and fetched one batch ahead into registers.  `so` = the stream's first line
// in the staged-line arrays.
template <int K, bool F32>
__device__ void near_sum(const AlphaArgs &A, int n, size_t so, int v, bool mixed, bool uni, const unsigned int *mask,
                         int nwords, int var, const int (&set)[K], const double (&x)[K], const float (&xf)[K],
                         double (&acc)[K], NearShared<K, F32> &sh) {
    const int tid = threadIdx.x, lane = tid % 64, wave = tid / 64;
    auto &st = sh.st[wave];
    const double lo = sh.rng.lo[var][0], hi = sh.rng.hi[var][0];
    const size_t pl = static_cast<size_t>(A.nl) * A.far_words * 32;   // FP64 coefficient plane
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
        // the staged form pass 1 stored for listed line e (past the list: a zero term):
        // FP32 (Ds, h, a s, c s), FP64 (D, a, c, g, G)
        auto fetch = [&](int e, double (&c5)[5], float2 &ac) {
            const int i = e < total ? w0 * 32 + sh.list[e] : n;
            c5[0] = 0.0; c5[1] = 0.0; c5[2] = 0.0; c5[3] = 0.0; c5[4] = 1.0;
            ac = make_float2(0.0f, 0.0f);
            if (i < n) {
                if constexpr (F32) {
                    const double2 dh = A.near_dh[so + i];
                    c5[0] = dh.x;
                    c5[1] = dh.y;
                    ac = A.near_ac[so + i];
                } else {
#pragma unroll
                    for (int m = 0; m < 5; ++m) c5[m] = A.near_c64[m * pl + so + i];
                }
            }
            if constexpr (F32) {
                if (!(i < n)) c5[1] = 1.0;   // padding h
            }
        };
        const int nb = (total + kWaveBatch - 1) / kWaveBatch;
        double c5[5];
        float2 ac;
        if (wave < nb) fetch(wave * kWaveBatch + lane, c5, ac);
        for (int b = wave; b < nb; b += kThreads / 64) {
            const int cntp = (min(kWaveBatch, total - b * kWaveBatch) + kSub - 1) / kSub * kSub;
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
            const int bn = b + kThreads / 64;
            double c5n[5] = {0.0, 1.0, 0.0, 0.0, 1.0};
            float2 ac2 = make_float2(0.0f, 0.0f);
            if (bn < nb) fetch(bn * kWaveBatch + lane, c5n, ac2);
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
#pragma unroll
            for (int m = 0; m < 5; ++m) c5[m] = c5n[m];
            ac = ac2;
        }
        __syncthreads();   // every wavefront is done with this window's list
    }
}

template <int K, bool F32>
__global__ __launch_bounds__(kThreads, K >= 4 ? 2 : 4) void alpha_near_kernel(AlphaArgs A) {
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
    const int one[K] = {};
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
    if (tid < kSegs * 2) sh.scale[tid / 2][tid % 2] = A.far_scale[static_cast<size_t>(layer) * kSegs * 2 + tid];
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
            const DevBlock &B = A.cat.blk[kSegKind[seg]];
            const bool inv = seg == 2;
            const int nwords = plan.nr[seg] * (kThreads / 32);
            const double s[2] = {sh.scale[seg][0], sh.scale[seg][1]};
            double a2[K] = {};
            const int var = kSegVar[seg];
            for (int v = 0; v < 2; ++v) {
                if (inv ? !((tsets >> v) & 1) : v > 0) continue;
                const int stream = kSegStream[seg] + (inv ? v : 0);
                const unsigned int *mk = masks + A.far_off[stream];
                const size_t so = static_cast<size_t>(layer) * A.far_words * 32 + static_cast<size_t>(A.far_off[stream]) * 32;
                if (inv)
                    near_sum<K, F32>(A, B.n, so, v, mixed, uni_set, mk, nwords, var, set, xn, xnf, a2, sh);
                else
                    near_sum<K, F32>(A, B.n, so, 0, false, true, mk, nwords, var, one, var ? xn : x, var ? xnf : xf,
                                     a2, sh);
            }
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const double sc = inv ? s[set[k]] : s[0];
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
