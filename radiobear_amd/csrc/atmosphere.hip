// Atmosphere construction on the device: computeProp's radii and refractive
// indices (atm_base.py:120-181), optionally after a hydrostatic recomputation of
// z, and its whole property table (atm_base.py:120-194), for many atmospheres laid
// end to end.  One thread per atmosphere walks its layers (a few hundred steps of a
// handful of FP64 operations); a retrieval batch of 1024 atmospheres is 16
// wavefronts.
#include "common.hpp"

namespace rbx {

// chemistry.py:7-46: gas constant, molar masses (g/mol) and microwave
// refractivity coefficients of the species in the rows of mix (RBX_X_* order);
// PH3 has no refractivity there (NULL), H2O's depends on T, and CO has neither
// (ConstituentProperties gives it the NULL entries, 0, chemistry.py:168-200).
constexpr double kRgas = 8.314462;
constexpr double kGravConst = 6.6738e-20;   // km^3 / kg / s^2 (chemistry.py:7)
__constant__ double kAmu[RBX_NUM_SPECIES] = {2.016, 4.003, 16.04, 17.030, 18.015, 34.076, 33.997, 0.0};
__constant__ double kRefr[RBX_NUM_SPECIES] = {124.43, 35.832, 413.0, 2700.0, 0.0, 2247.0, 0.0, 0.0};

__device__ inline double refractivity(const double *mix, size_t nlt, size_t l, double T) {
    double r = 0.0;
#pragma unroll
    for (int s = 0; s < RBX_NUM_SPECIES; ++s) {
        const double c = s == RBX_X_H2O ? 245.0 + 1.28E6 / T : kRefr[s];   // REFR_H2O, chemistry.py:33-41
        r += c * mix[s * nlt + l];
    }
    return r;
}

__device__ inline double mean_amu(const double *mix, size_t nlt, size_t l) {
    double m = 0.0;
#pragma unroll
    for (int s = 0; s < RBX_NUM_SPECIES; ++s) m += kAmu[s] * mix[s * nlt + l];
    return m;
}

// the layer nearest p_ref (first of equals), atm_base.py:135-144
__device__ inline int ref_layer(const double *P, int l0, int l1, double p_ref) {
    int iref = l0;
    double psep = 1.0E6;
    for (int l = l0; l < l1; ++l) {
        const double d = fabs(P[l] - p_ref);
        if (d < psep) {
            psep = d;
            iref = l;
        }
    }
    return iref;
}

struct AtmArgs {
    const double *P, *T, *mix, *Zin;
    int nl_total;
    const int *layer_off;
    int natm;
    double req, g, p_ref;
    double *Z, *R, *N;
};

__global__ __launch_bounds__(64) void atmosphere_kernel(AtmArgs A) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= A.natm) return;
    const int l0 = A.layer_off[a], l1 = A.layer_off[a + 1];
    if (l0 < 0 || l1 > A.nl_total || l1 - l0 < 1) return;   // bad offsets: rows left untouched
    const size_t nlt = A.nl_total;
    const double *P = A.P, *T = A.T;
    if (A.Zin) {
        for (int l = l0; l < l1; ++l) A.Z[l] = A.Zin[l];
    } else {
        // hydrostatic z (km), deepest layer z = 0: dz = (H_i + H_{i+1})/2 ln(P_{i+1}/P_i),
        // H = R T / (m g) with one g for the atmosphere
        double Hn = kRgas * T[l1 - 1] / (mean_amu(A.mix, nlt, l1 - 1) * A.g) / 1000.0;
        double z = 0.0;
        A.Z[l1 - 1] = 0.0;
        for (int l = l1 - 2; l >= l0; --l) {
            const double H = kRgas * T[l] / (mean_amu(A.mix, nlt, l) * A.g) / 1000.0;
            z = z + 0.5 * (H + Hn) * log(P[l + 1] / P[l]);
            A.Z[l] = z;
            Hn = H;
        }
    }
    const double zoff = A.Z[ref_layer(P, l0, l1, A.p_ref)];
    for (int l = l0; l < l1; ++l) {
        A.R[l] = A.req + A.Z[l] - zoff;                                               // atm_base.py:147
        A.N[l] = refractivity(A.mix, nlt, l, T[l]) * P[l] * (293.0 / T[l]) / 1.0E6 + 1.0;   // atm_base.py:176-181
    }
}

struct PropArgs {
    const double *P, *T, *mix, *Z;
    int nl_total;
    const int *layer_off;
    int natm;
    double req, gm_ref, p_ref;
    double *props;   // [RBX_NUM_PROPS][nl_total]
};

// computeProp's property table, atm_base.py:120-194: the mass above each layer enters GM
// as a running sum from the top (dM = 1e11 rho 4 pi r^2 |dz|), then GM is referred to
// GM_ref at the layer nearest p_ref; g = GM / R^2 and H = R T / (g m) / 1000 follow
__global__ __launch_bounds__(64) void atmosphere_props_kernel(PropArgs A) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= A.natm) return;
    const int l0 = A.layer_off[a], l1 = A.layer_off[a + 1];
    if (l0 < 0 || l1 > A.nl_total || l1 - l0 < 1) return;   // bad offsets: rows left untouched
    const size_t nlt = A.nl_total;
    const double *P = A.P, *T = A.T, *Z = A.Z;
    auto row = [&](int r) { return A.props + r * nlt; };
    const int iref = ref_layer(P, l0, l1, A.p_ref);
    const double zoff = Z[iref];
    double gm = 0.0, gm_ref_layer = 0.0;
    for (int l = l0; l < l1; ++l) {
        const double R = A.req + Z[l] - zoff;
        const double amu = mean_amu(A.mix, nlt, l);
        if (l > l0) {
            const double rho = (amu * P[l]) / (kRgas * T[l]);
            const double dr = fabs(Z[l] - Z[l - 1]);
            const double dM = 1.0e11 * rho * (4.0 * kPi * (R * R) * dr);
            gm = gm + kGravConst * dM;
            const double dT = fabs(T[l] - T[l - 1]);
            row(RBX_PROP_LAPSE)[l] = dT / dr;
            row(RBX_PROP_LAPSEP)[l] = dT / fabs(P[l] - P[l - 1]);
        } else {
            row(RBX_PROP_LAPSE)[l] = 0.0;
            row(RBX_PROP_LAPSEP)[l] = 0.0;
        }
        if (l == iref) gm_ref_layer = gm;
        const double refr = refractivity(A.mix, nlt, l, T[l]) * P[l] * (293.0 / T[l]);
        row(RBX_PROP_Z)[l] = Z[l];
        row(RBX_PROP_R)[l] = R;
        row(RBX_PROP_P)[l] = P[l];
        row(RBX_PROP_GM)[l] = gm;   // the running sum, referred to GM_ref below
        row(RBX_PROP_AMU)[l] = amu;
        row(RBX_PROP_REFR)[l] = refr;
        row(RBX_PROP_N)[l] = refr / 1.0E6 + 1.0;
    }
    for (int l = l0; l < l1; ++l) {
        const double g = A.gm_ref - (row(RBX_PROP_GM)[l] - gm_ref_layer);
        const double R = row(RBX_PROP_R)[l];
        const double little_g = g / (R * R);
        row(RBX_PROP_GM)[l] = g;
        row(RBX_PROP_G)[l] = little_g;
        row(RBX_PROP_H)[l] = (kRgas * T[l]) / (little_g * row(RBX_PROP_AMU)[l]) / 1000.0;
    }
}

}  // namespace rbx

extern "C" int rbx_atmosphere_props(const double *P, const double *T, const double *mix, const double *Z,
                                    int nl_total, const int *layer_off, int natm, double req, double gm_ref,
                                    double p_ref, double *props, void *stream) {
    RBX_CHECK_ARG(P && T && mix && Z && layer_off && props, "rbx_atmosphere_props: null device pointer");
    RBX_CHECK_ARG(nl_total > 0 && natm > 0, "rbx_atmosphere_props: nl_total=%d natm=%d", nl_total, natm);
    RBX_CHECK_ARG(req > 0.0 && gm_ref > 0.0, "rbx_atmosphere_props: req and GM must be positive");
    rbx::PropArgs A;
    A.P = P;
    A.T = T;
    A.mix = mix;
    A.Z = Z;
    A.nl_total = nl_total;
    A.layer_off = layer_off;
    A.natm = natm;
    A.req = req;
    A.gm_ref = gm_ref;
    A.p_ref = p_ref;
    A.props = props;
    hipLaunchKernelGGL(rbx::atmosphere_props_kernel, dim3((natm + 63) / 64), dim3(64), 0,
                       static_cast<hipStream_t>(stream), A);
    return rbx::launch_status("atmosphere_props_kernel");
}

extern "C" int rbx_atmosphere_batch(const double *P, const double *T, const double *mix, const double *Z_in,
                                    int nl_total, const int *layer_off, int natm, double req, double gm,
                                    double p_ref, double *Z, double *R, double *N, void *stream) {
    RBX_CHECK_ARG(P && T && mix && layer_off && Z && R && N, "rbx_atmosphere_batch: null device pointer");
    RBX_CHECK_ARG(nl_total > 0 && natm > 0, "rbx_atmosphere_batch: nl_total=%d natm=%d", nl_total, natm);
    RBX_CHECK_ARG(req > 0.0 && gm > 0.0, "rbx_atmosphere_batch: req and GM must be positive");
    rbx::AtmArgs A;
    A.P = P;
    A.T = T;
    A.mix = mix;
    A.Zin = Z_in;
    A.nl_total = nl_total;
    A.layer_off = layer_off;
    A.natm = natm;
    A.req = req;
    A.g = gm / (req * req);
    A.p_ref = p_ref;
    A.Z = Z;
    A.R = R;
    A.N = N;
    hipLaunchKernelGGL(rbx::atmosphere_kernel, dim3((natm + 63) / 64), dim3(64), 0,
                       static_cast<hipStream_t>(stream), A);
    return rbx::launch_status("atmosphere_kernel");
}
