// Atmosphere construction on the device: computeProp's radii and refractive
// indices (atm_base.py:120-181), optionally after a hydrostatic recomputation of
// z, for many atmospheres laid end to end.  One thread per atmosphere walks its
// layers (a few hundred steps of a handful of FP64 operations); a retrieval
// batch of 1024 atmospheres is 16 wavefronts.
#include "common.hpp"

namespace rbx {

// chemistry.py:7-46: gas constant, molar masses (g/mol) and microwave
// refractivity coefficients of the species in the rows of mix (RBX_X_* order);
// PH3 has no refractivity there (NULL), H2O's depends on T, and CO has neither
// (ConstituentProperties gives it the NULL entries, 0, chemistry.py:168-200).
constexpr double kRgas = 8.314462;
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
    // the layer nearest p_ref (first of equals), atm_base.py:135-144
    int iref = l0;
    double psep = 1.0E6;
    for (int l = l0; l < l1; ++l) {
        const double d = fabs(P[l] - A.p_ref);
        if (d < psep) {
            psep = d;
            iref = l;
        }
    }
    const double zoff = A.Z[iref];
    for (int l = l0; l < l1; ++l) {
        A.R[l] = A.req + A.Z[l] - zoff;                                               // atm_base.py:147
        A.N[l] = refractivity(A.mix, nlt, l, T[l]) * P[l] * (293.0 / T[l]) / 1.0E6 + 1.0;   // atm_base.py:176-181
    }
}

}  // namespace rbx

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
