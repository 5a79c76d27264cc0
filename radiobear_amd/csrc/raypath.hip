// Ray geometry in FP64: one thread per ray.
//
// Follows raypath.compute_ds / findEdge (raypath.py:60-273) and the ellipse
// branch of Shape (shape.py:223-298) step for step.  The reference converts
// latitude / longitude to degrees and back and rebuilds sines through
// arcsin/arccos; here the same quantities come from the vector components
// (sin lat = y/|v|, cos lat = sqrt(x^2+z^2)/|v|, cos lng = z/sqrt(x^2+z^2), ...),
// which is the same arithmetic up to rounding, with no transcendental calls in
// the loop.  Quirks kept: the latitude-0 substitution lat = 1e-6 rad
// (shape.py:231-233), the next-shell radius taken at the current point's
// latitude (raypath.py:182), refraction only at the first boundary
// (raypath.py:158, 257), the numpy arange trial grid and linear root of
// findEdge (raypath.py:64, 94-97), the ds < 0 truncation (raypath.py:212-216),
// NaN (not an exception) when a ray misses the next shell (raypath.py:190-209),
// and the loop end when layer + 1 runs past the last layer (raypath.py:255-264).
#include "common.hpp"

namespace rbx {

struct V3 {
    double x, y, z;
};
__device__ inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline double norm(V3 a) { return sqrt(dot(a, a)); }

struct Geo {
    double ratio;   // Rpol/Req for 'ellipse', 1 for 'sphere'
    double M[9];    // rotate2planet = rotX(rotate) * rotZ(tip), row-major
};

__device__ inline V3 mat(const double *M, V3 v) {
    return {M[0] * v.x + M[1] * v.y + M[2] * v.z, M[3] * v.x + M[4] * v.y + M[5] * v.z,
            M[6] * v.x + M[7] * v.y + M[8] * v.z};
}

// sines/cosines of (planetocentric latitude, delta longitude) of vector v
struct LatLng {
    double slat, clat, slng, clng;
};

__device__ inline LatLng latlng(V3 v) {
    LatLng o;
    const double r = norm(v);
    const double h = sqrt(v.x * v.x + v.z * v.z);
    if (v.y == 0.0) {                 // lat == 0 -> 1e-6 rad (shape.py:231-233)
        o.slat = 9.999999999998333e-07;   // sin(1e-6)
        o.clat = 0.9999999999995;         // cos(1e-6)
    } else {
        o.slat = v.y / r;
        o.clat = h / r;
    }
    if (h > 0.0) {
        o.slng = v.x / h;
        o.clng = v.z / h;
    } else {                           // atan2(0, 0) = 0
        o.slng = 0.0;
        o.clng = 1.0;
    }
    return o;
}

// Shape._calcEllipse for equatorial radius a at (lat, lng): |r|, r and normal n
struct Surf {
    double rmag;
    V3 r, n;
};

__device__ inline Surf ellipse(const Geo &G, double a, const LatLng &q) {
    const double b = G.ratio * a;
    V3 nn = {0.0, a * q.slat, b * q.clat};
    const double inv = 1.0 / sqrt(nn.y * nn.y + nn.z * nn.z);
    nn.y *= inv;
    nn.z *= inv;
    Surf s;
    s.n = {q.slng * nn.z, nn.y, q.clng * nn.z};                      // rotY(lng, norm)
    const double rc = a * q.clat;
    s.r = {q.slng * rc, b * q.slat, q.clng * rc};                    // rotY(lng, r_vec)
    s.rmag = norm(s.r);
    return s;
}

// the shell quantities at a path point v that depend only on its latitude: f, the radius of
// the unit-equatorial-radius ellipse there (a shell of equatorial radius a has a f), and the
// shell normal n.  With X = x^2 + z^2, Y = y^2 and sin lat = y/|v|, cos lat = sqrt(X)/|v|,
// ellipse() reduces to f = sqrt((X + ratio^2 Y) / (X + Y)) and
// n = (ratio x, y, ratio z) rsqrt(ratio^2 X + Y) (the same values up to rounding, with one
// division, one square root and one reciprocal square root in parallel instead of latlng's and
// ellipse's chain: ray paths 0.66 -> 0.41 ms on 16384 rays of 800 layers).  y == 0 takes the
// 1e-6 rad latitude substitution through latlng.
__device__ inline void shell_at(const Geo &G, V3 v, double &f, V3 &n) {
    if (v.y == 0.0) {
        const Surf s = ellipse(G, 1.0, latlng(v));
        f = s.rmag;
        n = s.n;
        return;
    }
    const double X = v.x * v.x + v.z * v.z, Y = v.y * v.y, r2 = G.ratio * G.ratio;
    f = sqrt(fma(r2, Y, X) / (X + Y));
    const double inv = rsqrt(fma(r2, X, Y));
    n = {G.ratio * v.x * inv, v.y * inv, G.ratio * v.z * inv};
}

// Rays per block.  One thread per ray: a ray's path is one dependent chain of FP64 steps, so
// the kernel is latency-bound and holds few wavefronts (256 for 16384 rays).  It runs beside the
// line sums (rbx_alpha) when a caller puts it on a second stream or graph branch; four
// wavefronts per block keep it on a quarter as many CUs as one per block, leaving the other
// CUs' slots whole for the line sums: 800 layers x 1024 frequencies x 16384 rays, graph step
// 3.71 -> 3.47 ms (the paths alone 0.42 -> 0.45 ms; 2 and 8 wavefronts per block: 3.59, 3.53 ms).
constexpr int kRayThreads = 256;

struct RayArgs {
    Geo G;
    const double *R, *N;
    int nl;
    const double *b;
    int nray;
    int units;   // 0 km, 1 half step in cm
    double *ds;
    int ld;
    int *nsteps;
    int limb;
    // several atmospheres laid end to end (null: one atmosphere of nl layers):
    // ray r uses layers [layer_off[a], layer_off[a + 1]) with a = ray_atm[r]
    const int *ray_atm, *layer_off;
    int natm;
};

__global__ __launch_bounds__(kRayThreads) void raypath_kernel(RayArgs A) {
    const int ray = blockIdx.x * blockDim.x + threadIdx.x;
    if (ray >= A.nray) return;
    const double b0 = A.b[2 * ray], b1 = A.b[2 * ray + 1];
    double *out = A.ds + static_cast<size_t>(ray) * A.ld;
    const Geo &G = A.G;
    const double *Rl = A.R, *Nl = A.N;
    int nl = A.nl;
    if (A.ray_atm) {
        const int a = A.ray_atm[ray];
        const int l0 = (a >= 0 && a < A.natm) ? A.layer_off[a] : -1;
        const int l1 = (a >= 0 && a < A.natm) ? A.layer_off[a + 1] : -1;
        if (l0 < 0 || l1 > A.nl || l1 - l0 < 2 || l1 - l0 - 1 > A.ld) {   // bad index or offsets
            A.nsteps[ray] = -2;
            return;
        }
        Rl += l0;
        Nl += l0;
        nl = l1 - l0;
    }
    const double rNorm = Rl[0];
    const double tmp = b0 * b0 + b1 * b1;
    if (!(tmp < 1.0)) {                                // raypath.py:126-127 (NaN b: off planet too)
        A.nsteps[ray] = -1;
        return;
    }
    const double mu = sqrt(1.0 - tmp);

    // ---- findEdge: trial zQ = arange(1.01 sqrt(1-b^2), 0, -0.005) ----
    const double start = sqrt(1.0 - tmp) * 1.01;
    const int ntrial = static_cast<int>(ceil((0.0 - start) / -0.005));
    const double z1 = start + -0.005;
    const double dz = z1 - start;
    double zprev = 0.0, dprev = 0.0, zhit = 0.0, dhit = 0.0;
    int hit = -1;
    for (int i = 0; i < ntrial; ++i) {
        const double zq = (i == 0) ? start : ((i == 1) ? z1 : start + i * dz);
        const V3 bv = mat(G.M, V3{b0, b1, zq});
        const double r1 = norm(bv) * rNorm;
        const Surf s = ellipse(G, rNorm, latlng(bv));
        const double dd = r1 - s.rmag;
        if (r1 < s.rmag) {
            hit = i;
            zhit = zq;
            dhit = dd;
            break;
        }
        zprev = zq;
        dprev = dd;
    }
    if (hit < 0) {
        A.nsteps[ray] = -1;
        return;
    }
    // np.interp(0, [dhit, dprev], [zhit, zprev]) on the two bracketing trials
    double zq;
    if (hit == 0 || dhit == 0.0) {
        zq = zhit;
    } else {
        const double slope = (zprev - zhit) / (dprev - dhit);
        zq = slope * (0.0 - dhit) + zhit;
        if (zq != zq) zq = slope * (0.0 - dprev) + zprev;
    }
    const V3 edge = rNorm * mat(G.M, V3{b0, b1, zq});

    // ---- compute_ds: start on the top shell at the edge's latitude ----
    const LatLng q = latlng(edge);
    const Surf geo = ellipse(G, rNorm, q);
    V3 s = mat(G.M, V3{0.0, 0.0, -1.0});
    V3 r = geo.r;
    V3 n = geo.n;
    double c = -dot(s, n);                            // cos(t_inc)
    double nratio = Nl[0] / Nl[1];
    double st = nratio * sqrt(1.0 - c * c);          // sin(t_tran)
    double ct = sqrt(1.0 - st * st);                 // cos(t_tran)
    const double cm_half = (A.units == 1);
    int layer = 0;
    int count = 0;
    double rNow = geo.rmag;                           // shell layer at the current point
    double rNext = ellipse(G, Rl[1], q).rmag;         // shell layer + 1 at the same latitude (raypath.py:182)
    // equatorial radii of shells layer + 1 and layer + 2, the latter loaded a step ahead (its
    // load latency hides behind one step of the chain)
    double R1 = Rl[1], R2 = nl > 2 ? Rl[2] : 0.0;
    while (true) {
        s = nratio * s + (nratio * c - ct) * n;      // raypath.py:179-180 (ingress)
        const double rdots = dot(r, s);
        const double disc = rdots * rdots + rNext * rNext - rNow * rNow;
        double ds = -rdots - sqrt(disc);             // NaN when the shell is missed
        if (ds < 0.0) break;                         // raypath.py:212-216
        if (A.limb == 1) ds = fabs(rNext - rNow) / mu;
        out[count] = cm_half ? (ds * 1.0E3 / 1.0E-2) * 0.5 : ds;
        ++count;
        const V3 rn = r + ds * s;
        double f;
        shell_at(G, rn, f, n);
        rNow = R1 * f;
        r = rn;
        ++layer;
        c = -dot(s, n);
        if (layer + 1 >= nl) break;                  // IndexError on nr[layer + 1]
        rNext = R2 * f;
        R1 = R2;
        R2 = layer + 2 < nl ? Rl[layer + 2] : 0.0;
        nratio = 1.0;
        st = sqrt(1.0 - c * c);
        ct = sqrt(1.0 - st * st);
    }
    A.nsteps[ray] = count;
}

}  // namespace rbx

// planet shape, view and sizes shared by rbx_raypaths and rbx_raypaths_batch
static int fill_ray_args(rbx::RayArgs &A, const rbx_planet *planet, const double *R, const double *N, int nl,
                         const double *b, int nray, int out_units, double *ds, int ld, int *nsteps, const char *who) {
    RBX_CHECK_ARG(planet && R && N && b && ds && nsteps, "%s: null argument", who);
    RBX_CHECK_ARG(nl >= 2, "%s: need at least 2 layers (nl=%d)", who, nl);
    RBX_CHECK_ARG(nray >= 0, "%s: negative ray count", who);
    RBX_CHECK_ARG(out_units == 0 || out_units == 1, "%s: bad out_units %d", who, out_units);
    RBX_CHECK_ARG(planet->gtype == 0 || planet->gtype == 1, "%s: bad gtype %d", who, planet->gtype);
    RBX_CHECK_ARG(planet->limb == 0 || planet->limb == 1, "%s: bad limb %d", who, planet->limb);
    RBX_CHECK_ARG(planet->req > 0.0 && planet->rpol > 0.0, "%s: radii must be positive", who);
    // computeAspect (raypath.py:39-44)
    const double f = 1.0 - planet->rpol / planet->req;
    const double tip = -planet->orientation[0] * rbx::kPi / 180.0;
    const double rot = -std::atan(std::tan(planet->orientation[1] * rbx::kPi / 180.0) * (1.0 - f) * (1.0 - f));
    const double ct = std::cos(tip), st = std::sin(tip), cr = std::cos(rot), sr = std::sin(rot);
    // rotX(rot) * rotZ(tip)
    const double M[9] = {ct, -st, 0.0, cr * st, cr * ct, -sr, sr * st, sr * ct, cr};
    for (int i = 0; i < 9; ++i) A.G.M[i] = M[i];
    A.G.ratio = planet->gtype == 0 ? planet->rpol / planet->req : 1.0;
    A.R = R;
    A.N = N;
    A.nl = nl;
    A.b = b;
    A.nray = nray;
    A.units = out_units;
    A.ds = ds;
    A.ld = ld;
    A.nsteps = nsteps;
    A.limb = planet->limb;
    A.ray_atm = nullptr;
    A.layer_off = nullptr;
    A.natm = 0;
    return RBX_OK;
}

static int launch_rays(const rbx::RayArgs &A, void *stream) {
    const int threads = rbx::kRayThreads;
    hipLaunchKernelGGL(rbx::raypath_kernel, dim3((A.nray + threads - 1) / threads), dim3(threads), 0,
                       static_cast<hipStream_t>(stream), A);
    return rbx::launch_status("raypath_kernel");
}

extern "C" int rbx_raypaths(const rbx_planet *planet, const double *R, const double *N, int nl, const double *b,
                            int nray, int out_units, double *ds, int ld, int *nsteps, void *stream) {
    rbx::RayArgs A;
    int rc = fill_ray_args(A, planet, R, N, nl, b, nray, out_units, ds, ld, nsteps, "rbx_raypaths");
    if (rc) return rc;
    RBX_CHECK_ARG(ld >= nl - 1, "rbx_raypaths: ld=%d < nl-1=%d", ld, nl - 1);
    if (nray == 0) return RBX_OK;
    return launch_rays(A, stream);
}

extern "C" int rbx_raypaths_batch(const rbx_planet *planet, const double *R, const double *N, int nl_total,
                                  const int *layer_off, int natm, const double *b, int nray, const int *ray_atm,
                                  int out_units, double *ds, int ld, int *nsteps, void *stream) {
    rbx::RayArgs A;
    int rc = fill_ray_args(A, planet, R, N, nl_total, b, nray, out_units, ds, ld, nsteps, "rbx_raypaths_batch");
    if (rc) return rc;
    RBX_CHECK_ARG(layer_off && ray_atm && natm > 0, "rbx_raypaths_batch: null offsets / ray map or natm=%d", natm);
    RBX_CHECK_ARG(ld >= 1, "rbx_raypaths_batch: ld=%d", ld);
    if (nray == 0) return RBX_OK;
    A.ray_atm = ray_atm;
    A.layer_off = layer_off;
    A.natm = natm;
    return launch_rays(A, stream);
}
