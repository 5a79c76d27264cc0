"""Seeded synthetic inputs for the benchmark workloads.

The five workloads (Jupiter-like atmospheres, synthetic line lists, ray sets)
are generated here so that tests, the benchmark and users build identical
inputs from a seed.  Nothing here runs on the GPU.

Choices not fixed by the workload description are recorded in
``WORKLOADS[...]['notes']`` and returned in ``meta`` of every generator.
"""
from dataclasses import dataclass, field

import numpy as np

SEED_BASE = 20261202
JUPITER_REQ = 71492.0          # km, equatorial radius at 1 bar
JUPITER_GM = 126686538.0       # km^3/s^2
ELLIPTICITY = 0.065            # (Req - Rpol) / Req of the synthetic oblate planet
MHZ_PER_TORR_TO_GHZ_PER_BAR = 0.75006
LGINT_TO_CM_PER_MOLECULE = 1.0 / 2.99792458e18   # JPL nm^2 MHz -> cm^-1/(molecule cm^-2)
LGINT_TO_HZ_CM2 = 1.0e-8                         # JPL nm^2 MHz -> cm^2 Hz (h2o_bk I_o units)
HCK = 1.438396                                   # K cm

# gas-table rows (the reference's Jupiter config.C ordering)
GAS_ROWS = {'Z': 0, 'T': 1, 'P': 2, 'H2': 3, 'HE': 4, 'CH4': 5, 'NH3': 6, 'H2O': 7,
            'H2S': 8, 'SOLN': 9, 'OTHER': 10, 'PH3': 11, 'CO': 12, 'CO13': 13,
            'HCN': 14, 'DZ': 15}
MIXING = {'H2': 0.86, 'HE': 0.136, 'NH3': 3.0e-4, 'H2O': 1.0e-3, 'H2S': 5.0e-5,
          'PH3': 6.0e-7, 'CH4': 2.0e-3}
AMU = {'H2': 2.016, 'HE': 4.003, 'H2S': 34.076, 'NH3': 17.030, 'H2O': 18.015,
       'CH4': 16.04, 'PH3': 33.997}
REFR = {'H2': 124.43, 'HE': 35.832, 'H2S': 2247.0, 'NH3': 2700.0, 'CH4': 413.0}
R_GAS = 8.314462


def psat_nh3_ice(T):
    """Briggs & Sackett NH3-over-ice saturation pressure (bar)."""
    return np.exp(-4122.0 / T + 27.8632 - 1.8163 * np.log(T))


def psat_h2o_ice(T):
    """Briggs & Sackett H2O-over-ice saturation pressure (bar)."""
    return np.exp(-5631.1206 / T - 22.1791 + 8.2312 * np.log(T)
                  - 3.861449e-2 * T + 2.77494e-5 * T * T)


def refractivity(key, T):
    if key == 'H2O':
        return 245.0 + 1.28E6 / T
    return REFR.get(key, 0.0)


@dataclass
class SynthAtmosphere:
    """A synthetic atmosphere laid out like the reference's ``atm``.

    gas: (16, nl) rows per GAS_ROWS, index 0 = top (lowest P);
    R, N: layer radii (km) and refractive indices (ray-path inputs).
    """
    gas: np.ndarray
    R: np.ndarray
    N: np.ndarray
    Req: float
    Rpol: float
    C: dict = field(default_factory=lambda: dict(GAS_ROWS))
    meta: dict = field(default_factory=dict)

    @property
    def nl(self):
        return self.gas.shape[1]

    @property
    def P(self):
        return self.gas[self.C['P']]

    @property
    def T(self):
        return self.gas[self.C['T']]


def jupiter_like(nl, T_offset=None, nh3_scale=None, ellipticity=ELLIPTICITY,
                 pmin=1e-3, pmax=100.0):
    """Jupiter-like atmosphere on nl log-spaced layers from pmin to pmax bar.

    T = max(110, 165 P^0.286) (+ T_offset per layer); NH3 3e-4 capped at
    Psat_ice/P above 0.7 bar (x nh3_scale per layer); H2O 1e-3 capped at
    Psat_ice/P; z by hydrostatic integration, R = Req + z - z(1 bar).
    """
    P = np.logspace(np.log10(pmin), np.log10(pmax), nl)
    T = np.maximum(110.0, 165.0 * P ** 0.286)
    if T_offset is not None:
        T = T + np.asarray(T_offset, dtype=np.float64)
    gas = np.zeros((len(GAS_ROWS), nl))
    gas[GAS_ROWS['P']] = P
    gas[GAS_ROWS['T']] = T
    for k in ('H2', 'HE', 'H2S', 'PH3', 'CH4'):
        gas[GAS_ROWS[k]] = MIXING[k]
    nh3 = np.where(P < 0.7, np.minimum(MIXING['NH3'], psat_nh3_ice(T) / P), MIXING['NH3'])
    if nh3_scale is not None:
        nh3 = nh3 * np.asarray(nh3_scale, dtype=np.float64)
    gas[GAS_ROWS['NH3']] = nh3
    gas[GAS_ROWS['H2O']] = np.minimum(MIXING['H2O'], psat_h2o_ice(T) / P)
    # hydrostatic z (km): dz = H dlnP, H = R T / (m g) with g = GM/Req^2
    mbar = sum(AMU[k] * gas[GAS_ROWS[k]] for k in AMU)
    g = JUPITER_GM / JUPITER_REQ ** 2
    H = R_GAS * T / (mbar * g) / 1000.0
    z = np.zeros(nl)
    for i in range(nl - 2, -1, -1):
        z[i] = z[i + 1] + 0.5 * (H[i] + H[i + 1]) * np.log(P[i + 1] / P[i])
    gas[GAS_ROWS['Z']] = z
    gas[GAS_ROWS['DZ']] = np.append(0.0, np.abs(np.diff(z)) * 1.0e5)
    i_ref = int(np.argmin(np.abs(P - 1.0)))
    R = JUPITER_REQ + z - z[i_ref]
    refr = np.zeros(nl)
    for k in ('H2', 'HE', 'CH4', 'NH3', 'H2O', 'H2S'):
        refr += refractivity(k, T) * gas[GAS_ROWS[k]]
    N = refr * P * (293.0 / T) / 1.0E6 + 1.0
    Req = JUPITER_REQ
    Rpol = JUPITER_REQ * (1.0 - ellipticity)
    return SynthAtmosphere(gas=gas, R=R, N=N, Req=Req, Rpol=Rpol,
                           meta=dict(nl=nl, pmin=pmin, pmax=pmax, ellipticity=ellipticity,
                                     CH4=MIXING['CH4']))


def _line_params(rng, n):
    f0 = np.exp(rng.uniform(np.log(0.5), np.log(1000.0), n))
    lgint = rng.uniform(-6.0, -1.0, n)
    E = rng.uniform(0.0, 2000.0, n)
    w = rng.uniform(1.0, 3.0, (3, n)) * MHZ_PER_TORR_TO_GHZ_PER_BAR   # H2, He, self
    return f0, lgint, E, w


def nh3_inversion_lines(n, seed):
    """Synthetic NH3 list in the nh3_hs inversion-block schema (ammonia.npz keys).

    fo GHz; Io = 10^LGINT / 2.99792458e18; Eo = E'' (cm^-1);
    gammaNH3o = self width (GHz/bar); H2HeBroad = H2 width (unused by nh3_hs).
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    f0, lgint, E, w = _line_params(rng, n)
    return {'fo': f0, 'Io': 10.0 ** lgint * LGINT_TO_CM_PER_MOLECULE, 'Eo': E,
            'gammaNH3o': w[2], 'H2HeBroad': w[0],
            'meta_width_h2_he': w[:2].copy(), 'meta_T_exponent': 0.67}


def h2o_lines(n, seed):
    """Synthetic H2O list in the h2o_bk schema (already in GHz/bar, no /1e-3)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    f0, lgint, E, w = _line_params(rng, n)
    return {'f_o': f0, 'I_o': 10.0 ** lgint * LGINT_TO_HZ_CM2, 'E_o': E * HCK / 300.0,
            'w_s': w[2], 'x_s': np.full(n, 0.67), 'w_h2': w[0], 'w_he': w[1],
            'x_h2': np.full(n, 0.67), 'x_he': np.full(n, 0.67), 'SR': np.zeros(n)}


def image_rays(n=64, extent=1.0):
    """n x n pixel-centre grid over [-extent, extent]^2 (row-major, y outer)."""
    c = -extent + (2.0 * np.arange(n) + 1.0) * extent / n
    bx, by = np.meshgrid(c, c)
    return np.stack([bx.ravel(), by.ravel()], axis=1)


def random_disc_rays(n, seed, rmax=1.02):
    rng = np.random.Generator(np.random.PCG64(seed))
    r = rng.uniform(0.0, rmax, n)
    th = rng.uniform(0.0, 2.0 * np.pi, n)
    return np.stack([r * np.cos(th), r * np.sin(th)], axis=1)


_GP_FACTORS = {}


def gp_perturbation(rng, logp, scale, length):
    """Squared-exponential Gaussian-process draw over log10 P (Cholesky factors cached
    per grid, scale and length)."""
    key = (len(logp), float(logp[0]), float(logp[-1]), scale, length)
    L = _GP_FACTORS.get(key)
    if L is None:
        d = logp[:, None] - logp[None, :]
        K = scale ** 2 * np.exp(-0.5 * (d / length) ** 2) + 1e-10 * np.eye(len(logp))
        L = _GP_FACTORS[key] = np.linalg.cholesky(K)
    return L @ rng.standard_normal(len(logp))


WORKLOADS = {
    'c1_plain_fp64': dict(nl=150, nh3_lines=2000, h2o_lines=0, freqs=('arange', 1.0, 96.0, 5.0),
                          rays='nadir', notes='1 nadir point ray'),
    'c2_image': dict(nl=400, nh3_lines=10000, h2o_lines=0, freqs=('logspace', 1.0, 100.0, 512),
                     rays=('image', 64), notes='512 log-spaced freqs; 64x64 pixel centres over [-1,1]^2'),
    'c3_large': dict(nl=800, nh3_lines=20000, h2o_lines=5000, freqs=('logspace', 0.3, 300.0, 1024),
                     rays=('random', 16384, 1.02), notes='radius U[0,1.02], angle U[0,2pi)'),
    'c4_jacobian': dict(nl=500, nh3_lines=10000, h2o_lines=0, freqs=('linspace', 1.0, 50.0, 64),
                        rays=('line', 256, 0.99), notes='10000 NH3 lines (unspecified; C2 value)'),
    'c5_retrieval': dict(nl=(200, 800), nh3_lines=10000, h2o_lines=0, freqs=('linspace', 1.0, 50.0, 128),
                         rays=('random', 16, 0.95), n_atm=1024,
                         notes='GP over log10 P, length 0.5 dex; T sigma 2.5 K clipped +-5; '
                               'NH3 x exp(GP sigma 0.55) clipped [0.3,3]'),
}


def frequencies(spec):
    kind = spec[0]
    if kind == 'arange':          # set_utils.proc_string 'a:b:step' (stop inclusive)
        a, b, s = spec[1:]
        return np.arange(a, b + s / 2.0, s)
    if kind == 'logspace':
        a, b, n = spec[1:]
        return np.logspace(np.log10(a), np.log10(b), int(n))
    if kind == 'linspace':
        a, b, n = spec[1:]
        return np.linspace(a, b, int(n))
    raise ValueError(kind)


def workload(name, index=None):
    """Build one workload: atmosphere, line lists, freqs, rays and meta.

    index is the workload's position (0..4), used in the seed."""
    spec = WORKLOADS[name]
    if index is None:
        index = list(WORKLOADS).index(name)
    seed = SEED_BASE + index
    ss = np.random.SeedSequence(seed)
    s_atm, s_lines, s_h2o, s_rays = [int(c.generate_state(1)[0]) for c in ss.spawn(4)]
    freqs = frequencies(spec['freqs'])
    nl = spec['nl']
    out = dict(name=name, seed=seed, freqs=freqs, meta=dict(spec))
    if isinstance(nl, tuple):
        out['atm'] = None
    else:
        out['atm'] = jupiter_like(nl)
    out['nh3_inv'] = nh3_inversion_lines(spec['nh3_lines'], s_lines)
    out['h2o'] = h2o_lines(spec['h2o_lines'], s_h2o) if spec['h2o_lines'] else None
    r = spec['rays']
    if r == 'nadir':
        rays = np.zeros((1, 2))
    elif r[0] == 'image':
        rays = image_rays(r[1])
    elif r[0] == 'random':
        rays = random_disc_rays(r[1], s_rays, r[2])
    elif r[0] == 'line':
        rays = np.stack([np.linspace(0.0, r[2], r[1]), np.zeros(r[1])], axis=1)
    out['rays'] = rays
    out['seeds'] = dict(atm=s_atm, lines=s_lines, h2o=s_h2o, rays=s_rays)
    return out


def retrieval_batch(n_atm=1024, seed=SEED_BASE + 4, nl_range=(200, 800)):
    """Perturbed atmospheres for the retrieval workload (T +-5 K, NH3 x[0.3,3])."""
    rng = np.random.Generator(np.random.PCG64(seed))
    atms = []
    for _ in range(n_atm):
        nl = int(rng.integers(nl_range[0], nl_range[1] + 1))
        logp = np.linspace(-3.0, 2.0, nl)
        dT = np.clip(gp_perturbation(rng, logp, 2.5, 0.5), -5.0, 5.0)
        snh3 = np.clip(np.exp(gp_perturbation(rng, logp, 0.55, 0.5)), 0.3, 3.0)
        atms.append(jupiter_like(nl, T_offset=dT, nh3_scale=snh3))
    return atms
