"""FP64 restatement of the RadioBEAR absorber plug-ins (TEST INFRASTRUCTURE ONLY).

Every public function has the reference plug-in signature
``alpha(freq, T, P, X, P_dict, other_dict, **kwargs)`` plus an explicit
``data=`` line-list dict (the reference keeps it in a module global, loaded on
first call; here the caller passes it, already truncated with ``truncate_*``).
Units follow ``constituents/parameters.py:4-8``: 'dBperkm' by default,
cm^-1 when ``units='invcm'``.
"""
import numpy as np

DB_PER_INVCM = 434294.5  # cm^-1 -> dB/km used by every plug-in


def _units(kwargs):
    # constituents/parameters.py:4-8 (default units 'dBperkm')
    return kwargs.get('units', 'dBperkm')


# ---------------------------------------------------------------- truncation
def truncate_ftable(data, truncate_strength=None, truncate_freq=None,
                    keys=('f0', 'I0', 'E', 'GH2S')):
    """Line cut of h2s_ddb.readInputFiles (constituents/h2s/h2s_ddb.py:23-38).

    Keeps I0 > strength, then f0 < freq.  ``keys`` lists the per-line columns.
    """
    out = {k: np.asarray(v, dtype=np.float64) for k, v in data.items()}
    if truncate_strength is not None:
        keep = out['I0'] > truncate_strength
        for k in keys:
            out[k] = out[k][keep]
    if truncate_freq is not None:
        keep = out['f0'] < truncate_freq
        for k in keys:
            out[k] = out[k][keep]
    return out


def truncate_ph3(lines, weights, truncate_strength=None, truncate_freq=None):
    """Line cut of ph3_jh.readInputFiles (constituents/ph3/ph3_jh.py:18-61)."""
    d = {k: np.asarray(v, dtype=np.float64) for k, v in lines.items()}
    w = {k: np.asarray(v, dtype=np.float64) for k, v in weights.items()}
    used_I0 = None
    if truncate_strength is not None:
        used_I0 = d['I0'] > truncate_strength
        for k in ('f0', 'I0', 'E'):
            d[k] = d[k][used_I0]
    test_length = len(d['f0'])
    used_f = None
    if truncate_freq is not None:
        used_f = d['f0'] < truncate_freq
        for k in ('f0', 'I0', 'E'):
            d[k] = d[k][used_f]
    if truncate_strength:
        for k in ('WgtI0', 'WgtFGB', 'WgtSB'):
            w[k] = w[k][used_I0]
    if len(w['WgtI0']) != test_length:
        raise ValueError("PH3 wgt vector wrong length.")
    if truncate_freq:
        for k in ('WgtI0', 'WgtFGB', 'WgtSB'):
            w[k] = w[k][used_f]
    d.update(w)
    return d


def truncate_h2o(data, truncate_strength=None, truncate_freq=None):
    """Line cut of h2o_bk.readInputFiles (constituents/h2o/h2o_bk.py:51-62)."""
    d = {k: np.asarray(v, dtype=np.float64) for k, v in data.items()}
    if truncate_strength:
        keep = d['I_o'] > truncate_strength
        d = {k: v[keep] for k, v in d.items()}
    if truncate_freq:
        keep = d['f_o'] < truncate_freq
        d = {k: v[keep] for k, v in d.items()}
    return d


def h2o_bk_lines():
    """The 15 Karpowicz/Rosenkranz lines typed in h2o_bk.readInputFiles
    (constituents/h2o/h2o_bk.py:23-49); w_s there is divided by 1e-3 (mbar->bar)."""
    f_o = [22.2351, 183.3101, 321.2256, 325.1529, 380.1974, 439.1508, 443.0183,
           448.0011, 470.8890, 474.6891, 488.4911, 556.9360, 620.7008, 752.0332, 916.1712]
    I_o = [0.1314E-13, 0.2279E-11, 0.8058E-13, 0.2701E-11, 0.2444E-10, 0.2185E-11,
           0.4637E-12, 0.2568E-10, 0.8392E-12, 0.3272E-11, 0.6676E-12, 0.1535E-08,
           0.1711E-10, 0.1014E-08, 0.4238E-10]
    E_o = [2.144, 0.668, 6.179, 1.541, 1.048, 3.595, 5.048, 1.405, 3.597, 2.379,
           2.852, 0.159, 2.391, 0.396, 1.441]
    w_s = np.array([0.01349, 0.01466, 0.01057, 0.01381, 0.01454, 0.009715, 0.00788,
                    0.01275, 0.00983, 0.01095, 0.01313, 0.01405, 0.011836, 0.01253,
                    0.01275]) / 0.001
    x_s = [0.61, 0.85, 0.54, 0.74, 0.89, 0.62, 0.50, 0.67, 0.65, 0.64, 0.72, 1.0,
           0.68, 0.84, 0.78]
    w_h2 = [2.395, 2.4000] + [2.395, 2.395, 2.390] + [2.395] * 10
    w_he = [0.67, 0.71, 0.67, 0.67, 0.63] + [0.67] * 10
    x_h2 = [0.900, 0.950, 0.900, 0.900, 0.850] + [0.900] * 10
    x_he = [0.515, 0.490, 0.515, 0.490, 0.540] + [0.515] * 10
    return {k: np.asarray(v, dtype=np.float64) for k, v in dict(
        f_o=f_o, I_o=I_o, E_o=E_o, w_s=w_s, x_s=x_s, w_h2=w_h2, w_he=w_he,
        x_h2=x_h2, x_he=x_he, SR=np.zeros(15)).items()}


# ------------------------------------------------------------------- h2_jj_ddb
def h2_jj_ddb(freq, T, P, X, P_dict, other_dict, **kwargs):
    """H2-H2/He/CH4 collision-induced absorption (constituents/h2/h2_jj_ddb.py:7-38)."""
    coef = 3.9522E-14
    T0 = 273.0
    P_h2 = P * X[P_dict['H2']]
    P_he = P * X[P_dict['HE']]
    P_ch4 = P * X[P_dict['CH4']]
    th = T0 / T
    state = other_dict['h2state']
    if state == 'e':                          # h2_jj_ddb.py:17-23
        pre = min((T / 55.0) ** 2.7, 1.0)
        pre = min(pre * (T / 120.0) ** 0.55, 1.0)
    elif state == 'n':                        # h2_jj_ddb.py:24-27
        pre = min((T / 40.0) ** 2.5, 1.0)
    else:                                     # h2_jj_ddb.py:28-30
        return 0.0
    out = []
    for f in freq:
        cf = coef * f ** 2 * P_h2 * pre
        a = cf * (P_h2 * pow(th, 3.12) + 1.382 * P_he * pow(th, 2.24)
                  + 9.322 * P_ch4 * pow(th, 3.34))
        if _units(kwargs) == 'dBperkm':
            a *= DB_PER_INVCM
        out.append(a)
    return out


# --------------------------------------------------------------------- h2o_bk
H2O_TO = 300.0
H2O_NA = 6.0221415e23
H2O_M_AMU = 8.314472 / 0.46151805
H2O_ISO = 0.997317


def _vvw_modified(f, fo, df, shift, S):
    """vvwlinecontribution_modified (constituents/h2o/h2o_bk.py:130-189)."""
    f = np.asarray(f, dtype=np.float64)[None, :]
    fo = fo[:, None]
    dfm = df[:, None]
    sh = shift[:, None]
    base = (df / (562500.0 + df ** 2))[:, None]
    A = np.square(f / fo) / np.pi
    B = dfm / (np.square(f - fo - sh) + np.square(dfm))
    C = dfm / (np.square(f + fo + sh) + np.square(dfm))
    F = A * (B - base + C - base)
    return np.sum(S[:, None] * F, axis=0)


def h2o_bk(freq, T, P, X, P_dict, other_dict, data=None, **kwargs):
    """Karpowicz/Rosenkranz H2O lines + continuum (constituents/h2o/h2o_bk.py:65-127)."""
    if data is None:
        data = h2o_bk_lines()
    P_h2 = P * X[P_dict['H2']]
    P_he = P * X[P_dict['HE']]
    P_h2o = P * X[P_dict['H2O']]
    Theta = H2O_TO / T
    density = (H2O_M_AMU * P_h2o) / (8.314472e-5 * T)          # h2o_bk.py:85
    density = H2O_ISO * (density / H2O_M_AMU) * H2O_NA * (1.0 / 1e6)  # h2o_bk.py:88
    S = data['I_o'] * (Theta ** 2.5) * np.exp(data['E_o'] * (1.0 - Theta))
    df = (data['w_s'] * P_h2o * np.power(Theta, data['x_s'])
          + data['w_h2'] * P_h2 * np.power(Theta, data['x_h2'])
          + data['w_he'] * P_he * np.power(Theta, data['x_he']))
    FSsum = _vvw_modified(freq, data['f_o'], df, data['SR'], S)
    lines = 4.342945 * 1e-4 * density * FSsum                      # h2o_bk.py:99
    Cf_he = ((1.0 / 0.001) ** 2) * 1.03562010226e-10               # h2o_bk.py:102-105
    Cf_h2 = ((1.0 / 0.001) ** 2) * 5.07722009423e-11
    Cs1 = 3.1e-07 * pow(Theta, 12.0)
    Cs2 = 0.0
    farr = np.asarray(freq, dtype=np.float64)
    foreign = (Cf_he * P_he * P_h2o * farr ** 2 * pow(Theta, 3.0)
               + Cf_h2 * P_h2 * P_h2o * farr ** 2 * pow(Theta, 3.0))
    self_c = Cs1 * ((P_h2o / 0.001) ** 2) * farr ** 2 + Cs2 * farr ** 2
    a = lines + 4.342945 * foreign + 4.342945 * self_c             # h2o_bk.py:121
    if _units(kwargs) != 'dBperkm':
        a = a / DB_PER_INVCM
    return list(np.ravel(a))


# ------------------------------------------------- Ben-Reuven sum (h2s/ph3/sjs)
GEISA_COEF = 7.244E+21
HCK = 1.438396
GHZ_PER_INVCM = 29.9792458


def _br_freq_sum(f, f0, gamma, zeta, delta, ITG):
    """One frequency of the shared Ben-Reuven sum, e.g. h2s_ddb.py:76-79."""
    f2 = f ** 2
    g2 = gamma ** 2
    z2 = zeta ** 2
    num = (gamma - zeta) * f2 + (gamma + zeta) * (np.power(f0 + delta, 2.0) + g2 - z2)
    den = np.power(f2 - np.power(f0 + delta, 2.0) - g2 + z2, 2.0) + 4.0 * f2 * g2
    shape = GHZ_PER_INVCM * 2.0 * np.power(f / f0, 2.0) * num / (np.pi * den)
    return np.sum(shape * ITG)


def h2s_ddb(freq, T, P, X, P_dict, other_dict, data=None, **kwargs):
    """H2S Ben-Reuven with zeta = gamma (constituents/h2s/h2s_ddb.py:42-87)."""
    T0 = 296.0
    P_h2 = P * X[P_dict['H2']]
    P_he = P * X[P_dict['HE']]
    P_h2s = P * X[P_dict['H2S']]
    f0, I0, E, GH2S = data['f0'], data['I0'], data['E'], data['GH2S']
    delta = 1.28 * P_h2s
    gamma = np.power(T0 / T, 0.7) * (1.960 * P_h2 + 1.200 * P_he + GH2S * P_h2s)
    ITG = I0 * np.exp(-((1.0 / T) - (1.0 / T0)) * E * HCK)
    out = np.array([_br_freq_sum(f, f0, gamma, gamma, delta, ITG) for f in freq])
    out = GEISA_COEF * (P_h2s / T0) * pow(T0 / T, 1.5 + 2) * out
    if _units(kwargs) == 'dBperkm':
        out *= DB_PER_INVCM
    return out


def ph3_jh(freq, T, P, X, P_dict, other_dict, data=None, **kwargs):
    """PH3 Ben-Reuven with zeta = delta = 0 and per-line weights
    (constituents/ph3/ph3_jh.py:64-108).  ``data`` holds f0,I0,E and the weights."""
    T0 = 300.0
    P_h2 = P * X[P_dict['H2']]
    P_he = P * X[P_dict['HE']]
    P_ph3 = P * X[P_dict['PH3']]
    gamma = (np.power(T0 / T, 2.0 / 3.0) * (3.2930 * P_h2 + 1.6803 * P_he) * data['WgtFGB']
             + np.power(T0 / T, 1.0) * 4.2157 * P_ph3 * data['WgtSB'])
    ITG = data['I0'] * data['WgtI0'] * np.exp(-((1.0 / T) - (1.0 / T0)) * data['E'] * HCK)
    out = np.array([_br_freq_sum(f, data['f0'], gamma, 0.0, 0.0, ITG) for f in freq])
    out = GEISA_COEF * (P_ph3 / T0) * np.power(T0 / T, 1.5 + 2) * out
    if _units(kwargs) == 'dBperkm':
        out *= DB_PER_INVCM
    return out


# ---------------------------------------------------------------------- nh3_sjs
def _sjs_coefficients(T, P_h2):
    """Joiner / Spilker coefficient sets (constituents/nh3/nh3_sjs.py:42-86).

    Returns lists indexed [Joiner, Spilker, Interp-placeholder]."""
    EPS = 1E-12
    c = dict(GH2=[1.690], GHe=[0.750], GNH3=[0.6], ZH2=[1.350], ZHe=[0.300],
             ZNH3=[0.200], C=[1.0], D=[-0.45])
    rexp = 8.79 * np.exp(-T / 83.0)
    GH2a = np.exp(9.024 - T / 20.3) - 0.9918 + P_h2
    with np.errstate(invalid='ignore'):
        GH2a = np.power(np.float64(GH2a), rexp)   # NaN (not ValueError) for a negative base
    if GH2a < EPS:
        for k, v in dict(GH2=1.690, GHe=0.750, GNH3=0.60, ZH2=1.35, ZHe=0.30,
                         ZNH3=0.20, C=1.00, D=-0.45).items():
            c[k].append(v)
    else:
        GH2a = 2.122 * np.exp(-T / 116.8) / GH2a
        GH2a = 2.34 * (1.0 - GH2a)
        c['GH2'].append(GH2a)
        c['GHe'].append(0.46 + T / 3000.0)
        c['GNH3'].append(0.74)
        c['ZH2'].append(5.7465 - 7.7644 * GH2a + 9.1931 * GH2a ** 2
                        - 5.6816 * GH2a ** 3 + 1.2307 * GH2a ** 4)
        c['ZHe'].append(0.28 - T / 1750.0)
        c['ZNH3'].append(0.50)
        c['C'].append(-0.337 + T / 110.4 - T ** 2 / 70600.0)
        c['D'].append(-0.45)
    for k in c:
        c[k].append(0.0)
    return c


def nh3_sjs(freq, T, P, X, P_dict, other_dict, data=None, **kwargs):
    """Spilker/Joiner Ben-Reuven NH3 (constituents/nh3/nh3_sjs.py:26-128)."""
    Joiner, Spilker, Interp = 0, 1, 2
    T0 = 296.0
    fLower, fHigher = 26.0, 34.0
    P_h2 = P * X[P_dict['H2']]
    P_he = P * X[P_dict['HE']]
    P_nh3 = P * X[P_dict['NH3']]
    Pscale = 1.0 + P / 1.0E5
    c = _sjs_coefficients(T, P_h2)
    f0, I0, E, G0 = data['f0'], data['I0'], data['E'], data['G0']
    n_dvl = 2.0 / 3.0
    ITG = I0 * np.exp(-((1.0 / T) - (1.0 / T0)) * E * HCK)
    out = []
    for f in freq:
        f2 = f ** 2
        if f <= fLower:
            use = Spilker
        elif f >= fHigher:
            use = Joiner
        else:
            use = Interp
            flfh = (fLower - fHigher) / (f - fLower)
            for k in c:
                c[k][Interp] = c[k][Spilker] + (c[k][Spilker] - c[k][Joiner]) / flfh
        delta = c['D'][use] * P_nh3
        tf = np.power(T0 / T, n_dvl)
        gamma = tf * (c['GH2'][use] * P_h2 + c['GHe'][use] * P_he + G0 * c['GNH3'][use] * P_nh3)
        zeta = tf * (c['ZH2'][use] * P_h2 + c['ZHe'][use] * P_he + G0 * c['ZNH3'][use] * P_nh3)
        g2 = gamma ** 2
        z2 = zeta ** 2
        num = (gamma - zeta) * f2 + (gamma + zeta) * (np.power(f0 + delta, 2.0) + g2 - z2)
        den = np.power(f2 - np.power(f0 + delta, 2.0) - g2 + z2, 2.0) + 4.0 * f2 * g2
        shape = GHZ_PER_INVCM * 2.0 * np.power(f / f0, 2.0) * num / (np.pi * den)
        out.append(np.sum(shape * ITG))
    out = GEISA_COEF * (P_nh3 / T0) * pow(T0 / T, 1.5 + 2) * np.array(out) * Pscale
    if _units(kwargs) == 'dBperkm':
        out *= DB_PER_INVCM
    return out


# ----------------------------------------------------------------------- nh3_hs
HS_GHZ_TO_INVCM = 1 / 29.9792458
HS_HC = 19.858252418E-24
HS_KB = 1.38E-23
HS_TO = 300.0
HS_COEF = 1.0E6 * 6.02297E23 / 8.31432E7
HS_F_SPLIT = 30.0
# Hanley/Bellotti inversion-line coefficient sets (constituents/nh3/nh3_hs.py:165-194):
#   (gnu_H2, gnu_He, gnu_NH3, GAMMA_H2, GAMMA_He, GAMMA_NH3,
#    zeta_H2, zeta_He, zeta_NH3, Z_H2, Z_He, Z_NH3, d, Con)
HS_INV_LO = (1.640, 0.75, 0.852, 0.7756, 0.666, 1.0,
             1.262, 0.3, 0.5296, 0.7964, 0.667, 1.554, -0.0498, 0.9301)
HS_INV_HI = (1.7465, 0.9779, 0.7298, 0.8202, 1.0, 1.0,
             1.2163, 0.0291, 0.5152, 0.8873, 0.8994, 2.0 / 3.0, -0.0627, 0.9862)
# nh3_dbs differs from nh3_hs only in its <= 30 GHz inversion set (the Bellotti/Steffes
# update, constituents/nh3/nh3_dbs.py:135-150); its > 30 GHz set and the rotational
# and v2 blocks are nh3_hs's (nh3_dbs.py:151-165, 205-297)
DBS_INV_LO = (1.6937, 0.6997, 0.7523, 0.8085, 1.0, 1.0,
              1.3263, 0.1607, 0.6162, 0.8199, 0.0, 1.3832, -0.0139, 0.9619)


def _gross(f, fo, gamma):
    """Gross line shape of nh3_hs.py:286-290 / 323-327 (lines x freqs)."""
    f2 = np.square(f)
    return (4.0 / np.pi) * f2 * gamma / (np.square(np.square(fo) - f2) + 4.0 * f2 * np.square(gamma))


def _nh3_hs_part(freq, T, P, X, P_dict, data, low_set=HS_INV_LO):
    """nh3_hs.__alpha__ (constituents/nh3/nh3_hs.py:117-340) for one side of 30 GHz;
    with low_set=DBS_INV_LO it is nh3_dbs.__alpha__ (constituents/nh3/nh3_dbs.py:92-313)."""
    freq = np.asarray(freq, dtype=np.float64)
    P_h2 = P * X[P_dict['H2']]
    P_he = P * X[P_dict['HE']]
    P_nh3 = P * X[P_dict['NH3']]
    Tdiv = HS_TO / T
    eta = 3.0 / 2.0
    if np.all(freq <= HS_F_SPLIT):
        cs = low_set
    elif np.all(freq > HS_F_SPLIT):
        cs = HS_INV_HI
    else:
        raise ValueError("Can't straddle 30 GHz in nh3_dbs.py.")
    (gnu_H2, gnu_He, gnu_NH3, G_H2, G_He, G_NH3,
     z_H2, z_He, z_NH3, Z_H2, Z_He, Z_NH3, d, Con) = cs
    g_o = data['gammaNH3o']
    gamma = (gnu_H2 * P_h2) * Tdiv ** G_H2 + (gnu_He * P_he) * Tdiv ** G_He \
        + (gnu_NH3 * P_nh3 * g_o) * (295.0 / T) ** G_NH3
    pst = d * gamma
    zeta = (z_H2 * P_h2) * Tdiv ** Z_H2 + (z_He * P_he) * Tdiv ** Z_He \
        + (z_NH3 * P_nh3 * g_o) * (295.0 / T) ** Z_NH3
    fo = data['fo'][:, None]
    f = freq[None, :]
    ST = data['Io'] * np.exp(-(1.0 / T - 1.0 / HS_TO) * data['Eo'] * HS_HC / HS_KB)
    noshape = Con * HS_COEF * (P_nh3 / HS_TO) * np.power(HS_TO / T, eta + 2.0) * ST
    g = gamma[:, None]
    z = zeta[:, None]
    p = pst[:, None]
    Aa = (2.0 / np.pi) * np.square(f / fo)
    Bb = (g - z) * np.square(f)
    Cc = g + z
    Dd = np.square(fo + p) + np.square(g) - np.square(z)
    Ff = Aa * (Bb + Cc * Dd) / (np.square(np.square(f) - np.square(fo + p) - np.square(g) + np.square(z))
                                + 4.0 * np.square(f) * np.square(g))
    a_inv = noshape[:, None] * ((1.0 / HS_GHZ_TO_INVCM) * Ff)
    # rotational lines, Gross shape (nh3_hs.py:257-291)
    ST_rot = data['Io_rot'] * np.exp((1.0 / HS_TO - 1.0 / T) * data['Eo_rot'] * HS_HC / HS_KB)
    g_rot = (0.2984 * P_h2 * data['gH2_rot']) * Tdiv ** 0.8730 \
        + (0.75 * P_he * data['gHe_rot']) * Tdiv ** (2.0 / 3.0) \
        + (3.1789 * P_nh3 * data['gNH3_rot']) * Tdiv ** 1.0
    F_rot = (1 / HS_GHZ_TO_INVCM) * _gross(f, data['fo_rot'][:, None], g_rot[:, None])
    a_rot = (2.4268 * HS_COEF * (P_nh3 / HS_TO) * (HS_TO / T) ** (eta + 2.0) * ST_rot)[:, None] * F_rot
    # v2 roto-vibrational lines, Gross shape (nh3_hs.py:293-328)
    ST_v2 = data['Io_v2'] * np.exp((1. / HS_TO - 1. / T) * data['Eo_v2'] * HS_HC / HS_KB)
    g_v2 = P_h2 * 1.4 * Tdiv ** 0.73 + P_he * 0.68 * Tdiv ** 0.5716 + P_nh3 * 9.5 * Tdiv ** 1.0
    F_v2 = (1.0 / HS_GHZ_TO_INVCM) * _gross(f, data['fo_v2'][:, None], g_v2)
    a_v2 = (1.1206 * HS_COEF * (P_nh3 / HS_TO) * (HS_TO / T) ** (eta + 2.0) * ST_v2)[:, None] * F_v2
    total = a_inv.sum(axis=0) + a_rot.sum(axis=0) + a_v2.sum(axis=0)   # nh3_hs.py:331
    return total


def nh3_hs(freq, T, P, X, P_dict, other_dict, data=None, low_set=HS_INV_LO, **kwargs):
    """nh3_hs.alpha: split at 30 GHz, low half first (nh3_hs.py:96-114).

    Units and the clamp alpha<0 -> 1e-8 are applied per half (nh3_hs.py:333-338).
    """
    freq = np.asarray(freq, dtype=np.float64)
    out = None
    for sel in (freq <= HS_F_SPLIT, freq > HS_F_SPLIT):
        frq = freq[sel]
        if len(frq):
            a = _nh3_hs_part(frq, T, P, X, P_dict, data, low_set)
            if _units(kwargs) == 'dBperkm':
                a = a * DB_PER_INVCM
            a[a < 0.0] = 1.0E-8
            out = a if out is None else np.concatenate((out, a))
    return out


def nh3_dbs(freq, T, P, X, P_dict, other_dict, data=None, **kwargs):
    """nh3_dbs.alpha (constituents/nh3/nh3_dbs.py:66-87): nh3_hs with the Bellotti/Steffes
    <= 30 GHz inversion set."""
    return nh3_hs(freq, T, P, X, P_dict, other_dict, data=data, low_set=DBS_INV_LO, **kwargs)


def nh3_hs_sjs(freq, T, P, X, P_dict, other_dict, data_hs=None, data_sjs=None, low_set=HS_INV_LO,
               **kwargs):
    """Pressure blend of nh3_hs and nh3_sjs (constituents/nh3/nh3_hs_sjs.py:6-26); with
    low_set=DBS_INV_LO the blend of nh3_dbs and nh3_sjs (nh3/nh3_dbs_sjs.py:6-26, the
    same 400 / 2000 bar linear weight)."""
    PLower, PHigher = 400.0, 2000.0
    if P < PLower:
        return nh3_hs(freq, T, P, X, P_dict, other_dict, data=data_hs, low_set=low_set, **kwargs)
    if P > PHigher:
        return nh3_sjs(freq, T, P, X, P_dict, other_dict, data=data_sjs, **kwargs)
    a2 = nh3_sjs(freq, T, P, X, P_dict, other_dict, data=data_sjs, **kwargs)
    a1 = nh3_hs(freq, T, P, X, P_dict, other_dict, data=data_hs, low_set=low_set, **kwargs)
    W = (P - PLower) / (PHigher - PLower)
    return W * a2 + (1.0 - W) * a1


def nh3_dbs_sjs(freq, T, P, X, P_dict, other_dict, data_hs=None, data_sjs=None, **kwargs):
    """nh3_dbs_sjs.alpha (constituents/nh3/nh3_dbs_sjs.py:6-26)."""
    return nh3_hs_sjs(freq, T, P, X, P_dict, other_dict, data_hs=data_hs, data_sjs=data_sjs,
                      low_set=DBS_INV_LO, **kwargs)
