"""Line catalogues shipped with the package (radiobear_amd/data).

The .npz line lists are the reference's, unchanged in content and keys
(constituents/nh3/ammonia.npz, nh3/nh3.npz, h2s/h2s.npz, ph3/ph3jh.npz,
ph3/PH3WGT.npz, co/co.npz), so that a user's own catalogue in the reference
schema drops in; orton_H2.npz holds h2_orton's CIA tables (see orton_to_npz.py).  The 15 H2O lines of h2o_bk live in ``h2o_bk_lines`` (the reference types
them in constituents/h2o/h2o_bk.py:23-49).
"""
import os

import numpy as np

DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'data')
CATALOGUES = ('ammonia', 'nh3', 'h2s', 'ph3jh', 'PH3WGT')


def load(name):
    """One catalogue as a dict of float64 arrays (pickle-free load)."""
    with np.load(os.path.join(DATA_DIR, name + '.npz'), allow_pickle=False) as d:
        return {k: np.asarray(d[k], dtype=np.float64) for k in d.files}


def raw_catalog():
    return {name: load(name) for name in CATALOGUES}


def load_orton():
    """Orton's H2 collision-induced absorption tables (h2_orton, constituents/h2/h2_orton.py:17-97),
    as orton_to_npz.py converts constituents/h2/orton_H2.tables: header [nTemp, Tmax, Tmin],
    ftab_invcm [nfreq] and logv [6 nfreq][nTemp] natural-log values."""
    return load('orton_H2')


ORTON_GHZ = 29.9792458   # cm^-1 -> GHz, h2_orton.py:11


def orton_columns(table):
    """The catalogue columns of the Orton table block: frequencies in GHz and the
    temperature grid computed as h2_orton.readInputFiles does (h2_orton.py:27-40,
    log-spaced from Tmin by repeated addition), and the log values row by row."""
    import math
    ntemp, tmax, tmin = int(table['header'][0]), float(table['header'][1]), float(table['header'][2])
    lTmx, lTmn = math.log(tmax), math.log(tmin)
    dlT = (lTmx - lTmn) / (ntemp - 1.0)
    ta = [lTmn]
    for i in range(ntemp - 1):
        ta.append(ta[i] + dlT)
    return {'ftab_ghz': np.asarray(table['ftab_invcm'], dtype=np.float64) * ORTON_GHZ,
            'logv': np.ascontiguousarray(table['logv'], dtype=np.float64).ravel(),
            'Ttab': np.array([math.exp(t) for t in ta])}


def truncate(data, keys, strength=None, freq=None):
    """Load-time line cut of h2s_ddb (constituents/h2s/h2s_ddb.py:23-38):
    keep I0 > strength, then f0 < freq, on the columns ``keys``."""
    out = {k: np.asarray(data[k], dtype=np.float64) for k in keys}
    if strength is not None:
        keep = out['I0'] > strength
        out = {k: v[keep] for k, v in out.items()}
    if freq is not None:
        keep = out['f0'] < freq
        out = {k: v[keep] for k, v in out.items()}
    return out


def truncate_ph3(lines, weights, strength=None, freq=None):
    """Load-time cut of ph3_jh (constituents/ph3/ph3_jh.py:28-60): lines and
    their weights are cut together; a cut that leaves them unequal raises."""
    d = {k: np.asarray(lines[k], dtype=np.float64) for k in ('f0', 'I0', 'E')}
    w = {k: np.asarray(weights[k], dtype=np.float64) for k in ('WgtI0', 'WgtFGB', 'WgtSB')}
    if len(w['WgtI0']) != len(d['f0']):
        raise ValueError("PH3 wgt vector wrong length.")
    keep = np.ones(len(d['f0']), dtype=bool)
    if strength is not None:
        keep &= d['I0'] > strength
    if freq is not None:
        keep &= d['f0'] < freq
    out = {k: v[keep] for k, v in d.items()}
    out.update({k: v[keep] for k, v in w.items()})
    return out


def truncate_h2o(data, strength=None, freq=None):
    """Load-time cut of h2o_bk (constituents/h2o/h2o_bk.py:51-62); a zero
    strength or frequency means no cut, as there."""
    d = {k: np.asarray(v, dtype=np.float64) for k, v in data.items()}
    if strength:
        keep = d['I_o'] > strength
        d = {k: v[keep] for k, v in d.items()}
    if freq:
        keep = d['f_o'] < freq
        d = {k: v[keep] for k, v in d.items()}
    return d


def h2o_bk_lines():
    """Karpowicz/Rosenkranz H2O lines in the h2o_bk schema, widths in GHz/bar."""
    f_o = [22.2351, 183.3101, 321.2256, 325.1529, 380.1974, 439.1508, 443.0183,
           448.0011, 470.8890, 474.6891, 488.4911, 556.9360, 620.7008, 752.0332, 916.1712]
    I_o = [0.1314E-13, 0.2279E-11, 0.8058E-13, 0.2701E-11, 0.2444E-10, 0.2185E-11,
           0.4637E-12, 0.2568E-10, 0.8392E-12, 0.3272E-11, 0.6676E-12, 0.1535E-08,
           0.1711E-10, 0.1014E-08, 0.4238E-10]
    E_o = [2.144, 0.668, 6.179, 1.541, 1.048, 3.595, 5.048, 1.405, 3.597, 2.379,
           2.852, 0.159, 2.391, 0.396, 1.441]
    w_s_mbar = [0.01349, 0.01466, 0.01057, 0.01381, 0.01454, 0.009715, 0.00788,
                0.01275, 0.00983, 0.01095, 0.01313, 0.01405, 0.011836, 0.01253, 0.01275]
    x_s = [0.61, 0.85, 0.54, 0.74, 0.89, 0.62, 0.50, 0.67, 0.65, 0.64, 0.72, 1.0,
           0.68, 0.84, 0.78]
    odd = {1: (2.4000, 0.71, 0.950, 0.490), 3: (2.395, 0.67, 0.900, 0.490),
           4: (2.390, 0.63, 0.850, 0.540)}
    w_h2, w_he, x_h2, x_he = [], [], [], []
    for i in range(15):
        a, b, c, d = odd.get(i, (2.395, 0.67, 0.900, 0.515))
        w_h2.append(a)
        w_he.append(b)
        x_h2.append(c)
        x_he.append(d)
    out = dict(f_o=f_o, I_o=I_o, E_o=E_o, w_s=np.asarray(w_s_mbar) / 0.001, x_s=x_s,
               w_h2=w_h2, w_he=w_he, x_h2=x_h2, x_he=x_he, SR=np.zeros(15))
    return {k: np.asarray(v, dtype=np.float64) for k, v in out.items()}
