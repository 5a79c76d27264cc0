"""Atmosphere on the host: the gas table and the derived layer properties.

The absorption and ray kernels read P, T and the mixing ratios of ``gas`` and
the layer radii R and refractive indices N of ``property``, laid out as in the
reference (``gas[C[name]]``, ``property[LP[name]]``, layer 0 at the top).
Two sources:

* ``Atmosphere.std()``: the reference's file pipeline (atm_base.readGas,
  regrid.regrid onto a log-P grid with linear interpolation in P, the tweak
  module, atm_base.computeProp; radiobear/atm_base.py:39-219,
  radiobear/regrid.py:11-164, radiobear/atmosphere.py:66-116);
* ``Atmosphere.from_synthetic()``: a ``synth.SynthAtmosphere``.
"""
import importlib
import os
import sys

import numpy as np

from . import utils

# chemistry.py:7-46: constants, amu (g/mol), refractivity (R = Rc P 293/T, N = R/1e6 + 1)
GRAV_CONST = 6.6738e-20
R_GAS = 8.314462
AMU = {'H2': 2.016, 'HE': 4.003, 'H2S': 34.076, 'NH3': 17.030, 'H2O': 18.015, 'CH4': 16.04,
       'PH3': 33.997, 'NH4SH': 51.110}
REFR = {'H2': 124.43, 'HE': 35.832, 'H2S': 2247.0, 'NH3': 2700.0, 'CH4': 413.0}


def refractivity(key, T):
    if key == 'H2O':
        return 245.0 + 1.28E6 / T
    return REFR.get(key, 0.0)


def jupiter_no_tweak(gas, cloud, C, Cl):
    """Jupiter/JupiterNoTweak.py: CO, HCN, SOLN, PH3 zero; CO13 = 1e-2 CO."""
    for k in ('CO', 'HCN', 'SOLN', 'PH3'):
        gas[C[k]][:] = 0.0
    gas[C['CO13']][:] = 1.0E-2 * gas[C['CO']]
    return "Jupiter no tweaking (except to add CO, CO13, HCN, SOLN, PH3 at zero)", gas, cloud


BUILTIN_TWEAKS = {'JupiterNoTweak': jupiter_no_tweak}


def _interp_linear(x, y, xnew, fill):
    """scipy interp1d(kind='linear', bounds_error=False, fill_value=fill) on ascending x:
    the bracketing segment's line, `fill` outside [x[0], x[-1]]."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    xnew = np.asarray(xnew, dtype=np.float64)
    hi = np.clip(np.searchsorted(x, xnew, side='left'), 1, len(x) - 1)
    lo = hi - 1
    slope = (y[hi] - y[lo]) / (x[hi] - x[lo])
    out = slope * (xnew - x[lo]) + y[lo]
    out[(xnew < x[0]) | (xnew > x[-1])] = fill
    return out


class Atmosphere:
    """gas (nC, nl), cloud, property (nLP, nl) and config, as the reference's Atmosphere."""

    def __init__(self, planet, idnum=0, config=None, log=None, verbose=False, **kwargs):
        from . import config as pcfg
        if config is None or isinstance(config, str):
            config = pcfg.planetConfig(planet, configFile=config)
            config.update_config(**kwargs)
        self.planet = planet.capitalize()
        self.config = config
        self.idnum = idnum
        self.verbose = verbose
        self.gas = None
        self.cloud = None
        self.property = None
        self.nAtm = 0
        self.tweakComment = None

    # ---------------------------------------------------------------- sources
    @classmethod
    def from_synthetic(cls, synth_atm, planet='jupiter', config=None):
        from . import config as pcfg
        if config is None:
            config = pcfg.planetConfig(planet)
        config.Req = float(synth_atm.Req)
        config.Rpol = float(synth_atm.Rpol)
        config.C = dict(synth_atm.C)
        atm = cls(planet, config=config)
        atm.gas = np.array(synth_atm.gas, dtype=np.float64)
        atm.cloud = np.zeros((len(config.Cl), synth_atm.nl))
        prop = np.zeros((len(config.LP), synth_atm.nl))
        prop[config.LP['R']] = synth_atm.R
        prop[config.LP['N']] = synth_atm.N
        prop[config.LP['P']] = synth_atm.P
        prop[config.LP['Z']] = synth_atm.gas[config.C['Z']]
        atm.property = prop
        atm.nAtm = synth_atm.nl
        return atm

    def std(self):
        """The file pipeline: read gas and cloud, regrid, tweak, derived properties
        (atmosphere.py:66-120).  A cloud file that is not there gives a cloud-free table
        (the reference stops with an IOError); cloud densities only matter to clouds_idp."""
        self.read_gas()
        cpath = os.path.join(self.config.path, self.config.cloudFile[self.idnum]) if self.config.cloudFile else None
        if cpath is not None and os.path.exists(cpath):
            self.read_cloud(cpath)
        else:
            if self.verbose:
                print('no cloud file {}: cloud densities are zero'.format(cpath))
            Cl = self.config.Cl
            self.cloud = np.zeros((len(Cl), self.gas.shape[1]))
            self.cloud[Cl['P']] = self.gas[self.config.C['P']]
        self.regrid()
        self.tweak()
        self.compute_prop()
        self.nAtm = self.gas.shape[1]
        return self.nAtm

    def read_gas(self, path=None):
        """atm_base.readGas (atm_base.py:39-78)."""
        cfg = self.config
        if path is None:
            path = os.path.join(cfg.path, cfg.gasFile[self.idnum])
        self.gas, cfg.C = self._read_table(path, cfg.C)
        self._renorm_z()

    def read_cloud(self, path=None):
        """atm_base.readCloud (atm_base.py:80-118): the cloud table laid out by config.Cl."""
        cfg = self.config
        if path is None:
            path = os.path.join(cfg.path, cfg.cloudFile[self.idnum])
        self.cloud, cfg.Cl = self._read_table(path, cfg.Cl)
        self._renorm_z('cloud')

    @staticmethod
    def _read_table(path, layout):
        """Rows of the most common entry count, missing columns zero, P ascending
        (flipped if the file runs deep to high); returns the table and its layout
        renumbered in column order (atm_base.py:39-118)."""
        with open(path) as fp:
            lines = fp.readlines()
        counts = {}
        for ln in lines:
            cv = utils.get_data_from(ln)
            if cv is not None:
                counts[len(cv)] = counts.get(len(cv), 0) + 1
        inv = {v: k for k, v in counts.items()}
        keys = sorted(inv, reverse=True)
        if len(keys) > 1 and 2 * sum(keys[1:]) > keys[0]:
            raise ValueError("Not enough data lines in file: {}".format(keys))
        expected = inv[keys[0]]
        nC = len(layout)
        rows = []
        for ln in lines:
            cv = utils.get_data_from(ln)
            if cv is None or len(cv) != expected:
                continue
            rows.append([cv[n] if n < len(cv) else 0.0 for n in range(nC)])
        tab = np.array(rows, dtype=np.float64).T
        order = sorted(layout, key=lambda k: layout[k])
        layout = {k: i for i, k in enumerate(order)}
        P = tab[layout['P']]
        if not np.all(np.diff(P) > 0.0):
            tab = np.fliplr(tab)
            if not np.all(np.diff(tab[layout['P']]) > 0.0):
                raise ValueError("Pressure not monotonically increasing in {}.".format(path))
        return np.ascontiguousarray(tab), layout

    def _renorm_z(self, which='gas'):
        """atm_base._renorm_z: deepest z = 0, DZ in cm (atm_base.py:206-219)."""
        ind = self.config.C if which == 'gas' else self.config.Cl
        tab = self.gas if which == 'gas' else self.cloud
        tab[ind['Z']] = tab[ind['Z']] - tab[ind['Z']][-1]
        tab[ind['DZ']] = np.append(0.0, np.abs(np.diff(tab[ind['Z']])) * 1.0E5)

    def regrid(self):
        """regrid.regrid with an integer regridType: logspace(pmin, pmax, n), linear in P.

        Points outside the file's pressure range are not extrapolated here
        (the reference's dry-adiabat extrapolation, regrid.py:121-164): such a
        request raises ValueError."""
        cfg = self.config
        rt = cfg.regridType
        if rt is None or (isinstance(rt, str) and rt.lower() == 'none'):
            return 0
        P = self.gas[cfg.C['P']]
        pmin = cfg.pmin if cfg.pmin not in (None, 'auto', 0) else P.min()
        pmax = cfg.pmax if cfg.pmax not in (None, 0, -1) else P.max()
        if isinstance(rt, str):
            try:
                rt = int(rt)
            except ValueError:
                Pgrid = np.loadtxt(os.path.join(cfg.path, rt))
                rt = None
        if rt is not None:
            Pgrid = np.logspace(np.log10(pmin), np.log10(pmax), int(rt))
        if Pgrid.min() < P.min() or Pgrid.max() > P.max():
            raise ValueError('regrid outside the file pressure range [{}, {}] is not supported'
                             .format(P.min(), P.max()))
        new = np.zeros((self.gas.shape[0], len(Pgrid)))
        new[cfg.C['P']] = Pgrid
        idx = np.clip(np.searchsorted(P, Pgrid), 1, len(P) - 1)
        lo, hi = idx - 1, idx
        for key, row in cfg.C.items():
            if key in ('P', 'DZ'):
                continue
            y = self.gas[row]
            slope = (y[hi] - y[lo]) / (P[hi] - P[lo])
            new[row] = slope * (Pgrid - P[lo]) + y[lo]
        self.gas = new
        # the cloud table: linear in P, zero outside the cloud file's range (regrid.py:74-77)
        Cl = cfg.Cl
        Pc = self.cloud[Cl['P']]
        cloud = np.zeros((self.cloud.shape[0], len(Pgrid)))
        cloud[Cl['P']] = Pgrid
        for key, row in Cl.items():
            if key in ('P', 'DZ'):
                continue
            cloud[row] = _interp_linear(Pc, self.cloud[row], Pgrid, 0.0)
        self.cloud = cloud
        self._renorm_z()
        self._renorm_z('cloud')
        return 1

    def tweak(self):
        """Run the config's tweak module modify(gas, cloud, C, Cl) (atm_modify.py:12-32)."""
        name = self.config.tweakmodule
        if name is None:
            return
        fn = BUILTIN_TWEAKS.get(name)
        if fn is None:
            if self.config.path not in sys.path:
                sys.path.insert(0, self.config.path)
            fn = importlib.import_module(name).modify
        self.tweakComment, self.gas, self.cloud = fn(self.gas, self.cloud, self.config.C, self.config.Cl)

    def compute_prop(self):
        """atm_base.computeProp (atm_base.py:120-194): R, N, GM, AMU, REFR, H, g, lapse rates."""
        cfg = self.config
        C, LP = cfg.C, cfg.LP
        P, T, Z = self.gas[C['P']], self.gas[C['T']], self.gas[C['Z']]
        nl = len(P)
        iref = 0
        psep = 1.0E6
        for i in range(nl):
            if abs(P[i] - cfg.p_ref) < psep:
                psep = abs(P[i] - cfg.p_ref)
                iref = i
        prop = np.zeros((len(LP), nl))
        R = cfg.Req + Z - Z[iref]
        keys = [k for k in C if k not in ('P', 'T', 'Z', 'DZ')]
        amu = np.zeros(nl)
        refr = np.zeros(nl)
        for k in keys:
            amu += AMU.get(k, 0.0) * self.gas[C[k]]
            refr += refractivity(k, T) * self.gas[C[k]]
        refr = refr * P * (293.0 / T)
        gm = np.zeros(nl)
        dr = np.abs(np.diff(Z))
        rho = amu[1:] * P[1:] / (R_GAS * T[1:])
        gm[1:] = np.cumsum(GRAV_CONST * 1.0e11 * rho * 4.0 * np.pi * R[1:] ** 2 * dr)
        gm = cfg.GM_ref - (gm - gm[iref])
        g = gm / R ** 2
        with np.errstate(divide='ignore', invalid='ignore'):
            prop[LP['LAPSE']][1:] = np.abs(np.diff(T)) / dr
            prop[LP['LAPSEP']][1:] = np.abs(np.diff(T)) / np.abs(np.diff(P))
            prop[LP['H']] = R_GAS * T / (g * amu) / 1000.0
        prop[LP['P']] = P
        prop[LP['Z']] = Z
        prop[LP['R']] = R
        prop[LP['AMU']] = amu
        prop[LP['GM']] = gm
        prop[LP['REFR']] = refr
        prop[LP['N']] = refr / 1.0E6 + 1.0
        prop[LP['g']] = g
        self.property = prop

    # ------------------------------------------------------------ accessors
    def layer_arrays(self):
        """(P, T, mix[7][nl], R, N) for the kernels."""
        from .kernels import mix_table
        C, LP = self.config.C, self.config.LP
        return (self.gas[C['P']], self.gas[C['T']], mix_table(self.gas, C),
                self.property[LP['R']], self.property[LP['N']])
