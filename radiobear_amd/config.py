"""Run configuration: per-planet defaults, token files and keyword updates.

Keeps the reference's attribute names and token-file format (radiobear/config.py,
default_config.json, default_state.json) for the settings the absorption and
ray-integration path reads: absorber selection and truncation, h2state, the
gas-table layout C, radii, orientation, shape and limb type, and the regrid /
tweak settings of the file-atmosphere pipeline.
"""
import copy
import os

from . import utils

_C = {'Z': 0, 'T': 1, 'P': 2, 'H2': 3, 'HE': 4, 'CH4': 5, 'NH3': 6, 'H2O': 7, 'H2S': 8, 'SOLN': 9,
      'OTHER': 10, 'PH3': 11, 'CO': 12, 'CO13': 13, 'HCN': 14, 'DZ': 15}
_CL = {'Z': 0, 'T': 1, 'P': 2, 'SOLN': 3, 'H2O': 4, 'NH4SH': 5, 'NH3': 6, 'H2S': 7, 'CH4': 8, 'AR': 9,
       'PH3': 10, 'DZ': 11}
_LP = {'Z': 0, 'R': 1, 'P': 2, 'GM': 3, 'AMU': 4, 'REFR': 5, 'N': 6, 'H': 7, 'LAPSE': 8, 'LAPSEP': 9, 'g': 10}

# per-planet values (default_config.json / default_state.json of the reference)
_PLANETS = {
    'Jupiter': dict(Req=71492.0, Rpol=66854.0, GM_ref=126686538.0, distance=5.2, omega_m=0.00017585,
                    orientation=[0.0, 0.0], nh3='nh3_hs_sjs', co=None, gasFile=['jupiter.junogas'],
                    cloudFile=['jupiter.junoclouds'], tweakmodule='JupiterTweak'),
    'Saturn': dict(Req=60268.0, Rpol=54364.0, GM_ref=37931272.0, distance=9.6, omega_m=0.00016378499,
                   orientation=[0.0, 0.0], nh3='nh3_dbs_sjs', co=None, gasFile=['saturn.Solar'],
                   cloudFile=['saturn.clSolar'], tweakmodule='SaturnTweak'),
    'Uranus': dict(Req=25559.0, Rpol=25559.0, GM_ref=5794000.0, distance=19.2, omega_m=10124.0,
                   orientation=[0.0, 0.0], nh3='nh3_dbs_sjs', co=None, gasFile=['uranus.paul10xsolar_wet.dat'],
                   cloudFile=['uranus.cloud10xsolar_wet.dat'], tweakmodule='UranusTweak'),
    'Neptune': dict(Req=24766.0, Rpol=24342.0, GM_ref=6835096.0, distance=29.7, omega_m=0.000101237195,
                    orientation=[348.8274, -28.97], nh3='nh3_dbs_sjs', co='co_ddb',
                    gasFile=['neptune.paulCO_cloud21_fletcher_best_dry'],
                    cloudFile=['nepcloud_CO.cloud21_fletcher_best_dry'], tweakmodule='NeptuneTweak'),
}

# token (as written in .par files) -> attribute, unit used for float values
TOKENS = {
    'constituents': ('C', None), 'clouds': ('Cl', None), 'other': ('LP', None),
    'alpha': ('constituent_alpha', None), 'truncate_freq': ('truncate_freq', 'GHz'),
    'truncate_strength': ('truncate_strength', None), 'truncate_method': ('truncate_method', None),
    'gasfile': ('gasFile', None), 'cloudfile': ('cloudFile', None), 'gastype': ('gasType', None),
    'cloudtype': ('cloudType', None), 'othertype': ('otherType', None),
    'req': ('Req', 'km'), 'rpol': ('Rpol', 'km'), 'p_ref': ('p_ref', 'bars'), 'gm': ('GM_ref', None),
    'pmin': ('pmin', 'bars'), 'pmax': ('pmax', 'bars'), 'distance': ('distance', 'AU'),
    'orientation': ('orientation', None), 'omega': ('omega_m', None), 'gtype': ('gtype', None),
    'limb': ('limb', None), 'h2state': ('h2state', None), 'doppler': ('Doppler', None),
    'regridtype': ('regridType', None), 'tweakmodule': ('tweakmodule', None),
    'bmapmodule': ('bmapmodule', None), 'write_output_files': ('write_output_files', None),
    'write_log_file': ('write_log_file', None), 'plot_atm': ('plot_atm', None),
    'plot_bright': ('plot_bright', None), 'scratch_directory': ('scratch_directory', None),
    'output_directory': ('output_directory', None), 'log_directory': ('log_directory', None),
    'normalize_weighting': ('normalize_weighting', None), 'atm_run_type': ('atm_run_type', None),
    'output_type': ('output_type', None), 'h2newset': ('h2newset', None), 'zonal': ('zonal', None),
    'water': ('water_p', 'um'), 'ice': ('ice_p', 'um'), 'h2sice': ('h2sice_p', 'um'), 'nh3ice': ('nh3ice_p', 'um'),
    'nh4sh': ('nh4sh_p', 'um'), 'ch4': ('ch4_p', 'um'), 'coshape': ('coshape', None),
}


def set_single_val(val, unit=None, special={'true': True, 'false': False, 'none': None, 'null': None}):
    """Type inference of one config value (config.py:11-32)."""
    if isinstance(val, str):
        if val.lower() in special:
            val = special[val.lower()]
        elif ',' in val:
            val = val.split(',')
        elif '.' in val or 'e' in val or 'E' in val:
            try:
                val = float(val)
            except ValueError:
                pass
        else:
            try:
                val = int(val)
            except ValueError:
                pass
    if isinstance(val, float):
        val = utils.convert_unit(val, unit)
    return val


class planetConfig:
    """Configuration of one planet: defaults, then a token file, then update_config."""

    def __init__(self, planet, configFile=None):
        planet = planet.capitalize()
        if planet not in _PLANETS:
            raise ValueError('{} not found (Jupiter, Saturn, Uranus, Neptune)'.format(planet))
        self.planet = planet
        self.filename = configFile
        self.path = planet
        p = _PLANETS[planet]
        self.C = dict(_C)
        self.Cl = dict(_CL)
        self.LP = dict(_LP)
        self.constituent_alpha = {'clouds': None, 'co': p['co'], 'h2': 'h2_jj_ddb', 'h2o': 'h2o_bk',
                                  'h2s': 'h2s_ddb', 'hcn': None, 'nh3': p['nh3'], 'ph3': 'ph3_jh'}
        self.truncate_freq = {k: None for k in ('co', 'h2o', 'h2s', 'hcn', 'nh3', 'ph3')}
        self.truncate_strength = {'co': None, 'h2o': None, 'h2s': 1e-22, 'hcn': None, 'nh3': None,
                                  'ph3': 1e-22}
        self.truncate_method = {'co': None, 'h2o': None, 'h2s': 'truncate_strength', 'hcn': None,
                                'nh3': None, 'ph3': 'truncate_strength'}
        self.gasFile = list(p['gasFile'])
        self.cloudFile = list(p['cloudFile'])
        self.gasType = 'read'
        self.cloudType = 'read'
        self.otherType = 'compute'
        self.Req = p['Req']
        self.Rpol = p['Rpol']
        self.p_ref = 1.0
        self.GM_ref = p['GM_ref']
        self.pmin = 0
        self.pmax = -1
        self.distance = p['distance'] * utils.Units['AU'] / utils.Units['km']
        self.orientation = list(p['orientation'])
        self.omega_m = p['omega_m']
        self.gtype = 'ellipse'
        self.limb = 'shape'
        self.h2state = 'e'
        self.h2newset = True
        # cloud particle sizes (um; clouds_idp only tests them > 0) and the CO line shape
        # (default_config.json 'water' .. 'ch4', 'coshape')
        for attr in ('water_p', 'ice_p', 'h2sice_p', 'nh3ice_p', 'nh4sh_p', 'ch4_p'):
            setattr(self, attr, utils.convert_unit(1.0E-4, 'um'))
        self.coshape = 'voigt'
        self.Doppler = False
        self.regridType = 'grid.dat'
        self.tweakmodule = p['tweakmodule']
        self.bmapmodule = 'nobmap'
        self.write_output_files = False
        self.write_log_file = False
        self.plot_atm = True
        self.plot_bright = True
        self.normalize_weighting = True
        self.output_type = 'frequency'
        self.atm_run_type = 'std'
        self.scratch_directory = 'Scratch'
        self.output_directory = 'Output'
        self.log_directory = 'Logs'
        self.zonal = None
        self.vwlat = [0.0, 90.0]
        self.vwdat = [0.0, 0.0]
        self.toks = {t: a for t, (a, _) in TOKENS.items()}
        self.setConfig(configFile)

    def setConfig(self, configFile):
        """Read a token file: 'token value [unit]  # comment' (config.py:66-139)."""
        if configFile is None:
            return 0
        if not os.path.exists(configFile):
            print(configFile, ' not found.  Using defaults.')
            return 0
        self.path = os.path.dirname(os.path.abspath(configFile))
        with open(configFile) as fp:
            for line in fp:
                if line[0] in utils.commentChars or len(line) < 4:
                    continue
                if '#' in line:
                    line = line[:line.index('#')]
                data = line.split()
                if not data:
                    continue
                tok = data[0].lower()
                del data[0]
                if tok not in TOKENS:
                    print('token {} not used by this build; ignored'.format(tok))
                    continue
                attr, unit_default = TOKENS[tok]
                preval = getattr(self, attr)
                if isinstance(preval, dict):
                    val = copy.copy(preval)
                    for i, v in enumerate(data):
                        if ':' in v:
                            k, m = v.split(':')
                            val[k.lower().strip()] = set_single_val(m.lower().strip())
                        else:
                            val[v.strip()] = i
                elif isinstance(preval, list):
                    if len(data) == 1 and ',' in data[0]:
                        data = data[0].split(',')
                    val = [set_single_val(x) for x in data]
                else:
                    if not data:
                        print("{} didn't have an associated argument in config file.".format(tok))
                        continue
                    unit = data[1] if len(data) == 2 else 'none'
                    val = set_single_val(data[0], unit)
                setattr(self, attr, val)
        if 'DZ' not in self.C:
            self.C['DZ'] = len(self.C)
        if 'DZ' not in self.Cl:
            self.Cl['DZ'] = len(self.Cl)
        return 1

    def update_config(self, key=None, value=None, **kwargs):
        """Set attributes by token or attribute name (config.py:141-189)."""
        pairs = []
        if isinstance(key, dict):
            pairs.extend(key.items())
        elif isinstance(key, list):
            if len(key) != len(value):
                print("key/value pairs not matched.")
            else:
                pairs.extend(zip(key, value))
        elif key is not None:
            pairs.append((key, value))
        pairs.extend(kwargs.items())
        for k, v in pairs:
            attr = self.toks.get(k, k)
            setattr(self, attr, set_single_val(v))

    def show(self, print_it=True):
        s = 'Run parameters:\n'
        for tok in sorted(self.toks):
            s += '\t{:20s}:  {}\n'.format(tok, str(getattr(self, self.toks[tok], None)))
        if print_it:
            print(s)
        return s
