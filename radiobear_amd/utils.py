"""Constants and small helpers of the reference call surface (radiobear/utils.py)."""
import numpy as np

# radiobear/utils.py:13-22 unit table (factors to SI) and processing units
Units = {'Frequency': 1, 'Hz': 1.0, 'kHz': 1.0E3, 'MHz': 1.0E6, 'GHz': 1.0E9,
         'Length': 1, 'm': 1.0, 'km': 1.0E3, 'cm': 1.0E-2, 'AU': 149597870691.0,
         'Pressure': 1, 'bars': 1.0, 'atm': 1.01325,
         'Time': 1, 'sec': 1.0, 'min': 60.0, 'hr': 3600.0, 'day': 86400.0, 'year': 31536000.0,
         'Acceleration': 1, 'mpersec2': 1.0, 'cmpersec2': 0.01}
processingUnits = {'GHz': ['GHz', 'Hz', 'kHz', 'MHz'],
                   'km': ['m', 'cm', 'AU', 'km'],
                   'bars': ['bars', 'atm'],
                   'sec': ['sec', 'min', 'hr', 'day', 'year'],
                   'mpersec2': ['mpersec2', 'cmpersec2']}
alphaUnit = 'invcm'
atmLayerUnit = 'km'
T_cmb = 2.725
commentChars = ['!', '#', '$', '%', '&', '*']


def isanynum(x):
    """True for anything float() accepts except bool, dict and list (utils.py:33-41)."""
    if isinstance(x, (bool, dict, list)):
        return False
    try:
        float(x)
    except (TypeError, ValueError):
        return False
    return True


def proc_unit(supplied_unit):
    for proc, units in processingUnits.items():
        if supplied_unit in units:
            return proc
    return None


def convert_unit(v, supplied_unit):
    """v in supplied_unit -> its processing unit (utils.py:59-68)."""
    if supplied_unit is None:
        return v
    for proc, units in processingUnits.items():
        if supplied_unit in units:
            return v * Units[supplied_unit] / Units[proc]
    return v


def r2d(a):
    return a * 180.0 / np.pi


def d2r(a):
    return a * np.pi / 180.0


def r2asec(a):
    return 3600.0 * r2d(a)


def b_type(b):
    """'disc', 'image', 'profile' or 'spectrum' from a b request (utils.py:139-149)."""
    if isinstance(b, str):
        return b.lower()
    if isinstance(b[0], str):
        return b[0].lower()
    if len(b) > 20:
        return 'image'
    if len(b) > 9:
        return 'profile'
    return 'spectrum'


def get_data_from(line):
    """Floats of a data line, None for comments/short/non-numeric lines (utils.py:152-160)."""
    if line[0] in commentChars or len(line) < 4:
        return None
    try:
        return [float(x) for x in line.split()]
    except ValueError:
        return None
