"""nh3_dbs_sjs plug-in: constituents/nh3/nh3_dbs_sjs.py:6-26 (pressure blend of nh3_dbs and nh3_sjs).

alpha(freq, T, P, X, P_dict, other_dict, **kwargs) -> len(freq) values
(dB/km, or cm^-1 with units='invcm'), computed by the rbx_alpha kernel.
"""
import numpy as np

from . import _Module

_module = _Module('nh3', 'nh3_dbs_sjs', ('NH3', 'H2', 'HE'))
data = None   # assign a line list before the first call to replace the shipped one


def alpha(freq, T, P, X, P_dict, other_dict, **kwargs):
    if data is not None and _module.data is not data:
        _module.data = data
        _module.reset()
    a = _module(freq, T, P, X, P_dict, other_dict, **kwargs)
    if np.isscalar(a):
        return a
    return np.asarray(a)
