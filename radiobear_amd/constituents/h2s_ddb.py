"""h2s_ddb plug-in: constituents/h2s/h2s_ddb.py:42-87 (Ben-Reuven with zeta = gamma).

alpha(freq, T, P, X, P_dict, other_dict, **kwargs) -> len(freq) values
(dB/km, or cm^-1 with units='invcm'), computed by the rbx_alpha kernel.
"""
import numpy as np

from . import _Module

_module = _Module('h2s', 'h2s_ddb', ('H2S', 'H2', 'HE'))
data = None   # assign a line list before the first call to replace the shipped one


def alpha(freq, T, P, X, P_dict, other_dict, **kwargs):
    if data is not None and _module.data is not data:
        _module.data = data
        _module.reset()
    a = _module(freq, T, P, X, P_dict, other_dict, **kwargs)
    if np.isscalar(a):
        return a
    return np.asarray(a)
