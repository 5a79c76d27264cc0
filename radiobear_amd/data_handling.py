"""Run results returned by Planet.run (radiobear/data_handling.py:11-85)."""
import copy

import numpy as np

from . import utils


class Data:
    """Holds f, b, Tb, header, ...; lists are stored as float32 arrays
    (data_handling.py:46-47), so Tb here is float32 as in the reference."""

    allowed_parameters = ['f', 'freqUnit', 'b', 'Tb', 'header', 'start', 'stop',
                          'log', 'type', 'logfile']

    def __init__(self):
        for a in self.allowed_parameters:
            setattr(self, a, None)

    def __repr__(self):
        s = ''
        for i, b in enumerate(self.b):
            s += 'b = {}:\n f = '.format(b)
            s += ' '.join('{:6.1f}'.format(x) for x in self.f) + '  GHz\nTb = '
            s += ' '.join('{:6.1f}'.format(x) for x in self.Tb[i]) + '  K\n'
        return s

    def set(self, par, val):
        if par not in self.allowed_parameters:
            print("{} not in valid data return list.".format(par))
            return
        val = copy.copy(val)
        if par == 'b' and utils.b_type(val).startswith('dis'):
            self.b = ['disc']
        elif isinstance(val, list):
            setattr(self, par, np.asarray(val, dtype=np.float32))
        else:
            setattr(self, par, val)
