"""Frequency and impact-parameter requests (radiobear/set_utils.py)."""
from argparse import Namespace

import numpy as np

from . import utils


def proc_string(srq):
    """'a,b,c' list; 'start:stop:step' arange with stop included; 'start;stop;n'
    logspace; a number; or a file name to loadtxt (set_utils.py:130-146).

    The reference passes n of 'start;stop;n' to numpy as a float, which numpy
    rejects; here n is made an int."""
    if ',' in srq:
        return [float(x) for x in srq.split(',')]
    if ':' in srq:
        start, stop, step = [float(x) for x in srq.split(':')]
        return list(np.arange(start, stop + step / 2.0, step))
    if ';' in srq:
        start, stop, n = [float(x) for x in srq.split(';')]
        return list(np.logspace(np.log10(start), np.log10(stop), int(n)))
    try:
        return [float(srq)]
    except ValueError:
        return list(np.loadtxt(srq))


def set_freq(freqs, freqUnit='GHz'):
    """Frequency request -> list in GHz (set_utils.py:91-129)."""
    if isinstance(freqs, list):
        freqs = list(freqs)
    elif isinstance(freqs, np.ndarray):
        freqs = list(freqs)
    elif isinstance(freqs, str):
        freqs = proc_string(freqs)
    elif utils.isanynum(freqs):
        freqs = [float(freqs)]
    else:
        raise ValueError('Invalid format for frequency request')
    freqs = [utils.convert_unit(f, freqUnit) for f in freqs]
    return freqs, freqUnit


def set_b(b, block=(1, 1), **kwargs):
    """Impact-parameter request -> Namespace(b, data_type, block, imSize) (set_utils.py:12-88).

    kwargs Rpol, Req are needed for 'start:stop:step[<angle]' profiles.  For a
    float grid the reference computes imSize from len() of the float (a
    TypeError); here imSize uses the number of grid rows instead."""
    block = list(block)
    rv = Namespace(b=None, data_type=None, block=block, imSize=None)
    if isinstance(b, str):
        b = b.lower()
        if b.startswith('dis'):
            rv.b = [b]
            rv.data_type = 'spectrum'
        elif b.startswith('stamp'):
            bres = float(b.split(':')[1])
            bext = [float(x) for x in b.split(':')[2].split(',')]
            rv.b = []
            for x in np.arange(bext[0], bext[1] + bres / 2.0, bres):
                for y in np.arange(bext[2], bext[3] + bres / 2.0, bres):
                    rv.b.append([y, x])
            xbr = len(np.arange(bext[2], bext[3] + bres / 2.0, bres))
            rv.imSize = [xbr, len(b) / xbr]
            rv.data_type = 'image'
        else:
            parts = b.split('<')
            angle_b = 0.0 if len(parts) == 1 else utils.d2r(float(parts[1]))
            mag_b = proc_string(parts[0])
            ab = kwargs['Rpol'] / kwargs['Req']
            rab = ab / np.sqrt(np.power(np.sin(angle_b), 2.0) + np.power(ab * np.cos(angle_b), 2.0))
            rv.b = [[v * np.cos(angle_b), v * np.sin(angle_b)] for v in mag_b if v < 0.995 * rab]
            rv.data_type = 'profile'
        return rv
    if isinstance(b, float):
        grid = -1.0 * np.flipud(np.arange(b, 1.5 + b, b))
        grid = np.concatenate((grid, np.arange(0.0, 1.5 + b, b)))
        bsplit = len(grid) / abs(block[1])
        last_row = block[0] / abs(block[1])
        if abs(block[1]) == 1:
            last_row = 0
        rv.b = []
        nrow = int(bsplit + last_row)
        for i in range(nrow):
            vrow = grid[i + int((block[0] - 1) * bsplit)]
            for vcol in grid:
                rv.b.append([vcol, vrow])
        rv.imSize = [len(grid), nrow]
        rv.data_type = 'image'
        return rv
    shape_b = np.shape(b)
    if len(shape_b) == 1:
        rv.data_type = 'spectrum'
        rv.b = [b]
    else:
        rv.data_type = 'spectrum' if shape_b[0] < 5 else 'profile'
        rv.b = b
    return rv
