"""Frequency and impact-parameter requests: the request grammar Planet.run accepts.

The grammar is the one radiobear/set_utils.py documents for set_b (lines 12-33)
and set_freq (lines 91-117) and its results are pinned by fixtures the
reference wrote (tests/golden/requests.json).  Numbers in a request string:

    'v1,v2,...'     the listed values
    'lo:hi:step'    lo, lo + step, ... through hi (numpy arange up to hi + step / 2)
    'lo;hi;n'       n values evenly spaced in log10 from lo to hi
    'v'             one value
    anything else   the name of a text file of numbers (numpy loadtxt)

Frequencies: a list or array (taken as is), one number, or a request string;
every value is converted from freqUnit to GHz.

Impact parameters, as fractions of the equatorial radius (x, y on the sky):

    'disc' / 'disk' ('dis...')        the disc average: a spectrum of b = ['disc']
    'stamp:res:x0,x1,y0,y1'          an image: for each x of the x range (step res)
                                     the points [y, x] of the y range
    'numbers[<deg]'                  a profile along position angle deg (default 0):
                                     [v cos, v sin] for each magnitude v of the numbers
                                     inside 0.995 of the limb radius along that angle
    spacing d (a float)              an image of the square grid -1.5..1.5 with step d
                                     (a block of its rows when block = [i, n])
    [x, y]                           one point: a spectrum
    [[x, y], ...]                    points: a spectrum (fewer than 5) or a profile

Deviations from the reference (each an exception or a meaningless value there):
  * 'lo;hi;n' passes n to numpy as an int (the reference passes a float, which
    numpy 2 refuses);
  * an image from a float spacing gets imSize [points per row, rows] (the
    reference takes len() of the float, a TypeError);
  * a stamp gets imSize [y points, x points] (the reference divides the length of
    the request string by the row length).
"""
from argparse import Namespace

import numpy as np

from . import utils

LIMB_FRACTION = 0.995   # profile points stop short of the limb by this factor (set_utils.py:56-61)
IMAGE_EXTENT = 1.5      # float-spacing images cover -1.5 .. 1.5 (set_utils.py:64-65)


def _inclusive_range(lo, hi, step):
    """lo, lo + step, ... through hi (the stop is hi + step / 2, so hi itself is kept
    whatever the rounding of the steps)."""
    return np.arange(lo, hi + step / 2.0, step)


def _value_list(text):
    """The numbers of one request string (module docstring grammar)."""
    if ',' in text:
        return [float(t) for t in text.split(',')]
    for sep, build in ((':', lambda lo, hi, step: _inclusive_range(lo, hi, step)),
                       (';', lambda lo, hi, n: np.logspace(np.log10(lo), np.log10(hi), int(n)))):
        if sep in text:
            lo, hi, third = (float(t) for t in text.split(sep))
            return list(build(lo, hi, third))
    try:
        return [float(text)]
    except ValueError:
        return list(np.loadtxt(text))


proc_string = _value_list     # the reference's name for it (set_utils.py:130)


def set_freq(freqs, freqUnit='GHz'):
    """Frequency request -> (list of GHz values, freqUnit)."""
    if isinstance(freqs, (list, np.ndarray)):
        values = list(freqs)
    elif isinstance(freqs, str):
        values = _value_list(freqs)
    elif utils.isanynum(freqs):
        values = [float(freqs)]
    else:
        raise ValueError('Invalid format for frequency request')
    return [utils.convert_unit(v, freqUnit) for v in values], freqUnit


def _limb_radius(angle, Rpol, Req):
    """Radius of the ellipse with axis ratio Rpol/Req along position angle `angle` (rad), in
    units of the equatorial radius."""
    q = Rpol / Req
    return q / np.sqrt(np.sin(angle) ** 2 + (q * np.cos(angle)) ** 2)


def _profile(text, Rpol, Req):
    magnitudes, _, angle = text.partition('<')
    theta = utils.d2r(float(angle)) if angle else 0.0
    edge = LIMB_FRACTION * _limb_radius(theta, Rpol, Req)
    c, s = np.cos(theta), np.sin(theta)
    return [[v * c, v * s] for v in _value_list(magnitudes) if v < edge]


def _stamp(text):
    """'stamp:res:x0,x1,y0,y1' -> (points, imSize): x outer, y inner, each point [y, x]."""
    _, res, extent = text.split(':')[:3]
    res = float(res)
    x0, x1, y0, y1 = (float(t) for t in extent.split(','))
    xs, ys = _inclusive_range(x0, x1, res), _inclusive_range(y0, y1, res)
    return [[y, x] for x in xs for y in ys], [len(ys), len(xs)]


def _image_grid(spacing, block):
    """The square image grid of a float spacing, rows of block [i, n] ('block i of n'):
    columns run -1.5..1.5 ([-flip(spacing..1.5), 0..1.5]); block i takes rows
    [(i - 1) len/n, (i - 1) len/n + len/n + i/n) (the last block takes one more when
    n > 1, as the reference's row split does)."""
    axis = np.concatenate((-1.0 * np.flipud(np.arange(spacing, IMAGE_EXTENT + spacing, spacing)),
                           np.arange(0.0, IMAGE_EXTENT + spacing, spacing)))
    per_block = len(axis) / abs(block[1])
    extra = 0 if abs(block[1]) == 1 else block[0] / abs(block[1])
    first = int((block[0] - 1) * per_block)
    nrows = int(per_block + extra)
    points = [[x, axis[first + r]] for r in range(nrows) for x in axis]
    return points, [len(axis), nrows]


def set_b(b, block=(1, 1), **kwargs):
    """Impact-parameter request -> Namespace(b, data_type, block, imSize).
    Profiles need kwargs Rpol and Req (the planet's radii)."""
    req = Namespace(b=None, data_type=None, block=list(block), imSize=None)
    if isinstance(b, str):
        text = b.lower()
        if text.startswith('dis'):
            req.b, req.data_type = [text], 'spectrum'
        elif text.startswith('stamp'):
            req.b, req.imSize = _stamp(text)
            req.data_type = 'image'
        else:
            req.b, req.data_type = _profile(text, kwargs['Rpol'], kwargs['Req']), 'profile'
    elif isinstance(b, float):
        req.b, req.imSize = _image_grid(b, req.block)
        req.data_type = 'image'
    elif len(np.shape(b)) == 1:
        req.b, req.data_type = [b], 'spectrum'
    else:
        req.b = b
        req.data_type = 'spectrum' if np.shape(b)[0] < 5 else 'profile'
    return req
