"""Convert the reference's Orton H2 CIA table (constituents/h2/orton_H2.tables) to
orton_H2.npz, the form radiobear_amd.lines loads.

    python radiobear_amd/data/orton_to_npz.py /path/to/orton_H2.tables

The text layout (as h2_orton.readInputFiles reads it, constituents/h2/h2_orton.py:17-97):
line 1 "nTemp Tmax Tmin", line 2 nfreq, then the nfreq table frequencies in
cm^-1 (several per line), then rows of nTemp natural-log absorption values, one
row per frequency, table after table in the order eh2h2, nh2h2, eh2he, nh2he,
eh2ch4, nh2ch4.  The reference reads only those six tables (the file holds
more); the npz keeps them, with the values exactly as parsed.
"""
import math
import sys

import numpy as np

N_TABLES = 6


def convert(src, dst):
    with open(src) as fp:
        lines = fp.read().split('\n')
    head = lines[0].split()
    ntemp, tmax, tmin = int(head[0]), float(head[1]), float(head[2])
    nfreq = int(lines[1])
    per_line = len(lines[2].split())
    nfline = int(math.ceil(1.0 * nfreq / per_line))
    ftab = [float(v) for ln in lines[2:2 + nfline] for v in ln.split()]
    if len(ftab) != nfreq:
        raise ValueError('frequency block has {} values, header says {}'.format(len(ftab), nfreq))
    first = 2 + nfline
    rows = [[float(v) for v in ln.split()] for ln in lines[first:first + N_TABLES * nfreq]]
    if any(len(r) != ntemp for r in rows):
        raise ValueError('table rows must hold {} values'.format(ntemp))
    np.savez_compressed(dst, header=np.array([ntemp, tmax, tmin], dtype=np.float64),
                        ftab_invcm=np.array(ftab), logv=np.array(rows))


if __name__ == '__main__':
    import os
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'orton_H2.npz')
    convert(sys.argv[1], out if len(sys.argv) < 3 else sys.argv[2])
    print('wrote', out if len(sys.argv) < 3 else sys.argv[2])
