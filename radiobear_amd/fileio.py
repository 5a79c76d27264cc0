"""Result files in the reference's text formats (radiobear/fileIO.py:10-107).

``FileIO.write`` writes a ``data_handling.Data`` as RadioBEAR does: the header
lines in sorted key order (fileIO.py:100-104), then the spectrum (one row per
frequency, one column per b, fileIO.py:40-68), profile (one row per b, one
column per frequency, fileIO.py:70-91) or image (rows of Tb, fileIO.py:93-99)
block, with the same format strings, so files written here and by the
reference compare byte for byte.  ``output_name`` is Planet.run's file name
(planet.py:140-144).
"""
import os

from . import utils

SPEED_OF_LIGHT = 2.9979E8   # m/s, utils.py:74 (the value the wavelength axis uses)


def output_name(planet, data_type, block, run_start, directory='Output'):
    """'{planet}_{type}{block}{YYYYmmdd_HHMMSS}.dat' in directory (planet.py:140-144)."""
    name = '{}_{}{}{}.dat'.format(planet, data_type, block, run_start.strftime("%Y%m%d_%H%M%S"))
    return os.path.join(directory, name)


class FileIO:
    """File output of Planet.run results."""

    def __init__(self, directory='Output', scratch_spec='Scratch/specoutputline.dat'):
        self.directory = directory
        self.scratch_spec = scratch_spec

    def write(self, output_file, data, xaxis='frequency'):
        """Write data (Data: header, type, f, freqUnit, b, Tb) to output_file; returns its name."""
        self.data = data
        xaxis = xaxis.lower()
        with open(output_file, 'w') as fp:
            self.writeHeader(data.header, fp)
            kind = data.type.lower()
            if kind == 'image':
                self.writeImage(fp)
            elif kind == 'spectrum':
                self.writeSpectrum(fp, xaxis=xaxis)
            elif kind == 'profile':
                self.writeProfile(fp, xaxis=xaxis)
            else:
                print("Invalid output type: {}".format(data.type))
        return output_file

    def _wavelength_cm(self, f):
        return 100.0 * SPEED_OF_LIGHT / (f * utils.Units[self.data.freqUnit])

    def writeSpectrum(self, fp, xaxis):
        """Frequency (or wavelength) rows, one Tb column per b; each row also goes to
        the scratch line file, as the reference does."""
        d = self.data
        s = '# cm   K@b  \t' if xaxis.startswith('wave') else '# {}  K@b  \t'.format(d.freqUnit)
        if utils.b_type(d.b) == 'disc':
            s += 'disc'
        else:
            for bv in d.b:
                s += '({:5.3f},{:5.3f})\t'.format(bv[0], bv[1])
        fp.write(s.strip('\t') + '\n')
        scratch = open(self.scratch_spec, 'w') if self.scratch_spec else None
        try:
            for i, f in enumerate(d.f):
                s = '{:.4f}\t  '.format(self._wavelength_cm(f)) if xaxis.startswith('wave') else \
                    '{:.2f}\t  '.format(f)
                for j in range(len(d.b)):
                    s += '  {:9.4f}  \t'.format(d.Tb[j][i])
                s = s.strip() + '\n'
                if scratch is not None:
                    scratch.write(s)
                fp.write(s)
        finally:
            if scratch is not None:
                scratch.close()

    def writeProfile(self, fp, xaxis):
        """One row per b, one Tb column per frequency."""
        d = self.data
        s = '# b  K@cm  \t' if xaxis.startswith('wave') else '# b  K@{} \t'.format(d.freqUnit)
        for f in d.f:
            s += ('  {:.4f}   \t'.format(self._wavelength_cm(f)) if xaxis.startswith('wave') else
                  '  {:9.4f}   \t'.format(f))
        fp.write(s.strip() + '\n')
        disc = utils.b_type(d.b) == 'disc'
        for i, bv in enumerate(d.b):
            s = 'disc' if disc else '{:5.3f} {:5.3f}\t'.format(bv[0], bv[1])
            for j in range(len(d.f)):
                s += ' {:7.2f}\t '.format(d.Tb[i][j])
            fp.write(s.strip() + '\n')

    def writeImage(self, fp):
        """Rows of Tb."""
        for row in self.data.Tb:
            s = ''
            for v in row:
                s += '{:7.2f}\t'.format(v)
            fp.write(s.strip() + '\n')

    def writeHeader(self, header, fp):
        """Header lines in sorted key order."""
        for key in sorted(header.keys()):
            fp.write(header[key] + '\n')
