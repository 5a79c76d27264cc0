"""Constituent plug-ins with the reference's module interface.

Each module exposes ``alpha(freq, T, P, X, P_dict, other_dict, **kwargs)`` for
one layer, as the reference's constituents/<c>/<module>.py do (called at
radiobear/alpha.py:210-213): X is the layer's gas column laid out by P_dict,
the result has len(freq) values in dB/km (default) or cm^-1 (units='invcm').
The line sum runs in the rbx_alpha kernel with only that constituent
selected.  As in the reference, the line list and its truncation
(truncate_strength / truncate_freq keywords) are fixed by a module's first call.
"""
import numpy as np

from .. import _lib, kernels, lines

DB_PER_INVCM = 434294.5
_SPECIES_KEYS = {'H2': 'H2', 'HE': 'HE', 'CH4': 'CH4', 'NH3': 'NH3', 'H2O': 'H2O', 'H2S': 'H2S', 'PH3': 'PH3'}


class _Module:
    """Per-module state: the frozen catalogue and the absorber selection."""

    def __init__(self, constituent, model, needs):
        self.constituent = constituent
        self.model = model
        self.needs = needs          # species the formalism reads from X (KeyError if absent)
        self.catalog = None
        self.data = None            # optional replacement line list (reference: module global `data`)

    def _catalog(self, kwargs):
        if self.catalog is None:
            ts = {self.constituent: kwargs.get('truncate_strength')}
            tf = {self.constituent: kwargs.get('truncate_freq')}
            extra = {}
            if self.data is not None:
                extra = {'nh3_inversion' if self.constituent == 'nh3' else 'h2o': self.data}
            self.catalog = kernels.jupiter_catalog(lines.raw_catalog(), ts, tf, **extra)
        return self.catalog

    def __call__(self, freq, T, P, X, P_dict, other_dict, **kwargs):
        units = kwargs.get('units', 'dBperkm')
        h2state = other_dict.get('h2state', 'e') if isinstance(other_dict, dict) else 'e'
        if self.constituent == 'h2' and h2state not in ('e', 'n'):
            print('INVALID H2STATE')
            return 0.0
        mix = np.zeros((len(_lib.SPECIES), 1))
        for s in self.needs:
            mix[_lib.SPECIES.index(s), 0] = X[P_dict[s]]
        ab = kernels.absorbers_struct({self.constituent: self.model}, h2state)
        f = np.atleast_1d(np.asarray(freq, dtype=np.float64))
        out = kernels.alpha(self._catalog(kwargs), ab, [float(P)], [float(T)], mix, f, 'fp64')
        a = out.cpu().numpy()[0]
        if units == 'dBperkm':
            a = a * DB_PER_INVCM
        return a

    def reset(self):
        self.catalog = None
