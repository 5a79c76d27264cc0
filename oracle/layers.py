"""FP64 restatement of Alpha.get_layers (TEST INFRASTRUCTURE ONLY).

Follows radiobear/alpha.py:151-305: per layer, every configured absorber is
called in ``sorted`` constituent order, the (nf, nc) table is scale-summed,
and the per-layer totals are stacked into ``layers[nf][nl]`` (cm^-1).
"""
import numpy as np

from . import absorbers


def isanynum(x):
    """utils.isanynum (radiobear/utils.py:33-41)."""
    if isinstance(x, (bool, dict, list)):
        return False
    try:
        float(x)
    except (TypeError, ValueError):
        return False
    return True


def layer_scale(scale, N, ordered):
    """Alpha.get_layer_scale (radiobear/alpha.py:235-259)."""
    if isinstance(scale, dict):
        for k, v in scale.items():
            if k not in ordered:
                raise ValueError("{} not found as constituent for alpha".format(k))
            if len(v) != N:
                raise ValueError("Incorrect scale for {}:  N {} vs {}".format(k, len(v), N))
        return [{k: v[lyr] for k, v in scale.items()} for lyr in range(N)]
    if isinstance(scale, (list, np.ndarray)):
        if len(scale) != N:
            raise ValueError("Incorrect number of scale layers.")
        return list(scale)
    val = float(scale) if isanynum(scale) else 1.0
    return [val] * N


def total_layer_alpha(absorb, lscale, ordered):
    """Alpha.total_layer_alpha (radiobear/alpha.py:151-192); absorb is (nf, nc)."""
    nf = absorb.shape[0]
    total = np.zeros(nf)
    if isanynum(lscale):
        s = float(lscale)
        for i in range(nf):
            total[i] = absorb[i].sum() * s
        return total, absorb * s
    saved = np.array(absorb, dtype=np.float64)
    for i in range(nf):
        for j, c in enumerate(ordered):
            v = absorb[i, j]
            if c in lscale:
                v *= lscale[c]
            saved[i, j] = v
            total[i] += v
    return total, saved


class AbsorberSet:
    """The configured absorbers: name -> callable(freq, T, P, X, P_dict, other) (cm^-1)."""

    def __init__(self, funcs, other_dict=None):
        self.funcs = dict(funcs)
        self.ordered = sorted(self.funcs)             # alpha.py:80
        self.other = other_dict or {}

    def absorb(self, freqs, T, P, X, P_dict):
        """Alpha.get_alpha_from_calc (radiobear/alpha.py:194-216) -> (nf, nc)."""
        cols = [self.funcs[c](freqs, T, P, X, P_dict, self.other.get(c, {}), units='invcm')
                for c in self.ordered]
        return np.array(cols).transpose()


def jupiter_absorbers(lines, h2state='e', nh3='nh3_hs_sjs'):
    """Absorber set of the Jupiter config (default_config.json 'alpha', config.par:23);
    nh3='nh3_dbs_sjs' gives the Saturn / Uranus / Neptune set (their config.par:23).

    ``lines`` maps 'nh3_hs', 'nh3_sjs', 'h2s', 'ph3', 'h2o' to already-truncated
    line dicts (h2s / ph3 at the Jupiter default I0 > 1e-22 cut).
    """
    nh3_fn = {'nh3_hs_sjs': absorbers.nh3_hs_sjs, 'nh3_dbs_sjs': absorbers.nh3_dbs_sjs}[nh3]
    def bind(fn, **data):
        return lambda f, T, P, X, D, o, **kw: fn(f, T, P, X, D, o, **data, **kw)
    funcs = {
        'h2': absorbers.h2_jj_ddb,
        'h2o': bind(absorbers.h2o_bk, data=lines['h2o']),
        'h2s': bind(absorbers.h2s_ddb, data=lines['h2s']),
        'nh3': bind(nh3_fn, data_hs=lines['nh3_hs'], data_sjs=lines['nh3_sjs']),
        'ph3': bind(absorbers.ph3_jh, data=lines['ph3']),
    }
    return AbsorberSet(funcs, other_dict={'h2': {'h2state': h2state}})


def get_layers(freqs, gas, P_dict, absorber_set, scale=False, return_saved=False):
    """Alpha.get_layers (radiobear/alpha.py:261-305) with get_alpha='calc'.

    gas: (nC, nl) array laid out by ``P_dict`` (the reference's atm.gas).
    Returns layers (nf, nl) and, if asked, the per-layer saved absorb tables.
    """
    nl = gas.shape[1]
    lscale = layer_scale(scale, nl, absorber_set.ordered)
    cols = []
    saved = []
    for layer in range(nl):
        P = gas[P_dict['P']][layer]
        T = gas[P_dict['T']][layer]
        X = gas[:, layer]
        absorb = absorber_set.absorb(freqs, T, P, X, P_dict)
        tot, sv = total_layer_alpha(absorb, lscale[layer], absorber_set.ordered)
        cols.append(tot)
        saved.append(sv)
    layers = np.array(cols).transpose()
    if return_saved:
        return layers, np.array(saved)
    return layers
