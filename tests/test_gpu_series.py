"""Far-line series (launch.alpha_series 2): the per-tile near sums plus the series of the far
lines' moments against the direct sums over every line, in both precisions, and against the
oracle.  Series against direct: 1e-6 in FP32 (16 terms at pole ratio 1/3, plus FP32 rounding),
1e-11 in FP64 (16 terms at ratio 0.15 leave at most ~8e-12 of a far line's value); against the
oracle the precision's own bar (1e-5 FP32, 1e-9 FP64)."""
import numpy as np
import pytest

from oracle import layers
from radiobear_amd import synth

pytestmark = pytest.mark.gpu
C = synth.GAS_ROWS
JUPITER_ALPHA = {'h2': 'h2_jj_ddb', 'h2o': 'h2o_bk', 'h2s': 'h2s_ddb', 'nh3': 'nh3_hs_sjs', 'ph3': 'ph3_jh'}


@pytest.fixture(scope='module')
def K():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no HIP device')
    from radiobear_amd import kernels
    return kernels


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


SERIES_TOL = {'fp32': 1e-6, 'fp64': 1e-11}
ORACLE_TOL = {'fp32': 1e-5, 'fp64': 1e-9}
PRECISIONS = pytest.mark.parametrize('prec', ['fp32', 'fp64'])


def both_paths(K, cat, sel, atm, f, prec='fp32', launch=None, **kw):
    out = {}
    for mode in (1, 2):
        r = K.alpha(cat, K.absorbers_struct(sel), atm.P, atm.T, K.mix_table(atm.gas, C), f, prec,
                    launch=dict(launch or {}, alpha_series=mode), **kw)
        out[mode] = [t.double().cpu().numpy() for t in r] if isinstance(r, tuple) else r.double().cpu().numpy()
    return out[1], out[2]


@PRECISIONS
@pytest.mark.parametrize('tile_k', [1, 2, 4])
def test_series_matches_direct_and_oracle(K, raw_lines, jupiter_lines, tile_k, prec):
    """330 log-spaced frequencies 1-200 GHz (a tile straddles the 30 GHz switch, the last tile
    is partial), 41 layers to 100 bar, 300 synthetic H2O lines: series against the direct sums
    at every tile width, both against the FP64 oracle."""
    atm = synth.jupiter_like(41)
    h2o = synth.h2o_lines(300, 5)
    cat = K.jupiter_catalog(raw_lines, h2o=h2o)
    lines = dict(jupiter_lines)
    lines['h2o'] = h2o
    f = np.logspace(0.0, np.log10(200.0), 330)
    direct, series = both_paths(K, cat, JUPITER_ALPHA, atm, f, prec, launch={'alpha_k': tile_k})
    ref = layers.get_layers(list(f), atm.gas, C, layers.jupiter_absorbers(lines)).T
    assert rel(series, direct) < SERIES_TOL[prec]
    assert rel(series, ref) < ORACLE_TOL[prec]
    assert rel(direct, ref) < ORACLE_TOL[prec]


@PRECISIONS
def test_series_per_constituent_and_deep_layers(K, raw_lines, jupiter_lines, prec):
    """Layers to 3000 bar (nh3_hs blended into nh3_sjs, then off), every constituent's row."""
    atm = synth.jupiter_like(37, pmax=3000.0)
    h2o = synth.h2o_lines(700, 9)
    cat = K.jupiter_catalog(raw_lines, h2o=h2o)
    f = np.logspace(np.log10(0.5), np.log10(250.0), 190)
    (da, dbyc), (sa, sbyc) = both_paths(K, cat, JUPITER_ALPHA, atm, f, prec, by_constituent=True)
    assert rel(sa, da) < SERIES_TOL[prec]
    for c in range(dbyc.shape[0]):
        if np.any(dbyc[c] != 0.0):
            assert rel(sbyc[c], dbyc[c]) < SERIES_TOL[prec], c


@PRECISIONS
def test_series_unordered_request_and_extreme_frequencies(K, raw_lines, jupiter_lines, prec):
    """A shuffled request (the NH3 columns are permuted, so tiles hold scattered x ranges)
    with frequencies from 0.05 to 2000 GHz: lines fall below, above and across tiles."""
    atm = synth.jupiter_like(23)
    cat = K.jupiter_catalog(raw_lines)
    rng = np.random.default_rng(3)
    f = rng.permutation(np.concatenate([np.geomspace(0.05, 2000.0, 150), [22.2, 23.7, 30.0, 30.01]]))
    direct, series = both_paths(K, cat, JUPITER_ALPHA, atm, f, prec)
    ref = layers.get_layers(list(f), atm.gas, C, layers.jupiter_absorbers(jupiter_lines)).T
    assert rel(series, direct) < SERIES_TOL[prec]
    assert rel(series, ref) < ORACLE_TOL[prec]


@PRECISIONS
@pytest.mark.parametrize('nf', [1, 63, 65])
def test_series_small_requests(K, raw_lines, jupiter_lines, nf, prec):
    atm = synth.jupiter_like(9)
    cat = K.jupiter_catalog(raw_lines)
    f = np.linspace(1.5, 95.0, nf)
    direct, series = both_paths(K, cat, JUPITER_ALPHA, atm, f, prec)
    assert rel(series, direct) < SERIES_TOL[prec]


@PRECISIONS
def test_series_large_workload_sample(K, prec):
    """The 800-layer, 25753-line, 0.3-300 GHz workload's catalogue on 40 of its layers and
    all 1024 frequencies: series against the direct sums."""
    from radiobear_amd import lines
    w = synth.workload('c3_large')
    atm = w['atm']
    keep = np.arange(0, atm.nl, 20)
    cat = K.jupiter_catalog(lines.raw_catalog(), nh3_inversion=w['nh3_inv'], h2o=w['h2o'])
    sel = K.absorbers_struct(JUPITER_ALPHA)
    P, T, mix = atm.P[keep], atm.T[keep], K.mix_table(atm.gas[:, keep], C)
    direct = K.alpha(cat, sel, P, T, mix, w['freqs'], prec, launch={'alpha_series': 1}).double().cpu().numpy()
    series = K.alpha(cat, sel, P, T, mix, w['freqs'], prec, launch={'alpha_series': 2}).double().cpu().numpy()
    assert rel(series, direct) < SERIES_TOL[prec]


@pytest.mark.parametrize('precision', ['FP32', 'FP64'])
def test_series_workspace_cap(K, raw_lines, precision):
    """The series path stores every line's staged form per layer (24 B FP32, 40 B FP64): for a
    request whose workspace would pass 4 GiB the built-in choice is the direct sums; forcing it
    still works.  Below 4 wavefront tiles (here 20 frequencies, one tile) the built-in choice is
    the direct sums too.  (Sizes from rbx_alpha_workspace_size only: nothing this large is
    allocated.)"""
    import ctypes
    from radiobear_amd import _lib
    cat = K.jupiter_catalog(raw_lines)
    ab = K.absorbers_struct(JUPITER_ALPHA)

    def size(nl, mode, nf=1024):
        nb = ctypes.c_size_t(0)
        _lib.check(_lib.lib().rbx_alpha_workspace_size(cat.handle, ctypes.byref(ab), nl, nf, nf // 2,
                                                       getattr(_lib, precision),
                                                       _lib.launch_ref({'alpha_series': mode}), ctypes.byref(nb)))
        return nb.value

    big = 400000
    assert size(big, 2) > 4 << 30                  # the series workspace of this request
    assert size(big, 0) == size(big, 1)            # built in: the direct sums
    assert size(40, 0) == size(40, 2) != size(40, 1)  # below the cap: the series
    assert size(40, 0, 20) == size(40, 1, 20) != size(40, 2, 20)   # one tile: the direct sums


@pytest.mark.parametrize('band', [(1.0, 25.0), (35.0, 400.0)])
@PRECISIONS
def test_series_one_inversion_set(K, raw_lines, band, prec):
    """Requests entirely below or above the 30 GHz switch: one coefficient set of the inversion
    lines is never used (its stream is neither written nor read)."""
    atm = synth.jupiter_like(17)
    cat = K.jupiter_catalog(raw_lines)
    f = np.geomspace(band[0], band[1], 300)
    direct, series = both_paths(K, cat, JUPITER_ALPHA, atm, f, prec)
    assert rel(series, direct) < SERIES_TOL[prec]


@pytest.mark.parametrize('sel', [{'nh3': 'nh3_hs'}, {'h2o': 'h2o_bk', 'h2': 'h2_jj_ddb'}, {'ph3': 'ph3_jh'},
                                 {'h2s': 'h2s_ddb', 'nh3': 'nh3_dbs'}])
@PRECISIONS
def test_series_absorber_subsets(K, raw_lines, sel, prec):
    """Absorber selections that leave line segments empty."""
    atm = synth.jupiter_like(13)
    h2o = synth.h2o_lines(200, 3)
    cat = K.jupiter_catalog(raw_lines, h2o=h2o)
    f = np.geomspace(0.5, 300.0, 257)
    direct, series = both_paths(K, cat, sel, atm, f, prec)
    assert rel(series, direct) < SERIES_TOL[prec]
