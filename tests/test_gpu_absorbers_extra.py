"""GPU parity of the other planets' and the optional absorbers (nh3_dbs, nh3_dbs_sjs,
co_ddb, clouds_idp, h2_orton) through rbx_alpha, against the reference's own outputs
(tests/golden/other_absorbers.npz) and the pinned oracle."""
import numpy as np
import pytest

from oracle import absorbers as A
from oracle import extra_absorbers as E
from oracle import layers
from radiobear_amd import synth

pytestmark = pytest.mark.gpu
C = synth.GAS_ROWS
CLOUD_DENSITY = {'H2O': 1e-4, 'SOLN': 5e-4, 'NH4SH': 2e-5, 'NH3': 3e-5, 'H2S': 1e-5, 'CH4': 1e-5}
# alpha in FP32 mode is the FP64 closed forms rounded once to float32 at the end
TOL = {'fp64': 1e-9, 'fp32': 1e-6}


@pytest.fixture(scope='module')
def K():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no HIP device')
    from radiobear_amd import kernels
    return kernels


@pytest.fixture(scope='module')
def fx(golden):
    return golden('other_absorbers.npz')


@pytest.fixture(scope='module')
def cat(K, raw_lines):
    return K.jupiter_catalog(raw_lines, co=True, orton=True)


def columns(TP, co=0.0):
    """(nC, n) gas table with one layer per (T, P) point, synthetic mixing ratios."""
    g = np.zeros((len(C), len(TP)))
    g[C['T']], g[C['P']] = TP[:, 0], TP[:, 1]
    for k, v in synth.MIXING.items():
        g[C[k]] = v
    g[C['CO']] = co
    return g


def row_err(got, ref):
    """max over rows of |got - ref| / max|ref| of that row."""
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    scale = np.maximum(np.max(np.abs(ref), axis=1, keepdims=True), 1e-300)
    return float(np.max(np.abs(got - ref) / scale))


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


def run(K, cat, sel, g, freqs, precision, other=None, cloud=None, h2state='e'):
    ab = K.absorbers_struct(sel, h2state, other)
    out = K.alpha(cat, ab, g[C['P']], g[C['T']], K.mix_table(g, C), freqs, precision, cloud=cloud)
    return out.double().cpu().numpy()


@pytest.mark.parametrize('precision', ['fp64', 'fp32'])
@pytest.mark.parametrize('model', ['nh3_dbs', 'nh3_dbs_sjs'])
def test_nh3_dbs_matches_reference(K, cat, fx, model, precision):
    """All regimes of the 12 (T, P) points: 30 GHz switch, blend, deep nh3_sjs."""
    g = columns(fx['TP'])
    got = run(K, cat, {'nh3': model}, g, fx['freqs'], precision)
    tol = 1e-9 if precision == 'fp64' else 1e-4
    assert rel(got, fx['alpha_' + model]) < tol


@pytest.mark.parametrize('precision', ['fp64', 'fp32'])
@pytest.mark.parametrize('shape', ['voigt', 'vvw', 'diff', 'std'])
def test_co_matches_reference(K, cat, fx, shape, precision):
    g = columns(fx['TP_co'], co=1.0e-6)
    got = run(K, cat, {'co': 'co_ddb'}, g, fx['freqs'], precision, other={'co': {'coshape': shape}})
    assert row_err(got, fx['alpha_co_' + shape]) < TOL[precision]


@pytest.mark.parametrize('precision', ['fp64', 'fp32'])
def test_clouds_match_reference(K, cat, fx, precision):
    from radiobear_amd import _lib
    TP = fx['TP_cloud']
    g = columns(TP)
    cloud = np.stack([np.full(len(TP), CLOUD_DENSITY[k]) for k in _lib.CLOUDS])
    for ci, case in enumerate(fx['cloud_cases']):
        flags = {f: 1e-4 for f in str(case).split('+')}
        got = run(K, cat, {'clouds': 'clouds_idp'}, g, fx['freqs'], precision, other={'clouds': flags},
                  cloud=cloud)
        ref = fx['alpha_clouds_{}'.format(ci)]
        assert row_err(got, ref) < TOL[precision], case


@pytest.mark.parametrize('precision', ['fp64', 'fp32'])
@pytest.mark.parametrize('state', ['e', 'n'])
def test_h2_orton_matches_reference(K, cat, fx, state, precision):
    """Below (30 K), inside (cubic spline) and above (h2_jj scaled) the 40-400 K table."""
    g = columns(fx['TP_orton'])
    got = run(K, cat, {'h2': 'h2_orton'}, g, fx['freqs'], precision, h2state=state)
    assert row_err(got, fx['alpha_h2_orton_' + state]) < TOL[precision]


def test_h2_orton_dense_temperatures_vs_oracle(K, cat):
    """Many temperatures inside every spline interval and at the table nodes."""
    from radiobear_amd import lines
    tab = lines.load_orton()
    T = np.concatenate([np.linspace(20.0, 39.9, 7), np.geomspace(40.0, 400.0, 37), [400.0, 401.0, 800.0]])
    TP = np.stack([T, np.geomspace(0.01, 50.0, len(T))], axis=1)
    g = columns(TP)
    f = np.array([0.5, 1.0, 7.5, 22.0, 60.0, 150.0, 600.0])
    got = run(K, cat, {'h2': 'h2_orton'}, g, f, 'fp64')
    prep = E.orton_prepare(tab, f)
    ref = np.array([E.h2_orton(f, Tl, Pl, g[:, i], C, {'h2state': 'e'}, prepared=prep, units='invcm')
                    for i, (Tl, Pl) in enumerate(TP)])
    assert rel(got, ref) < 1e-9


def test_h2_orton_rejects_out_of_table_frequencies(K, cat):
    g = columns(np.array([[165.0, 1.0]]))
    with pytest.raises(ValueError):
        run(K, cat, {'h2': 'h2_orton'}, g, np.array([1.0, 80000.0]), 'fp64')


def test_neptune_absorber_set_vs_oracle(K, cat, jupiter_lines):
    """The Neptune defaults (co_ddb with nh3_dbs_sjs and the shared absorbers) over a
    synthetic atmosphere with CO, total and per constituent, against the oracle's
    get_layers (alpha.py:261-305) with the same absorber set."""
    from radiobear_amd import _lib, lines
    atm = synth.jupiter_like(60)
    gas = np.array(atm.gas)
    gas[C['CO']] = 1.0e-6
    freqs = np.array([0.5, 2.0, 10.0, 25.0, 30.0, 30.5, 90.0, 115.0, 230.0])
    co = lines.load('co')
    aset = layers.jupiter_absorbers(jupiter_lines, nh3='nh3_dbs_sjs')
    aset.funcs['co'] = lambda f, T, P, X, D, o, **kw: E.co_ddb(f, T, P, X, D, o, data=co, **kw)
    aset.ordered = sorted(aset.funcs)
    aset.other['co'] = {'coshape': 'voigt'}
    ref, saved = layers.get_layers(list(freqs), gas, C, aset, return_saved=True)
    sel = {'co': 'co_ddb', 'h2': 'h2_jj_ddb', 'h2o': 'h2o_bk', 'h2s': 'h2s_ddb', 'nh3': 'nh3_dbs_sjs',
           'ph3': 'ph3_jh'}
    ab = K.absorbers_struct(sel, 'e', {'co': {'coshape': 'voigt'}})
    out, byc = K.alpha(cat, ab, gas[C['P']], gas[C['T']], K.mix_table(gas, C), freqs, 'fp64', by_constituent=True)
    assert rel(out.cpu().numpy().T, ref) < 1e-9
    byc = byc.cpu().numpy()
    for j, c in enumerate(aset.ordered):
        assert rel(byc[_lib.CONSTITUENTS.index(c)], saved[:, :, j]) < 1e-9, c
