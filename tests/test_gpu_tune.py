"""Every launch shape the autotuner can pick gives the same results."""
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


@pytest.fixture(scope='module')
def at(K):
    from radiobear_amd import autotune
    return autotune


@pytest.mark.parametrize('precision,tol', [('fp64', 1e-9), ('fp32', 1e-5)])
def test_alpha_every_tile_width(K, at, raw_lines, jupiter_lines, precision, tol):
    """1, 2 and 4 frequencies per lane, all matching the oracle, with the direct line sums and
    with the far-line series.  FP64 direct sums: identical tables (the per-frequency line order
    does not depend on the tile width).  FP32: a chunk of 8 staged lines takes the all-FP32 path
    when it is far from every column of the wavefront, and with the series which lines are far
    depends on the tile's frequency range: tables agree to 1e-6 (FP32) and 1e-11 (FP64 series).
    330 freqs straddle the 30 GHz regime switch inside a tile; 41 layers leave a partial last
    tile."""
    atm = synth.jupiter_like(41)
    h2o = synth.h2o_lines(300, 5)
    cat = K.jupiter_catalog(raw_lines, h2o=h2o)
    lines = dict(jupiter_lines)
    lines['h2o'] = h2o
    f = np.logspace(np.log10(1.0), np.log10(200.0), 330)
    ref = layers.get_layers(list(f), atm.gas, C, layers.jupiter_absorbers(lines))
    for series in (1, 2):
        out = {}
        for k in at.CANDIDATES['alpha_k']:
            out[k] = K.alpha(cat, K.absorbers_struct(JUPITER_ALPHA), atm.P, atm.T, K.mix_table(atm.gas, C), f,
                             precision, launch={'alpha_series': series, 'alpha_k': k}).double().cpu().numpy()
            got = out[k].T
            assert np.max(np.abs(got - ref) / np.abs(ref)) < tol, (series, k)
        for k in out:
            if precision == 'fp64' and series == 1:
                assert np.array_equal(out[k], out[1]), k
            else:
                same = 1e-11 if precision == 'fp64' else 1e-6
                assert np.max(np.abs(out[k] - out[1]) / np.abs(out[1])) < same, (series, k)


@pytest.mark.parametrize('precision,tol', [('fp64', 1e-9), ('fp32', 1e-5)])
def test_alpha_every_split(K, at, raw_lines, jupiter_lines, precision, tol):
    """Line sums split into 1..16 parts per layer (part boundaries fall inside the h2o and
    nh3 inversion blocks and between segments): the parts add up to the unsplit table up to
    FP64 reassociation, and every split matches the oracle.  Deep layers (to 3000 bar) take
    the nh3_sjs / blend branches of nh3_hs_sjs in the finish kernel."""
    atm = synth.jupiter_like(37, pmax=3000.0)
    h2o = synth.h2o_lines(700, 9)
    cat = K.jupiter_catalog(raw_lines, h2o=h2o)
    lines = dict(jupiter_lines)
    lines['h2o'] = h2o
    f = np.logspace(np.log10(0.5), np.log10(250.0), 190)
    ref = layers.get_layers(list(f), atm.gas, C, layers.jupiter_absorbers(lines))
    out = {}
    for n in (1, 2, 3, 7, 16):
        a, byc = K.alpha(cat, K.absorbers_struct(JUPITER_ALPHA), atm.P, atm.T, K.mix_table(atm.gas, C), f,
                         precision, by_constituent=True, launch={'alpha_series': 1, 'alpha_split': n})
        out[n] = a.double().cpu().numpy()
        assert np.max(np.abs(out[n].T - ref) / np.abs(ref)) < tol, n
        assert np.allclose(byc.sum(axis=0).cpu().numpy(), out[n], rtol=1e-6 if precision == 'fp32' else 1e-12)
    for n in out:
        assert np.max(np.abs(out[n] - out[1]) / np.abs(out[1])) < (1e-12 if precision == 'fp64' else 1e-6), n


@pytest.mark.parametrize('precision', ['fp64', 'fp32'])
def test_integrate_every_unroll(K, at, precision):
    """4 and 8 path steps per load batch: the same Tb to rounding (the compiler may contract
    multiply-adds differently in the two loop bodies; in FP32 mode the FP32 trapezoid sums
    are added into the FP64 totals once per batch) without the tau cut, and with the cut
    (checked once per batch, so the two stop at different steps) within e^-50 of it."""
    import torch
    atm = synth.jupiter_like(120)
    rng = np.random.default_rng(3)
    nf = 70
    tab = rng.uniform(1e-9, 2e-6, size=(atm.nl, nf))
    dt = torch.float64 if precision == 'fp64' else torch.float32
    tab = torch.as_tensor(tab, device='cuda').to(dt)
    pl = K.planet_desc(atm.Req, atm.Rpol)
    rays = synth.random_disc_rays(300, 11)
    hds, ns = K.raypaths(pl, atm.R, atm.N, rays)
    res = {}
    for u in at.CANDIDATES['int_unroll']:
        L = {'int_unroll': u}
        a = K.integrate(tab, atm.T, hds, ns, precision=precision, tau_cut=0.0, launch=L).cpu().numpy()
        b = K.integrate(tab, atm.T, hds, ns, precision=precision, tau_cut=50.0, launch=L).cpu().numpy()
        res[u] = (a, b)
    base = res[at.CANDIDATES['int_unroll'][0]][0]
    ok = ~np.isnan(base)
    assert ok.any()
    for u, (a, b) in res.items():
        assert np.array_equal(np.isnan(a), ~ok)
        scale = np.max(np.abs(base[ok]))
        tol = 1e-13 if precision == 'fp64' else 2e-6
        assert np.max(np.abs(a[ok] - base[ok])) < tol * scale, u
        assert np.max(np.abs(b[ok] - base[ok])) < max(tol, 1e-12) * scale, u


@pytest.mark.parametrize('precision', ['fp64', 'fp32'])
def test_integrate_every_block_order(K, at, precision):
    """Frequency-group-major and ray-major (XCD-aware) block orders give bit-identical Tb:
    only which block integrates which (ray, frequency group) changes.  Ray counts that are
    not a multiple of 4 x 8 leave padding blocks with no ray."""
    import torch
    atm = synth.jupiter_like(90)
    rng = np.random.default_rng(4)
    nf = 130                                        # 3 frequency groups, the last partial
    tab = torch.as_tensor(rng.uniform(1e-9, 2e-6, size=(atm.nl, nf)), device='cuda')
    if precision == 'fp32':
        tab = tab.float()
    pl = K.planet_desc(atm.Req, atm.Rpol)
    for nray in (1, 37, 200):
        rays = synth.random_disc_rays(nray, 12 + nray)
        hds, ns = K.raypaths(pl, atm.R, atm.N, rays)
        out = {}
        for m in at.CANDIDATES['int_map']:
            out[m] = K.integrate(tab, atm.T, hds, ns, precision=precision, tau_cut=50.0,
                                 launch={'int_map': m}).cpu().numpy()
        a, b = out[1], out[2]
        assert np.array_equal(np.isnan(a), np.isnan(b))
        assert np.array_equal(a[~np.isnan(a)], b[~np.isnan(b)]), nray


def test_tune_picks_a_candidate_and_caches(K, at, tmp_path):
    calls = []
    L, rec = at.tune('int_unroll', 'test-shape', lambda L: calls.append(dict(L)), {'int_map': 2}, reps=1,
                     path=str(tmp_path / 't.json'))
    assert rec['best'] in at.CANDIDATES['int_unroll']
    assert L == {'int_map': 2, 'int_unroll': rec['best']}
    assert {c['int_unroll'] for c in calls} == set(at.CANDIDATES['int_unroll'])
    assert all(c['int_map'] == 2 for c in calls)
    n = len(calls)
    L2, again = at.tune('int_unroll', 'test-shape', lambda L: calls.append(dict(L)), {'int_map': 2}, reps=1,
                        path=str(tmp_path / 't.json'))
    assert again == rec and L2 == L and len(calls) == n


def test_far_split_parts(K, at, raw_lines):
    """The series' classification pass split into 1..16 parts per layer: the same table up to
    the order the parts' moments are added in (FP64 tables to 1e-12, FP32 to 1e-6)."""
    atm = synth.jupiter_like(29)
    cat = K.jupiter_catalog(raw_lines, h2o=synth.h2o_lines(500, 4))
    f = np.geomspace(0.5, 300.0, 300)
    ab = K.absorbers_struct(JUPITER_ALPHA)
    for prec, tol in (('fp64', 1e-12), ('fp32', 1e-6)):
        out = {n: K.alpha(cat, ab, atm.P, atm.T, K.mix_table(atm.gas, C), f, prec,
                          launch={'alpha_series': 2, 'far_split': n}).double().cpu().numpy() for n in (1, 3, 16)}
        for n in out:
            assert np.max(np.abs(out[n] - out[1]) / np.abs(out[1])) < tol, (prec, n)


def test_two_threads_with_different_launch_shapes(K, raw_lines):
    """Launch shapes are per call: two threads, each on its own stream, computing tables with
    different shapes (direct sums split 3 ways at 1 frequency per lane against the far-line
    series at 4 per lane, and different integration block orders) get the results each shape
    gives when run alone, call after call."""
    import threading

    import torch
    atm = synth.jupiter_like(33)
    cat = K.jupiter_catalog(raw_lines, h2o=synth.h2o_lines(400, 6))
    ab = K.absorbers_struct(JUPITER_ALPHA)
    f = np.geomspace(1.0, 200.0, 520)
    mix = K.mix_table(atm.gas, C)
    pl = K.planet_desc(atm.Req, atm.Rpol)
    hds, ns = K.raypaths(pl, atm.R, atm.N, synth.random_disc_rays(90, 21))
    shapes = [{'alpha_series': 1, 'alpha_split': 3, 'alpha_k': 1, 'int_map': 1, 'int_unroll': 4},
              {'alpha_series': 2, 'alpha_k': 4, 'far_split': 5, 'int_map': 2, 'int_unroll': 8}]

    def run(L):
        a = K.alpha(cat, ab, atm.P, atm.T, mix, f, 'fp64', launch=L)
        return a, K.integrate(a, atm.T, hds, ns, precision='fp64', tau_cut=40.0, launch=L)

    alone = []
    for L in shapes:
        a, tb = run(L)
        alone.append((a.cpu().numpy(), tb.cpu().numpy()))
    torch.cuda.synchronize()
    errors = []

    def worker(i):
        try:
            s = torch.cuda.Stream()
            with torch.cuda.stream(s):
                for _ in range(6):
                    a, tb = run(shapes[i])
                    s.synchronize()
                    if not (np.array_equal(a.cpu().numpy(), alone[i][0]) and
                            np.array_equal(tb.cpu().numpy(), alone[i][1], equal_nan=True)):
                        errors.append(i)
        except Exception as e:   # reported below: a thread's exception would be lost
            errors.append(repr(e))

    th = [threading.Thread(target=worker, args=(i,)) for i in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert errors == []
    assert np.max(np.abs(alone[0][0] - alone[1][0]) / np.abs(alone[1][0])) < 1e-11
