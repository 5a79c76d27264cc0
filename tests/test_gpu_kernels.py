"""GPU parity of the HIP kernels against the FP64 oracle (through the C ABI)."""
import numpy as np
import pytest

from oracle import absorbers as A
from oracle import geometry, layers, pipeline, rt
from radiobear_amd import synth

pytestmark = pytest.mark.gpu
C = synth.GAS_ROWS
FREQS = np.array([0.3, 0.6, 1.0, 2.6, 5.2, 10.0, 21.9, 22.235, 25.0, 26.0, 27.5, 30.0,
                  30.5, 33.0, 34.0, 40.0, 60.0, 96.0, 150.0, 183.31, 300.0])


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


@pytest.fixture(scope='module')
def bench_atm(bench_gas_text):
    return pipeline.benchmark_atmosphere(bench_gas_text)


@pytest.fixture(scope='module')
def jcat(K, raw_lines):
    return K.jupiter_catalog(raw_lines)


JUPITER_ALPHA = {'h2': 'h2_jj_ddb', 'h2o': 'h2o_bk', 'h2s': 'h2s_ddb', 'nh3': 'nh3_hs_sjs', 'ph3': 'ph3_jh'}


def oracle_alpha(gas, freqs, jupiter_lines, return_saved=False):
    aset = layers.jupiter_absorbers(jupiter_lines)
    return layers.get_layers(list(freqs), gas, C, aset, return_saved=return_saved)


@pytest.mark.parametrize('precision,tol', [('fp64', 1e-9), ('fp32', 1e-5)])
def test_alpha_benchmark_layers(K, jcat, bench_atm, jupiter_lines, precision, tol):
    """Benchmark atmosphere (0.01-1e4 bar: nh3_hs, blend and nh3_sjs regimes), 21 freqs
    straddling the 30 GHz and 26/34 GHz switches."""
    gas, _ = bench_atm
    sel = np.unique(np.concatenate([np.arange(0, 1000, 23), [561, 600, 650, 700, 760, 999]]))
    g = gas[:, sel]
    ref, saved = oracle_alpha(g, FREQS, jupiter_lines, return_saved=True)
    ab = K.absorbers_struct(JUPITER_ALPHA, 'e')
    out, byc = K.alpha(jcat, ab, g[C['P']], g[C['T']], K.mix_table(g, C), FREQS, precision, by_constituent=True)
    got = out.double().cpu().numpy().T
    assert rel(got, ref) < tol
    # per-constituent tables match the reference absorb columns (sorted h2, h2o, h2s, nh3, ph3)
    b = byc.cpu().numpy()            # (7, nl, nf): RBX_C_* rows, the first five these
    for c in range(5):
        r = saved[:, :, c]           # (nl, nf)
        m = np.abs(r) > 1e-30
        if m.any():
            assert rel(b[c][m], r[m]) < max(tol, 1e-9)


def test_alpha_unsorted_frequencies(K, jcat, bench_atm, jupiter_lines):
    gas, _ = bench_atm
    g = gas[:, [10, 300, 500]]
    f = FREQS[[20, 3, 12, 0, 9, 15, 6]]
    ref = oracle_alpha(g, f, jupiter_lines)
    out = K.alpha(jcat, K.absorbers_struct(JUPITER_ALPHA), g[C['P']], g[C['T']], K.mix_table(g, C), f, 'fp64')
    assert rel(out.cpu().numpy().T, ref) < 1e-9


@pytest.mark.parametrize('precision,tol', [('fp64', 1e-9), ('fp32', 1e-5)])
def test_alpha_synthetic_c1(K, raw_lines, jupiter_lines, precision, tol):
    """Workload 1: 150 layers, 2000 synthetic NH3 inversion lines, 20 freqs, one nadir ray."""
    w = synth.workload('c1_plain_fp64')
    atm = w['atm']
    cat = K.jupiter_catalog(raw_lines, nh3_inversion=w['nh3_inv'])
    lines = dict(jupiter_lines)
    inj = dict(lines['nh3_hs'])
    for k in ('fo', 'Io', 'Eo', 'gammaNH3o'):
        inj[k] = w['nh3_inv'][k]
    lines['nh3_hs'] = inj
    ref = oracle_alpha(atm.gas, w['freqs'], lines)
    out = K.alpha(cat, K.absorbers_struct(JUPITER_ALPHA), atm.P, atm.T, K.mix_table(atm.gas, C), w['freqs'],
                  precision)
    assert rel(out.double().cpu().numpy().T, ref) < tol
    # its nadir ray: Tb against the oracle's geometry and integration over the oracle's table
    # (FP64 with the FP64 cut at 40; FP32 with its cut at 20, the 0.01 K bar)
    pl = K.planet_desc(atm.Req, atm.Rpol)
    hds, ns = K.raypaths(pl, atm.R, atm.N, w['rays'])
    cut = K.FP64_TAU_CUT if precision == 'fp64' else K.FP32_TAU_CUT
    Tb = K.integrate(out, atm.T, hds, ns, precision=precision, tau_cut=cut).cpu().numpy()
    for i, b in enumerate(w['rays']):
        r = geometry.compute_ds(atm.R, atm.N, b, atm.Req, atm.Rpol)
        tb = np.array(rt.single(ref, atm.T, r.ds, r.layer4ds))
        assert np.max(np.abs(Tb[i] - tb)) < (1e-6 if precision == 'fp64' else 0.01), i


@pytest.mark.parametrize('precision,tol', [('fp64', 1e-9), ('fp32', 1e-5)])
def test_alpha_synthetic_h2o_lines(K, raw_lines, jupiter_lines, precision, tol):
    """Synthetic H2O list (h2o_bk schema) over 0.3-300 GHz, mixed-sign VVW terms."""
    atm = synth.jupiter_like(40)
    h2o = synth.h2o_lines(600, 77)
    cat = K.jupiter_catalog(raw_lines, h2o=h2o)
    lines = dict(jupiter_lines)
    lines['h2o'] = h2o
    f = np.logspace(np.log10(0.3), np.log10(300.0), 300)
    ref = oracle_alpha(atm.gas, f, lines)
    out = K.alpha(cat, K.absorbers_struct(JUPITER_ALPHA), atm.P, atm.T, K.mix_table(atm.gas, C), f, precision)
    assert rel(out.double().cpu().numpy().T, ref) < tol


def test_alpha_scale_by_constituent(K, jcat, bench_atm, jupiter_lines):
    gas, _ = bench_atm
    g = gas[:, ::50]
    nl = g.shape[1]
    aset = layers.jupiter_absorbers(jupiter_lines)
    scale = {'nh3': list(np.linspace(0.5, 2.0, nl)), 'h2o': list(np.linspace(3.0, 1.0, nl))}
    ref = layers.get_layers(list(FREQS), g, C, aset, scale=scale)
    from radiobear_amd import _lib
    sc = np.ones((len(_lib.CONSTITUENTS), nl))
    sc[3] = scale['nh3']
    sc[1] = scale['h2o']
    out = K.alpha(jcat, K.absorbers_struct(JUPITER_ALPHA), g[C['P']], g[C['T']], K.mix_table(g, C), FREQS,
                  'fp64', scale=sc)
    assert rel(out.cpu().numpy().T, ref) < 1e-9


CASES = {'ell': ('ellipse', (0.0, 0.0), 'shape'), 'tilt': ('ellipse', (20.0, 30.0), 'shape'),
         'sph': ('sphere', (0.0, 0.0), 'shape'), 'sec': ('ellipse', (0.0, 0.0), 'sec')}


@pytest.mark.parametrize('tag', sorted(CASES))
def test_raypaths_fixture(K, golden, tag):
    G = golden('geometry.npz')
    gt, orient, limb = CASES[tag]
    pl = K.planet_desc(float(G['Req']), float(G['Rpol']), orient, gt, limb)
    ds, ns = K.raypaths(pl, G['R'], G['N'], G['rays'], units='km')
    ds = ds.cpu().numpy()
    ns = ns.cpu().numpy()
    for i in range(len(G['rays'])):
        n = int(G['nds_' + tag][i])
        assert ns[i] == n, (tag, i, ns[i], n)
        if n > 0:
            ref = G['ds_' + tag][i, :n]
            got = ds[i, :n]
            assert np.array_equal(np.isnan(got), np.isnan(ref)), (tag, i)
            ok = ~np.isnan(ref)
            assert rel(got[ok], ref[ok]) < 1e-9, (tag, i)


@pytest.mark.parametrize('tag', sorted(CASES))
def test_integrate_fixture(K, golden, tag):
    G = golden('geometry.npz')
    gt, orient, limb = CASES[tag]
    pl = K.planet_desc(float(G['Req']), float(G['Rpol']), orient, gt, limb)
    hds, ns = K.raypaths(pl, G['R'], G['N'], G['rays'])
    import torch
    tab = torch.as_tensor(np.ascontiguousarray(G['alpha_layers'].T), device='cuda')
    Tb = K.integrate(tab, G['gas'][C['T']], hds, ns, precision='fp64').cpu().numpy()
    ref = G['Tb_' + tag]
    assert np.array_equal(np.isnan(Tb), np.isnan(ref))
    ok = ~np.isnan(ref)
    assert np.max(np.abs(Tb[ok] - ref[ok])) < 1e-8
    Tb32 = K.integrate(tab.float(), G['gas'][C['T']], hds, ns, precision='fp32',
                       tau_cut=K.FP32_TAU_CUT).cpu().numpy()
    assert np.max(np.abs(Tb32[ok] - ref[ok])) < 1e-3


def test_integrate_disc_fixture(K, golden):
    G = golden('geometry.npz')
    pl = K.planet_desc(float(G['Req']), float(G['Rpol']))
    hds, ns = K.raypaths(pl, G['R'], G['N'], [[0.0, 0.0]])
    import torch
    tab = torch.as_tensor(np.ascontiguousarray(G['alpha_layers'].T), device='cuda')
    Tb = K.integrate(tab, G['gas'][C['T']], hds, ns, disc=True).cpu().numpy()[0]
    assert np.max(np.abs(Tb - G['Tb_disc'])) < 1e-8


T_RB = np.array([[805.38385, 424.58337, 298.74915, 227.0338, 179.9148, 147.3627],
                 [790.66736, 420.24728, 296.62033, 225.69223, 179.22531, 147.24104],
                 [745.55975, 407.28714, 290.07492, 221.56863, 177.13437, 146.8625],
                 [668.4528, 385.59015, 278.51672, 214.30847, 173.56506, 146.17966]])


@pytest.mark.parametrize('precision', ['fp64', 'fp32'])
def test_benchmark_T_rb_on_gpu(K, jcat, bench_atm, golden, precision):
    """The reference's golden vector (scripts/benchmark.py:20-24) through the HIP path."""
    gas, prop = bench_atm
    ab = K.absorbers_struct(JUPITER_ALPHA, 'e')
    tab = K.alpha(jcat, ab, gas[C['P']], gas[C['T']], K.mix_table(gas, C), pipeline.BENCH_FREQS, precision)
    pl = K.planet_desc(pipeline.JUPITER_REQ, pipeline.JUPITER_RPOL)
    hds, ns = K.raypaths(pl, prop['R'], prop['N'], pipeline.BENCH_B)
    Tb = K.integrate(tab, gas[C['T']], hds, ns, precision=precision,
                     tau_cut=K.FP32_TAU_CUT if precision == 'fp32' else K.FP64_TAU_CUT).cpu().numpy()
    assert np.max(np.abs(Tb - T_RB)) < 0.01
    ref = golden('benchmark.npz')['Tb']
    assert np.max(np.abs(Tb - ref)) < (1e-6 if precision == 'fp64' else 2e-3)


def test_alpha_tangent_vs_central_differences(K, jcat, bench_atm):
    """dalpha/dT, dalpha/dq_NH3, dalpha/dq_H2O from dual numbers vs central FD (h = 1e-4 x)."""
    gas, _ = bench_atm
    g = gas[:, [5, 150, 300, 450, 520, 600, 700, 800]]
    ab = K.absorbers_struct(JUPITER_ALPHA, 'e')
    f = FREQS
    P, T, mix = g[C['P']], g[C['T']], K.mix_table(g, C)
    d = K.alpha_tangent(jcat, ab, P, T, mix, f).cpu().numpy()
    hT = 1e-4 * T
    ap = K.alpha(jcat, ab, P, T + hT, mix, f, 'fp64').cpu().numpy()
    am = K.alpha(jcat, ab, P, T - hT, mix, f, 'fp64').cpu().numpy()
    fdT = (ap - am) / (2 * hT[:, None])
    assert rel(d[0], fdT) < 1e-5
    for slot, row in ((1, 3), (2, 4)):
        h = 1e-4 * mix[row]
        mp = mix.copy()
        mm = mix.copy()
        mp[row] += h
        mm[row] -= h
        ap = K.alpha(jcat, ab, P, T, mp, f, 'fp64').cpu().numpy()
        am = K.alpha(jcat, ab, P, T, mm, f, 'fp64').cpu().numpy()
        fd = (ap - am) / (2 * h[:, None])
        scale = np.max(np.abs(fd), axis=1, keepdims=True) + 1e-300
        assert np.max(np.abs(d[slot] - fd) / scale) < 1e-5


def test_jacobian_vs_central_differences(K, raw_lines):
    """dTb/dT and dTb/dq_NH3 rows vs central FD of the FP64 forward (h = 1e-4 x)."""
    import torch
    atm = synth.jupiter_like(60)
    cat = K.jupiter_catalog(raw_lines)
    ab = K.absorbers_struct(JUPITER_ALPHA, 'e')
    f = np.linspace(1.0, 50.0, 7)
    rays = np.array([[0.0, 0.0], [0.5, 0.0], [0.9, 0.1]])
    pl = K.planet_desc(atm.Req, atm.Rpol)
    hds, ns = K.raypaths(pl, atm.R, atm.N, rays)
    P, T, mix = atm.P, atm.T, K.mix_table(atm.gas, C)

    def fwd(T_, mix_):
        tab = K.alpha(cat, ab, P, T_, mix_, f, 'fp64')
        return K.integrate(tab, T_, hds, ns, precision='fp64').cpu().numpy()

    tab = K.alpha(cat, ab, P, T, mix, f, 'fp64')
    dal = K.alpha_tangent(cat, ab, P, T, mix, f)
    J, Tb = K.jacobian(tab, dal, T, hds, ns)
    J = J.cpu().numpy()
    assert np.max(np.abs(Tb.cpu().numpy() - fwd(T, mix))) < 1e-9
    layers_to_check = [1, 5, 12, 20, 30, 40, 50]
    for l in layers_to_check:
        h = 1e-4 * T[l]
        Tp = T.copy()
        Tm = T.copy()
        Tp[l] += h
        Tm[l] -= h
        fdT = (fwd(Tp, mix) - fwd(Tm, mix)) / (2 * h)
        scale = np.max(np.abs(J[:, 0]), axis=1) + 1e-30       # (nray, nf) row scale
        assert np.max(np.abs(J[:, 0, l, :] - fdT) / scale) < 1e-3, l
        hq = 1e-4 * mix[3, l]
        mp = mix.copy()
        mm = mix.copy()
        mp[3, l] += hq
        mm[3, l] -= hq
        fdq = (fwd(T, mp) - fwd(T, mm)) / (2 * hq)
        scale = np.max(np.abs(J[:, 1]), axis=1) + 1e-30
        assert np.max(np.abs(J[:, 1, l, :] - fdq) / scale) < 1e-3, l
    del torch


@pytest.mark.parametrize('alpha_precision', ['fp64', 'fp32'])
def test_jacobian_vs_torch_autograd(K, raw_lines, jupiter_lines, alpha_precision):
    """GPU Jacobian rows vs float64 torch autograd of the restated forward model
    (oracle.jacobian), with the oracle's own ray paths: <= 1e-3 of each row's scale."""
    from oracle import jacobian as OJ
    atm = synth.jupiter_like(60)
    cat = K.jupiter_catalog(raw_lines)
    ab = K.absorbers_struct(JUPITER_ALPHA, 'e')
    f = np.array([1.5, 8.0, 22.0, 29.0, 31.0, 40.0, 50.0])
    rays = np.array([[0.0, 0.0], [0.55, 0.2], [0.93, -0.1]])
    mix = K.mix_table(atm.gas, C)
    paths = []
    for b in rays:
        r = geometry.compute_ds(atm.R, atm.N, b, atm.Req, atm.Rpol)
        paths.append((r.ds, r.layer4ds))
    Tb_ref, J_ref = OJ.jacobian(f, atm.T, atm.P, mix, jupiter_lines, paths, planes=3)
    pl = K.planet_desc(atm.Req, atm.Rpol)
    hds, ns = K.raypaths(pl, atm.R, atm.N, rays)
    tab = K.alpha(cat, ab, atm.P, atm.T, mix, f, alpha_precision)
    dal = K.alpha_tangent(cat, ab, atm.P, atm.T, mix, f)
    J, Tb = K.jacobian(tab, dal, atm.T, hds, ns)
    J = J.double().cpu().numpy()
    assert np.max(np.abs(Tb.cpu().numpy() - Tb_ref)) < (1e-8 if alpha_precision == 'fp64' else 1e-3)
    for slot in (0, 1, 2):
        scale = np.max(np.abs(J_ref[:, slot]), axis=1, keepdims=True) + 1e-30    # (nray, 1, nf)
        err = np.max(np.abs(J[:, slot] - J_ref[:, slot]) / scale)
        assert err < 1e-3, (slot, err)


# the launch shape bench.py's autotuner picks on the large workload (both precisions)
BENCH_LAUNCH = {'alpha_series': 2, 'alpha_k': 2, 'far_split': 4, 'int_unroll': 8, 'int_map': 2}


@pytest.mark.parametrize('name', ['c2_image', 'c3_large'])
def test_workload_full_shape_sampled(K, raw_lines, jupiter_lines, name):
    """The image and large workloads at their full shapes (all layers, lines, freqs, rays on
    the GPU) with the launch shapes the bench's autotuner picks for the large workload and the
    production optical-depth cuts (FP32 20, FP64 40), checked on a sample the oracle finishes
    quickly: 3 layers of alpha at every frequency (both precisions, from the FP32 and FP64
    tables the bench times), and 4 rays integrated over those tables with the oracle's
    geometry, in request order and in cut order (the bench's); the two orders within the cut's
    own bound on every ray.  Also the built-in launch shapes, FP32 table within 1e-6 of the
    tuned one."""
    import torch
    w = synth.workload(name)
    atm = w['atm']
    cat = K.jupiter_catalog(raw_lines, nh3_inversion=w['nh3_inv'], h2o=w['h2o'])
    lines = dict(jupiter_lines)
    inj = dict(lines['nh3_hs'])
    for k in ('fo', 'Io', 'Eo', 'gammaNH3o'):
        inj[k] = w['nh3_inv'][k]
    lines['nh3_hs'] = inj
    if w['h2o'] is not None:
        lines['h2o'] = w['h2o']
    ab = K.absorbers_struct(JUPITER_ALPHA, 'e')
    mix = K.mix_table(atm.gas, C)
    f = w['freqs']
    L = BENCH_LAUNCH
    a64 = K.alpha(cat, ab, atm.P, atm.T, mix, f, 'fp64', launch=L)
    a32 = K.alpha(cat, ab, atm.P, atm.T, mix, f, 'fp32', launch=L)
    built_in = K.alpha(cat, ab, atm.P, atm.T, mix, f, 'fp32').double().cpu().numpy()
    assert rel(built_in, a32.double().cpu().numpy()) < 1e-6
    sel = [0, atm.nl // 2, atm.nl - 1]
    ref = oracle_alpha(atm.gas[:, sel], f, lines).T                      # (3, nf)
    assert rel(a64.cpu().numpy()[sel], ref) < 1e-9
    assert rel(a32.double().cpu().numpy()[sel], ref) < 1e-5
    rays = w['rays']
    pl = K.planet_desc(atm.Req, atm.Rpol)
    hds, ns = K.raypaths(pl, atm.R, atm.N, rays)
    Tb = K.integrate(a64, atm.T, hds, ns, precision='fp64', tau_cut=K.FP64_TAU_CUT, launch=L).cpu().numpy()
    Tb32 = K.integrate(a32, atm.T, hds, ns, precision='fp32', tau_cut=K.FP32_TAU_CUT, launch=L).cpu().numpy()
    # the bench integrates with the frequencies in the order they pass the cut
    Tbo = K.integrate(a64, atm.T, hds, ns, precision='fp64', tau_cut=K.FP64_TAU_CUT, launch=L,
                      ordered=True).cpu().numpy()
    Tb32o = K.integrate(a32, atm.T, hds, ns, precision='fp32', tau_cut=K.FP32_TAU_CUT, launch=L,
                        ordered=True).cpu().numpy()
    tab = a64.cpu().numpy().T                                            # (nf, nl)
    nsn = ns.cpu().numpy()
    on = np.flatnonzero(nsn > 0)
    Tmax = float(np.max(atm.T))
    for got, want, cut in ((Tbo, Tb, K.FP64_TAU_CUT), (Tb32o, Tb32, K.FP32_TAU_CUT)):
        fin = np.isfinite(want[on])
        assert np.array_equal(fin, np.isfinite(got[on]))
        assert np.max(np.abs(got[on][fin] - want[on][fin])) < 1e-9 + 2.0 * Tmax * np.exp(-cut)
    picks = [on[0], on[len(on) // 3], on[-1], int(np.flatnonzero(nsn < 0)[0])]
    for i in picks:
        r = geometry.compute_ds(atm.R, atm.N, rays[i], atm.Req, atm.Rpol)
        tb = np.array(rt.single(tab, atm.T, r.ds, r.layer4ds))
        assert np.max(np.abs(Tb[i] - tb)) < 1e-6, i
        assert np.max(np.abs(Tb32[i] - tb)) < 0.01, i
        assert np.max(np.abs(Tbo[i] - tb)) < 1e-6, i
        assert np.max(np.abs(Tb32o[i] - tb)) < 0.01, i
    del torch


def test_alpha_rescale_reuses_saved_constituents(K, jcat, bench_atm):
    """Saved per-constituent absorption (Alpha.alpha_data layout, nl x nf x nc) rescaled on
    the device equals rbx_alpha run with the same per-constituent scales, and numpy."""
    gas, _ = bench_atm
    g = gas[:, [3, 200, 400, 520]]
    ab = K.absorbers_struct(JUPITER_ALPHA, 'e')
    P, T, mix = g[C['P']], g[C['T']], K.mix_table(g, C)
    rng = np.random.default_rng(5)
    from radiobear_amd import _lib
    scale = rng.uniform(0.3, 3.0, size=(len(_lib.CONSTITUENTS), g.shape[1]))
    ref = K.alpha(jcat, ab, P, T, mix, FREQS, 'fp64', scale=scale).cpu().numpy()
    _, byc = K.alpha(jcat, ab, P, T, mix, FREQS, 'fp64', by_constituent=True)
    absorb = np.transpose(byc.cpu().numpy(), (1, 2, 0))          # (nl, nf, nc), sorted constituents
    got = K.alpha_rescale(absorb, scale, 'fp64').cpu().numpy()
    assert rel(got, ref) < 1e-14
    assert rel(got, np.einsum('lfc,cl->lf', absorb, scale)) < 1e-14
    got32 = K.alpha_rescale(absorb, None, 'fp32').double().cpu().numpy()
    assert rel(got32, absorb.sum(axis=2)) < 1e-7
