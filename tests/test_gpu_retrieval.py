"""Many atmospheres in one launch: the batched path against one atmosphere at a time,
the fused tangent pass against rbx_alpha, and the FP32 tangents against FP64."""
import ctypes

import numpy as np
import pytest

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
def small_batch(K, raw_lines):
    from radiobear_amd.retrieval import RetrievalBatch
    atms = synth.retrieval_batch(4, seed=7, nl_range=(25, 70))
    inv = synth.nh3_inversion_lines(700, 3)
    cat = K.jupiter_catalog(raw_lines, nh3_inversion=inv)
    ab = K.absorbers_struct(JUPITER_ALPHA, 'e')
    f = np.linspace(1.0, 50.0, 75)                     # straddles 30 GHz inside a 64-column tile
    rays = np.array([[0.0, 0.0], [0.4, 0.3], [0.8, -0.2], [0.97, 0.05], [1.2, 0.0]])
    a0 = atms[0]
    out = {}
    for prec in ('fp64', 'fp32'):
        # one launch shape for the batch and the single-atmosphere calls it is compared with
        out[prec] = RetrievalBatch(atms, cat, ab, f, rays, a0.Req, a0.Rpol, C, precision=prec,
                                   launch={'tan_split': 3})
    return atms, cat, ab, f, rays, out


@pytest.mark.parametrize('prec', ['fp64', 'fp32'])
def test_batch_equals_one_atmosphere_at_a_time(K, small_batch, prec):
    atms, cat, ab, f, rays, batches = small_batch
    rb = batches[prec]
    Tb, J = rb.run()
    Tb, J = Tb.cpu().numpy(), J.cpu().numpy()
    pl = K.planet_desc(atms[0].Req, atms[0].Rpol)
    for a, atm in enumerate(atms):
        rows, _ = rb.atmosphere_slice(a)
        mix = K.mix_table(atm.gas, C)
        alpha, dal = K.alpha_tangent(cat, ab, atm.P, atm.T, mix, f, precision=prec, with_alpha=True,
                                     launch=rb.launch)
        hds, ns = K.raypaths(pl, atm.R, atm.N, rays)
        J1, Tb1 = K.jacobian(alpha.float() if prec == 'fp32' else alpha, dal, atm.T, hds, ns,
                             tau_cut=rb.tau_cut)
        if prec == 'fp64':
            assert np.array_equal(Tb[rows], Tb1.cpu().numpy()), a
            assert np.array_equal(J[rows, :, :atm.nl], J1.cpu().numpy()), a
        else:   # the single-atmosphere call reads FP32 alpha; the batch reads the FP64 value
            assert np.max(np.abs(Tb[rows] - Tb1.cpu().numpy())) < 1e-3, a
            j1 = J1.cpu().numpy()
            scale = np.max(np.abs(j1), axis=2, keepdims=True) + 1e-30
            assert np.max(np.abs(J[rows, :, :atm.nl] - j1) / scale) < 1e-4, a
        assert not np.any(J[rows, :, atm.nl:]), a              # rows past the atmosphere
    assert np.all(Tb[4::5] == 2.725)                           # off-disc rays: T_cmb


def test_fused_alpha_value_and_fp32_tangents(K, small_batch):
    atms, cat, ab, f, rays, batches = small_batch
    atm = atms[1]
    mix = K.mix_table(atm.gas, C)
    ref = K.alpha(cat, ab, atm.P, atm.T, mix, f, 'fp64').cpu().numpy()
    a64, d64 = K.alpha_tangent(cat, ab, atm.P, atm.T, mix, f, precision='fp64', with_alpha=True)
    a32, d32 = K.alpha_tangent(cat, ab, atm.P, atm.T, mix, f, precision='fp32', with_alpha=True)
    assert np.max(np.abs(a64.cpu().numpy() - ref) / ref) < 1e-12
    assert np.max(np.abs(a32.cpu().numpy() - ref) / ref) < 1e-5
    d64, d32 = d64.cpu().numpy(), d32.cpu().numpy()
    for m in range(3):
        scale = np.max(np.abs(d64[m]), axis=1, keepdims=True) + 1e-300
        assert np.max(np.abs(d32[m] - d64[m]) / scale) < 1e-5, m


def deep_h2s_case(K, raw_lines):
    """Layers from 150 to 3000 bar with only h2s_ddb (Ben-Reuven with zeta = gamma): its lines'
    widths reach several times their centres (168 GHz and up), where the shifted denominator
    (x - Ds)^2 + h cancels at every column of 1-50 GHz (h < 0, Ds^2 + h << Ds^2)."""
    from radiobear_amd import _lib
    cat = K.jupiter_catalog(raw_lines)
    ab = K.absorbers_struct({'h2s': 'h2s_ddb'}, 'e')
    nl = 16
    P = np.geomspace(150.0, 3000.0, nl)
    T = np.linspace(450.0, 1500.0, nl)
    mix = np.zeros((len(_lib.SPECIES), nl))
    mix[_lib.SPECIES.index('H2')] = 0.86
    mix[_lib.SPECIES.index('HE')] = 0.135
    mix[_lib.SPECIES.index('H2S')] = 3.0e-5
    f = np.linspace(1.0, 50.0, 70)
    return cat, ab, P, T, mix, f


def test_fp32_tangent_denominator_guard(K, raw_lines):
    """FP32 tangent mode forms (x - Ds)^2 + h in FP32 only where its error stays within 32
    roundings; these lines cancel far beyond that (a million-fold amplification at the deepest
    layer), so the pass takes their denominators in FP64, and its value and tangents stay within
    1e-5 of FP64 mode (in FP32 the denominators would be wrong in their leading digits)."""
    cat, ab, P, T, mix, f = deep_h2s_case(K, raw_lines)
    a64, d64 = K.alpha_tangent(cat, ab, P, T, mix, f, precision='fp64', with_alpha=True)
    a32, d32 = K.alpha_tangent(cat, ab, P, T, mix, f, precision='fp32', with_alpha=True)
    a64, a32 = a64.cpu().numpy(), a32.cpu().numpy()
    assert np.all(a64 > 0.0)
    assert np.max(np.abs(a32 - a64) / a64) < 1e-5
    d64, d32 = d64.cpu().numpy(), d32.cpu().numpy()
    scale = np.max(np.abs(d64[0]), axis=1, keepdims=True)
    assert np.all(scale > 0.0)
    assert np.max(np.abs(d32[0] - d64[0]) / scale) < 1e-5


def test_bad_atmosphere_index_is_flagged_not_read(K, small_batch):
    """An out-of-range atmosphere index in the device ray map gives nsteps -2 and Tb NaN
    (the kernels check it before any load); the host wrapper refuses it up front."""
    import torch
    from radiobear_amd import _lib
    atms, cat, ab, f, rays, batches = small_batch
    rb = batches['fp64']
    with pytest.raises(ValueError):
        K.raypaths_batch(rb.planet, rb.R, rb.N, rb.layer_off, np.full(rb.nray, rb.natm), rb.b, rb.ld)
    bad = rb.ray_atm.copy()
    bad[3] = rb.natm + 5
    bad[7] = -1
    bad_d = K._i32(bad)
    ds = torch.empty((rb.nray, rb.ld), dtype=torch.float64, device='cuda')
    ns = torch.empty((rb.nray,), dtype=torch.int32, device='cuda')
    _lib.check(_lib.lib().rbx_raypaths_batch(ctypes.byref(rb.planet), K._ptr(rb.R), K._ptr(rb.N), rb.nl_total,
                                             K._ptr(rb.off_d), rb.natm, K._ptr(rb.b), rb.nray, K._ptr(bad_d), 1,
                                             K._ptr(ds), rb.ld, K._ptr(ns), K._stream()))
    nsn = ns.cpu().numpy()
    assert nsn[3] == -2 and nsn[7] == -2 and np.all(nsn[[0, 1, 2]] >= 0)
    alpha, dal = rb.absorption()
    Tb = K.integrate_batch(alpha, rb.T, rb.off_d, bad_d, ds, ns).cpu().numpy()
    assert np.all(np.isnan(Tb[[3, 7]])) and not np.any(np.isnan(Tb[[0, 1, 2]]))
    J, Tb2 = K.jacobian_batch(alpha, dal, rb.T, rb.off_d, bad_d, ds, ns, rb.ldj)
    assert np.all(np.isnan(Tb2.cpu().numpy()[[3, 7]]))
    assert not torch.any(J[[3, 7]]).item()


def test_retrieval_rows_vs_torch_autograd(K, raw_lines, jupiter_lines):
    """Two atmospheres of a batch against float64 autograd of the restated forward model."""
    from oracle import geometry, jacobian as OJ
    from radiobear_amd.retrieval import RetrievalBatch
    atms = synth.retrieval_batch(2, seed=11, nl_range=(30, 45))
    cat = K.jupiter_catalog(raw_lines)
    ab = K.absorbers_struct(JUPITER_ALPHA, 'e')
    f = np.array([2.0, 12.0, 23.0, 29.5, 31.0, 44.0])
    rays = np.array([[0.1, 0.0], [0.7, 0.4]])
    a0 = atms[0]
    for prec in ('fp64', 'fp32'):
        rb = RetrievalBatch(atms, cat, ab, f, rays, a0.Req, a0.Rpol, C, precision=prec)
        Tb, J = rb.run()
        Tb, J = Tb.cpu().numpy(), J.double().cpu().numpy()
        for a, atm in enumerate(atms):
            rows, _ = rb.atmosphere_slice(a)
            paths = []
            for b in rays:
                r = geometry.compute_ds(atm.R, atm.N, b, atm.Req, atm.Rpol)
                paths.append((r.ds, r.layer4ds))
            Tr, Jr = OJ.jacobian(f, atm.T, atm.P, K.mix_table(atm.gas, C), jupiter_lines, paths, planes=3)
            assert np.max(np.abs(Tb[rows] - Tr)) < (1e-8 if prec == 'fp64' else 1e-3)
            for slot in (0, 1, 2):
                scale = np.max(np.abs(Jr[:, slot]), axis=1, keepdims=True) + 1e-30
                err = np.max(np.abs(J[rows, slot, :atm.nl] - Jr[:, slot]) / scale)
                assert err < 1e-3, (prec, a, slot, err)


def test_device_atmosphere_matches_host_construction(K):
    """Hydrostatic z, R and N on the device vs the host construction of the same
    perturbed atmospheres (synth.jupiter_like's integration), to rounding."""
    atms = synth.retrieval_batch(5, seed=3, nl_range=(40, 90))
    off = np.concatenate([[0], np.cumsum([a.nl for a in atms])])
    cat = lambda f: np.concatenate([f(a) for a in atms])
    mix = np.concatenate([K.mix_table(a.gas, C) for a in atms], axis=1)
    Z, R, N = K.atmosphere_batch(cat(lambda a: a.P), cat(lambda a: a.T), mix, off, synth.JUPITER_REQ,
                                 synth.JUPITER_GM)
    Z, R, N = Z.cpu().numpy(), R.cpu().numpy(), N.cpu().numpy()
    assert np.max(np.abs(Z - cat(lambda a: a.gas[C['Z']]))) < 1e-9               # km
    assert np.max(np.abs(R - cat(lambda a: a.R)) / cat(lambda a: a.R)) < 1e-14
    assert np.max(np.abs(N - cat(lambda a: a.N))) < 1e-15


def test_device_atmosphere_given_z_matches_oracle_compute_prop(K, bench_gas_text):
    """With z from the benchmark atmosphere file, R and N equal the oracle's computeProp
    (itself pinned to the reference's benchmark fixture)."""
    from oracle import pipeline
    from oracle import atmosphere as OA
    gas, prop = pipeline.benchmark_atmosphere(bench_gas_text)
    mix = K.mix_table(gas, C)
    P, T = gas[C['P']], gas[C['T']]
    Z, R, N = K.atmosphere_batch(P, T, mix, [0, gas.shape[1]], pipeline.JUPITER_REQ, synth.JUPITER_GM,
                                 Z=gas[C['Z']])
    ref = OA.compute_prop(gas, C, pipeline.JUPITER_REQ)
    assert np.max(np.abs(R.cpu().numpy() - ref['R']) / ref['R']) < 1e-14
    assert np.max(np.abs(N.cpu().numpy() - ref['N'])) < 1e-15
    assert np.max(np.abs(R.cpu().numpy() - prop['R']) / prop['R']) < 1e-14


def test_device_property_table_matches_oracle(K, bench_gas_text, golden):
    """rbx_atmosphere_props on the benchmark atmosphere (its own z) against the oracle's
    computeProp (pinned to the reference-written table), and on a batch of perturbed
    atmospheres with the hydrostatic z built on the device, each against the oracle."""
    from oracle import pipeline
    from oracle import atmosphere as OA
    F = golden('props.npz')
    req, gm, pref = float(F['Req']), float(F['GM_ref']), float(F['p_ref'])
    gas, _ = pipeline.benchmark_atmosphere(bench_gas_text)
    got = K.atmosphere_props(gas[C['P']], gas[C['T']], K.mix_table(gas, C), gas[C['Z']], [0, gas.shape[1]],
                             req, gm, pref)
    ref = OA.compute_prop(gas, C, req, pref, GM_ref=gm)
    for r in OA.PROP_ROWS:
        g, e = got[r].cpu().numpy(), ref[r]
        assert np.max(np.abs(g - e) / np.maximum(np.abs(e), 1e-300)) < 1e-13, r
        assert np.max(np.abs(g - F['prop'][OA.PROP_ROWS.index(r)]) / np.maximum(np.abs(e), 1e-300)) < 1e-13, r
    atms = synth.retrieval_batch(7, seed=5, nl_range=(30, 120))
    off = np.concatenate([[0], np.cumsum([a.nl for a in atms])])
    cat = lambda f: np.concatenate([f(a) for a in atms])
    mix = np.concatenate([K.mix_table(a.gas, C) for a in atms], axis=1)
    P, T = cat(lambda a: a.P), cat(lambda a: a.T)
    Z, _, _ = K.atmosphere_batch(P, T, mix, off, req, gm)
    got = K.atmosphere_props(P, T, mix, Z, off, req, gm)
    Zh = Z.cpu().numpy()
    for a, atm in enumerate(atms):
        sl = slice(off[a], off[a + 1])
        g2 = atm.gas.copy()
        g2[C['Z']] = Zh[sl]
        ref = OA.compute_prop(g2, C, req, 1.0, GM_ref=gm)
        for r in OA.PROP_ROWS:
            e = ref[r]
            d = got[r].cpu().numpy()[sl]
            assert np.max(np.abs(d - e) / np.maximum(np.abs(e), 1e-300)) < 1e-13, (a, r)


def test_retrieval_with_device_atmosphere(K, small_batch):
    """The batch rebuilt on the device (hydrostatic z) integrates to the same Tb and J as
    with the host-built shells."""
    import torch
    from radiobear_amd.retrieval import RetrievalBatch
    atms, cat, ab, f, rays, batches = small_batch
    a0 = atms[0]
    rb = RetrievalBatch(atms, cat, ab, f, rays, a0.Req, a0.Rpol, C, precision='fp64', gm=synth.JUPITER_GM,
                        device_props=True, launch=batches['fp64'].launch)
    Tb, J = rb.run()
    Tb0, J0 = batches['fp64'].run()
    assert torch.max(torch.abs(Tb - Tb0)).item() < 1e-9
    assert torch.max(torch.abs(J - J0)).item() <= 1e-6 * torch.max(torch.abs(J0)).item()


@pytest.mark.parametrize('prec,tol', [('fp64', 1e-12), ('fp32', 1e-6)])
def test_tangent_every_split(K, small_batch, prec, tol):
    """The tangent pass with each layer's lines in 1, 2, 5 and 16 parts (partial sums in a
    workspace, then a finish pass): the same alpha and tangents up to the order of the
    parts' additions (and, in FP32, the FP32 blocks of lines each part sums)."""
    atms, cat, ab, f, rays, batches = small_batch
    rb = batches[prec]
    out = {}
    for n in (1, 2, 5, 16):
        L = {'tan_split': n}
        a, d = K.alpha_tangent(cat, ab, rb.P, rb.T, rb.mix, f, precision=prec, with_alpha=True, launch=L)
        out[n] = (a.cpu().numpy(), d.cpu().numpy())
    a1, d1 = out[1]
    for n, (a, d) in out.items():
        assert np.max(np.abs(a - a1) / np.abs(a1)) < tol, n
        for m in range(3):
            scale = np.max(np.abs(d1[m]), axis=1, keepdims=True) + 1e-300
            assert np.max(np.abs(d[m] - d1[m]) / scale) < tol, (n, m)
