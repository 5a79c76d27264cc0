"""The forward + adjoint workloads at their full shapes, checked against the oracle.

Workload 4: one atmosphere of 500 layers to 100 bar, 10000 synthetic NH3 lines plus the
shipped lists, 64 frequencies 1-50 GHz, 256 rays b = 0..0.99 at angle 0.
Workload 5: 1024 perturbed atmospheres of 200-800 layers (T +-5 K, NH3 x[0.3, 3]), the same
lines, 128 frequencies 1-50 GHz, 16 disc rays each, shells rebuilt on the device.

Every kernel runs at the full shape; the oracle checks a sample it finishes in seconds: the
Jacobian rows of 4 rays at 8 frequencies over every layer (oracle.jacobian: forward-mode
tangents of the restated absorption chained with reverse-mode derivatives of the
restated integration, both float64 autograd), central differences of the FP64 forward on 5
layers, and Tb against the numpy oracle.  Rows: <= 1e-3 of each row's largest entry (the
task's gradient gate), as sensitivities with a 1e-6 K noise floor (rows_err); Tb: <= 0.01 K."""
import numpy as np
import pytest

from radiobear_amd import synth

pytestmark = pytest.mark.gpu
JUPITER_ALPHA = {'h2': 'h2_jj_ddb', 'h2o': 'h2o_bk', 'h2s': 'h2s_ddb', 'nh3': 'nh3_hs_sjs', 'ph3': 'ph3_jh'}


@pytest.fixture(scope='module')
def K():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no HIP device')
    from radiobear_amd import kernels
    return kernels


def oracle_lines(jupiter_lines, inv):
    lines = dict(jupiter_lines)
    lines['nh3_hs'] = dict(lines['nh3_hs'])
    for k in ('fo', 'Io', 'Eo', 'gammaNH3o'):
        lines['nh3_hs'][k] = inv[k]
    return lines


def oracle_rows(K, atm, lines, freqs, rays, cols):
    """Oracle Tb and J rows [ray][3][nl][col] of one atmosphere at the frequency columns cols."""
    from oracle import geometry, jacobian as OJ
    f = np.asarray(freqs)[cols]
    alpha, dal = OJ.alpha_tangents(f, atm.T, atm.P, K.mix_table(atm.gas, atm.C), lines)
    paths = [(r.ds, r.layer4ds) for r in (geometry.compute_ds(atm.R, atm.N, b, atm.Req, atm.Rpol) for b in rays)]
    Tb, Ja, JT = OJ.rt_jacobian(alpha, atm.T, paths)
    J = np.stack([JT + Ja * dal[0][None], Ja * dal[1][None], Ja * dal[2][None]], axis=1)
    return Tb, J


def rows_err(J, ref, x):
    """Largest difference over the rows J [ray][3][nl][col], relative to each row's largest
    entry, with both sides taken as sensitivities: J[ray][p][l] x[p][l], the change of Tb (K)
    per unit relative change of T_l, q_NH3,l or q_H2O,l (x: (3, nl) the layer values).  A row
    whose exact value vanishes (an isothermal layer stack above the emitting level gives
    dTb/dalpha = 0 exactly, so the q rows are rounding noise, ~1e-20 K here and ~1e-16 on the
    GPU) is compared against a floor of 1e-6 K per unit relative change instead."""
    sj, sr = J * x[None, :, :, None], ref * x[None, :, :, None]
    scale = np.maximum(np.max(np.abs(sr), axis=2, keepdims=True), 1e-6)
    return float(np.max(np.abs(sj - sr) / scale))


def layer_values(K, atm):
    mix = K.mix_table(atm.gas, atm.C)
    return np.stack([atm.T, mix[3], mix[4]])


@pytest.fixture(scope='module')
def c4(K, raw_lines):
    from radiobear_amd.retrieval import RetrievalBatch
    w = synth.workload('c4_jacobian')
    atm = w['atm']
    cat = K.jupiter_catalog(raw_lines, nh3_inversion=w['nh3_inv'])
    ab = K.absorbers_struct(JUPITER_ALPHA, 'e')
    batches = {p: RetrievalBatch([atm], cat, ab, w['freqs'], w['rays'], atm.Req, atm.Rpol, atm.C, precision=p)
               for p in ('fp32', 'fp64')}
    return w, batches


@pytest.mark.parametrize('prec', ['fp32', 'fp64'])
def test_c4_rows_vs_oracle_autograd(K, c4, jupiter_lines, prec):
    """4 rays (nadir to b = 0.99) x 8 frequencies x 500 layers x (T, q_NH3, q_H2O), the
    production tau cuts (20 FP32, 40 FP64) in force; Tb of every ray at those frequencies."""
    w, batches = c4
    rb = batches[prec]
    Tb, J = rb.run()
    Tb, J = Tb.cpu().numpy(), J.double().cpu().numpy()
    assert J.shape == (256, 3, 500, 64) and np.all(np.isfinite(J)) and np.all(np.isfinite(Tb))
    cols = np.arange(0, 64, 8)
    pick = [0, 85, 170, 255]
    Tr, Jr = oracle_rows(K, w['atm'], oracle_lines(jupiter_lines, w['nh3_inv']), w['freqs'], w['rays'][pick], cols)
    assert np.max(np.abs(Tb[pick][:, cols] - Tr)) < 0.01
    err = rows_err(J[pick][:, :, :, cols], Jr, layer_values(K, w['atm']))
    assert err < 1e-3, (prec, err)


def test_c4_rows_vs_central_differences_of_fp64_forward(K, c4):
    """dTb/dT on 5 layers at every ray and frequency against central differences
    (h = 1e-4 T) of the FP64 forward of the same batch (the FP64 Jacobian and the FP32 one)."""
    import torch
    w, batches = c4
    rb = batches['fp64']
    _, J64 = rb.run()
    _, J32 = batches['fp32'].run()
    J64, J32 = J64.double().cpu().numpy(), J32.double().cpu().numpy()
    T0 = rb.T.clone()
    for l in (20, 120, 240, 360, 480):
        h = 1e-4 * float(T0[l])
        tb = []
        for sgn in (1.0, -1.0):
            rb.T = T0.clone()
            rb.T[l] += sgn * h
            tb.append(rb.run()[0].cpu().numpy())
        rb.T = T0
        fd = (tb[0] - tb[1]) / (2 * h)                                  # (ray, f)
        scale = np.max(np.abs(J64[:, 0]), axis=1) + 1e-30                # row scale (ray, f)
        assert np.max(np.abs(J64[:, 0, l] - fd) / scale) < 1e-3, l
        assert np.max(np.abs(J32[:, 0, l] - fd) / scale) < 1e-3, l
    del torch


@pytest.fixture(scope='module')
def c5(K):
    from radiobear_amd import retrieval
    return {p: retrieval.synthetic(n_atm=1024, precision=p, device_props=True) for p in ('fp32', 'fp64')}


def test_c5_full_batch_finite_and_sampled_rows(K, c5, jupiter_lines):
    """All 16384 rows of the FP32 batch finite (the rays stay inside 0.95 of the disc); the
    shortest and the longest atmosphere's rows (2 rays x 8 frequencies x every layer) against
    the oracle, with the shells the device rebuilt."""
    import torch
    rb = c5['fp32']
    rb.update()
    Tb, J = rb.run()
    assert bool(torch.isfinite(Tb).all()) and bool(torch.isfinite(J).all())
    assert J.shape == (1024 * 16, 3, rb.ldj, 128)
    R, N = rb.R.cpu().numpy(), rb.N.cpu().numpy()
    nls = rb.nl
    lines = oracle_lines(jupiter_lines, rb.nh3_inv)
    cols = np.arange(0, 128, 16)
    for a in (int(np.argmin(nls)), int(np.argmax(nls))):
        rows, lay = rb.atmosphere_slice(a)
        atm = rb.atmospheres[a]
        atm.R, atm.N = R[lay], N[lay]                                      # the device-built shells
        rr = [rows.start, rows.start + 9]
        Tr, Jr = oracle_rows(K, atm, lines, rb.freqs, rb.b.cpu().numpy()[rr], cols)
        Tg = Tb[rr].cpu().numpy()[:, cols]
        Jg = J[rr].double().cpu().numpy()[:, :, :atm.nl][:, :, :, cols]
        assert np.max(np.abs(Tg - Tr)) < 0.01, a
        err = rows_err(Jg, Jr, layer_values(K, atm))
        assert err < 1e-3, (a, nls[a], err)


def test_c5_batch_equals_single_atmospheres_fp64(K, c5):
    """The FP64 batch of 1024 against 3 of its atmospheres run alone with the same launch
    shape and shells: Tb and J bit-identical."""
    from radiobear_amd.retrieval import RetrievalBatch
    rb = c5['fp64']
    rb.launch = {'tan_split': 1}
    rb.update()
    Tb, J = rb.run()
    for a in (0, 511, 1023):
        rows, lay = rb.atmosphere_slice(a)
        atm = rb.atmospheres[a]
        one = RetrievalBatch([atm], rb.catalog, rb.absorbers, rb.freqs, rb.b.cpu().numpy()[rows], atm.Req, atm.Rpol,
                             synth.GAS_ROWS, precision='fp64', gm=rb.gm, device_props=True, launch=rb.launch)
        T1, J1 = one.run()
        assert np.array_equal(Tb[rows].cpu().numpy(), T1.cpu().numpy()), a
        assert np.array_equal(J[rows][:, :, :atm.nl].cpu().numpy(), J1.cpu().numpy()), a
