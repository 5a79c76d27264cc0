"""The hand-written adjoint (rbx_integrate_adjoint + rbx_alpha_vjp behind torch.autograd.Function)
against float64 torch autograd of the restated forward model (oracle.jacobian.vjp), against
central differences of the FP64 forward, and against the contraction of the Jacobian rows.
Gradients: <= 1e-6 (FP64) and 1e-3 (FP32) of each gradient row's largest entry against
autograd; <= 1e-4 against central differences (h = 1e-4 x)."""
import numpy as np
import pytest

from oracle import geometry
from radiobear_amd import synth

pytestmark = pytest.mark.gpu
C = synth.GAS_ROWS
JUPITER_ALPHA = {'h2': 'h2_jj_ddb', 'h2o': 'h2o_bk', 'h2s': 'h2s_ddb', 'nh3': 'nh3_hs_sjs', 'ph3': 'ph3_jh'}
NH3, H2O = 3, 4     # rows of the mix table


@pytest.fixture(scope='module')
def K():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no HIP device')
    from radiobear_amd import kernels
    return kernels


@pytest.fixture(scope='module')
def case(K, raw_lines):
    """48 layers to 100 bar, 7 frequencies across the 30 GHz switch, 6 rays from nadir to
    the limb and one off the planet, random upstream gradients."""
    atm = synth.jupiter_like(48)
    cat = K.jupiter_catalog(raw_lines)
    ab = K.absorbers_struct(JUPITER_ALPHA, 'e')
    f = np.array([1.5, 8.0, 22.0, 29.0, 31.0, 40.0, 50.0])
    rays = np.array([[0.0, 0.0], [0.35, 0.1], [0.6, -0.3], [0.85, 0.2], [0.97, 0.0], [1.1, 0.0]])
    pl = K.planet_desc(atm.Req, atm.Rpol)
    hds, ns = K.raypaths(pl, atm.R, atm.N, rays)
    g = np.random.default_rng(17).normal(size=(len(rays), len(f)))
    return atm, cat, ab, f, rays, hds, ns, g


def model_of(K, case, prec, **kw):
    from radiobear_amd.adjoint import ForwardModel
    atm, cat, ab, f, rays, hds, ns, g = case
    return ForwardModel(cat, ab, atm.P, K.mix_table(atm.gas, C), f, hds, ns, precision=prec, **kw)


def grads(K, model, atm, g, T=None, mix=None):
    """Tb and dL/d(T, q_NH3, q_H2O) of L = sum g Tb through torch autograd."""
    import torch
    mix = K.mix_table(atm.gas, C) if mix is None else mix
    x = [torch.tensor(v, dtype=torch.float64, device='cuda', requires_grad=True)
         for v in (atm.T if T is None else T, mix[NH3], mix[H2O])]
    Tb = model(*x)
    (Tb * torch.as_tensor(g, device='cuda')).sum().backward()
    return Tb.detach().cpu().numpy(), np.stack([v.grad.cpu().numpy() for v in x])


def oracle_paths(atm, rays):
    return [(r.ds, r.layer4ds) for r in (geometry.compute_ds(atm.R, atm.N, b, atm.Req, atm.Rpol) for b in rays)]


def row_err(a, b):
    return float(np.max(np.abs(a - b) / (np.max(np.abs(b), axis=-1, keepdims=True) + 1e-300)))


@pytest.mark.parametrize('prec,tol_tb,tol', [('fp64', 1e-8, 1e-6), ('fp32', 1e-3, 1e-3)])
def test_vjp_vs_oracle_autograd(K, case, jupiter_lines, prec, tol_tb, tol):
    """The production optical-depth cuts (20 FP32, 40 FP64) are in force; the oracle has none."""
    from oracle import jacobian as OJ
    atm, cat, ab, f, rays, hds, ns, g = case
    Tb, gr = grads(K, model_of(K, case, prec), atm, g)
    Tr, ref = OJ.vjp(f, atm.T, atm.P, K.mix_table(atm.gas, C), jupiter_lines, oracle_paths(atm, rays), g)
    assert np.max(np.abs(Tb - Tr)) < tol_tb
    assert np.all(Tb[-1] == 2.725)                      # off the planet: T_cmb, no gradient
    for m in range(3):
        assert np.any(ref[m] != 0.0)
        assert row_err(gr[m], ref[m]) < tol, (prec, m, row_err(gr[m], ref[m]))


def test_vjp_equals_jacobian_rows_contracted(K, case):
    """sum_{ray,f} g dTb/d(T, q) from rbx_jacobian's rows (float32) = the adjoint's gradient."""
    atm, cat, ab, f, rays, hds, ns, g = case
    model = model_of(K, case, 'fp64')
    _, gr = grads(K, model, atm, g)
    alpha, dal = K.alpha_tangent(cat, ab, atm.P, atm.T, K.mix_table(atm.gas, C), f, with_alpha=True)
    J, _ = K.jacobian(alpha, dal, atm.T, hds, ns, tau_cut=model.tau_cut)
    contracted = np.einsum('rf,rplf->pl', g, J.double().cpu().numpy())
    for m in range(3):
        assert row_err(gr[m], contracted[m]) < 1e-5, m


def test_vjp_vs_central_differences(K, case):
    """dL/dT and dL/dq_NH3 on 5 layers, dL/dq_H2O on the 3 wettest, against central differences
    of the FP64 forward model (h = 1e-4 x)."""
    atm, cat, ab, f, rays, hds, ns, g = case
    model = model_of(K, case, 'fp64')
    _, gr = grads(K, model, atm, g)
    mix0 = K.mix_table(atm.gas, C)

    def loss(T, mix):
        import torch
        Tb = model(torch.as_tensor(T, device='cuda'), torch.as_tensor(mix[NH3], device='cuda'),
                   torch.as_tensor(mix[H2O], device='cuda'))
        return float(np.sum(Tb.cpu().numpy()[:-1] * g[:-1]))   # the off-planet ray is constant

    wet = [int(i) for i in np.argsort(mix0[H2O])[-3:]]
    checks = [(0, l) for l in (2, 9, 20, 33, 45)] + [(1, l) for l in (2, 9, 20, 33, 45)] + [(2, l) for l in wet]
    for m, l in checks:
        T, mix = atm.T.copy(), mix0.copy()
        if m == 0:
            h = 1e-4 * T[l]
            Tp, Tm = T.copy(), T.copy()
            Tp[l] += h
            Tm[l] -= h
            fd = (loss(Tp, mix) - loss(Tm, mix)) / (2 * h)
        else:
            row = NH3 if m == 1 else H2O
            h = 1e-4 * mix[row, l]
            mp, mm = mix.copy(), mix.copy()
            mp[row, l] += h
            mm[row, l] -= h
            fd = (loss(T, mp) - loss(T, mm)) / (2 * h)
        assert abs(gr[m, l] - fd) < 1e-4 * np.max(np.abs(gr[m])), (m, l, gr[m, l], fd)


def test_masked_rays_and_nan_paths(K, case):
    """A ray whose path holds a NaN half step has a NaN Tb: with gTb = 0 on its row it adds
    nothing (the gradient equals that of the other rays alone); with gTb != 0 the gradient is
    NaN, as autograd through the reference arithmetic would make it."""
    import torch
    from radiobear_amd.adjoint import ForwardModel
    atm, cat, ab, f, rays, hds, ns, g = case
    bad = hds.clone()
    bad[2, 5] = float('nan')
    mix = K.mix_table(atm.gas, C)
    keep = [0, 1, 3, 4, 5]
    g0 = g.copy()
    g0[2] = 0.0
    m_all = ForwardModel(cat, ab, atm.P, mix, f, bad, ns, precision='fp64')
    m_keep = ForwardModel(cat, ab, atm.P, mix, f, hds[keep].contiguous(), ns[keep].contiguous(), precision='fp64')
    Tb, gr = grads(K, m_all, atm, g0)
    assert np.all(np.isnan(Tb[2])) and not np.any(np.isnan(Tb[keep]))
    _, gk = grads(K, m_keep, atm, g0[keep])
    assert np.all(np.isfinite(gr))
    assert np.allclose(gr, gk, rtol=1e-12, atol=1e-300)
    _, gn = grads(K, m_all, atm, g)
    # every layer the NaN ray reaches (all but the deepest, which no path uses) is NaN
    assert np.all(np.isnan(gn[:2, :-1])) and np.all(gn[:, -1] == 0.0)
    del torch


@pytest.mark.parametrize('prec', ['fp64', 'fp32'])
def test_rays_per_wavefront_agree(K, case, prec):
    """1, 2, 4 or 8 rays summed per wavefront before the atomic adds, rays in path order or
    in input order, the adds spread over 1, 3 or 32 copies of the accumulators: the same
    gradient up to the order of FP64 additions."""
    import torch
    atm, cat, ab, f, rays, hds, ns, g = case
    out = []
    for nr in (1, 2, 4, 8):
        for order in (None, torch.arange(len(rays), dtype=torch.int32, device='cuda')):
            out.append(grads(K, model_of(K, case, prec, launch={'adj_rays': nr}, order=order), atm, g)[1])
    for copies in (1, 3, 32):                      # accumulator copies: summed at the end
        out.append(grads(K, model_of(K, case, prec, launch={'adj_rays': 1, 'adj_copies': copies}), atm, g)[1])
    for o in out[1:]:
        for m in range(3):
            assert row_err(o[m], out[0][m]) < 1e-12, m


def test_batch_adjoint_equals_single_atmospheres(K, raw_lines):
    """Three atmospheres of uneven layer counts in one batch: the batch gradient of
    sum g Tb, split by atmosphere, equals each atmosphere's own (FP64, up to the order of
    the atomic additions); 5 rays per atmosphere put rays of two atmospheres in one
    wavefront (the one-ray-at-a-time branch)."""
    import torch
    from radiobear_amd.adjoint import ForwardModel
    from radiobear_amd.retrieval import RetrievalBatch
    atms = synth.retrieval_batch(3, seed=21, nl_range=(30, 55))
    cat = K.jupiter_catalog(raw_lines)
    ab = K.absorbers_struct(JUPITER_ALPHA, 'e')
    f = np.linspace(2.0, 48.0, 70)
    rays = np.array([[0.0, 0.0], [0.3, 0.2], [0.6, 0.1], [0.9, -0.1], [1.05, 0.0]])
    a0 = atms[0]
    rb = RetrievalBatch(atms, cat, ab, f, rays, a0.Req, a0.Rpol, C, precision='fp64')
    g = np.random.default_rng(5).normal(size=(rb.nray, len(f)))
    model = rb.model()
    x = [rb.T.clone().requires_grad_(True), rb.mix[NH3].clone().requires_grad_(True),
         rb.mix[H2O].clone().requires_grad_(True)]
    Tb = model(*x)
    (Tb * torch.as_tensor(g, device='cuda')).sum().backward()
    gb = np.stack([v.grad.cpu().numpy() for v in x])
    pl = K.planet_desc(a0.Req, a0.Rpol)
    for a, atm in enumerate(atms):
        rows, lay = rb.atmosphere_slice(a)
        hds, ns = K.raypaths(pl, atm.R, atm.N, rays)
        m1 = ForwardModel(cat, ab, atm.P, K.mix_table(atm.gas, C), f, hds, ns, precision='fp64')
        Tb1, g1 = grads(K, m1, atm, g[rows])
        assert np.max(np.abs(Tb.detach().cpu().numpy()[rows] - Tb1)) < 1e-9
        for m in range(3):
            assert row_err(gb[m, lay], g1[m]) < 1e-12, (a, m)
