"""The optical-depth cut of the integrators keeps the reference's NaN results: a ray whose
path holds a NaN half step (raypath.py:190-209 gives one to a ray that misses a shell) must
end in a NaN Tb, as the uncut recurrence (brightness.py:93-101) does, even when the NaN lies
past the depth where every frequency is opaque."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def K():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no HIP device')
    from radiobear_amd import kernels
    return kernels


def paths(nl, nray, bad_ray, bad_step):
    import torch
    hds = np.full((nray, nl), 5.0e5)               # half steps (cm): tau grows ~1 per step
    hds[bad_ray, bad_step] = np.nan
    ns = np.full(nray, nl, dtype=np.int32)
    return (torch.as_tensor(hds, device='cuda'), torch.as_tensor(ns, device='cuda'))


@pytest.mark.parametrize('precision,cut', [('fp64', 40.0), ('fp64', 50.0), ('fp32', 20.0), ('fp32', 30.0),
                                           ('fp32', 50.0)])
def test_cut_keeps_nan_rays(K, precision, cut):
    import torch
    nl, nf, nray = 200, 130, 3
    rng = np.random.default_rng(4)
    alpha = 1.0e-6 * np.exp(rng.uniform(-0.5, 0.5, (nl, nf)))   # cm^-1: opaque after ~50 steps
    T = np.linspace(120.0, 600.0, nl)
    hds, ns = paths(nl, nray, bad_ray=1, bad_step=120)
    tab = torch.as_tensor(alpha, device='cuda')
    if precision == 'fp32':
        tab = tab.float()
    cutTb = K.integrate(tab, T, hds, ns, precision=precision, tau_cut=cut).cpu().numpy()
    full = K.integrate(tab, T, hds, ns, precision=precision, tau_cut=0.0).cpu().numpy()
    assert np.all(np.isnan(full[1])) and np.all(np.isnan(cutTb[1]))
    good = [0, 2]
    assert np.all(np.isfinite(cutTb[good]))
    # what the cut leaves out weighs < e^-cut of the weight sum and of the T-weighted sum
    # < T_max e^-cut: Tb moves by at most ~2 T_max e^-cut (2.5e-6 K at a cut of 20)
    assert np.max(np.abs(cutTb[good] - full[good])) < 1e-9 + 2.0 * T.max() * np.exp(-cut)


def test_jacobian_cut_keeps_nan_rays(K):
    import torch
    nl, nf, nray = 200, 64, 3
    alpha = torch.full((nl, nf), 1.0e-6, dtype=torch.float64, device='cuda')
    dal = torch.zeros((3, nl, nf), dtype=torch.float64, device='cuda')
    T = np.linspace(120.0, 600.0, nl)
    hds, ns = paths(nl, nray, bad_ray=2, bad_step=150)
    _, Tb_cut = K.jacobian(alpha, dal, T, hds, ns, tau_cut=50.0)
    _, Tb_full = K.jacobian(alpha, dal, T, hds, ns, tau_cut=0.0)
    Tb_cut, Tb_full = Tb_cut.cpu().numpy(), Tb_full.cpu().numpy()
    assert np.all(np.isnan(Tb_full[2])) and np.all(np.isnan(Tb_cut[2]))
    assert np.max(np.abs(Tb_cut[:2] - Tb_full[:2])) < 1e-9
