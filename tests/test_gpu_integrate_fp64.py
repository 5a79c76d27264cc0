"""FP64 point-ray integration against a closed form.

With alpha constant along the path and equal half steps h, tau at path point p is 2 alpha h p,
so the integrated weight of brightness.py:93-101 (W_0 = 0 at the top point, W_p = alpha e^-tau_p,
trapezoids sum (W_i + W_{i+1}) h) is the geometric sum
    intW = alpha h (2 sum_{p=1}^{n-1} q^p + q^n),  q = e^{-2 alpha h},
evaluated here with 40-digit decimals.  It pins the kernel's e^-tau (a 64-entry table of
2^(-j/64) times a degree-6 polynomial) to about an ulp per term: a wrong table entry or
polynomial coefficient moves intW by 1e-10 or more, against the 1e-13 tolerance (n rounding
steps of 1.1e-16 each)."""
from decimal import Decimal, getcontext

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


def closed_form(alpha, h, n):
    getcontext().prec = 40
    a, hh = Decimal(repr(float(alpha))), Decimal(repr(float(h)))
    q = (-2 * a * hh).exp()
    head = q * (1 - q ** (n - 1)) / (1 - q)
    return float(a * hh * (2 * head + q ** n))


@pytest.mark.parametrize('nl', [160, 801])
def test_fp64_weight_sum_closed_form(K, nl):
    import torch
    rng = np.random.default_rng(nl)
    nf, nray, h = 96, 2, 2.5e5
    alpha_col = 10.0 ** rng.uniform(-9.0, -5.0, nf)           # total tau from ~1e-4 to ~800
    alpha_col[:4] = [0.0, 1e-30, 2e-6, 1.3e-7]
    tab = torch.as_tensor(np.tile(alpha_col, (nl, 1)), device='cuda')
    T = np.full(nl, 300.0)
    hds = torch.full((nray, nl), h, dtype=torch.float64, device='cuda')
    ns = torch.full((nray,), nl - 1, dtype=torch.int32, device='cuda')
    Tb, iw = K.integrate(tab, T, hds, ns, precision='fp64', tau_cut=0.0, with_intW=True)
    iw = iw.cpu().numpy()
    n = nl - 2   # path steps: nsteps = nl - 1 half steps give nl - 2 trapezoids
    want = np.array([closed_form(a, h, n) if a > 0 else 0.0 for a in alpha_col])
    for r in range(nray):
        ok = want > 0
        assert np.max(np.abs(iw[r, ok] - want[ok]) / want[ok]) < 1e-13
        assert np.all(iw[r, ~ok] == 0.0)
    # constant T: Tb is T wherever the weight sum is above the T_cmb floor's reach
    Tb = Tb.cpu().numpy()
    big = want * 300.0 > 2.725 * 2
    assert np.max(np.abs(Tb[:, big] - 300.0)) < 1e-10
