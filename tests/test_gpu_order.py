"""Frequency order for the optical-depth cut (rbx_integrate_order + rbx_integrate_cols).

The order changes only which frequencies share a wavefront, never a lane's own arithmetic:
with a cut that no frequency reaches the reordered integration is bit-identical to the plain
one (brightness.py:93-101 per frequency), and with a binding cut the two differ only by what
the cut itself leaves out (< 2 T_max e^-cut)."""
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


def case(nl, nf, nray, seed, dtype):
    """Alpha whose columns span three decades of opacity, rays of several path lengths (one
    off the planet), half steps of a few km."""
    import torch
    rng = np.random.default_rng(seed)
    scale = 10.0 ** rng.uniform(-8.0, -5.0, nf)
    alpha = scale[None, :] * np.exp(rng.uniform(-0.3, 0.3, (nl, nf)))
    T = np.linspace(110.0, 650.0, nl)
    hds = rng.uniform(2.0e5, 6.0e5, (nray, nl))
    ns = rng.integers(nl // 2, nl + 1, nray).astype(np.int32)
    ns[nray // 2] = -1
    tab = torch.as_tensor(alpha, device='cuda').to(dtype)
    return (tab, T, torch.as_tensor(hds, device='cuda'), torch.as_tensor(ns, device='cuda'),
            tab.double().cpu().numpy(), hds, ns)


def order(K, tab, hds, ns, cut):
    import torch
    from radiobear_amd import _lib
    nl, nf = tab.shape
    nray, ld = hds.shape
    aprec = _lib.FP32 if tab.dtype == torch.float32 else _lib.FP64
    cols = torch.full((nf,), -7, dtype=torch.int32, device='cuda')
    a_cols = torch.empty_like(tab)
    _lib.check(_lib.lib().rbx_integrate_order(K._ptr(tab), aprec, nl, nf, K._ptr(hds), ld, K._ptr(ns), nray,
                                              float(cut), K._ptr(cols), K._ptr(a_cols), K._stream()))
    return cols.cpu().numpy(), a_cols


def cut_steps(alpha, hds, ns, cut):
    """Each column's first step at which the longest path (lowest index among equals) passes
    the cut, the path's step count when it never does."""
    r = int(np.argmax(ns))
    S = max(min(int(ns[r]), alpha.shape[0]) - 1, 0)
    tau = np.cumsum((alpha[:-1] + alpha[1:])[:S] * hds[r, :S, None], axis=0)
    past = tau > cut
    return np.where(past.any(axis=0), np.argmax(past, axis=0), S)


@pytest.mark.parametrize('dtype_name,nf', [('float64', 300), ('float32', 300), ('float32', 1), ('float32', 5000),
                                             ('float64', 16384)])
def test_order_is_stable_sort_by_cut_step(K, dtype_name, nf):
    import torch
    dtype = getattr(torch, dtype_name)
    tab, T, hds, ns, alpha, hds_np, ns_np = case(150, nf, 9, 11, dtype)
    cut = 30.0
    cols, a_cols = order(K, tab, hds, ns, cut)
    assert sorted(cols.tolist()) == list(range(nf))
    key = cut_steps(alpha, hds_np, ns_np, cut)
    want = np.argsort(key, kind='stable')
    assert np.array_equal(cols, want)
    assert torch.equal(a_cols.cpu(), tab.cpu()[:, torch.as_tensor(cols, dtype=torch.long)])
    assert nf < 100 or len(np.unique(key)) > 10   # the case spreads the cut steps


@pytest.mark.parametrize('dtype_name', ['float32', 'float64'])
def test_order_long_path(K, dtype_name):
    """A reference path longer than the half steps the order kernel stages in LDS (8192): the
    same stable sort by cut step, with columns passing the cut anywhere from a few thousand
    steps to never."""
    import torch
    rng = np.random.default_rng(17)
    nl, nf, nray = 9000, 70, 3
    scale = 10.0 ** rng.uniform(-8.6, -7.9, nf)
    alpha = scale[None, :] * np.exp(rng.uniform(-0.3, 0.3, (nl, nf)))
    hds = rng.uniform(2.0e5, 6.0e5, (nray, nl))
    ns = np.array([nl - 200, nl, 50], dtype=np.int32)
    tab = torch.as_tensor(alpha, device='cuda').to(getattr(torch, dtype_name))
    hg, ng = torch.as_tensor(hds, device='cuda'), torch.as_tensor(ns, device='cuda')
    cols, a_cols = order(K, tab, hg, ng, 30.0)
    key = cut_steps(tab.double().cpu().numpy(), hds, ns, 30.0)
    assert np.array_equal(cols, np.argsort(key, kind='stable'))
    assert key.min() > 2000 and (key == nl - 1).sum() > 5 and len(np.unique(key)) > 20


@pytest.mark.parametrize('precision,dtype_name,disc', [('fp64', 'float64', False), ('fp32', 'float32', False),
                                                       ('fp64', 'float32', True), ('fp32', 'float32', True)])
def test_identity_order_is_bit_identical(K, precision, dtype_name, disc):
    """rbx_integrate_cols with cols = 0..nf-1 is rbx_integrate."""
    import torch
    from radiobear_amd import _lib
    tab, T, hds, ns, *_ = case(120, 200, 7, 5, getattr(torch, dtype_name))
    nl, nf = tab.shape
    nray, ld = hds.shape
    aprec = _lib.FP32 if tab.dtype == torch.float32 else _lib.FP64
    prec = _lib.PRECISIONS[precision]
    Tgpu = K.as_device(T)
    ident = torch.arange(nf, dtype=torch.int32, device='cuda')
    fac = torch.full((nray, ld, 4), np.nan, dtype=torch.float32, device='cuda')
    _lib.check(_lib.lib().rbx_integrate_factors(K._ptr(Tgpu), nl, K._ptr(hds), ld, K._ptr(ns), nray, K._ptr(fac),
                                                K._stream()))
    for cut in (0.0, 25.0):
        Tb0, iw0 = K.integrate(tab, T, hds, ns, disc=disc, precision=precision, tau_cut=cut, with_intW=True)
        on = ns.cpu().numpy() >= 0
        for factors in (None, fac):   # the factor table (FP32 point rays) or LDS staging: the same sums
            Tb1 = torch.full_like(Tb0, np.nan)
            iw1 = torch.full_like(iw0, np.nan)
            _lib.check(_lib.lib().rbx_integrate_cols(K._ptr(tab), aprec, K._ptr(ident), K._ptr(Tgpu), nl, nf,
                                                     K._ptr(hds), ld, K._ptr(ns), nray, 1 if disc else 0, prec, cut,
                                                     K._ptr(Tb1), K._ptr(iw1), K._ptr(factors), None, K._stream()))
            assert torch.equal(Tb0[on], Tb1[on]) and torch.equal(iw0[on], iw1[on])


def test_factor_table_nan_path(K):
    """A ray whose path holds a NaN half step keeps the NaN in Tb (no cut for it, as the staged
    sums do), while the other rays in the factor table integrate as usual; bit-identical to
    rbx_integrate."""
    import torch
    from radiobear_amd import _lib
    tab, T, hds, ns, *_ = case(90, 150, 6, 8, torch.float32)
    hds[2, 5] = float('nan')
    nl, nf = tab.shape
    nray, ld = hds.shape
    ident = torch.arange(nf, dtype=torch.int32, device='cuda')
    fac = torch.full((nray, ld, 4), 7.0, dtype=torch.float32, device='cuda')
    Tgpu = K.as_device(T)
    _lib.check(_lib.lib().rbx_integrate_factors(K._ptr(Tgpu), nl, K._ptr(hds), ld, K._ptr(ns), nray, K._ptr(fac),
                                                K._stream()))
    Tb0 = K.integrate(tab, T, hds, ns, precision='fp32', tau_cut=15.0)
    Tb1 = torch.full_like(Tb0, -1.0)
    _lib.check(_lib.lib().rbx_integrate_cols(K._ptr(tab), _lib.FP32, K._ptr(ident), K._ptr(Tgpu), nl, nf,
                                             K._ptr(hds), ld, K._ptr(ns), nray, 0, _lib.FP32, 15.0, K._ptr(Tb1),
                                             None, K._ptr(fac), None, K._stream()))
    on = ns.cpu().numpy() >= 0
    assert np.array_equal(Tb0[on].cpu().numpy(), Tb1[on].cpu().numpy(), equal_nan=True)
    assert on[2] and ns[2].item() > 7 and torch.isnan(Tb1[2]).all()
    assert not torch.isnan(Tb1[on][[r != 2 for r in np.flatnonzero(on)]]).any()


@pytest.mark.parametrize('precision,dtype_name', [('fp64', 'float64'), ('fp32', 'float32')])
def test_ordered_integration(K, precision, dtype_name):
    import torch
    tab, T, hds, ns, *_ = case(160, 700, 11, 23, getattr(torch, dtype_name))
    on = ns.cpu().numpy() >= 0
    # a cut no frequency reaches: every lane runs its whole path, so the order changes nothing
    far = 1.0e30
    Tb0 = K.integrate(tab, T, hds, ns, precision=precision, tau_cut=far).cpu().numpy()
    Tb1 = K.integrate(tab, T, hds, ns, precision=precision, tau_cut=far, ordered=True).cpu().numpy()
    assert np.array_equal(Tb0[on], Tb1[on])
    # a binding cut: both within what the cut leaves out of the uncut sums
    full = K.integrate(tab, T, hds, ns, precision=precision, tau_cut=0.0).cpu().numpy()
    for cut in (20.0, 40.0):
        Tbp = K.integrate(tab, T, hds, ns, precision=precision, tau_cut=cut).cpu().numpy()
        Tbo = K.integrate(tab, T, hds, ns, precision=precision, tau_cut=cut, ordered=True).cpu().numpy()
        bound = 1e-9 + 2.0 * T.max() * np.exp(-cut)
        assert np.max(np.abs(Tbo[on] - full[on])) < bound
        assert np.max(np.abs(Tbo[on] - Tbp[on])) < 2.0 * bound


def test_order_rejects_bad_arguments(K):
    import torch
    from radiobear_amd import _lib
    tab, T, hds, ns, *_ = case(20, 16, 3, 1, torch.float64)
    cols = torch.empty((16,), dtype=torch.int32, device='cuda')
    a_cols = torch.empty_like(tab)
    L = _lib.lib()
    args = [K._ptr(tab), _lib.FP64, 20, 16, K._ptr(hds), hds.shape[1], K._ptr(ns), 3]
    assert L.rbx_integrate_order(*args, 0.0, K._ptr(cols), K._ptr(a_cols), K._stream()) == -1
    assert L.rbx_integrate_order(*args, 30.0, None, K._ptr(a_cols), K._stream()) == -1
    big = [K._ptr(tab), _lib.FP64, 1, K.MAX_ORDER_FREQS + 1, K._ptr(hds), hds.shape[1], K._ptr(ns), 3]
    assert L.rbx_integrate_order(*big, 30.0, K._ptr(cols), K._ptr(a_cols), K._stream()) == -1


def test_order_without_on_planet_rays(K):
    """No ray on the planet: every column's cut step is 0, so the order is the request order,
    and the ordered integration returns what the plain one does (T_cmb rows)."""
    import torch
    tab, T, hds, ns, *_ = case(40, 130, 5, 3, torch.float32)
    ns.fill_(-1)
    cols, a_cols = order(K, tab, hds, ns, 20.0)
    assert np.array_equal(cols, np.arange(130))
    assert torch.equal(a_cols, tab)
    plain = K.integrate(tab, T, hds, ns, precision='fp32', tau_cut=20.0).cpu().numpy()
    ordered = K.integrate(tab, T, hds, ns, precision='fp32', tau_cut=20.0, ordered=True).cpu().numpy()
    assert np.array_equal(plain, ordered) and np.all(plain == 2.725)


def test_factor_table_values(K):
    """rbx_integrate_factors: 16-byte records of (h_k log2 e in FP64, then in FP32 h_k + h_{k+1} and
    T_{k+1} (h_k + h_{k+1})) for k < n = nsteps - 1 (h_n = 0); record n holds (0, NaN if the path
    holds a NaN else 0, 0); records past n untouched; nothing written for rays off the planet."""
    import torch
    from radiobear_amd import _lib
    rng = np.random.default_rng(4)
    nl, nray = 70, 5
    ld = nl + 3
    T = rng.uniform(100.0, 600.0, nl)
    hds = rng.uniform(1.0e5, 9.0e5, (nray, ld))
    ns = np.array([nl - 1, 40, -1, 2, nl - 1], dtype=np.int32)
    hds[4, 17] = np.nan
    fac = torch.full((nray, ld, 4), 5.0, dtype=torch.float32, device='cuda')
    Tg, hg, ng = K.as_device(T), torch.as_tensor(hds, device='cuda'), torch.as_tensor(ns, device='cuda')
    _lib.check(_lib.lib().rbx_integrate_factors(K._ptr(Tg), nl, K._ptr(hg), ld, K._ptr(ng), nray, K._ptr(fac),
                                                K._stream()))
    raw = fac.cpu().numpy()
    rec = raw.view(np.dtype([('hl', '<f8'), ('hh', '<f4'), ('thh', '<f4')]))[..., 0]
    log2e = np.float64(1.4426950408889634)
    assert log2e == np.log2(np.e)
    for r in range(nray):
        S = min(int(ns[r]), nl - 1)
        if S <= 1:
            assert np.all(raw[r] == 5.0)
            continue
        n = S - 1
        h = hds[r, :n]
        hh = h + np.append(hds[r, 1:n], 0.0)
        assert np.array_equal(rec['hl'][r, :n], h * log2e, equal_nan=True)
        assert np.array_equal(rec['hh'][r, :n], hh.astype(np.float32), equal_nan=True)
        assert np.array_equal(rec['thh'][r, :n], (T[1:n + 1] * hh).astype(np.float32), equal_nan=True)
        assert rec['hl'][r, n] == 0.0 and rec['thh'][r, n] == 0.0
        assert np.isnan(rec['hh'][r, n]) == (r == 4)
        assert np.all(raw[r, n + 1:] == 5.0)
    assert np.isnan(rec['hh'][4, 16:18]).all()


def test_factor_table_rejects_misaligned(K):
    import torch
    from radiobear_amd import _lib
    tab, T, hds, ns, *_ = case(20, 16, 3, 1, torch.float32)
    fac = torch.zeros((3 * hds.shape[1] * 4 + 1,), dtype=torch.float32, device='cuda')
    bad = ctypes_ptr_offset(fac, 4)
    L = _lib.lib()
    assert L.rbx_integrate_factors(K._ptr(K.as_device(T)), 20, K._ptr(hds), hds.shape[1], K._ptr(ns), 3, bad,
                                   K._stream()) == -1


def ctypes_ptr_offset(t, nbytes):
    import ctypes
    return ctypes.c_void_p(t.data_ptr() + nbytes)
