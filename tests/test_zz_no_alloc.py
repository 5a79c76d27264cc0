"""rbx_alpha_noalloc (the static-scratch variant of rbx_alpha) against rbx_alpha, through the C ABI.

With the same launch shape both run the same plan, so their tables are bit-identical in
either precision; two calls in a row (the second smaller) check that the shared scratch keeps
nothing a later call depends on.  A request whose built-in plan needs more than the scratch
maximum takes the direct line sums (FP32 within 1e-6 and FP64 within 1e-11 of the series), and
one that cannot fit at all is refused before any GPU work."""
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
def cat(K, raw_lines):
    return K.jupiter_catalog(raw_lines, h2o=synth.h2o_lines(400, 8))


@pytest.fixture(scope='module')
def big_cat(K, raw_lines):
    """20000 synthetic NH3 inversion lines: the series' workspace (partial sums, moments and
    near masks) of 320 layers x 1024 frequencies then takes about 87 MB, more than the scratch
    maximum."""
    return K.jupiter_catalog(raw_lines, nh3_inversion=synth.nh3_inversion_lines(20000, 5))


def both(K, cat, atm, f, prec, launch=None, by_constituent=False):
    ab = K.absorbers_struct(JUPITER_ALPHA)
    args = (cat, ab, atm.P, atm.T, K.mix_table(atm.gas, C), f, prec)
    ref = K.alpha(*args, launch=launch, by_constituent=by_constituent)
    got = K.alpha_noalloc(*args, launch=launch, by_constituent=by_constituent)
    if by_constituent:
        return [t.double().cpu().numpy() for t in ref], [t.double().cpu().numpy() for t in got]
    return ref.double().cpu().numpy(), got.double().cpu().numpy()


@pytest.mark.parametrize('prec', ['fp32', 'fp64'])
@pytest.mark.parametrize('launch', [None, {'alpha_series': 1, 'alpha_split': 3}, {'alpha_series': 2, 'alpha_k': 1}])
def test_noalloc_matches_alpha_twice(K, cat, prec, launch):
    """Two calls in a row, the second with fewer layers and frequencies: each bit-identical to
    rbx_alpha with the same launch shape (the built-in one for None)."""
    big = synth.jupiter_like(47, pmax=3000.0)
    small = synth.jupiter_like(19)
    f_big = np.geomspace(0.8, 250.0, 330)
    f_small = np.geomspace(2.0, 90.0, 70)
    for atm, f in ((big, f_big), (small, f_small)):
        ref, got = both(K, cat, atm, f, prec, launch)
        assert got.shape == (atm.nl, len(f))
        assert np.all(np.isfinite(got))
        assert np.array_equal(got, ref), (prec, launch, atm.nl)


def test_noalloc_per_constituent_rows(K, cat):
    atm = synth.jupiter_like(31)
    (ra, rb), (ga, gb) = both(K, cat, atm, np.geomspace(1.0, 120.0, 200), 'fp64', by_constituent=True)
    assert np.array_equal(ga, ra) and np.array_equal(gb, rb)


@pytest.mark.parametrize('prec,tol', [('fp32', 1e-6), ('fp64', 1e-11)])
def test_noalloc_falls_back_to_direct_sums(K, big_cat, prec, tol):
    """The built-in plan of 320 layers x 1024 frequencies is the far-line series, whose workspace
    needs more scratch than the maximum: the call takes the direct sums instead (bit-identical
    to rbx_alpha with the direct sums split into the parts the fallback chose, and within the
    series' own error of the built-in table)."""
    from radiobear_amd import _lib
    cat = big_cat
    atm = synth.jupiter_like(320)
    f = np.geomspace(0.3, 300.0, 1024)
    order, nf_low = K.nh3_order(f)
    ab = K.absorbers_struct(JUPITER_ALPHA)
    nb = ctypes.c_size_t(0)
    _lib.check(_lib.lib().rbx_alpha_workspace_size(cat.handle, ctypes.byref(ab), atm.nl, len(f), nf_low,
                                                   _lib.PRECISIONS[prec], None, ctypes.byref(nb)))
    assert nb.value > _lib.NOALLOC_SCRATCH_BYTES
    built_in, got = both(K, cat, atm, f, prec)
    assert np.max(np.abs(got - built_in) / np.abs(built_in)) < tol
    fits = []
    for n in range(16, 0, -1):
        L = {'alpha_series': 1, 'alpha_split': n}
        _lib.check(_lib.lib().rbx_alpha_workspace_size(cat.handle, ctypes.byref(ab), atm.nl, len(f), nf_low,
                                                       _lib.PRECISIONS[prec], _lib.launch_ref(L), ctypes.byref(nb)))
        if nb.value <= _lib.NOALLOC_SCRATCH_BYTES:
            fits.append(n)
    direct = {n: K.alpha(cat, ab, atm.P, atm.T, K.mix_table(atm.gas, C), f, prec,
                         launch={'alpha_series': 1, 'alpha_split': n}).double().cpu().numpy() for n in fits}
    assert any(np.array_equal(got, d) for d in direct.values())


def test_noalloc_refuses_what_cannot_fit(K, cat, big_cat):
    """Forcing the series on a request whose workspace exceeds the scratch maximum, or a
    request too large even for one part of the direct sums, returns RBX_ERR_ARG before any
    kernel runs; the next call still works."""
    from radiobear_amd import _lib
    atm = synth.jupiter_like(320)
    f = np.geomspace(0.3, 300.0, 1024)
    with pytest.raises(_lib.RbxError) as e:
        K.alpha_noalloc(big_cat, K.absorbers_struct(JUPITER_ALPHA), atm.P, atm.T, K.mix_table(atm.gas, C), f,
                        'fp32', launch={'alpha_series': 2})
    assert e.value.status == -1 and 'scratch' in str(e.value)
    deep = synth.jupiter_like(2100)
    with pytest.raises(_lib.RbxError) as e:
        K.alpha_noalloc(cat, K.absorbers_struct(JUPITER_ALPHA), deep.P, deep.T, K.mix_table(deep.gas, C), f,
                        'fp64', launch={'alpha_series': 1})
    assert e.value.status == -1 and 'scratch' in str(e.value)
    small = synth.jupiter_like(12)
    ref, got = both(K, cat, small, np.geomspace(1.0, 40.0, 30), 'fp64')
    assert np.array_equal(got, ref)
