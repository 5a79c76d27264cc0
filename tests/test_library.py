"""CPU checks of the C-ABI library and the host-side logic (no GPU compute)."""
import ctypes
import os
import re

import numpy as np
import pytest

from oracle import absorbers as OA
from radiobear_amd import _lib, kernels, lines

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), 'include', 'rbx.h')


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r'^int\s+(rbx_\w+)\s*\(', text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    lib = _lib.lib()
    names = declared()
    assert len(names) >= 12
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.EXPORTS)
    assert lib.rbx_abi_version() == _lib.ABI_VERSION


def test_argument_errors_are_reported_without_gpu_work():
    lib = _lib.lib()
    pl = kernels.planet_desc(71492.0, 66854.0)
    rc = lib.rbx_raypaths(ctypes.byref(pl), None, None, 1, None, 1, 0, None, 0, None, None)
    assert rc == -1
    assert 'null' in _lib.last_error() or 'layers' in _lib.last_error()
    with pytest.raises(_lib.RbxError) as e:
        _lib.check(lib.rbx_integrate(None, 0, None, 10, 5, None, 9, None, 1, 0, 0, 0.0, None, None, None, None))
    assert e.value.status == -1
    h = ctypes.c_void_p()
    assert lib.rbx_catalog_create(ctypes.byref(h)) == 0
    assert lib.rbx_catalog_set_lines(h, 99, 0, None) == -1
    assert 'bad kind' in _lib.last_error()
    assert lib.rbx_catalog_set_lines(h, 0, 0, None) == 0          # empty block: no device work
    n = ctypes.c_int(-5)
    assert lib.rbx_catalog_lines(h, 0, ctypes.byref(n)) == 0 and n.value == 0
    ab = kernels.absorbers_struct({'nh3': 'nh3_hs'})
    dummy = ctypes.c_void_p(8)
    rc = lib.rbx_alpha(h, ctypes.byref(ab), dummy, dummy, dummy, None, 3, dummy, None, 2, 0, None, 0, dummy, None,
                       None, dummy, 0, None)
    assert rc == -4 and 'nh3_hs' in _lib.last_error()             # empty inversion block
    for sel, what in (({'co': 'co_ddb'}, 'co_ddb'), ({'h2': 'h2_orton'}, 'h2_orton')):
        rc = lib.rbx_alpha(h, ctypes.byref(kernels.absorbers_struct(sel)), dummy, dummy, dummy, None, 3, dummy,
                           None, 2, 0, None, 0, dummy, None, None, dummy, 0, None)
        assert rc == -4 and what in _lib.last_error()             # its block is empty
    cl = kernels.absorbers_struct({'clouds': 'clouds_idp'})
    rc = lib.rbx_alpha(h, ctypes.byref(cl), dummy, dummy, dummy, None, 3, dummy, None, 2, 0, None, 0, dummy, None,
                       None, dummy, 0, None)
    assert rc == -1 and 'cloud table' in _lib.last_error()        # clouds need the cloud densities
    rc = lib.rbx_alpha_tangent(h, ctypes.byref(cl), dummy, dummy, dummy, 3, dummy, None, 2, 0, None, 0, dummy,
                               dummy, None, None, 0, None)
    assert rc == -1 and 'no tangent' in _lib.last_error()
    # tangent workspace: the derivers' table, 146 doubles per layer, plus with P > 1 parts the
    # partial sums [P][4 slots][3][nl][nf] and h2o base terms [P][3][nl]
    nb = ctypes.c_size_t(0)
    h2t = kernels.absorbers_struct({'h2': 'h2_jj_ddb'})
    one = _lib.launch_struct({'tan_split': 1})
    assert lib.rbx_alpha_tangent_workspace_size(h, ctypes.byref(h2t), 10, 7, 0, 1, _lib.launch_ref(one),
                                                ctypes.byref(nb)) == 0
    assert nb.value == 146 * 10 * 8
    # workspace: partial sums [4 slots][parts][nl][nf] + h2o base [parts][nl] doubles
    h2 = kernels.absorbers_struct({'h2': 'h2_jj_ddb'})
    assert lib.rbx_alpha_workspace_size(h, ctypes.byref(h2), 10, 7, 0, 1, None, ctypes.byref(nb)) == 0
    assert nb.value == (4 * 10 * 7 + 10) * 8                      # no lines: one part
    # the launch shape is an argument of each call: nothing persists between calls
    three = _lib.launch_ref({'alpha_split': 3})
    assert lib.rbx_alpha_workspace_size(h, ctypes.byref(h2), 10, 7, 0, 1, three, ctypes.byref(nb)) == 0
    assert nb.value == (4 * 3 * 10 * 7 + 3 * 10) * 8
    assert lib.rbx_alpha_workspace_size(h, ctypes.byref(h2), 10, 7, 0, 1, None, ctypes.byref(nb)) == 0
    assert nb.value == (4 * 10 * 7 + 10) * 8
    assert lib.rbx_alpha_workspace_size(h, ctypes.byref(h2), 0, 7, 0, 1, None, ctypes.byref(nb)) == -1
    for bad in ({'alpha_k': 3}, {'alpha_split': 17}, {'int_unroll': 5}, {'alpha_series': 3}, {'tan_split': -1}):
        rc = lib.rbx_alpha_workspace_size(h, ctypes.byref(h2), 10, 7, 0, 1, _lib.launch_ref(bad), ctypes.byref(nb))
        assert rc == -1 and 'launch.' in _lib.last_error(), bad
    res = _lib.Launch()
    res.reserved[2] = 1
    assert lib.rbx_alpha_workspace_size(h, ctypes.byref(h2), 10, 7, 0, 1, ctypes.byref(res), ctypes.byref(nb)) == -1
    with pytest.raises(ValueError):
        _lib.launch_struct({'no_such_knob': 1})
    assert lib.rbx_catalog_destroy(h) == 0


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, '_lib', None)
    monkeypatch.setattr(_lib, 'LIB_PATH', str(tmp_path / 'nope.so'))
    with pytest.raises(_lib.RbxLibraryError):
        _lib.lib()


def test_nh3_order_and_absorber_selection():
    order, nlow = kernels.nh3_order([40.0, 1.0, 30.0, 31.0, 2.0])
    assert list(order) == [1, 2, 4, 0, 3] and nlow == 3
    ab = kernels.absorbers_struct({'h2': 'h2_jj_ddb', 'nh3': 'nh3_hs_sjs', 'ph3': None}, h2state='n')
    assert (ab.h2, ab.h2o, ab.h2s, ab.nh3, ab.ph3) == (2, 0, 0, 1, 0)
    with pytest.raises(ValueError):
        kernels.absorbers_struct({'h2': 'h2_jj_ddb'}, h2state='x')
    with pytest.raises(ValueError):
        kernels.absorbers_struct({'nh3': 'nh3_bg'})
    with pytest.raises(ValueError):
        kernels.absorbers_struct({'hcn': 'hcn_ddb'})
    with pytest.raises(ValueError):
        kernels.planet_desc(1.0, 1.0, gtype='gravity')
    # the other planets' NH3 models and the optional absorbers with their options
    assert kernels.absorbers_struct({'nh3': 'nh3_dbs_sjs'}).nh3 == 4
    assert kernels.absorbers_struct({'nh3': 'nh3_dbs'}).nh3 == 5
    assert kernels.absorbers_struct({'h2': 'h2_orton'}, h2state='n').h2 == 4
    co = kernels.absorbers_struct({'co': 'co_ddb'})
    assert (co.co, co.co_shape) == (1, 0)                          # coshape defaults to 'voigt'
    for shape, code in (('vvw', 1), ('diff', 2), ('std', 3), (' ', 3)):
        assert kernels.absorbers_struct({'co': 'co_ddb'}, other_dict={'co': {'coshape': shape}}).co_shape == code
    flags = {'water_p': 1e-4, 'ice_p': 1e-4, 'nh4sh_p': 0.0, 'nh3ice_p': 1e-4, 'h2sice_p': 1e-4, 'ch4_p': 1e-4}
    cl = kernels.absorbers_struct({'clouds': 'clouds_idp'}, other_dict={'clouds': flags})
    # bits in RBX_CL_* order: ice (H2O) 0, water (SOLN) 1, nh4sh 2, nh3 3, h2s 4, ch4 5; the
    # module tests the key 'ch4', which Alpha never copies ('ch4_p'), so CH4 stays off
    assert (cl.clouds, cl.cloud_types) == (1, 0b011011)


def test_truncation_matches_oracle(raw_lines):
    a = lines.truncate(raw_lines['h2s'], ('f0', 'I0', 'E', 'GH2S'), 1e-22, 900.0)
    b = OA.truncate_ftable(raw_lines['h2s'], 1e-22, 900.0)
    for k in a:
        assert np.array_equal(a[k], b[k])
    a = lines.truncate_ph3(raw_lines['ph3jh'], raw_lines['PH3WGT'], 1e-22, 1200.0)
    b = OA.truncate_ph3(raw_lines['ph3jh'], raw_lines['PH3WGT'], 1e-22, 1200.0)
    for k in a:
        assert np.array_equal(a[k], b[k])
    a = lines.truncate_h2o(lines.h2o_bk_lines(), 1e-12, 600.0)
    b = OA.truncate_h2o(OA.h2o_bk_lines(), 1e-12, 600.0)
    for k in a:
        assert np.array_equal(a[k], b[k])
    assert len(lines.truncate(raw_lines['h2s'], ('f0', 'I0', 'E', 'GH2S'), 1e-22)['f0']) == 121
    assert len(lines.truncate_ph3(raw_lines['ph3jh'], raw_lines['PH3WGT'], 1e-22)['f0']) == 33


def test_adjoint_workspace_plan():
    """The adjoint's accumulator copies (host planning, no GPU work): a large image of one
    atmosphere spreads its atomic adds over 32 copies, a retrieval batch of 16 rays per
    atmosphere over one; a single 64-column frequency group needs no workspace at one copy."""
    lib = _lib.lib()
    nb = ctypes.c_size_t(0)

    def size(nl, nf, nray, natm, launch=None):
        assert lib.rbx_integrate_adjoint_workspace_size(nl, nf, nray, natm, _lib.launch_ref(launch),
                                                        ctypes.byref(nb)) == 0
        return nb.value

    assert size(800, 1024, 16384, 1) == (32 * 800 * 1024 + 32 * 16 * 800) * 8
    assert size(500, 128, 1024 * 16, 1024) == (1 * 2 * 500) * 8          # one copy; gT per frequency group
    assert size(100, 64, 8, 1) == 0                                       # one copy, one group
    assert size(100, 64, 8, 1, {'adj_copies': 4}) == (4 * 100 * 64 + 4 * 100) * 8
    assert lib.rbx_integrate_adjoint_workspace_size(1, 64, 8, 1, None, ctypes.byref(nb)) == -1
