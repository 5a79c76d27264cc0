"""CPU tests of the host-side logic: requests, config, data return, atmosphere pipeline."""
import os

import numpy as np
import pytest

from oracle import pipeline
from radiobear_amd import atmosphere, config, data_handling, set_utils, synth, utils

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')


def test_set_freq_forms():
    assert set_utils.set_freq('1:96:5')[0] == list(np.arange(1.0, 98.5, 5.0))
    assert len(set_utils.set_freq('1:96:5')[0]) == 20
    f, _ = set_utils.set_freq('1;100;512')
    assert len(f) == 512 and abs(f[0] - 1.0) < 1e-12 and abs(f[-1] - 100.0) < 1e-9
    assert set_utils.set_freq('1.5,2.5')[0] == [1.5, 2.5]
    assert set_utils.set_freq(500, 'MHz')[0] == [0.5]
    assert set_utils.set_freq(np.array([1.0, 2.0]))[0] == [1.0, 2.0]
    with pytest.raises(ValueError):
        set_utils.set_freq({'a': 1})


def test_set_b_forms():
    rv = set_utils.set_b([0.1, 0.2])
    assert rv.data_type == 'spectrum' and rv.b == [[0.1, 0.2]]
    rv = set_utils.set_b([[0, 0]] * 6)
    assert rv.data_type == 'profile'
    rv = set_utils.set_b('Disc')
    assert rv.b == ['disc'] and rv.data_type == 'spectrum'
    rv = set_utils.set_b('0:1:0.1<90', Rpol=0.9, Req=1.0)      # pole: rab = Rpol/Req
    assert rv.data_type == 'profile' and len(rv.b) == 9        # v < 0.995*0.9
    assert abs(rv.b[-1][1] - 0.8) < 1e-12 and abs(rv.b[-1][0]) < 1e-12
    rv = set_utils.set_b(0.5)
    assert rv.data_type == 'image' and rv.imSize == [7, 7] and len(rv.b) == 49
    rv = set_utils.set_b(0.5, block=[2, 7])
    assert len(rv.b) == 7 and rv.b[0][1] == -1.0
    rv = set_utils.set_b('stamp:0.5:-0.5,0.5,0,0.5')
    assert rv.data_type == 'image' and len(rv.b) == 6


def test_utils():
    assert utils.b_type('DISC') == 'disc'
    assert utils.b_type([[0, 0]] * 21) == 'image'
    assert utils.b_type([[0, 0]] * 10) == 'profile'
    assert utils.convert_unit(2.0, 'MHz') == 2.0e-3
    assert utils.isanynum('3.5') and not utils.isanynum(True) and not utils.isanynum([1])


def test_config_token_file():
    c = config.planetConfig('jupiter', os.path.join(GOLDEN, 'config_benchmark.par'))
    assert c.constituent_alpha['nh3'] == 'nh3_hs_sjs' and c.constituent_alpha['cloud'] is None
    assert c.regridType == 1000 and c.pmin == 0.01 and c.pmax == 10000.0
    assert c.tweakmodule == 'JupiterNoTweak' and c.gasFile == ['jupiter.benchmark_gas']
    assert c.orientation == [0.0, 0.0] and c.gtype == 'ellipse' and c.h2state == 'e'
    assert abs(c.distance - 5.2 * utils.Units['AU'] / 1e3) < 1e-3
    c.update_config(h2state='n', Req=70000.0)
    assert c.h2state == 'n' and c.Req == 70000.0
    with pytest.raises(ValueError):
        config.planetConfig('pluto')


def test_data_return_float32():
    d = data_handling.Data()
    d.set('Tb', [[1.0, 2.0]])
    assert d.Tb.dtype == np.float32
    d.set('b', ['disc'])
    assert d.b == ['disc']


def test_atmosphere_pipeline_matches_oracle(bench_gas_text):
    c = config.planetConfig('jupiter', os.path.join(GOLDEN, 'config_benchmark.par'))
    a = atmosphere.Atmosphere('jupiter', config=c)
    a.std()
    gas, prop = pipeline.benchmark_atmosphere(bench_gas_text)
    assert np.array_equal(a.gas, gas)
    assert np.array_equal(a.property[c.LP['R']], prop['R'])
    assert np.array_equal(a.property[c.LP['N']], prop['N'])
    g = a.property[c.LP['g']]
    i1 = int(np.argmin(np.abs(a.gas[c.C['P']] - 1.0)))
    assert abs(g[i1] - c.GM_ref / c.Req ** 2) < 1e-12          # GM normalised at p_ref


def test_synthetic_atmosphere_generator():
    atm = synth.jupiter_like(200)
    P, T = atm.P, atm.T
    assert np.all(np.diff(P) > 0) and abs(P[0] - 1e-3) < 1e-15 and abs(P[-1] - 100.0) < 1e-12
    assert np.allclose(T, np.maximum(110.0, 165.0 * P ** 0.286))
    assert np.all(np.diff(atm.R) < 0)                           # radius decreases with depth
    i1 = int(np.argmin(np.abs(P - 1.0)))
    assert atm.R[i1] == synth.JUPITER_REQ
    nh3 = atm.gas[atm.C['NH3']]
    assert np.all(nh3 <= 3e-4) and nh3[-1] == 3e-4
    w1 = synth.workload('c3_large')
    w2 = synth.workload('c3_large')
    assert np.array_equal(w1['rays'], w2['rays']) and np.array_equal(w1['nh3_inv']['fo'], w2['nh3_inv']['fo'])
    assert len(w1['freqs']) == 1024 and len(w1['rays']) == 16384 and len(w1['h2o']['f_o']) == 5000
