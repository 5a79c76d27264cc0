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


def test_cloud_file_read_and_regrid(tmp_path, bench_gas_text):
    """readCloud + the cloud regrid (atm_base.py:80-118, regrid.py:74-81, 97-118): the
    cloud table on the gas grid, linear in P, zero outside the cloud file's range,
    deepest z = 0, checked against scipy's interp1d as the reference calls it."""
    from scipy.interpolate import interp1d
    (tmp_path / 'gas.dat').write_text(bench_gas_text)
    # a deck written deep to high (the reader flips it), 3 bar .. 0.05 bar, clouds near 1 bar
    P = np.geomspace(3.0, 0.05, 40)
    z = np.linspace(0.0, 120.0, 40)
    T = 150.0 + 40.0 * np.log10(P)
    rows = []
    for i in range(40):
        soln = 5e-4 * np.exp(-((np.log(P[i]) - 0.7) / 0.3) ** 2)
        ice = 2e-4 * np.exp(-((np.log(P[i]) + 0.5) / 0.4) ** 2)
        nh3 = 3e-5 * np.exp(-((np.log(P[i]) + 1.2) / 0.3) ** 2)
        rows.append('{:8.2f} {:9.3f} {:11.4e} {:11.4e} {:11.4e} {:11.4e} {:11.4e} {:11.4e} {:11.4e} {:11.4e}'
                    .format(z[i], T[i], P[i], soln, ice, 1e-6, nh3, 0.0, 0.0, 0.0))
    (tmp_path / 'cloud.dat').write_text('   Z T P SOLN H2O NH4SH NH3 H2S CH4 AR\n' + '\n'.join(rows) + '\n')
    par = (tmp_path / 'config.par')
    par.write_text(open(os.path.join(GOLDEN, 'config_benchmark.par')).read()
                   .replace('gasfile jupiter.benchmark_gas', 'gasfile gas.dat')
                   .replace('cloudfile jupiter.benchmark_nocloud', 'cloudfile cloud.dat'))
    c = config.planetConfig('jupiter', str(par))
    a = atmosphere.Atmosphere('jupiter', config=c)
    a.std()
    gas, _ = pipeline.benchmark_atmosphere(bench_gas_text)
    assert np.array_equal(a.gas, gas)                          # clouds do not touch the gas table
    Cl = c.Cl
    Pg = a.gas[c.C['P']]
    assert np.array_equal(a.cloud[Cl['P']], Pg)
    Pin = np.array([float(r.split()[2]) for r in rows])[::-1]     # as written (4 digits)
    for key, col in (('SOLN', 3), ('H2O', 4), ('NH3', 6)):
        src = np.array([float(r.split()[col]) for r in rows])[::-1]
        want = interp1d(Pin, src, kind='linear', fill_value=0.0, bounds_error=False)(Pg)
        assert np.array_equal(a.cloud[Cl[key]], want), key
    assert a.cloud[Cl['Z']][-1] == 0.0                         # deepest z = 0 after regrid
    assert np.max(a.cloud[Cl['SOLN']]) > 1e-4 and np.all(a.cloud[Cl['SOLN']][Pg > 3.0] == 0.0)


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


def test_result_files_in_reference_formats(tmp_path):
    """FileIO writes the reference's text layouts (fileIO.py:19-104): sorted header lines,
    spectrum rows '{:.2f}' + '  {:9.4f}' columns, profile rows, image rows, disc label."""
    import datetime
    from radiobear_amd.data_handling import Data
    from radiobear_amd.fileio import FileIO, output_name
    d = Data()
    d.set('f', [1.5, 22.235])
    d.set('freqUnit', 'GHz')
    d.set('b', [[0.0, 0.0], [0.25, 0.5]])
    d.set('Tb', [[200.25, 150.5], [180.125, 140.0]])
    d.set('header', {'zz': '# last', 'aa': '# first'})
    d.set('type', 'spectrum')
    io = FileIO(str(tmp_path), str(tmp_path / 'line.dat'))
    out = io.write(str(tmp_path / 's.dat'), d)
    lines = open(out).read().split('\n')
    assert lines[0] == '# first' and lines[1] == '# last'
    assert lines[2] == '# GHz  K@b  \t(0.000,0.000)\t(0.250,0.500)'
    f32 = [float(np.float32(v)) for v in (1.5, 22.235, 200.25, 150.5, 180.125, 140.0)]
    assert lines[3] == '{:.2f}\t    {:9.4f}  \t  {:9.4f}'.format(f32[0], f32[2], f32[4])
    assert lines[4] == '22.24\t    {:9.4f}  \t  {:9.4f}'.format(f32[3], f32[5])
    assert open(tmp_path / 'line.dat').read().split('\n')[:2] == lines[3:5]
    d.set('type', 'profile')
    lines = open(io.write(str(tmp_path / 'p.dat'), d)).read().split('\n')
    assert lines[2] == '# b  K@GHz \t     1.5000   \t    22.2350'
    assert lines[3] == '0.000 0.000\t  200.25\t   150.50'
    d.set('type', 'image')
    lines = open(io.write(str(tmp_path / 'i.dat'), d)).read().split('\n')
    assert lines[2] == '200.25\t 150.50' and lines[3] == '180.12\t 140.00'
    d.set('b', 'disc')
    d.set('Tb', [[200.0, 150.0]])
    d.set('type', 'spectrum')
    lines = open(io.write(str(tmp_path / 'd.dat'), d)).read().split('\n')
    assert lines[2] == '# GHz  K@b  \tdisc'
    t = datetime.datetime(2024, 5, 6, 7, 8, 9)
    assert output_name('Jupiter', 'spectrum', '_', t, 'Out') == os.path.join('Out', 'Jupiter_spectrum_20240506_070809.dat')


def test_doppler_is_refused_before_any_device_work():
    """The reference's Doppler branch cannot run (brightness.py:79-90 calls a missing
    Alpha.get_alpha); here a Doppler config is refused up front."""
    from argparse import Namespace
    from radiobear_amd.brightness import Brightness
    c = config.planetConfig('jupiter')
    c.Doppler = True
    br = Brightness(config=c, verbose=False)
    with pytest.raises(NotImplementedError):
        br.batch([[0.0, 0.0]], [1.0], None, Namespace(config=c, table=None))
    with pytest.raises(NotImplementedError):
        br.single([0.0, 0.0], [1.0], None, Namespace(config=c, table=None))
