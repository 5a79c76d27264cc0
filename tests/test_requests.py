"""Request parsing (set_freq, set_b) and result files (FileIO) against what the reference
wrote for the same inputs (tests/golden/requests.json, made by tests/golden/make_fixtures.py):
values equal, files byte for byte.  Where the reference raises, the build's documented
behaviour is checked instead (module docstring of radiobear_amd.set_utils)."""
import json
import os

import numpy as np
import pytest

from radiobear_amd import data_handling, fileio, set_utils

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')


@pytest.fixture(scope='module')
def fixture():
    with open(os.path.join(GOLDEN, 'requests.json')) as fp:
        return json.load(fp)


def test_set_freq_matches_reference(fixture):
    checked = 0
    for rec in fixture['set_freq']:
        if 'raises' in rec:
            continue
        f, unit = set_utils.set_freq(rec['request'], rec['unit'])
        assert [float(x) for x in f] == rec['freqs'], rec['request']
        assert unit == rec['out_unit']
        checked += 1
    assert checked >= 9


def test_set_freq_log_request_takes_an_integer_count(fixture):
    """'lo;hi;n': the reference hands n to numpy as a float (a TypeError in numpy 2); here n
    is an int and the values are numpy's logspace."""
    rec = [r for r in fixture['set_freq'] if r['request'] == '0.3;300;7'][0]
    assert rec['raises'] == 'TypeError'
    f, _ = set_utils.set_freq('0.3;300;7')
    assert np.array_equal(f, np.logspace(np.log10(0.3), np.log10(300.0), 7))


def test_set_b_matches_reference(fixture):
    checked = 0
    for rec in fixture['set_b']:
        if 'raises' in rec:
            continue
        r = set_utils.set_b(rec['request'], block=rec['block'], Rpol=66854.0, Req=71492.0)
        got = [[float(v) for v in p] if not isinstance(p, str) else p for p in r.b]
        assert got == rec['b'], rec['request']
        assert r.data_type == rec['data_type'], rec['request']
        if str(rec['request']).startswith('stamp'):
            # the reference divides the request string's length by the row length; here
            # imSize is [points per row, rows]
            xs = len(rec['b']) // rec['imSize'][0]
            assert r.imSize == [rec['imSize'][0], xs]
        else:
            assert r.imSize == rec['imSize']
        checked += 1
    assert checked >= 12


def test_set_b_image_spacing(fixture):
    """A float spacing: the reference builds the grid, then fails on len() of the float; here
    imSize counts the grid's points per row and rows, and blocks split the rows."""
    assert [r for r in fixture['set_b'] if r['request'] == 0.25][0]['raises'] == 'TypeError'
    r = set_utils.set_b(0.25)
    axis = np.concatenate((-np.flipud(np.arange(0.25, 1.75, 0.25)), np.arange(0.0, 1.75, 0.25)))
    assert r.imSize == [len(axis), len(axis)] and r.data_type == 'image'
    assert r.b == [[x, y] for y in axis for x in axis]
    parts = [set_utils.set_b(0.25, block=[i, 3]).b for i in (1, 2, 3)]
    assert parts[0][0] == [axis[0], axis[0]] and parts[1][0][1] == axis[len(axis) // 3]


@pytest.mark.parametrize('name', ['spectrum_frequency.dat', 'spectrum_wavelength.dat', 'profile_frequency.dat',
                                  'profile_wavelength.dat', 'image_frequency.dat', 'image_wavelength.dat',
                                  'spectrum-disc_frequency.dat', 'spectrum-disc_wavelength.dat'])
def test_result_files_byte_for_byte(fixture, tmp_path, name, monkeypatch):
    """The same Data written by the reference's FileIO and by fileio.FileIO: identical text."""
    kind, xaxis = name[:-4].split('_')
    monkeypatch.chdir(tmp_path)
    os.mkdir('Scratch')
    tb = np.array([[134.2567, 151.5, 172.25001], [133.0, 150.75, 171.5], [131.125, 149.875, 170.0625]])
    d = data_handling.Data()
    d.set('header', {'a_planet': '# jupiter', 'b_freqs': '# freqs: 1.5 22.2 40', 'c_x': '# x'})
    d.set('type', kind.split('-')[0])
    d.set('f', [1.5, 22.2, 40.0])
    d.set('freqUnit', 'GHz')
    d.set('b', 'disc' if kind == 'spectrum-disc' else [[0.0, 0.0], [0.45, 0.1], [0.7, -0.2]])
    d.set('Tb', tb[:1].tolist() if kind == 'spectrum-disc' else tb.tolist())
    fileio.FileIO().write(name, d, xaxis=xaxis)
    with open(name) as fp:
        assert fp.read() == fixture['files'][name]
