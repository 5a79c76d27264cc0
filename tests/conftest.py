import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(REPO, 'tests', 'golden')
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built library")


@pytest.fixture(scope='session')
def golden():
    def load(name):
        return np.load(os.path.join(GOLDEN, name))
    return load


@pytest.fixture(scope='session')
def raw_lines():
    from radiobear_amd import lines
    return lines.raw_catalog()


@pytest.fixture(scope='session')
def jupiter_lines(raw_lines):
    from oracle import pipeline
    return pipeline.jupiter_lines(raw_lines)


@pytest.fixture(scope='session')
def bench_gas_text():
    with open(os.path.join(GOLDEN, 'jupiter.benchmark_gas')) as fp:
        return fp.read()
