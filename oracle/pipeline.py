"""End-to-end FP64 oracle runs (TEST INFRASTRUCTURE ONLY).

``jupiter_benchmark`` rebuilds scripts/benchmark.py:9-16 (Planet('jupiter',
config_file='config_benchmark.par').run(freqs, b)) from the restated pieces:
the benchmark atmosphere, Alpha.get_layers and Brightness.single per ray.
"""
import numpy as np

from . import absorbers, atmosphere, geometry, layers, rt

JUPITER_REQ = 71492.0      # default_config.json 'req' (Jupiter)
JUPITER_RPOL = 66854.0     # default_config.json 'rpol' (Jupiter)
BENCH_FREQS = [0.6, 1.25, 2.6, 5.2, 10, 21.9]                     # benchmark.py:15
BENCH_B = [[0.0, 0.0], [np.sin(np.radians(15)), 0.0],
           [np.sin(np.radians(30)), 0.0], [np.sin(np.radians(45)), 0.0]]   # benchmark.py:16


def jupiter_lines(raw):
    """Jupiter line lists with the default truncation (default_config.json:48-71).

    raw: dict with 'ammonia', 'nh3', 'h2s', 'ph3jh', 'PH3WGT' npz contents."""
    return {
        'nh3_hs': {k: np.asarray(v, dtype=np.float64) for k, v in raw['ammonia'].items()},
        'nh3_sjs': {k: np.asarray(v, dtype=np.float64) for k, v in raw['nh3'].items()},
        'h2s': absorbers.truncate_ftable(raw['h2s'], truncate_strength=1e-22),
        'ph3': absorbers.truncate_ph3(raw['ph3jh'], raw['PH3WGT'], truncate_strength=1e-22),
        'h2o': absorbers.h2o_bk_lines(),
    }


def benchmark_atmosphere(gas_text):
    """Jupiter/config_benchmark.par: regrid 1000 layers over 0.01-10000 bar, no tweak."""
    C = atmosphere.C_JUPITER
    gas = atmosphere.read_gas(gas_text, C)
    gas = atmosphere.regrid_logp(gas, C, 0.01, 10000.0, 1000)
    gas = atmosphere.jupiter_no_tweak(gas, C)
    prop = atmosphere.compute_prop(gas, C, JUPITER_REQ)
    return gas, prop


def run(gas, prop, C, absorber_set, freqs, bs, Req, Rpol, orientation=(0.0, 0.0),
        gtype='ellipse', disc=False):
    """Planet.run for a list of rays: alpha once, then one Brightness.single per ray."""
    alpha = layers.get_layers(freqs, gas, C, absorber_set)
    Tb = []
    rays = []
    for b in bs:
        ray = geometry.compute_ds(prop['R'], prop['N'], b, Req, Rpol, orientation, gtype)
        rays.append(ray)
        Tb.append(rt.single(alpha, gas[C['T']], ray.ds, ray.layer4ds, disc=disc))
    return np.array(Tb), alpha, rays


def jupiter_benchmark(gas_text, raw_lines):
    gas, prop = benchmark_atmosphere(gas_text)
    aset = layers.jupiter_absorbers(jupiter_lines(raw_lines), h2state='e')
    return run(gas, prop, atmosphere.C_JUPITER, aset, BENCH_FREQS, BENCH_B,
               JUPITER_REQ, JUPITER_RPOL)
