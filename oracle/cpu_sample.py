"""Timing sample of the FP64 oracle on the host cores (BASELINE INFRASTRUCTURE ONLY).

The CPU baseline beside the GPU figures: the oracle (the vectorised numpy
restatement of the reference's absorption and Brightness.single) timed on a
bounded sample of a workload -- a few layers of absorption at the full line
and frequency counts, a few on-planet rays of geometry + integration at the
full frequency count -- and extrapolated linearly to the whole workload.

single_core() runs the sample in this process.  all_core() forks `workers`
processes (before the calling process has touched the GPU: fork only, no
exec) that each time their own share of layers and rays at the same time, so
contention for memory bandwidth is in the figure; the whole workload's time
on the pool is the summed per-item times divided by the worker count.
"""
import multiprocessing as mp
import time

import numpy as np

from . import geometry, layers, pipeline, rt

_STATE = None   # (atm, freqs, rays, on_planet indices, absorber set), set before forking


def setup(name):
    """The workload's atmosphere, frequencies, rays and the oracle absorber set with the
    workload's synthetic line lists injected (as the bench's catalogue holds them)."""
    from radiobear_amd import lines, synth
    w = synth.workload(name)
    atm = w['atm']
    jl = pipeline.jupiter_lines(lines.raw_catalog())
    inj = dict(jl['nh3_hs'])
    for k in ('fo', 'Io', 'Eo', 'gammaNH3o'):
        inj[k] = w['nh3_inv'][k]
    jl['nh3_hs'] = inj
    if w['h2o'] is not None:
        jl['h2o'] = w['h2o']
    aset = layers.jupiter_absorbers(jl)
    rays = np.asarray(w['rays'], dtype=np.float64)
    onp = [i for i in range(len(rays)) if np.hypot(*rays[i]) < 0.99]
    return atm, np.asarray(w['freqs'], dtype=np.float64), rays, onp, aset


def _time_items(state, layer_idx, ray_idx):
    atm, freqs, rays, _, aset = state
    t0 = time.perf_counter()
    for l in layer_idx:
        aset.absorb(list(freqs), atm.T[l], atm.P[l], atm.gas[:, l], atm.C)
    t1 = time.perf_counter()
    # integration cost does not depend on the alpha values (no early exit in the oracle)
    alpha_tab = np.full((len(freqs), atm.nl), 1e-7)
    for i in ray_idx:
        ray = geometry.compute_ds(atm.R, atm.N, rays[i], atm.Req, atm.Rpol)
        rt.single(alpha_tab, atm.T, ray.ds, ray.layer4ds)
    t2 = time.perf_counter()
    return t1 - t0, len(layer_idx), t2 - t1, len(ray_idx)


def _worker(task):
    return _time_items(_STATE, *task)


def single_core(state, n_layers=2, n_rays=3, seed=0):
    """Per-item times of the sample run in this process."""
    atm, _, _, onp, _ = state
    sel_l = [0, atm.nl // 2][:n_layers]
    rng = np.random.default_rng(seed)
    sel_r = list(rng.choice(onp, size=min(n_rays, len(onp)), replace=False))
    t0 = time.perf_counter()
    tl, nl_, tr, nr = _time_items(state, sel_l, sel_r)
    return {'t_layer': tl / nl_, 't_ray': tr / nr, 'n_layers': nl_, 'n_rays': nr, 'cores': 1,
            'wall': time.perf_counter() - t0, 'cpu_seconds': tl + tr}


def all_core(state, workers, layers_per=2, rays_per=4, seed=1):
    """Per-item times with `workers` processes each timing layers_per layers (spread over
    the depth) and rays_per random on-planet rays, all at once.  Must run before the
    calling process initialises the GPU (the pool forks it)."""
    global _STATE
    atm, _, _, onp, _ = state
    rng = np.random.default_rng(seed)
    lay = np.linspace(0, atm.nl - 1, workers * layers_per).astype(int)
    ray = rng.choice(onp, size=min(workers * rays_per, len(onp)), replace=False)
    tasks = [(list(lay[w::workers]), list(ray[w::workers])) for w in range(workers)]
    _STATE = state
    ctx = mp.get_context('fork')
    t0 = time.perf_counter()
    with ctx.Pool(workers) as pool:
        res = pool.map(_worker, tasks)
    wall = time.perf_counter() - t0
    _STATE = None
    nl_, nr = sum(r[1] for r in res), sum(r[3] for r in res)
    return {'t_layer': sum(r[0] for r in res) / nl_, 't_ray': sum(r[2] for r in res) / nr,
            'n_layers': nl_, 'n_rays': nr, 'cores': workers, 'wall': wall,
            'cpu_seconds': sum(r[0] + r[2] for r in res)}


def baseline(meas, nl, nf, on_planet):
    """The cpu_baseline record: the workload's on-planet rays x frequencies over the
    extrapolated time (layers and rays shared evenly over the cores)."""
    total = (meas['t_layer'] * nl + meas['t_ray'] * on_planet) / meas['cores']
    sample = ('{} process(es) timing {} layers of alpha at full lines x freqs ({:.2f} s/layer) and {} rays '
              'of geometry + integration at full freqs x steps ({:.3f} s/ray) at once ({:.1f} s of CPU work, '
              '{:.1f} s wall), extrapolated to {} layers and {} on-planet rays shared over the {} '
              'process(es)').format(meas['cores'], meas['n_layers'], meas['t_layer'], meas['n_rays'],
                                    meas['t_ray'], meas['cpu_seconds'], meas['wall'], nl, on_planet,
                                    meas['cores'])
    return {'value': on_planet * nf / total, 'unit': 'ray*freq Tb/s', 'cores': meas['cores'], 'kind': 'port',
            'sample': sample, 'seconds_extrapolated': total}
