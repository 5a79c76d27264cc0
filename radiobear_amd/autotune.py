"""Launch-shape autotuner for the absorption and integration kernels.

For a problem shape it times each candidate launch configuration on the
device (HIP events on the launch stream, after a warm-up), keeps the fastest
and pins it with ``rbx_tune``.  Results are cached per shape in memory and,
optionally, in a JSON file so a production run tunes once.

Knobs:
* ``alpha_k``     frequencies per lane of rbx_alpha (1, 2, 4): wider tiles
                  amortise the LDS reads of each line over more frequencies,
                  narrower tiles waste fewer lanes and give more blocks for
                  the tail of the grid;
* ``alpha_split`` parts each layer's line sums are split into (0 = the
                  library's model of the grid tail, or 1, 2, 3, 4, 6, 8): more parts
                  give more, shorter blocks, so the last wave of blocks idles less;
* ``int_unroll``  path steps per load batch of rbx_integrate (4, 8);
* ``int_map``     block order of rbx_integrate: 1 frequency group major, 2 ray major
                  (every frequency group of a ray block on one XCD);
* ``alpha_series`` line sums of rbx_alpha: 1 every line at every frequency, 2 per
                  wavefront tile only the lines near it, the others as power series
                  (the built-in choice takes 2 from 4 tiles up while the workspace
                  stays under 4 GiB).
"""
import json
import os

import torch

from . import _lib

CANDIDATES = {'alpha_k': (1, 2, 4), 'alpha_split': (0, 1, 2, 3, 4, 6, 8), 'int_unroll': (4, 8), 'int_map': (1, 2),
              'alpha_series': (1, 2)}
_KEYS = {'alpha_k': _lib.TUNE_ALPHA_K, 'alpha_split': _lib.TUNE_ALPHA_SPLIT, 'int_unroll': _lib.TUNE_INT_UNROLL,
         'int_map': _lib.TUNE_INT_MAP, 'alpha_series': _lib.TUNE_ALPHA_SERIES}
_cache = {}


def set_knob(name, value):
    _lib.check(_lib.lib().rbx_tune(_KEYS[name], int(value)))


def reset():
    for name in _KEYS:
        set_knob(name, 0)


def time_call(fn, reps=3, warmup=1):
    """Average ms of fn() over reps launches (events on the current stream)."""
    for _ in range(warmup):
        fn()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(reps):
        fn()
    stop.record()
    torch.cuda.synchronize()
    return start.elapsed_time(stop) / reps


def tune(name, shape_key, fn, reps=3, path=None):
    """Pick the fastest candidate of knob `name` for fn() (which launches the kernel)."""
    key = '{}:{}'.format(name, shape_key)
    if key in _cache:
        set_knob(name, _cache[key]['best'])
        return _cache[key]
    if path and os.path.exists(path):
        with open(path) as fp:
            disk = json.load(fp)
        if key in disk:
            _cache[key] = disk[key]
            set_knob(name, disk[key]['best'])
            return disk[key]
    times = {}
    for v in CANDIDATES[name]:
        set_knob(name, v)
        times[v] = time_call(fn, reps=reps)
    best = min(times, key=times.get)
    set_knob(name, best)
    rec = {'best': best, 'ms': {str(k): t for k, t in times.items()}}
    _cache[key] = rec
    if path:
        disk = {}
        if os.path.exists(path):
            with open(path) as fp:
                disk = json.load(fp)
        disk[key] = rec
        with open(path, 'w') as fp:
            json.dump(disk, fp, indent=1, sort_keys=True)
    return rec
