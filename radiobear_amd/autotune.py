"""Launch-shape autotuner for the absorption, tangent and integration kernels.

For a problem shape it times each candidate value of one launch knob on the
device (HIP events on the launch stream, after a warm-up) and returns the
launch shape with the fastest one.  Nothing is set in the library: a launch
shape is a plain {field: value} dict (the fields of ``rbx_launch``) that the
caller passes to every call it is meant for, so calls with different shapes on
other threads or streams are not affected.  Results are cached per shape in
memory and, optionally, in a JSON file so a production run tunes once.

Knobs (0 = the library's built-in choice):
* ``alpha_k``      frequencies per lane of rbx_alpha (1, 2, 4): wider tiles
                   amortise the LDS reads of each line over more frequencies,
                   narrower tiles waste fewer lanes and give more blocks for
                   the tail of the grid;
* ``alpha_series`` line sums of rbx_alpha: 1 every line at every frequency, 2 per
                   wavefront tile only the lines near it, the others as power series;
* ``alpha_split``  parts each layer's direct line sums are split into: more parts
                   give more, shorter blocks, so the last wave of blocks idles less;
* ``far_split``    parts of each layer's lines in the series' classification pass;
* ``tan_split``    parts of each layer's lines in rbx_alpha_tangent (small layer
                   counts need several parts to fill the device);
* ``int_unroll``   path steps per load batch of the integration (4, 8);
* ``int_map``      block order of the integration: 1 frequency group major, 2 ray
                   major (every frequency group of a ray block on one XCD).
"""
import json
import os

import torch

from . import _lib

CANDIDATES = {'alpha_k': (1, 2, 4), 'alpha_series': (1, 2), 'alpha_split': (0, 1, 2, 3, 4, 6, 8),
              'far_split': (0, 1, 2, 4, 8), 'tan_split': (0, 1, 2, 4, 8, 16), 'int_unroll': (4, 8),
              'int_map': (1, 2)}
_cache = {}


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


def _load(path):
    if path and os.path.exists(path):
        with open(path) as fp:
            return json.load(fp)
    return {}


def tune(name, shape_key, fn, launch=None, reps=3, path=None, candidates=None):
    """Time fn(launch) for every candidate value of knob `name` (the other fields as in
    `launch`) and return (launch with the fastest value of `name`, record).

    fn(launch) must launch the kernel with the given {field: value} dict.  The record
    holds 'best' and the time of each candidate in ms ('ms')."""
    if name not in CANDIDATES:
        raise ValueError('unknown launch knob {}'.format(name))
    base = dict(launch or {})
    key = '{}:{}:{}'.format(name, shape_key, json.dumps(base, sort_keys=True))
    rec = _cache.get(key) or _load(path).get(key)
    if rec is None:
        times = {}
        for v in (candidates or CANDIDATES[name]):
            trial = dict(base, **{name: v})
            times[v] = time_call(lambda: fn(trial), reps=reps)
        best = min(times, key=times.get)
        rec = {'best': best, 'ms': {str(k): t for k, t in times.items()}}
        if path:
            disk = _load(path)
            disk[key] = rec
            with open(path, 'w') as fp:
                json.dump(disk, fp, indent=1, sort_keys=True)
    _cache[key] = rec
    return dict(base, **{name: rec['best']}), rec


def tune_all(names, shape_key, fn, launch=None, reps=3, path=None):
    """tune() each knob of `names` in turn, each starting from the shape the previous ones
    chose: (launch, {name: record})."""
    recs = {}
    for n in names:
        launch, recs[n] = tune(n, shape_key, fn, launch, reps=reps, path=path)
    return launch, recs


def launch_struct(launch):
    """The ctypes rbx_launch of a {field: value} dict (None: built in)."""
    return _lib.launch_struct(launch)
