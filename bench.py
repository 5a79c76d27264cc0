"""Benchmark: batched forward model (absorption + ray paths + integration) on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c3_large]

One step = the whole hot path over one batch of synthetic input already resident
in HBM: rbx_alpha (line sums), rbx_raypaths (FP64 geometry), rbx_integrate
(FP64 tau scan), captured once in a HIP graph and replayed.  Prints ONE JSON
line.  ``value`` is on-planet ray x frequency Tb values per second in the
FP32-line-sum mode; the plain FP64 mode, the Jacobian workload (rows/s), the
per-kernel times, the roofline of the dominant kernel and the CPU baseline
(the FP64 oracle on a bounded sample) are reported beside it.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

T_CMB = 2.725   # K (utils.py:77): the floor Brightness.single puts under a ray's sum
PEAK = {'fp32': 157.3, 'fp64': 78.6, 'hbm': 8000.0}   # TFLOP/s, TFLOP/s, GB/s (MI355X)
# algorithmic FLOP per line x layer x frequency (division = 1 FLOP)
FLOP_BR, FLOP_GROSS, FLOP_VVW = 9, 7, 12
FLOP_RT_STEP = 11         # 10 FLOP + 1 exp per ray x freq x step
CLOSED_FORM_FLOP = 20     # h2 CIA / h2o continuum per layer x frequency


def parse():
    p = argparse.ArgumentParser()
    p.add_argument('--gpus', type=int, default=1)
    p.add_argument('--steps', type=int, default=20)
    p.add_argument('--warmup', type=int, default=3)
    p.add_argument('--workload', default='c3_large')
    p.add_argument('--no-cpu-baseline', action='store_true')
    p.add_argument('--no-secondary', action='store_true')
    p.add_argument('--quiet', action='store_true')
    p.add_argument('--no-tune', action='store_true', help='use the built-in launch shapes (no autotuning)')
    p.add_argument('--tune-cache', default=None,
                   help='JSON file of tuned launch shapes: read if it holds this shape, else written')
    p.add_argument('--copy-calibration', action='store_true',
                   help='run only a plain device copy of 1 GiB (calibrates FETCH_SIZE / WRITE_SIZE)')
    p.add_argument('--only', choices=('jacobian', 'retrieval', 'adjoint'), default=None,
                   help='run only that record (for profiling its kernels) and print it')
    return p.parse_args()


def copy_calibration():
    """A plain copy kernel of known size: 2^27 float64 = 1 GiB read + 1 GiB written."""
    import torch
    a = torch.ones(2 ** 27, dtype=torch.float64, device='cuda')
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({'copy_bytes': a.numel() * 8, 'launches': 3}))


def log(args, *a):
    if not args.quiet:
        print(*a, file=sys.stderr, flush=True)


class Problem:
    """Device-resident inputs and outputs of one workload."""

    def __init__(self, name):
        import torch
        from radiobear_amd import kernels as K, lines, synth
        self.K = K
        w = synth.workload(name)
        self.w = w
        self.w_name = name
        atm = w['atm']
        self.atm = atm
        raw = lines.raw_catalog()
        self.cat = K.jupiter_catalog(raw, nh3_inversion=w['nh3_inv'], h2o=w['h2o'])
        self.ab = K.absorbers_struct({'h2': 'h2_jj_ddb', 'h2o': 'h2o_bk', 'h2s': 'h2s_ddb',
                                      'nh3': 'nh3_hs_sjs', 'ph3': 'ph3_jh'}, 'e')
        dev = K.device()
        self.freqs = np.asarray(w['freqs'], dtype=np.float64)
        self.nf = len(self.freqs)
        self.nl = atm.nl
        self.rays = w['rays']
        self.nray = len(self.rays)
        C = atm.C
        self.P = K.as_device(atm.P)
        self.T = K.as_device(atm.T)
        self.mix = K.as_device(K.mix_table(atm.gas, C))
        self.R = K.as_device(atm.R)
        self.N = K.as_device(atm.N)
        self.b = K.as_device(self.rays)
        order, self.nf_low = K.nh3_order(self.freqs)
        assert np.array_equal(order, np.arange(self.nf))      # ascending request
        self.f = K.as_device(self.freqs)
        self.planet = K.planet_desc(atm.Req, atm.Rpol)
        self.alpha32 = torch.empty((self.nl, self.nf), dtype=torch.float32, device=dev)
        self.alpha64 = torch.empty((self.nl, self.nf), dtype=torch.float64, device=dev)
        self.ld = self.nl - 1
        self.hds = torch.empty((self.nray, self.ld), dtype=torch.float64, device=dev)
        self.ns = torch.empty((self.nray,), dtype=torch.int32, device=dev)
        self.Tb = torch.empty((self.nray, self.nf), dtype=torch.float64, device=dev)
        self.Tb64 = torch.empty((self.nray, self.nf), dtype=torch.float64, device=dev)
        self.nlines = {k: self.cat.lines(k) for k in range(7)}
        self.ws = None
        self.side = None
        self.launch = {}    # precision -> launch shape ({field: value}) used by step()
        # integration in the order the frequencies pass the tau cut (rbx_integrate_order):
        # precision -> on / off, and its buffers
        self.ordered = {}
        self.cols = torch.empty((self.nf,), dtype=torch.int32, device=dev)
        self.alpha32_cols = torch.empty_like(self.alpha32)
        self.alpha64_cols = torch.empty_like(self.alpha64)
        self.factors = torch.empty((self.nray, self.ld, 4), dtype=torch.float32, device=dev)   # FP32 point rays

    # the three launches of one step, on torch's current stream
    def launch_alpha(self, prec, launch=None):
        import ctypes
        from radiobear_amd import _lib
        K = self.K
        out = self.alpha32 if prec == _lib.FP32 else self.alpha64
        ls = _lib.launch_struct(self.launch.get(prec) if launch is None else launch)
        self.ws, nbytes = K.alpha_workspace(self.cat, self.ab, self.nl, self.nf, self.nf_low, prec, reuse=self.ws,
                                            launch=ls)
        _lib.check(_lib.lib().rbx_alpha(self.cat.handle, ctypes.byref(self.ab), K._ptr(self.P), K._ptr(self.T),
                                        K._ptr(self.mix), None, self.nl, K._ptr(self.f), None, self.nf, self.nf_low,
                                        None, prec, K._ptr(out), None, _lib.launch_ref(ls), K._ptr(self.ws), nbytes,
                                        K._stream()))

    def launch_paths(self):
        import ctypes
        from radiobear_amd import _lib
        K = self.K
        _lib.check(_lib.lib().rbx_raypaths(ctypes.byref(self.planet), K._ptr(self.R), K._ptr(self.N), self.nl,
                                           K._ptr(self.b), self.nray, 1, K._ptr(self.hds), self.ld,
                                           K._ptr(self.ns), K._stream()))

    def launch_factors(self):
        """Each ray's trapezoid factors for the FP32 cut-order integration (paths and T only)."""
        from radiobear_amd import _lib
        K = self.K
        _lib.check(_lib.lib().rbx_integrate_factors(K._ptr(self.T), self.nl, K._ptr(self.hds), self.ld,
                                                    K._ptr(self.ns), self.nray, K._ptr(self.factors), K._stream()))

    def uses_factors(self, prec):
        from radiobear_amd import _lib
        return prec == _lib.FP32 and self.ordered.get(prec, False)

    def launch_integrate(self, prec, launch=None, tau_cut=None, Tb=None, ordered=None, factors_ready=False):
        from radiobear_amd import _lib
        K = self.K
        a = self.alpha32 if prec == _lib.FP32 else self.alpha64
        if Tb is None:
            Tb = self.Tb if prec == _lib.FP32 else self.Tb64
        if tau_cut is None:
            tau_cut = K.FP32_TAU_CUT if prec == _lib.FP32 else K.FP64_TAU_CUT
        if ordered is None:
            ordered = self.ordered.get(prec, False)
        ls = _lib.launch_struct(self.launch.get(prec) if launch is None else launch)
        if ordered and tau_cut > 0.0:
            ac = self.alpha32_cols if prec == _lib.FP32 else self.alpha64_cols
            _lib.check(_lib.lib().rbx_integrate_order(K._ptr(a), prec, self.nl, self.nf, K._ptr(self.hds), self.ld,
                                                      K._ptr(self.ns), self.nray, tau_cut, K._ptr(self.cols),
                                                      K._ptr(ac), K._stream()))
            if prec == _lib.FP32 and not factors_ready:
                self.launch_factors()
            _lib.check(_lib.lib().rbx_integrate_cols(K._ptr(ac), prec, K._ptr(self.cols), K._ptr(self.T), self.nl,
                                                     self.nf, K._ptr(self.hds), self.ld, K._ptr(self.ns), self.nray,
                                                     0, prec, tau_cut, K._ptr(Tb), None, K._ptr(self.factors),
                                                     _lib.launch_ref(ls), K._stream()))
            return
        _lib.check(_lib.lib().rbx_integrate(K._ptr(a), prec, K._ptr(self.T), self.nl, self.nf, K._ptr(self.hds),
                                            self.ld, K._ptr(self.ns), self.nray, 0, prec, tau_cut, K._ptr(Tb), None,
                                            _lib.launch_ref(ls), K._stream()))

    def launch_integrate_nocut(self, prec):
        """The integration without the optical-depth cut, into the FP32 Tb buffer."""
        self.launch_integrate(prec, tau_cut=0.0, Tb=self.Tb)

    def step(self, prec):
        """Absorption and ray paths are independent: the paths (a few hundred
        wavefronts, latency-bound) and, for the FP32 cut-order integration, each ray's
        trapezoid factors run on a side stream beside the line sums, and the
        integration waits for both."""
        import torch
        main = torch.cuda.current_stream()
        if self.side is None:
            self.side = torch.cuda.Stream()
        fac = self.uses_factors(prec)
        self.side.wait_stream(main)
        with torch.cuda.stream(self.side):
            self.launch_paths()
            if fac:
                self.launch_factors()
        self.launch_alpha(prec)
        main.wait_stream(self.side)
        self.launch_integrate(prec, factors_ready=fac)

    def alpha_flop(self):
        n = self.nlines
        per_lf = (FLOP_BR * (n[0] + n[4] + n[5]) + FLOP_GROSS * (n[1] + n[2]) + FLOP_VVW * n[6]
                  + 2 * CLOSED_FORM_FLOP)
        return float(per_lf) * self.nl * self.nf


def autotune(prob, prec, args):
    """Pick the launch shapes of rbx_alpha / rbx_integrate for this shape (outside the timed
    region) and keep them in prob.launch[prec] for step()."""
    import torch
    from radiobear_amd import autotune as at
    if args.no_tune:
        prob.launch[prec] = {}
        return None
    key = '{}:{}x{}x{}:p{}'.format(args.workload, prob.nl, prob.nf, prob.nray, prec)
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        prob.launch_paths()
        # far-line series (1 off, 2 on), then the tile width and the splits for that path
        la, rec = at.tune_all(('alpha_series', 'alpha_k'), key, lambda L: prob.launch_alpha(prec, L),
                              path=args.tune_cache)
        split = 'far_split' if la['alpha_series'] == 2 else 'alpha_split'
        la, rec_s = at.tune(split, key, lambda L: prob.launch_alpha(prec, L), la, path=args.tune_cache)
        li, rec_i = at.tune_all(('int_unroll', 'int_map'), key, lambda L: prob.launch_integrate(prec, L),
                                path=args.tune_cache)
        # the integration in request order or in cut order (order + reordered table included),
        # on this step's own table and paths (a tune cache skips the trial launches above)
        prob.launch[prec] = dict(la, **li)
        prob.launch_paths()
        prob.launch_alpha(prec)
        order_ms = {}
        for on in (False, True):
            prob.launch_integrate(prec, ordered=on)
            ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
            ev[0].record()
            for _ in range(5):
                prob.launch_integrate(prec, ordered=on)
            ev[1].record()
            torch.cuda.synchronize()
            order_ms[str(int(on))] = ev[0].elapsed_time(ev[1]) / 5
    torch.cuda.synchronize()
    prob.ordered[prec] = order_ms['1'] < order_ms['0']
    return {'alpha_path': 'far-line series' if la['alpha_series'] == 2 else 'direct line sums',
            'launch': prob.launch[prec], 'alpha_series_ms': rec['alpha_series']['ms'],
            'alpha_k_ms': rec['alpha_k']['ms'], split + '_ms': rec_s['ms'],
            'int_unroll_ms': rec_i['int_unroll']['ms'], 'int_map_ms': rec_i['int_map']['ms'],
            'int_order': int(prob.ordered[prec]), 'int_order_ms': order_ms}


def other_workloads(args):
    """The nadir-ray case (workload 1: 150 layers, 2000 NH3 lines, 20 freqs, FP64 plain; launch
    bound, so its absolute time is the figure) and the image case (workload 2: 400 layers,
    10000 lines, 512 freqs, 64x64 pixels; FP64 and FP32), one graph-replayed step each, with the
    library's built-in launch choices (the knobs tuned for workload 3 are cleared first)."""
    from radiobear_amd import _lib
    out = {}
    for name, precs in (('c1_plain_fp64', ('fp64',)), ('c2_image', ('fp64', 'fp32'))):
        prob = Problem(name)
        rec = {'layers': prob.nl, 'freqs': prob.nf, 'rays': prob.nray}
        for p in precs:
            prec = _lib.PRECISIONS[p]
            dt, _ = timed_graph(prob, prec, 20, 3)
            on = int((prob.ns.cpu().numpy() >= 0).sum())
            rec[p] = {'ms_per_step': dt * 1e3, 'ray_freq_Tb_per_s': on * prob.nf / dt, 'on_planet_rays': on}
        out[name] = rec
    return out


def timed_graph(prob, prec, steps, warmup):
    import torch
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        prob.step(prec)                       # eager warm-up (also the reference result)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=stream):
        prob.step(prec)
    for _ in range(warmup):
        g.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        g.replay()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps, g


def kernel_times(prob, prec, steps):
    """Average duration of each launch, HIP events on the launch stream."""
    import torch
    fac = prob.uses_factors(prec)
    names = ('rbx_alpha', 'rbx_raypaths') + (('rbx_integrate_factors',) if fac else ()) + ('rbx_integrate',)
    acc = {n: 0.0 for n in names}
    for _ in range(steps):
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(len(names) + 1)]
        ev[0].record()
        prob.launch_alpha(prec)
        ev[1].record()
        prob.launch_paths()
        ev[2].record()
        if fac:
            prob.launch_factors()
            ev[3].record()
        prob.launch_integrate(prec, factors_ready=fac)
        ev[-1].record()
        torch.cuda.synchronize()
        for i, n in enumerate(names):
            acc[n] += ev[i].elapsed_time(ev[i + 1])
    return {n: acc[n] / steps for n in names}


def cpu_workers():
    """Host processes for the all-core CPU baseline: the cores this process may run on,
    at most 16 (a one-GPU box's CPU share)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_measure(args):
    """Per-item oracle times on one core and on all cores (taken before the GPU is touched:
    the all-core pool forks this process)."""
    from oracle import cpu_sample
    state = cpu_sample.setup(args.workload)
    one = cpu_sample.single_core(state)
    allc = cpu_sample.all_core(state, cpu_workers())
    return one, allc


def cpu_baseline(prob, meas):
    """The FP64 oracle on a bounded sample, extrapolated to the whole workload: all cores
    (the value) and one core."""
    from oracle import cpu_sample
    one, allc = meas
    rec = cpu_sample.baseline(allc, prob.nl, prob.nf, prob.on_planet)
    rec['single_core'] = cpu_sample.baseline(one, prob.nl, prob.nf, prob.on_planet)
    return rec


def tangent_flop(cat, nl, nf):
    """Algorithmic FLOP of one tangent pass: per line, layer and frequency the forward count
    (Ben-Reuven 9, Gross 7, VVW 12) plus 6 per tangent for the two tangents each line has
    (d/dT and d/dq of its own species: dN 2, dQ 2, (dN - t dQ) / Q 2 with the reciprocal
    shared), plus the closed forms."""
    n = {k: cat.lines(k) for k in range(7)}
    per = (FLOP_BR * (n[0] + n[4] + n[5]) + FLOP_GROSS * (n[1] + n[2]) + FLOP_VVW * n[6]
           + 12 * (n[0] + n[1] + n[2] + n[4] + n[5] + n[6]) + 2 * CLOSED_FORM_FLOP)
    return float(per) * nl * nf


def _rows_rate(rb, reps, rebuild=False, tune_key=None, args=None):
    """Forward + adjoint Jacobian rows/s of a RetrievalBatch: the step (with rebuild, the
    shells rebuilt on the device first) captured in a HIP graph and replayed; per-stage
    times from HIP events on the launch stream; the tangent pass's split autotuned first."""
    import torch
    from radiobear_amd import autotune as at, kernels as K
    rec = {}
    if tune_key is not None and not args.no_tune:
        cand = (1, 2, 4, 8, 16)

        def tangent(L):
            K.alpha_tangent(rb.catalog, rb.absorbers, rb.P, rb.T, rb.mix, rb.freqs, precision=rb.precision,
                            with_alpha=True, launch=L)
        rb.launch, r = at.tune('tan_split', tune_key, tangent, rb.launch, reps=2, path=args.tune_cache,
                               candidates=cand)
        rec['tan_split_ms'] = r['ms']
    graph, (Tb, J) = rb.capture(rebuild)
    graph.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        graph.replay()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    # one eager pass split by stage
    st = [torch.cuda.Event(enable_timing=True) for _ in range(5)]
    st[0].record()
    if rebuild:
        rb.update()
    st[1].record()
    alpha, dal = rb.absorption()
    st[2].record()
    hds, ns = rb.paths()
    st[3].record()
    K.jacobian_batch(alpha, dal, rb.T, rb.off_d, rb.ray_atm_d, hds, ns, rb.ldj, order=rb.order_d,
                     tau_cut=rb.tau_cut, precision=rb.precision)
    st[4].record()
    torch.cuda.synchronize()
    stages = {'rbx_atmosphere_batch': st[0].elapsed_time(st[1]), 'rbx_alpha_tangent': st[1].elapsed_time(st[2]),
              'rbx_raypaths_batch': st[2].elapsed_time(st[3]), 'rbx_jacobian_batch': st[3].elapsed_time(st[4])}
    if not rebuild:
        stages.pop('rbx_atmosphere_batch')
    flop = tangent_flop(rb.catalog, rb.nl_total, len(rb.freqs))
    tan_s = stages['rbx_alpha_tangent'] * 1e-3
    roof = {'kernel': 'rbx_alpha_tangent', 'bound': 'valu', 'achieved': flop / tan_s / 1e12, 'peak': PEAK['fp32'],
            'unit': 'TFLOP/s', 'frac': flop / tan_s / 1e12 / PEAK['fp32'], 'algorithmic_flop_per_launch': flop,
            'basis': 'forward line-shape FLOP plus 6 per tangent, two tangents per line, every (line, layer, '
                     'frequency)'}
    rec.update({'rows_per_s': rb.rows() / dt, 'ms_per_step': dt * 1e3, 'rows': rb.rows(), 'stage_ms': stages,
                'graph': 'hipGraph replay of the whole step', 'launch': rb.launch, 'roofline': roof,
                'nan_rows': int(torch.isnan(Tb).any(dim=1).sum().item())})
    return rec


def jacobian_rate(args):
    """Workload 4 (500 layers, 10000 NH3 lines, 64 freqs, 256 rays): forward + adjoint
    Jacobian rows/s (dTb/dT, dTb/dq_NH3, dTb/dq_H2O per layer), fused FP32 alpha+tangent pass."""
    from radiobear_amd import kernels as K, lines, synth
    from radiobear_amd.retrieval import RetrievalBatch
    w = synth.workload('c4_jacobian')
    atm = w['atm']
    cat = K.jupiter_catalog(lines.raw_catalog(), nh3_inversion=w['nh3_inv'])
    ab = K.absorbers_struct({'h2': 'h2_jj_ddb', 'h2o': 'h2o_bk', 'h2s': 'h2s_ddb', 'nh3': 'nh3_hs_sjs',
                             'ph3': 'ph3_jh'}, 'e')
    rb = RetrievalBatch([atm], cat, ab, w['freqs'], w['rays'], atm.Req, atm.Rpol, atm.C, precision='fp32')
    out = _rows_rate(rb, 20, tune_key='c4_jacobian:fp32', args=args)
    out.update({'workload': 'c4_jacobian', 'entries_per_row': 3 * atm.nl,
                'precision': 'fp32 line sums and tangents, FP64 tau and adjoint sweep'})
    return out


def retrieval_rate(args):
    """Workload 5: 1024 perturbed atmospheres (200-800 layers each), 10000 NH3 lines,
    128 freqs, 16 rays each, forward + adjoint in one batch: Jacobian rows/s."""
    from radiobear_amd import retrieval
    rb = retrieval.synthetic(n_atm=1024, precision='fp32', device_props=True)
    # 1024 atmospheres x 2 tiles are a million tangent wavefronts: one part per layer (built in)
    out = _rows_rate(rb, 3, rebuild=True, args=args)
    out.update({'workload': 'c5_retrieval', 'atmospheres': rb.natm, 'layers_total': rb.nl_total,
                'rays': rb.nray, 'freqs': len(rb.freqs), 'precision': 'fp32 line sums and tangents',
                'step': 'atmosphere shells (hydrostatic z, R, N) + alpha and tangents + ray paths + '
                        'forward/adjoint Jacobian, all on the device'})
    return out


def adjoint_rate(prob, dt64, args):
    """The north-star comparison on the large workload: FP32-line-sum forward + adjoint
    (the gradient of a scalar loss sum g Tb with respect to T, q_NH3 and q_H2O of every layer,
    through adjoint.ForwardModel: rbx_alpha_tangent + rbx_integrate, then rbx_integrate_adjoint
    + rbx_alpha_vjp) against the plain FP64 forward + central-difference Jacobian (2 forwards
    per layer and variable, 3 variables, plus the base forward), priced at the measured FP64
    graph step.  dL/dT from central differences of the FP64 forward on two layers is compared
    with the adjoint's, (ray, frequency) pairs on the T_cmb floor left out of both."""
    import torch
    from radiobear_amd import _lib
    from radiobear_amd.adjoint import ForwardModel
    FP64 = _lib.FP64
    model = ForwardModel(prob.cat, prob.ab, prob.P, prob.mix, prob.freqs, prob.hds, prob.ns, precision='fp32')
    gen = torch.Generator(device='cuda').manual_seed(3)
    g = torch.randn((prob.nray, prob.nf), generator=gen, device='cuda', dtype=torch.float64)
    g[prob.ns < 0] = 0.0
    nan_rows = torch.isnan(prob.Tb64).any(dim=1)
    g[nan_rows] = 0.0                        # the loss leaves out the rays the geometry loses (NaN Tb)
    NH3, H2O = _lib.SPECIES.index('NH3'), _lib.SPECIES.index('H2O')

    def step(weights):
        x = [prob.T.clone().requires_grad_(True), prob.mix[NH3].clone().requires_grad_(True),
             prob.mix[H2O].clone().requires_grad_(True)]
        Tb = model(*x)
        (Tb * weights).nan_to_num(0.0).sum().backward()
        return Tb, [v.grad for v in x]

    step(g)
    torch.cuda.synchronize()
    reps = 3
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record()
    for _ in range(reps):
        Tb32, gr = step(g)
    ev[1].record()
    torch.cuda.synchronize()
    t_adj = ev[0].elapsed_time(ev[1]) / reps * 1e-3
    # stages of one pass (HIP events on the launch stream)
    T0 = prob.T.clone()
    se = [torch.cuda.Event(enable_timing=True) for _ in range(5)]
    se[0].record()
    alpha, dal = model.absorption(T0, prob.mix[NH3], prob.mix[H2O])
    se[1].record()
    model.integrate(alpha, T0)
    se[2].record()
    galpha, gT = model.adjoint(alpha, T0, g)
    se[3].record()
    from radiobear_amd import kernels as K
    K.alpha_vjp(dal, galpha, gT)
    se[4].record()
    torch.cuda.synchronize()
    stages = {'rbx_alpha_tangent': se[0].elapsed_time(se[1]), 'rbx_integrate': se[1].elapsed_time(se[2]),
              'rbx_integrate_adjoint': se[2].elapsed_time(se[3]), 'rbx_alpha_vjp': se[3].elapsed_time(se[4])}
    # FD side: graph-replayed FP64 forwards with T[l] +- h (prob.T is read by the graph in place),
    # on the layer of the largest |dL/dT| and the one nearest half of it.  Tb is discontinuous
    # where a ray's unnormalised sum crosses T_cmb (the floor of brightness.py:110-111): (ray,
    # frequency) pairs on the floor in the base or a perturbed forward are left out of both
    # sides of the comparison.
    _, graph = timed_graph(prob, FP64, 1, 1)
    graph.replay()
    tb0 = prob.Tb64.clone()
    gT_ad = gr[0].detach().abs().cpu().numpy()
    l_max = int(np.argmax(gT_ad))
    l_half = int(np.argmin(np.abs(gT_ad - 0.5 * gT_ad[l_max])))
    fd_layers = (l_max, l_half)
    tbs, hs, n_fwd = {}, {}, 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for l in fd_layers:
        hs[l] = 1e-4 * float(T0[l])
        for sgn in (1.0, -1.0):
            prob.T.copy_(T0)
            prob.T[l] += sgn * hs[l]
            graph.replay()
            tbs[(l, sgn)] = prob.Tb64.clone()
            n_fwd += 1
    torch.cuda.synchronize()
    t_loop = (time.perf_counter() - t0) / n_fwd
    prob.T.copy_(T0)
    floor = tb0 == T_CMB
    for v in tbs.values():
        floor |= v == T_CMB
    gm = g.clone()
    gm[floor] = 0.0
    _, grm = step(gm)
    fd = {l: float(((tbs[(l, 1.0)] - tbs[(l, -1.0)]) * gm).nan_to_num(0.0).sum()) / (2 * hs[l]) for l in fd_layers}
    n_total = 6 * prob.nl + 1
    # the FD Jacobian's cost: its forwards at the FP64 graph step the bench measured (the
    # perturbation loop above adds a host synchronisation per layer; its own time is reported)
    t_fwd = dt64
    t_fd = n_total * t_fwd
    gT = grm[0].cpu().numpy()
    scale = float(np.max(np.abs(gT)))
    agree = {str(l): {'adjoint_fp32': float(gT[l]), 'central_fd_fp64': fd[l],
                      'rel_to_row_max': abs(float(gT[l]) - fd[l]) / scale} for l in fd_layers}
    ok = ~torch.isnan(prob.Tb64)
    dtb = float(torch.max(torch.abs(Tb32.detach() - prob.Tb64)[ok]))
    return {'workload': prob.w_name, 'fp32_forward_adjoint_ms': t_adj * 1e3, 'stage_ms': stages,
            'fp64_forward_ms': t_fwd * 1e3, 'fd_forwards': n_total, 'fp64_forward_fd_jacobian_s': t_fd,
            'speedup_vs_fd_jacobian': t_fd / t_adj, 'target_speedup': 10.0,
            'max_abs_dTb_fp32_vs_fp64_K': dtb, 'dL_dT_adjoint_vs_fd': agree,
            'fd_ray_freqs_left_out_on_tcmb_floor': int(floor.sum()),
            'fd_timing': '{} forwards at the FP64 graph step ({:.3f} ms); the perturbation loop on {} layers '
                         'ran {:.3f} ms per forward with its host synchronisations'.format(
                             n_total, t_fwd * 1e3, len(fd_layers), t_loop * 1e3)}


def main():
    args = parse()
    if args.gpus != 1:
        print('bench.py measures one GPU only: this build has no multi-GPU path '
              '(every GPU box here has one GPU); run with --gpus 1', file=sys.stderr)
        sys.exit(2)
    cpu_meas = None
    if not (args.no_cpu_baseline or args.copy_calibration):
        log(args, 'cpu baseline (oracle sample, one core and all cores)...')
        cpu_meas = cpu_measure(args)
    import torch
    from radiobear_amd import _lib
    if not torch.cuda.is_available():
        print('no HIP device', file=sys.stderr)
        sys.exit(1)
    if args.copy_calibration:
        copy_calibration()
        return
    if args.only in ('jacobian', 'retrieval'):
        print(json.dumps((jacobian_rate if args.only == 'jacobian' else retrieval_rate)(args)), flush=True)
        return
    log(args, 'building inputs for', args.workload)
    prob = Problem(args.workload)
    FP32, FP64 = _lib.FP32, _lib.FP64

    if args.only == 'adjoint':
        autotune(prob, FP64, args)
        dt64, _ = timed_graph(prob, FP64, 3, 1)
        autotune(prob, FP32, args)
        print(json.dumps(adjoint_rate(prob, dt64, args)), flush=True)
        return
    tune32 = autotune(prob, FP32, args)
    log(args, 'fp32 launch shapes', tune32)
    dt32, _ = timed_graph(prob, FP32, args.steps, args.warmup)
    log(args, 'fp32 graph step {:.3f} ms'.format(dt32 * 1e3))
    ns = prob.ns.cpu().numpy()
    prob.on_planet = int(np.sum(ns >= 0))
    kt32 = kernel_times(prob, FP32, max(3, min(args.steps, 10)))
    log(args, 'fp32 kernels', kt32)
    res = {}
    if not args.no_secondary:
        tune64 = autotune(prob, FP64, args)
        dt64, _ = timed_graph(prob, FP64, max(3, args.steps // 4), 1)
        kt64 = kernel_times(prob, FP64, 3)
        log(args, 'fp64 graph step {:.3f} ms'.format(dt64 * 1e3), kt64)
        Tb32 = prob.Tb.cpu().numpy()
        Tb64 = prob.Tb64.cpu().numpy()
        ok = ~np.isnan(Tb64)
        a32 = prob.alpha32.double().cpu().numpy()
        a64 = prob.alpha64.cpu().numpy()
        res['fp64_plain'] = {'value': prob.on_planet * prob.nf / dt64, 'ms_per_step': dt64 * 1e3,
                             'kernel_ms': kt64, 'launch_shapes': tune64,
                             'fp32_vs_fp64_max_abs_dTb_K': float(np.max(np.abs(Tb32[ok] - Tb64[ok]))),
                             'fp32_vs_fp64_max_rel_dalpha': float(np.max(np.abs(a32 - a64) / np.abs(a64))),
                             'nan_rays': int(np.sum(np.isnan(Tb64).any(axis=1)))}
        # the tau cuts change Tb below each mode's resolution: show it
        prob.launch_integrate_nocut(FP64)
        torch.cuda.synchronize()
        d64 = np.abs(prob.Tb.cpu().numpy()[ok] - Tb64[ok])
        res['fp64_plain']['tau_cut_max_abs_dTb_K'] = float(np.max(d64))
        autotune(prob, FP32, args)
        prob.launch_integrate(FP32)
        cut = prob.Tb.clone()
        prob.launch_integrate_nocut(FP32)
        torch.cuda.synchronize()
        d = (cut - prob.Tb).abs()
        res['tau_cut_max_abs_dTb_K'] = float(torch.nan_to_num(d, nan=0.0).max())
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ev[0].record()
        for _ in range(3):
            prob.launch_integrate_nocut(FP32)
        ev[1].record()
        torch.cuda.synchronize()
        res['integrate_fp32_without_tau_cut_ms'] = ev[0].elapsed_time(ev[1]) / 3
        res['ms_per_step_without_tau_cut'] = dt32 * 1e3 - kt32['rbx_integrate'] + res['integrate_fp32_without_tau_cut_ms']
        res['speedup_fp32_vs_fp64'] = dt64 / dt32
        res['adjoint'] = adjoint_rate(prob, dt64, args)
        log(args, 'adjoint', res['adjoint'])
        res['jacobian'] = jacobian_rate(args)
        log(args, 'jacobian', res['jacobian'])
        res['retrieval'] = retrieval_rate(args)
        log(args, 'retrieval', res['retrieval'])
        res['other_workloads'] = other_workloads(args)
        log(args, 'other workloads', res['other_workloads'])

    # roofline of the dominant kernel (by duration)
    dom = max(kt32, key=kt32.get)
    if dom == 'rbx_alpha':
        flop = prob.alpha_flop()
        achieved = flop / (kt32[dom] * 1e-3) / 1e12
        roof = {'kernel': dom, 'bound': 'valu', 'achieved': achieved, 'peak': PEAK['fp32'], 'unit': 'TFLOP/s',
                'frac': achieved / PEAK['fp32'], 'traffic': None,
                'algorithmic_flop_per_launch': flop,
                'basis': 'FLOP of the direct sum over every (line, layer, frequency); the far-line series path '
                         'sums about a third of those pairs and adds the rest as per-tile power series'}
    else:
        steps = float(np.sum(np.maximum(ns - 1, 0)))
        flop = steps * prob.nf * FLOP_RT_STEP
        achieved = flop / (kt32[dom] * 1e-3) / 1e12
        roof = {'kernel': dom, 'bound': 'valu', 'achieved': achieved, 'peak': PEAK['fp64'], 'unit': 'TFLOP/s',
                'frac': achieved / PEAK['fp64'], 'traffic': None, 'algorithmic_flop_per_launch': flop}
    tr_path = os.path.join(REPO, 'profiles', '{}_fp32_traffic.json'.format(args.workload))
    if os.path.exists(tr_path):
        with open(tr_path) as fp:
            tr = json.load(fp)['bytes_per_launch'].get(dom)
        if tr:
            roof['traffic'] = tr['read'] + tr['write']
            roof['traffic_source'] = 'rocprofv3 FETCH_SIZE/WRITE_SIZE, copy-calibrated (profiles/)'
            # the HBM side of the same kernel: measured bytes over its live duration, vs 8 TB/s
            roof['hbm_GBps'] = roof['traffic'] / (kt32[dom] * 1e-3) / 1e9
            roof['hbm_frac'] = roof['hbm_GBps'] / PEAK['hbm']
    # the integration's side: its HBM rate against 8 TB/s (it is bound by VALU issue; its
    # compulsory bytes -- the FP32 table, the ds rows, T, Tb -- would move in tens of
    # microseconds), and its rate in FLOP of the uncut paths
    steps_all = float(np.sum(np.maximum(np.minimum(ns, prob.nl - 1) - 1, 0)))
    roof_int = {'kernel': 'rbx_integrate', 'bound': 'valu', 'ms': kt32['rbx_integrate'],
                'algorithmic_bytes_per_launch': float(prob.nl * prob.nf * 4 + prob.nray * prob.ld * 8 + prob.nl * 8
                                                      + prob.nray * prob.nf * 8),
                'algorithmic_flop_per_launch': steps_all * prob.nf * FLOP_RT_STEP,
                'basis': '{} FLOP per ray x frequency x step of the uncut paths; the tau cut skips '
                         'the steps past it'.format(FLOP_RT_STEP)}
    roof_int['achieved_TFLOPs_in_uncut_unit'] = roof_int['algorithmic_flop_per_launch'] / (kt32['rbx_integrate'] * 1e-3) / 1e12
    if os.path.exists(tr_path):
        with open(tr_path) as fp:
            tri = json.load(fp)['bytes_per_launch'].get('rbx_integrate')
        if tri:
            roof_int['traffic'] = tri['read'] + tri['write']
            roof_int['hbm_GBps'] = roof_int['traffic'] / (kt32['rbx_integrate'] * 1e-3) / 1e9
            roof_int['hbm_frac'] = roof_int['hbm_GBps'] / PEAK['hbm']
    cpu = cpu_baseline(prob, cpu_meas) if cpu_meas is not None else None
    value = prob.on_planet * prob.nf / dt32
    line = {
        'metric': 'ray*freq Tb/s (FP32 batched forward; Jacobian rows/s and FP64 plain alongside)',
        'value': value, 'unit': 'ray*freq Tb/s', 'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': dt32 * 1e3, 'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': 'f32', 'data': 'synthetic (seeded generator)',
        'config': {'workload': args.workload, 'layers': prob.nl, 'freqs': prob.nf, 'rays': prob.nray,
                   'on_planet_rays': prob.on_planet, 'lines': {k: v for k, v in prob.nlines.items()},
                   'graph': 'hipGraph replay of alpha || raypaths (side stream), then integrate',
                   'parallelism': 'single GPU',
                   'precision': 'FP32 line sums (FP64 outer accumulation), FP64 detuning, geometry, half '
                                'steps and tau prefix sum; FP32 weights and trapezoid sums added into FP64 '
                                'every batch of steps; tau cut at {:g} (FP64 mode {:g})'.format(
                                    prob.K.FP32_TAU_CUT, prob.K.FP64_TAU_CUT)},
        'kernel_ms': kt32, 'launch_shapes': tune32, 'roofline': roof, 'roofline_integrate': roof_int,
        'cpu_baseline': cpu,
    }
    line.update(res)
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
