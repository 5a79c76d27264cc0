"""Forward + adjoint over a batch of atmospheres in one pass (retrieval batches).

The reference runs ``Planet.run`` once per atmosphere of a list
(planet_base.py:209-221 loops ``alpha_layers`` over ``atmos``) and has no
derivatives.  Here a batch of atmospheres, each with its own layer count, is
laid end to end along the layer axis and every kernel runs once for the whole
batch:

1. ``rbx_alpha_tangent``  alpha and d alpha / d(T, q_NH3, q_H2O) for every layer
   of every atmosphere (layers are independent);
2. ``rbx_raypaths_batch`` the ray paths of each ray through its atmosphere's
   shells (T perturbations move the radii, so geometry is per atmosphere);
3. ``rbx_jacobian_batch`` Tb and the Jacobian rows dTb/dT[l], dTb/dq_NH3[l],
   dTb/dq_H2O[l] of every (ray, frequency), forward and adjoint sweeps in one
   kernel; or, for a scalar loss, ``model()`` (a differentiable torch function
   whose backward pass is rbx_integrate_adjoint_batch + rbx_alpha_vjp).

Atmospheres have uneven layer counts, so ray wavefronts are scheduled longest
path first (``order``), which keeps the last wave of the grid short.
"""
import numpy as np

from . import _lib, kernels as K


class RetrievalBatch:
    """Device-resident inputs of many atmospheres observed along the same rays.

    atmospheres: objects with gas (C-indexed rows), P, T, R, N, nl (the synthetic
    atmospheres or Atmosphere.layer_arrays); rays: (nr, 2) impact parameters used
    for every atmosphere.  Ray i of atmosphere a is row a * nr + i of the outputs.
    """

    def __init__(self, atmospheres, catalog, absorbers, freqs, rays, req, rpol, C, precision='fp32',
                 tau_cut=None, orientation=(0.0, 0.0), gm=None, p_ref=1.0, device_props=False, launch=None):
        if precision not in _lib.PRECISIONS:
            raise ValueError('precision must be fp32 or fp64')
        self.precision = precision
        self.tau_cut = (K.FP32_TAU_CUT if precision == 'fp32' else K.FP64_TAU_CUT) if tau_cut is None else float(tau_cut)
        self.catalog = catalog
        self.absorbers = absorbers
        self.freqs = np.asarray(freqs, dtype=np.float64)
        if np.any(np.diff(self.freqs) < 0):
            raise ValueError('freqs must be ascending (nh3_hs returns its <= 30 GHz half first)')
        nls = np.array([a.nl for a in atmospheres], dtype=np.int64)
        self.natm = len(nls)
        self.nl = nls
        self.layer_off = np.concatenate([[0], np.cumsum(nls)])
        self.nl_total = int(self.layer_off[-1])
        self.ldj = int(nls.max())
        self.ld = self.ldj - 1
        rays = np.asarray(rays, dtype=np.float64).reshape(-1, 2)
        self.nr = len(rays)
        self.nray = self.natm * self.nr
        self.ray_atm = np.repeat(np.arange(self.natm), self.nr)
        # longest paths first: rays of the atmospheres with the most layers
        self.order = np.argsort(-nls[self.ray_atm], kind='stable')
        cat = lambda f: np.concatenate([np.asarray(f(a), dtype=np.float64) for a in atmospheres])
        self.P = K.as_device(cat(lambda a: a.P))
        self.T = K.as_device(cat(lambda a: a.T))
        self.R = K.as_device(cat(lambda a: a.R))
        self.N = K.as_device(cat(lambda a: a.N))
        self.mix = K.as_device(np.concatenate([K.mix_table(a.gas, C) for a in atmospheres], axis=1))
        self.b = K.as_device(np.tile(rays, (self.natm, 1)))
        self.off_d = K._i32(self.layer_off)
        self.ray_atm_d = K._i32(self.ray_atm)
        self.order_d = K._i32(self.order)
        self.planet = K.planet_desc(req, rpol, orientation)
        self.req, self.p_ref = float(req), float(p_ref)
        self.gm = None if gm is None else float(gm)
        self._side = None   # side stream of the ray paths (created on first run)
        self.launch = launch   # launch shape of every call ({field: value}; None: built in)
        if device_props:
            self.update()

    def update(self, T=None, mix=None):
        """New layer temperatures and/or mixing ratios (concatenated layout, host or device):
        the shells move with them — hydrostatic z, R and N are rebuilt on the device
        (rbx_atmosphere_batch; needs gm, the planet's GM in km^3/s^2)."""
        if self.gm is None:
            raise ValueError('RetrievalBatch needs gm (km^3/s^2) to rebuild the atmosphere')
        if T is not None:
            self.T = K.as_device(T)
        if mix is not None:
            self.mix = K.as_device(mix)
        self.Z, self.R, self.N = K.atmosphere_batch(self.P, self.T, self.mix, self.layer_off, self.req, self.gm,
                                                    self.p_ref, off_dev=self.off_d)

    def absorption(self):
        """alpha (nl_total, nf) and dalpha (3, nl_total, nf), float64, one fused pass."""
        return K.alpha_tangent(self.catalog, self.absorbers, self.P, self.T, self.mix, self.freqs,
                               precision=self.precision, with_alpha=True, launch=self.launch)

    def paths(self):
        return K.raypaths_batch(self.planet, self.R, self.N, self.layer_off, self.ray_atm, self.b, self.ld,
                                off_dev=self.off_d, ray_atm_dev=self.ray_atm_d)

    def run(self):
        """Tb (nray, nf) float64 and J (nray, 3, ldj, nf) float32 (rows past an atmosphere's
        layers are 0).  The ray paths (a few hundred latency-bound wavefronts) run on a side
        stream beside the absorption pass; the Jacobian pass waits for both."""
        import torch
        main = torch.cuda.current_stream()
        if self._side is None:
            self._side = torch.cuda.Stream()
        self._side.wait_stream(main)
        with torch.cuda.stream(self._side):
            hds, ns = self.paths()
        alpha, dalpha = self.absorption()
        main.wait_stream(self._side)
        hds.record_stream(main)   # allocated on the side stream, read on this one
        ns.record_stream(main)
        J, Tb = K.jacobian_batch(alpha, dalpha, self.T, self.off_d, self.ray_atm_d, hds, ns, self.ldj,
                                 order=self.order_d, tau_cut=self.tau_cut, precision=self.precision)
        return Tb, J

    def step(self, rebuild=False):
        """One retrieval step: (optionally) the shells rebuilt on the device for the current
        T and mix, then run().  Returns (Tb, J)."""
        if rebuild:
            self.update()
        return self.run()

    def capture(self, rebuild=False):
        """The whole step captured once in a HIP graph (torch.cuda.CUDAGraph): the kernels of
        step() replay with no host work per launch.  Returns (graph, (Tb, J)); each
        graph.replay() recomputes Tb and J in place from the batch's current device inputs
        (T, mix and, with rebuild, the shells), which a sampler updates with copy_()."""
        import torch
        stream = torch.cuda.Stream()
        stream.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(stream):
            self.step(rebuild)                     # eager warm-up: allocations and the side stream
        torch.cuda.current_stream().wait_stream(stream)
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=stream):
            out = self.step(rebuild)
        return graph, out

    def model(self, hds=None, nsteps=None):
        """A differentiable forward model of the batch (adjoint.ForwardModel): Tb (nray, nf)
        as a function of the concatenated T, q_NH3 and q_H2O, its backward pass the adjoint.
        The ray paths (computed now unless given) are held fixed."""
        from .adjoint import ForwardModel
        if hds is None:
            hds, nsteps = self.paths()
        return ForwardModel(self.catalog, self.absorbers, self.P, self.mix, self.freqs, hds, nsteps,
                            precision=self.precision, tau_cut=self.tau_cut, layer_off=self.off_d,
                            ray_atm=self.ray_atm_d, order=self.order_d, launch=self.launch)

    def rows(self):
        return self.nray * len(self.freqs)

    def atmosphere_slice(self, a):
        """Rows of the outputs and layers of the inputs that belong to atmosphere a."""
        return slice(a * self.nr, (a + 1) * self.nr), slice(int(self.layer_off[a]), int(self.layer_off[a + 1]))


def synthetic(n_atm=1024, nr=16, seed_rays=None, precision='fp32', nh3_lines=10000, nf=128, **kw):
    """The retrieval workload: n_atm perturbed Jupiter-like atmospheres (200-800 layers,
    T +-5 K and NH3 x[0.3, 3] Gaussian-process perturbations), nh3_lines synthetic NH3
    lines, nf frequencies over 1-50 GHz and nr random disc rays (radius <= 0.95)."""
    from . import lines, synth
    w = synth.WORKLOADS['c5_retrieval']
    idx = list(synth.WORKLOADS).index('c5_retrieval')
    ss = np.random.SeedSequence(synth.SEED_BASE + idx)
    s_atm, s_lines, _, s_rays = [int(c.generate_state(1)[0]) for c in ss.spawn(4)]
    atms = synth.retrieval_batch(n_atm, seed=s_atm, nl_range=w['nl'])
    inv = synth.nh3_inversion_lines(nh3_lines, s_lines)
    cat = K.jupiter_catalog(lines.raw_catalog(), nh3_inversion=inv)
    ab = K.absorbers_struct({'h2': 'h2_jj_ddb', 'h2o': 'h2o_bk', 'h2s': 'h2s_ddb', 'nh3': 'nh3_hs_sjs',
                             'ph3': 'ph3_jh'}, 'e')
    freqs = np.linspace(1.0, 50.0, nf)
    rays = synth.random_disc_rays(nr, s_rays if seed_rays is None else seed_rays, 0.95)
    a0 = atms[0]
    kw.setdefault('gm', synth.JUPITER_GM)
    batch = RetrievalBatch(atms, cat, ab, freqs, rays, a0.Req, a0.Rpol, synth.GAS_ROWS, precision=precision, **kw)
    batch.atmospheres = atms
    batch.nh3_inv = inv
    return batch


__all__ = ['RetrievalBatch', 'synthetic']
