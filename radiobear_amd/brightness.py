"""Brightness temperature along ray paths (integration kernel).

Mirrors radiobear/brightness.py:13-126.  ``single`` keeps the reference's
signature and side attributes (travel, tau, W, Tb_lyr, P, z, integrated_W);
``batch`` integrates many rays in one launch, which is what Planet.run uses.
The arithmetic (trapezoid tau, W = alpha e^-tau or 2 alpha E2(tau) for the disc
average, normalisation by the integrated weight, T_cmb floor) runs in the
rbx_integrate / rbx_integrate_profiles kernels.
"""
import numpy as np
import torch

from . import kernels, raypath, utils

ORDER_MIN_RAY_FREQS = 1 << 23   # point ray x frequency pairs from which batch() orders for the cut


class Brightness:
    """Brightness temperature of the planet."""

    def __init__(self, config=None, log=None, verbose=True, precision=None, tau_cut=None, **kwargs):
        from . import config as pcfg
        if config is None or isinstance(config, str):
            config = pcfg.planetConfig('jupiter', configFile=config)
            config.update_config(**kwargs)
        self.config = config
        self.verbose = verbose
        self.precision = precision
        self.tau_cut = tau_cut

    @staticmethod
    def _no_doppler(config):
        # the reference's Doppler branch calls Alpha.get_alpha, which does not exist, and
        # announces itself as broken (brightness.py:79-90): with Doppler on it cannot run
        if getattr(config, 'Doppler', False):
            raise NotImplementedError('Doppler is not supported: the reference branch is broken '
                                      '(brightness.py:79-90 calls a missing Alpha.get_alpha)')

    def _modes(self, alpha):
        prec = self.precision or ('fp32' if alpha.table.dtype == torch.float32 else 'fp64')
        tau_cut = self.tau_cut if self.tau_cut is not None else (kernels.FP32_TAU_CUT if prec == 'fp32' else kernels.FP64_TAU_CUT)
        return prec, tau_cut

    def start_paths(self, bs, atm, orientation=None):
        """Launch the ray paths of bs on a side stream (they do not depend on the
        absorption, so they run beside it); batch(..., paths=) joins them."""
        self._no_doppler(self.config)
        disc = [isinstance(b, str) or (not np.isscalar(b) and isinstance(b[0], str)) for b in bs]
        rays = np.array([[0.0, 0.0] if d else [float(b[0]), float(b[1])] for b, d in zip(bs, disc)])
        main = torch.cuda.current_stream()
        side = torch.cuda.Stream()
        side.wait_stream(main)                    # inputs uploaded on the main stream
        with torch.cuda.stream(side):
            hds, ns = raypath.batch(atm, rays, orientation)
        return disc, hds, ns, side

    def batch(self, bs, freqs, atm, alpha, orientation=None, paths=None):
        """Tb (nray, nf) numpy float64 for the rays bs; 'disc' rays use the disc average.
        paths: what start_paths returned for these rays, or None to compute them here."""
        self._no_doppler(alpha.config)
        if paths is None:
            paths = self.start_paths(bs, atm, orientation)
        disc, hds, ns, side = paths
        if alpha.table is None:
            alpha.get_layers(freqs, atm)
        torch.cuda.current_stream().wait_stream(side)
        hds.record_stream(torch.cuda.current_stream())
        ns.record_stream(torch.cuda.current_stream())
        rays = hds                                    # (row count only)
        T = atm.gas[atm.config.C['T']]
        prec, tau_cut = self._modes(alpha)
        Tb = np.empty((rays.shape[0], alpha.table.shape[1]))
        point = [i for i, d in enumerate(disc) if not d]
        if point:
            sel = torch.as_tensor(point, device=hds.device)
            # frequencies in the order they pass the cut (a fixed ~0.07 ms that saves ~10 % of
            # the integration): for large batches only
            big = len(point) * alpha.table.shape[1] >= ORDER_MIN_RAY_FREQS
            Tb[point] = kernels.integrate(alpha.table, T, hds[sel], ns[sel], precision=prec,
                                          tau_cut=tau_cut, ordered=big).cpu().numpy()
        discs = [i for i, d in enumerate(disc) if d]
        if discs:
            sel = torch.as_tensor(discs, device=hds.device)
            Tb[discs] = kernels.integrate(alpha.table, T, hds[sel], ns[sel], disc=True,
                                          precision=prec).cpu().numpy()
        self.nsteps = ns.cpu().numpy()
        return Tb

    def single(self, b, freqs, atm, alpha, orientation=None, taulimit=20.0):
        """Tb (list over freqs) for one ray, with the reference's side attributes."""
        self._no_doppler(alpha.config)
        disc_average = utils.b_type(b).startswith('dis')
        if disc_average:
            b = [0.0, 0.0]
        self.alpha = alpha
        self.freqs = freqs
        self.b = b
        if alpha.table is None:
            alpha.get_layers(freqs, atm)
        travel = raypath.compute_ds(atm, b, orientation)
        self.travel = travel
        nf = alpha.table.shape[1]
        if travel.ds is None:
            print('Off planet')
            self.Tb = [utils.T_cmb] * nf
            return self.Tb
        hds, ns = raypath.batch(atm, [b], orientation)
        T = atm.gas[atm.config.C['T']]
        prec, _ = self._modes(alpha)
        Tb, intW = kernels.integrate(alpha.table, T, hds, ns, disc=disc_average, precision=prec, with_intW=True)
        tau, W, Tbl = kernels.integrate_profiles(alpha.table, T, hds, ns, disc=disc_average, precision=prec)
        n = len(travel.ds)
        P = atm.gas[atm.config.C['P']]
        z = atm.gas[atm.config.C['Z']]
        self.Tb = list(Tb.cpu().numpy()[0])
        self.integrated_W = intW.cpu().numpy()[0]
        # profiles at points 0..n-1 (the reference stores n entries: the top and n-1 steps)
        self.tau = tau.cpu().numpy()[0, :n].T
        self.W = W.cpu().numpy()[0, :n].T
        self.Tb_lyr = Tbl.cpu().numpy()[0, :n].T
        self.P = np.array([P[0]] + [(P[i] + P[i + 1]) / 2.0 for i in range(n - 1)])
        self.z = np.array([z[0]] + [(z[i] + z[i + 1]) / 2.0 for i in range(n - 1)])
        if self.verbose:
            for j, f in enumerate(freqs):
                if not self.Tb[j] == utils.T_cmb and self.integrated_W[j] < 0.96:
                    print("Weight correction at {:.2f} is {:.4f} (showing below 0.96)".format(f, self.integrated_W[j]))
        return self.Tb
