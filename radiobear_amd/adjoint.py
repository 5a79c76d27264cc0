"""The forward model as a differentiable torch function, with a hand-written adjoint.

Tb(T, q_NH3, q_H2O) for fixed rays: the absorption (Alpha.get_layers,
alpha.py:151-305) and the brightness integration along the ray paths
(Brightness.single, brightness.py:30-126).  The reference has no derivatives;
retrieval and MCMC samplers (mcmc/__init__.py) need dTb/dT and dTb/dq per
layer.  ``torch.autograd.Function`` carries them:

* forward:  rbx_alpha_tangent (alpha and d alpha / d(T, q_NH3, q_H2O) in one
  fused pass, kept for the backward) then rbx_integrate;
* backward: rbx_integrate_adjoint (the vector-Jacobian product over all rays,
  e^-tau recomputed, not stored) then rbx_alpha_vjp (the chain through the
  tangents).

The ray geometry is held fixed: a perturbed temperature does not move the
shells here (RetrievalBatch.update rebuilds them for a new state).  In FP32
mode the line sums and tangents are FP32 (FP64 outer accumulation) and the
integration takes its FP32 exponential; tau, the sums along the path and the
adjoint are FP64.
"""
import torch

from . import _lib, kernels as K

_NH3 = _lib.SPECIES.index('NH3')
_H2O = _lib.SPECIES.index('H2O')


class ForwardModel:
    """Tb of fixed rays through one atmosphere, or through several laid end to end.

    catalog, absorbers: the line catalogue and absorber selection (no co_ddb, clouds_idp
    or h2_orton: they have no tangents); P (nl,) bar and mix (8, nl) the layer pressures
    and mixing ratios (the NH3 and H2O rows are replaced by the q inputs of each call);
    freqs ascending (GHz); hds, nsteps the ray paths (rbx_raypaths, half steps in cm).
    For a batch: layer_off (natm + 1,), ray_atm (nray,) and order (nray,), device int32,
    as RetrievalBatch holds them."""

    def __init__(self, catalog, absorbers, P, mix, freqs, hds, nsteps, precision='fp32', tau_cut=None,
                 layer_off=None, ray_atm=None, order=None, launch=None):
        if precision not in _lib.PRECISIONS:
            raise ValueError('precision must be fp32 or fp64')
        self.catalog, self.absorbers = catalog, absorbers
        self.P = K.as_device(P)
        self.mix = K.as_device(mix)
        self.freqs = freqs
        self.hds, self.nsteps = hds.contiguous(), nsteps.contiguous()
        self.precision = precision
        default_cut = K.FP32_TAU_CUT if precision == 'fp32' else K.FP64_TAU_CUT
        self.tau_cut = default_cut if tau_cut is None else float(tau_cut)
        self.batch = layer_off is not None
        self.layer_off, self.ray_atm = layer_off, ray_atm
        self.order = order if (order is not None or self.batch) else K.path_order(self.nsteps)
        self.launch = launch

    def absorption(self, T, q_nh3, q_h2o):
        mix = self.mix.clone()
        mix[_NH3] = q_nh3
        mix[_H2O] = q_h2o
        return K.alpha_tangent(self.catalog, self.absorbers, self.P, T, mix, self.freqs, precision=self.precision,
                               with_alpha=True, launch=self.launch)

    def integrate(self, alpha, T):
        if self.batch:
            return K.integrate_batch(alpha, T, self.layer_off, self.ray_atm, self.hds, self.nsteps,
                                     order=self.order, precision=self.precision, tau_cut=self.tau_cut,
                                     launch=self.launch)
        return K.integrate(alpha, T, self.hds, self.nsteps, precision=self.precision, tau_cut=self.tau_cut,
                           launch=self.launch)

    def adjoint(self, alpha, T, gTb):
        if self.batch:
            return K.integrate_adjoint_batch(alpha, T, self.layer_off, self.ray_atm, self.hds, self.nsteps, gTb,
                                             order=self.order, precision=self.precision, tau_cut=self.tau_cut,
                                             launch=self.launch)
        return K.integrate_adjoint(alpha, T, self.hds, self.nsteps, gTb, order=self.order,
                                   precision=self.precision, tau_cut=self.tau_cut, launch=self.launch)

    def __call__(self, T, q_nh3=None, q_h2o=None):
        """Tb (nray, nf) float64; differentiable in T, q_nh3 and q_h2o ((nl,) float64 device
        tensors; None keeps the model's own row)."""
        if q_nh3 is None:
            q_nh3 = self.mix[_NH3]
        if q_h2o is None:
            q_h2o = self.mix[_H2O]
        return BrightnessTemperature.apply(T, q_nh3, q_h2o, self)


class BrightnessTemperature(torch.autograd.Function):
    """Tb = model(T, q_NH3, q_H2O) with the backward pass of rbx_integrate_adjoint +
    rbx_alpha_vjp."""

    @staticmethod
    def forward(ctx, T, q_nh3, q_h2o, model):
        for name, t in (('T', T), ('q_nh3', q_nh3), ('q_h2o', q_h2o)):
            if not isinstance(t, torch.Tensor) or t.dtype != torch.float64 or t.dim() != 1:
                raise ValueError('{} must be a 1-D float64 tensor'.format(name))
        T_d = K.as_device(T.detach())
        alpha, dalpha = model.absorption(T_d, K.as_device(q_nh3.detach()), K.as_device(q_h2o.detach()))
        Tb = model.integrate(alpha, T_d)
        ctx.model = model
        ctx.devices = (T.device, q_nh3.device, q_h2o.device)
        ctx.save_for_backward(alpha, dalpha, T_d)
        return Tb

    @staticmethod
    def backward(ctx, gTb):
        alpha, dalpha, T_d = ctx.saved_tensors
        galpha, gT = ctx.model.adjoint(alpha, T_d, gTb.contiguous())
        grad = K.alpha_vjp(dalpha, galpha, gT)
        dT, dq, dw = (grad[i].to(d) for i, d in enumerate(ctx.devices))
        return dT, dq, dw, None


__all__ = ['ForwardModel', 'BrightnessTemperature']
