"""Ray paths through the layered atmosphere (FP64 geometry kernel).

Mirrors radiobear/raypath.py:20-273: ``compute_ds`` returns a ``Ray`` with the
step lengths ``ds`` (km) and the layer of each step, or ``ds = None`` when the
ray misses the planet.  The stepping runs in rbx_raypaths (one GPU thread per
ray); ``batch`` traces many rays in one launch and keeps the result on the
device for the integration kernel.
"""
import numpy as np

from . import kernels


class Ray:
    """Ray parameters (raypath.py:20-36)."""

    allowed_parameters = ['ds', 'layer4ds', 'r4ds', 'P4ds', 'doppler', 'tip', 'rotate', 'rNorm']

    def __init__(self):
        for k in self.allowed_parameters:
            setattr(self, k, None)

    def update(self, **kwargs):
        for k, v in kwargs.items():
            if k in self.allowed_parameters:
                setattr(self, k, v)
            else:
                raise ValueError('{} not allowed Ray parameter'.format(k))


def computeAspect(Q, f=1.0):
    """[position angle, planetographic latitude] -> (tip, rotate) (raypath.py:39-44)."""
    tip = -Q[0] * np.pi / 180.0
    rotate = -np.arctan(np.tan(Q[1] * np.pi / 180.0) * (1.0 - f) ** 2)
    return tip, rotate


def planet_of(atm, orientation=None, gtype=None):
    cfg = atm.config
    return kernels.planet_desc(cfg.Req, cfg.Rpol, cfg.orientation if orientation is None else orientation,
                               cfg.gtype if gtype is None else gtype, cfg.limb)


def batch(atm, bs, orientation=None, gtype=None, units='half_cm'):
    """All rays of bs (nray, 2) in one launch: device (ds, nsteps)."""
    LP = atm.config.LP
    b = np.asarray(bs, dtype=np.float64).reshape(-1, 2)
    return kernels.raypaths(planet_of(atm, orientation, gtype), atm.property[LP['R']], atm.property[LP['N']], b,
                            units=units)


def compute_ds(atm, b, orientation=None, gtype=None, verbose=False):
    """Path of one ray (raypath.py:108-273).  Ray.doppler is not produced
    (only the reference's disabled Doppler branch reads it)."""
    path = Ray()
    if (b[0] ** 2 + b[1] ** 2) >= 1.0:
        return path
    cfg = atm.config
    f = 1.0 - cfg.Rpol / cfg.Req
    tip, rotate = computeAspect(cfg.orientation if orientation is None else orientation, f)
    ds, ns = batch(atm, [b], orientation, gtype, units='km')
    n = int(ns.cpu().numpy()[0])
    if n < 0:
        return path
    dsl = list(ds.cpu().numpy()[0, :n])
    P = atm.gas[cfg.C['P']]
    path.update(ds=dsl, layer4ds=list(range(n)), P4ds=[P[k] for k in range(n)],
                tip=tip, rotate=rotate, rNorm=atm.property[cfg.LP['R']][0])
    return path
