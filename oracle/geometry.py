"""FP64 restatement of the ray geometry (TEST INFRASTRUCTURE ONLY).

Scalar Python loops that follow radiobear/raypath.py:39-273 and the ellipse
branch of radiobear/shape.py:223-298 operation by operation, on numpy float64
scalars so that sqrt of a negative number yields NaN exactly as the reference
does (its ``except ValueError`` branches never fire).  Small inputs only.
"""
import math

import numpy as np

_DIR = {'egress': -1, 'ingress': 1, 'tangent': 0}
_XH = np.array([1.0, 0.0, 0.0])
_YH = np.array([0.0, 1.0, 0.0])
_ZH = np.array([0.0, 0.0, 1.0])


def r2d(a):
    return a * 180.0 / np.pi          # utils.py:90-92


def d2r(a):
    return a * np.pi / 180.0          # utils.py:95-97


def rot_x(x, V):
    R = np.array([[1.0, 0.0, 0.0], [0.0, np.cos(x), -np.sin(x)], [0.0, np.sin(x), np.cos(x)]])
    return np.dot(R, V)               # shape.py:277-282


def rot_y(y, V):
    R = np.array([[np.cos(y), 0.0, np.sin(y)], [0.0, 1.0, 0.0], [-np.sin(y), 0.0, np.cos(y)]])
    return np.dot(R, V)               # shape.py:285-290


def rot_z(z, V):
    R = np.array([[np.cos(z), -np.sin(z), 0.0], [np.sin(z), np.cos(z), 0.0], [0.0, 0.0, 1.0]])
    return np.dot(R, V)               # shape.py:293-298


def compute_aspect(Q, f=1.0):
    """raypath.computeAspect (raypath.py:39-44)."""
    tip = -Q[0] * np.pi / 180.0
    rotate = -np.arctan(np.tan(Q[1] * np.pi / 180.0) * (1.0 - f) ** 2)
    return tip, rotate


def rotate2planet(rotate, tip, b):
    return rot_x(rotate, rot_z(tip, b))   # raypath.py:47-51


class Ellipse:
    """State of Shape for gtype 'ellipse'/'sphere' (shape.py:223-274).

    Only the fields the ray tracer reads are kept: rmag, r, n."""

    def __init__(self, gtype, rpol_over_req):
        self.gtype = gtype
        self.ratio = rpol_over_req
        self.rmag = 0.0
        self.r = np.zeros(3)
        self.n = np.zeros(3)

    def calc(self, r, pclat, delta_lng):
        a = r
        b = self.ratio * r if self.gtype == 'ellipse' else r
        lat = d2r(pclat)
        lng = d2r(delta_lng)
        if lat == 0.0:
            lat = 1.0E-6
        norm = np.array([0.0, a * np.sin(lat), b * np.cos(lat)])
        norm = norm / np.linalg.norm(norm)
        norm = rot_y(lng, norm)
        r_vec = rot_y(lng, np.array([0.0, b * np.sin(lat), a * np.cos(lat)]))
        self.rmag = np.linalg.norm(r_vec)
        self.r = r_vec
        self.n = norm
        return self.rmag


def _lat_lng(v):
    pclat = r2d(np.arcsin(np.dot(v, _YH) / np.linalg.norm(v)))
    dlng = r2d(np.arctan2(np.dot(v, _XH), np.dot(v, _ZH)))
    return pclat, dlng


def np_arange(start, stop, step):
    """numpy.arange for doubles: x_i = start + i*((start+step)-start)."""
    return np.arange(start, stop, step)


def find_edge(b, rNorm, tip, rotate, geoid):
    """raypath.findEdge (raypath.py:60-105)."""
    tmp = b[0] ** 2 + b[1] ** 2
    zQ_trial = np.arange(np.sqrt(1.0 - tmp) * 1.01, 0.0, -0.005)
    r_trial = []
    r_shape = []
    hit = False
    for zQ in zQ_trial:
        bv = rotate2planet(rotate, tip, np.array([b[0], b[1], zQ]))
        r1 = np.linalg.norm(bv) * rNorm
        r_trial.append(r1)
        pclat, dlng = _lat_lng(bv)
        r2 = geoid.calc(rNorm, pclat, dlng)
        r_shape.append(r2)
        if r1 < r2:
            hit = True
            break
    if not hit:
        return None, None
    xx = np.flipud(np.array(r_trial) - np.array(r_shape))
    yy = np.flipud(np.array(zQ_trial[0:len(r_trial)]))
    zQ = np.interp(0.0, xx, yy)
    bq = np.array([b[0], b[1], zQ])
    edge = rNorm * rotate2planet(rotate, tip, bq)
    return edge, bq


class Ray:
    def __init__(self):
        self.ds = None
        self.layer4ds = None
        self.r4ds = None
        self.tip = None
        self.rotate = None
        self.rNorm = None


def compute_ds(R, N, b, Req, Rpol, orientation=(0.0, 0.0), gtype='ellipse', limb='shape',
               max_steps=None):
    """raypath.compute_ds (raypath.py:108-273), Doppler list omitted (unused by Tb).

    R, N: layer radii (km) and refractive indices, top layer first.
    Returns a Ray with ds (km) and layer4ds, or ds=None when off planet.
    """
    path = Ray()
    req = np.asarray(R, dtype=np.float64)
    nr = np.asarray(N, dtype=np.float64)
    rNorm = req[0]
    if (b[0] ** 2 + b[1] ** 2) >= 1.0:
        return path
    mu = np.sqrt(1.0 - b[0] ** 2 - b[1] ** 2)
    f = 1.0 - Rpol / Req
    tip, rotate = compute_aspect(orientation, f)
    geoid = Ellipse(gtype, Rpol / Req)
    edge, bq = find_edge(b, rNorm, tip, rotate, geoid)
    if edge is None:
        return path
    pclat, dlng = _lat_lng(edge)
    geoid.calc(rNorm, pclat, dlng)

    start = rotate2planet(rotate, tip, np.array([0.0, 0.0, -1.0]))
    s = [start]
    n = [geoid.n]
    r = [geoid.r]
    t_inc = [np.arccos(-np.dot(s[-1], n[-1]))]
    nratio = nr[0] / nr[1]
    t_tran = [np.arcsin(nratio * np.sin(t_inc[-1]))]
    ds_list = [0.0]
    layers = [0]
    r4ds = [0.0]

    i = 0
    layer = 0
    direction = 'ingress'
    with np.errstate(invalid='ignore'):
        while True:
            d = _DIR[direction]
            s.append(nratio * s[i] + d * (nratio * np.cos(t_inc[i]) * n[i] - np.cos(t_tran[i]) * n[i]))
            rNow = geoid.rmag
            rNext = geoid.calc(req[layer + d], pclat, dlng)
            rdots = np.dot(r[i], s[i + 1])
            disc = rdots ** 2.0 + rNext ** 2.0 - rNow ** 2.0
            dsp = -rdots + np.sqrt(disc)
            dsm = -rdots - np.sqrt(disc)
            ds_step = dsm if direction == 'ingress' else dsp
            if ds_step < 0.0:                        # raypath.py:212-216
                break
            if limb == 'sec':                        # raypath.py:218-219
                ds_step = abs(rNext - rNow) / mu
            ds_list.append(ds_step)
            layers.append(layer)
            r4ds.append(rNow)
            rnext = r[i] + ds_list[i + 1] * s[i + 1]
            pclat, dlng = _lat_lng(rnext)
            geoid.calc(req[layer + d], pclat, dlng)
            r.append(rnext)
            n.append(geoid.n)
            layer += d
            t_inc.append(np.arccos(-d * np.dot(s[i + 1], n[i + 1])))
            i += 1
            if layer + d >= len(nr) or layer + d < -len(nr):   # IndexError at raypath.py:256
                break
            nratio = 1.0                             # raypath.py:257
            t_tran.append(np.arcsin(nratio * np.sin(t_inc[-1])))
            if max_steps is not None and i >= max_steps:
                break
    path.ds = ds_list[1:]
    path.layer4ds = layers[1:]
    path.r4ds = r4ds[1:]
    path.tip, path.rotate, path.rNorm = tip, rotate, rNorm
    return path


def nan_safe_isclose(a, b):
    return (math.isnan(a) and math.isnan(b)) or a == b
