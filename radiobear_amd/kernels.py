"""Python entry points of the HIP kernels (torch tensors as device buffers).

Each function checks shapes and dtypes on the host, then calls the C ABI of
librbx.so on torch's current HIP stream.  Inputs may be numpy arrays or torch
tensors; outputs are torch tensors on the GPU.  There is no CPU path: without
a HIP device (or without the library) these functions raise.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import check, lib

F_SPLIT = 30.0          # nh3_hs switches coefficient sets above 30 GHz (nh3_hs.py:84)
NH3_MODELS = {'nh3_hs_sjs': 1, 'nh3_hs': 2, 'nh3_sjs': 3, 'nh3_dbs_sjs': 4, 'nh3_dbs': 5}


def device():
    if not torch.cuda.is_available():
        raise RuntimeError('radiobear_amd needs a HIP device (torch.cuda.is_available() is False)')
    return torch.device('cuda', torch.cuda.current_device())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def as_device(x, dtype=torch.float64):
    """Contiguous device tensor of dtype from numpy / list / tensor."""
    if isinstance(x, torch.Tensor):
        t = x.to(device=device(), dtype=dtype)
    else:
        t = torch.as_tensor(np.asarray(x), dtype=dtype, device=device())
    return t.contiguous()


_CONSTS = {}          # small host arrays already copied to the device (request frequencies)
_CONSTS_MAX = 64


def device_const(a):
    """A float64 device copy of a small host array, kept and reused for equal contents, so
    that repeated calls (and calls under HIP graph capture after a first eager one) copy
    nothing from the host."""
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    key = (torch.cuda.current_device(), a.shape, a.tobytes())
    t = _CONSTS.get(key)
    if t is None:
        if len(_CONSTS) >= _CONSTS_MAX:
            _CONSTS.pop(next(iter(_CONSTS)))
        t = torch.as_tensor(a, device=device())
        _CONSTS[key] = t
    return t


# ------------------------------------------------------------------ catalogue
class Catalog:
    """Device-resident line lists (rbx_catalog)."""

    def __init__(self):
        h = ctypes.c_void_p()
        check(lib().rbx_catalog_create(ctypes.byref(h)))
        self._h = h
        self.counts = {}

    @property
    def handle(self):
        return self._h

    def set_lines(self, kind, data):
        """Upload block ``kind`` from a dict holding its columns (see _lib.LINE_COLUMNS)."""
        names = _lib.LINE_COLUMNS[kind]
        cols = [np.ascontiguousarray(np.asarray(data[n], dtype=np.float64).ravel()) for n in names]
        n = len(cols[0])
        if kind == _lib.TABLE_H2_ORTON:
            nt = _lib.ORTON_NTEMP
            if len(cols[1]) != 6 * n * nt or len(cols[2]) != nt:
                raise ValueError('Orton table: {} frequencies need {} values and {} temperatures'.format(
                    n, 6 * n * nt, nt))
            self.orton_ftab = cols[0].copy()
        elif any(len(c) != n for c in cols):
            raise ValueError('line block {}: columns of unequal length'.format(kind))
        arr = (ctypes.POINTER(ctypes.c_double) * len(cols))(
            *[c.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) for c in cols])
        check(lib().rbx_catalog_set_lines(self._h, kind, n, arr))
        self.counts[kind] = n

    def lines(self, kind):
        n = ctypes.c_int()
        check(lib().rbx_catalog_lines(self._h, kind, ctypes.byref(n)))
        return n.value

    def close(self):
        if getattr(self, '_h', None) is not None and _lib._lib is not None:
            _lib._lib.rbx_catalog_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def absorbers_struct(constituent_alpha, h2state='e', other_dict=None):
    """rbx_absorbers from a reference-style {'nh3': 'nh3_hs_sjs', ...} mapping.

    other_dict: the per-constituent options Alpha.setup_formalisms copies from the
    config (alpha.py:93-102): other_dict['co']['coshape'] (default 'voigt',
    co_ddb.py:37-40) and the clouds_idp flags in other_dict['clouds']."""
    other_dict = other_dict or {}
    ab = _lib.Absorbers()
    for c, model in constituent_alpha.items():
        if model is None:
            continue
        if c == 'h2':
            if model not in ('h2_jj_ddb', 'h2_orton'):
                raise ValueError('unsupported h2 absorber {}'.format(model))
            if h2state not in ('e', 'n'):
                raise ValueError('INVALID H2STATE {!r}'.format(h2state))
            ab.h2 = (1 if h2state == 'e' else 2) + (2 if model == 'h2_orton' else 0)
        elif c == 'h2o':
            if model != 'h2o_bk':
                raise ValueError('unsupported h2o absorber {}'.format(model))
            ab.h2o = 1
        elif c == 'h2s':
            if model != 'h2s_ddb':
                raise ValueError('unsupported h2s absorber {}'.format(model))
            ab.h2s = 1
        elif c == 'nh3':
            if model not in NH3_MODELS:
                raise ValueError('unsupported nh3 absorber {}'.format(model))
            ab.nh3 = NH3_MODELS[model]
        elif c == 'ph3':
            if model != 'ph3_jh':
                raise ValueError('unsupported ph3 absorber {}'.format(model))
            ab.ph3 = 1
        elif c == 'co':
            if model != 'co_ddb':
                raise ValueError('unsupported co absorber {}'.format(model))
            ab.co = 1
            shape = (other_dict.get('co') or {}).get('coshape', 'voigt')
            ab.co_shape = _lib.CO_SHAPES.get(shape, 3)
        elif c == 'clouds':
            if model != 'clouds_idp':
                raise ValueError('unsupported clouds absorber {}'.format(model))
            ab.clouds = 1
            flags = other_dict.get('clouds') or {}
            ab.cloud_types = sum(1 << i for i, k in enumerate(_lib.CLOUD_FLAGS)
                                 if k in flags and flags[k] is not None and flags[k] > 0.0)
        else:
            raise ValueError('unsupported constituent {}'.format(c))
    return ab


def _check_orton_freqs(catalog, freqs):
    """h2_orton brackets each frequency by the first tabulated one above it and the next
    one up (h2_orton.py:77-92): the reference fails (IndexError) at or beyond the
    second-to-last table frequency, and so does this."""
    ft = getattr(catalog, 'orton_ftab', None)
    if ft is None:
        raise ValueError('h2_orton selected but the catalogue has no Orton table')
    f = np.asarray(freqs, dtype=np.float64)
    if np.any(f >= ft[-2]):
        raise ValueError('h2_orton: frequencies must be below {} GHz (the table range)'.format(ft[-2]))


def nh3_order(freqs):
    """Column order of nh3_hs's output: request indices <= 30 GHz, then > 30 GHz."""
    f = np.asarray(freqs, dtype=np.float64)
    lo = np.where(f <= F_SPLIT)[0]
    hi = np.where(f > F_SPLIT)[0]
    return np.concatenate([lo, hi]), len(lo)


def mix_table(gas, C):
    """(8, nl) mixing ratios in RBX_X_* order from a reference gas table (a species
    the table has no row for, such as CO in some gas files, is zero)."""
    nl = np.asarray(gas).shape[1]
    return np.stack([np.asarray(gas[C[s]], dtype=np.float64) if s in C else np.zeros(nl)
                     for s in _lib.SPECIES])


def cloud_table(cloud, Cl):
    """(6, nl) cloud densities (g/cm^3) in RBX_CL_* order from a reference cloud table
    laid out by Cl (config.Cl); a cloud the table has no row for is zero."""
    cloud = np.asarray(cloud, dtype=np.float64)
    nl = cloud.shape[1]
    return np.stack([cloud[Cl[k]] if k in Cl and Cl[k] < cloud.shape[0] else np.zeros(nl)
                     for k in _lib.CLOUDS])


def _alpha_inputs(P, T, mix, freqs, scale):
    f = np.asarray(freqs, dtype=np.float64)
    order, nf_low = nh3_order(f)
    P = as_device(P)
    T = as_device(T)
    mix = as_device(mix)
    nl = P.shape[0]
    if T.shape != (nl,) or mix.shape != (len(_lib.SPECIES), nl):
        raise ValueError('shape mismatch: P {}, T {}, mix {}'.format(tuple(P.shape), tuple(T.shape),
                                                                     tuple(mix.shape)))
    fd = device_const(f)
    fn = None if np.array_equal(order, np.arange(len(f))) else device_const(f[order])
    sc = None
    if scale is not None:
        sc = as_device(scale)
        if sc.shape != (len(_lib.CONSTITUENTS), nl):
            raise ValueError('scale must be ({}, {})'.format(len(_lib.CONSTITUENTS), nl))
    return P, T, mix, fd, fn, nf_low, sc, nl


def alpha(catalog, absorbers, P, T, mix, freqs, precision='fp64', scale=None, by_constituent=False, cloud=None,
          launch=None):
    """alpha[nl][nf] (cm^-1) on the device; optionally also [7][nl][nf] per constituent
    (RBX_C_* rows).  cloud: (6, nl) densities (cloud_table), needed by clouds_idp.
    launch: launch shape ({field: value} of _lib.LAUNCH_FIELDS or None for the built-in one)."""
    prec = _lib.PRECISIONS[precision]
    P, T, mix, fd, fn, nf_low, sc, nl = _alpha_inputs(P, T, mix, freqs, scale)
    nf = fd.shape[0]
    cl = None
    if absorbers.clouds:
        if cloud is None:
            raise ValueError('clouds_idp needs the cloud densities (cloud=)')
        cl = as_device(cloud)
        if cl.shape != (len(_lib.CLOUDS), nl):
            raise ValueError('cloud must be ({}, {})'.format(len(_lib.CLOUDS), nl))
    if absorbers.h2 >= 3:
        _check_orton_freqs(catalog, freqs)
    out = torch.empty((nl, nf), dtype=torch.float32 if prec == _lib.FP32 else torch.float64, device=device())
    byc = torch.zeros((len(_lib.CONSTITUENTS), nl, nf), dtype=torch.float64, device=device()) \
        if by_constituent else None
    ls = _lib.launch_struct(launch)
    ws, nbytes = alpha_workspace(catalog, absorbers, nl, nf, nf_low, prec, launch=ls)
    check(lib().rbx_alpha(catalog.handle, ctypes.byref(absorbers), _ptr(P), _ptr(T), _ptr(mix), _ptr(cl), nl, _ptr(fd),
                          _ptr(fn), nf, nf_low, _ptr(sc), prec, _ptr(out), _ptr(byc), _lib.launch_ref(ls), _ptr(ws),
                          nbytes, _stream()))
    return (out, byc) if by_constituent else out


def alpha_noalloc(catalog, absorbers, P, T, mix, freqs, precision='fp64', scale=None, by_constituent=False,
                  cloud=None, launch=None, out=None, by_c=None):
    """alpha() through rbx_alpha_noalloc: the library's static scratch instead of a workspace.
    out / by_c: preallocated outputs ((nl, nf) of the precision's dtype / (7, nl, nf) float64),
    so that the call itself allocates nothing on the device."""
    prec = _lib.PRECISIONS[precision]
    P, T, mix, fd, fn, nf_low, sc, nl = _alpha_inputs(P, T, mix, freqs, scale)
    nf = fd.shape[0]
    cl = None
    if absorbers.clouds:
        if cloud is None:
            raise ValueError('clouds_idp needs the cloud densities (cloud=)')
        cl = as_device(cloud)
    dt = torch.float32 if prec == _lib.FP32 else torch.float64
    if out is None:
        out = torch.empty((nl, nf), dtype=dt, device=device())
    if out.shape != (nl, nf) or out.dtype != dt or not out.is_contiguous():
        raise ValueError('out must be a contiguous ({}, {}) {} tensor'.format(nl, nf, dt))
    if by_constituent and by_c is None:
        by_c = torch.zeros((len(_lib.CONSTITUENTS), nl, nf), dtype=torch.float64, device=device())
    check(lib().rbx_alpha_noalloc(catalog.handle, ctypes.byref(absorbers), _ptr(P), _ptr(T), _ptr(mix), _ptr(cl), nl,
                                  _ptr(fd), _ptr(fn), nf, nf_low, _ptr(sc), prec, _ptr(out),
                                  _ptr(by_c if by_constituent else None), _lib.launch_ref(launch), _stream()))
    return (out, by_c) if by_constituent else out


def alpha_rescale(absorb, scale=None, precision='fp64'):
    """alpha[nl][nf] = sum_c scale[c][l] absorb[l][f][c] (rbx_alpha_rescale) on the device."""
    prec = _lib.PRECISIONS[precision]
    a = as_device(absorb)
    nl, nf, nc = a.shape
    sc = None
    if scale is not None:
        sc = as_device(scale)
        if sc.shape != (nc, nl):
            raise ValueError('scale must be ({}, {})'.format(nc, nl))
    out = torch.empty((nl, nf), dtype=torch.float32 if prec == _lib.FP32 else torch.float64, device=device())
    check(lib().rbx_alpha_rescale(_ptr(a), nl, nf, nc, _ptr(sc), prec, _ptr(out), _stream()))
    return out


def alpha_workspace(catalog, absorbers, nl, nf, nf_low, prec, reuse=None, launch=None):
    """Device scratch for rbx_alpha with this launch shape (torch's stream-ordered allocator):
    (tensor, bytes).  `reuse` is returned instead of a new tensor when it is large enough."""
    nb = ctypes.c_size_t(0)
    check(lib().rbx_alpha_workspace_size(catalog.handle, ctypes.byref(absorbers), nl, nf, nf_low, prec,
                                         _lib.launch_ref(launch), ctypes.byref(nb)))
    if reuse is not None and reuse.numel() * reuse.element_size() >= nb.value:
        return reuse, nb.value
    return torch.empty((max(1, (nb.value + 7) // 8),), dtype=torch.float64, device=device()), nb.value


def alpha_tangent(catalog, absorbers, P, T, mix, freqs, scale=None, precision='fp64', with_alpha=False,
                  launch=None):
    """d alpha / d(T, q_NH3, q_H2O): (3, nl, nf) float64 on the device; with_alpha also
    returns alpha (nl, nf) float64 from the same pass: (alpha, dalpha).  launch.tan_split
    splits each layer's lines into parts (a workspace from torch's allocator)."""
    prec = _lib.PRECISIONS[precision]
    P, T, mix, fd, fn, nf_low, sc, nl = _alpha_inputs(P, T, mix, freqs, scale)
    nf = fd.shape[0]
    ls = _lib.launch_struct(launch)
    nb = ctypes.c_size_t(0)
    check(lib().rbx_alpha_tangent_workspace_size(catalog.handle, ctypes.byref(absorbers), nl, nf, nf_low, prec,
                                                 _lib.launch_ref(ls), ctypes.byref(nb)))
    ws = torch.empty((max(1, (nb.value + 7) // 8),), dtype=torch.float64, device=device()) if nb.value else None
    out = torch.empty((3, nl, nf), dtype=torch.float64, device=device())
    val = torch.empty((nl, nf), dtype=torch.float64, device=device()) if with_alpha else None
    check(lib().rbx_alpha_tangent(catalog.handle, ctypes.byref(absorbers), _ptr(P), _ptr(T), _ptr(mix), nl,
                                  _ptr(fd), _ptr(fn), nf, nf_low, _ptr(sc), prec, _ptr(val), _ptr(out),
                                  _lib.launch_ref(ls), _ptr(ws), nb.value, _stream()))
    return (val, out) if with_alpha else out


def planet_desc(req, rpol, orientation=(0.0, 0.0), gtype='ellipse', limb='shape'):
    gt = {'ellipse': 0, 'sphere': 1, 'circle': 1}
    lm = {'shape': 0, 'sec': 1}
    if gtype not in gt:
        raise ValueError('{} not a supported planet shape (ellipse, sphere, circle)'.format(gtype))
    if limb not in lm:
        raise ValueError('unsupported limb {}'.format(limb))
    d = _lib.PlanetDesc()
    d.req = float(req)
    d.rpol = float(rpol)
    d.orientation[0] = float(orientation[0])
    d.orientation[1] = float(orientation[1])
    d.gtype = gt[gtype]
    d.limb = lm[limb]
    return d


def _rays(b):
    """(nray, 2) float64 device tensor of impact parameters."""
    if isinstance(b, torch.Tensor):
        return as_device(b).reshape(-1, 2)
    return as_device(np.asarray(b, dtype=np.float64).reshape(-1, 2))


def raypaths(planet, R, N, b, units='half_cm'):
    """Ray paths for impact parameters b (nray, 2): (ds (nray, nl-1) float64, nsteps (nray,) int32)."""
    R = as_device(R)
    N = as_device(N)
    b = _rays(b)
    nl = R.shape[0]
    nray = b.shape[0]
    ld = max(nl - 1, 1)
    ds = torch.empty((nray, ld), dtype=torch.float64, device=device())
    ns = torch.empty((nray,), dtype=torch.int32, device=device())
    check(lib().rbx_raypaths(ctypes.byref(planet), _ptr(R), _ptr(N), nl, _ptr(b), nray,
                             1 if units == 'half_cm' else 0, _ptr(ds), ld, _ptr(ns), _stream()))
    return ds, ns


# FP32-mode optical-depth cut: a ray's integration stops once tau > 20 at every frequency of a
# wavefront (batch granularity).  What is left weighs < e^-20 ~ 2e-9 of the weight sum, and of
# the T-weighted sum at most T_max / Tb times that (~1e-8 with deep layers at 7 Tb): below the
# FP32 sums' unit roundoff 6e-8.  On 800 layers x 1024 frequencies it moves Tb by 3e-13 K; the
# integration takes 1.04 ms against 1.19 ms at a cut of 30 and 1.43 ms at 50.
FP32_TAU_CUT = 20.0
# FP64 mode: tau > 40 (what is left weighs < e^-40 ~ 4e-18 of the weight sum and ~3e-17 of the
# T-weighted sum, below the FP64 unit roundoff 1.1e-16)
FP64_TAU_CUT = 40.0


MAX_ORDER_FREQS = 16384   # rbx_integrate_order's limits
MAX_ORDER_LAYERS = 131072


def integrate(alpha_table, T, hds, nsteps, disc=False, precision=None, tau_cut=0.0, with_intW=False, launch=None,
              ordered=False):
    """Tb (nray, nf) float64 from an alpha table (nl, nf) and ray paths.  ordered: integrate the
    frequencies in the order in which a reference ray passes tau_cut (rbx_integrate_order +
    rbx_integrate_cols; with a cut only): the wavefronts then stop together, and Tb differs
    from the request order's only by what the cut itself changes.  FP32 point rays in that
    order read each ray's path factors from a table written once per ray (16 nray ld
    bytes) rather than staging them in LDS per frequency group."""
    aprec = _lib.FP32 if alpha_table.dtype == torch.float32 else _lib.FP64
    prec = aprec if precision is None else _lib.PRECISIONS[precision]
    a = alpha_table.contiguous()
    T = as_device(T)
    nl, nf = a.shape
    nray, ld = hds.shape
    hds, nsteps = hds.contiguous(), nsteps.contiguous()   # held here: the C call reads their memory
    Tb = torch.empty((nray, nf), dtype=torch.float64, device=device())
    iw = torch.empty((nray, nf), dtype=torch.float64, device=device()) if with_intW else None
    if ordered and tau_cut > 0.0 and nf <= MAX_ORDER_FREQS and nl <= MAX_ORDER_LAYERS and nray > 0:
        cols = torch.empty((nf,), dtype=torch.int32, device=device())
        a_cols = torch.empty_like(a)
        fac = None
        if not disc and aprec == _lib.FP32 and prec == _lib.FP32:
            fac = torch.empty((nray, ld, 4), dtype=torch.float32, device=device())   # 16-byte records
        check(lib().rbx_integrate_order(_ptr(a), aprec, nl, nf, _ptr(hds), ld, _ptr(nsteps), nray, float(tau_cut),
                                        _ptr(cols), _ptr(a_cols), _stream()))
        if fac is not None:
            check(lib().rbx_integrate_factors(_ptr(T), nl, _ptr(hds), ld, _ptr(nsteps), nray, _ptr(fac), _stream()))
        check(lib().rbx_integrate_cols(_ptr(a_cols), aprec, _ptr(cols), _ptr(T), nl, nf, _ptr(hds), ld,
                                       _ptr(nsteps), nray, 1 if disc else 0, prec, float(tau_cut),
                                       _ptr(Tb), _ptr(iw), _ptr(fac), _lib.launch_ref(launch), _stream()))
        return (Tb, iw) if with_intW else Tb
    check(lib().rbx_integrate(_ptr(a), aprec, _ptr(T), nl, nf, _ptr(hds), ld,
                              _ptr(nsteps), nray, 1 if disc else 0, prec, float(tau_cut),
                              _ptr(Tb), _ptr(iw), _lib.launch_ref(launch), _stream()))
    return (Tb, iw) if with_intW else Tb


def integrate_profiles(alpha_table, T, hds, nsteps, disc=False, precision=None):
    """tau, W, Tb_lyr at every path point: three (nray, nl, nf) float64 tensors."""
    aprec = _lib.FP32 if alpha_table.dtype == torch.float32 else _lib.FP64
    prec = aprec if precision is None else _lib.PRECISIONS[precision]
    a = alpha_table.contiguous()
    T = as_device(T)
    nl, nf = a.shape
    nray, ld = hds.shape
    hds, nsteps = hds.contiguous(), nsteps.contiguous()   # held here: the C call reads their memory
    out = [torch.empty((nray, nl, nf), dtype=torch.float64, device=device()) for _ in range(3)]
    check(lib().rbx_integrate_profiles(_ptr(a), aprec, _ptr(T), nl, nf, _ptr(hds), ld,
                                       _ptr(nsteps), nray, 1 if disc else 0, prec,
                                       *[_ptr(t) for t in out], _stream()))
    return tuple(out)


def jacobian(alpha_table, dalpha, T, hds, nsteps, tau_cut=0.0):
    """J (nray, 3, nl, nf) float32: dTb/dT, dTb/dq_NH3 and dTb/dq_H2O per layer; and Tb (nray, nf)."""
    aprec = _lib.FP32 if alpha_table.dtype == torch.float32 else _lib.FP64
    a = alpha_table.contiguous()
    T = as_device(T)
    nl, nf = a.shape
    nray, ld = hds.shape
    hds, nsteps = hds.contiguous(), nsteps.contiguous()   # held here: the C call reads their memory
    dalpha = dalpha.contiguous()
    J = torch.empty((nray, 3, nl, nf), dtype=torch.float32, device=device())
    Tb = torch.empty((nray, nf), dtype=torch.float64, device=device())
    check(lib().rbx_jacobian(_ptr(a), aprec, _ptr(dalpha), _ptr(T), nl, nf,
                             _ptr(hds), ld, _ptr(nsteps), nray, float(tau_cut),
                             _ptr(J), _ptr(Tb), _stream()))
    return J, Tb


def _grad_input(gTb, nray, nf):
    g = gTb if isinstance(gTb, torch.Tensor) else torch.as_tensor(np.asarray(gTb))
    g = g.to(device=device(), dtype=torch.float64).contiguous()
    if g.shape != (nray, nf):
        raise ValueError('gTb must be ({}, {}), got {}'.format(nray, nf, tuple(g.shape)))
    return g


def path_order(nsteps):
    """Rays by path length, longest first (int32 on the device): the schedule that lets the
    adjoint's wavefronts hold rays of similar length."""
    return torch.argsort(nsteps.to(torch.int64), descending=True, stable=True).to(torch.int32)


def _adjoint_workspace(nl, nf, nray, natm, launch):
    nb = ctypes.c_size_t(0)
    check(lib().rbx_integrate_adjoint_workspace_size(nl, nf, nray, natm, _lib.launch_ref(launch), ctypes.byref(nb)))
    if nb.value == 0:
        return None, 0
    return torch.empty(((nb.value + 7) // 8,), dtype=torch.float64, device=device()), nb.value


def integrate_adjoint(alpha_table, T, hds, nsteps, gTb, order=None, precision=None, tau_cut=0.0, launch=None):
    """Vector-Jacobian product of integrate(): (galpha (nl, nf), gT (nl,)) float64, the gradient
    of L = sum gTb * Tb with respect to the alpha table and (directly) the layer temperatures."""
    aprec = _lib.FP32 if alpha_table.dtype == torch.float32 else _lib.FP64
    prec = aprec if precision is None else _lib.PRECISIONS[precision]
    a = alpha_table.contiguous()
    T = as_device(T)
    nl, nf = a.shape
    nray, ld = hds.shape
    hds, nsteps = hds.contiguous(), nsteps.contiguous()
    g = _grad_input(gTb, nray, nf)
    od = path_order(nsteps) if order is None else (order if isinstance(order, torch.Tensor) else _i32(order))
    galpha = torch.empty((nl, nf), dtype=torch.float64, device=device())
    gT = torch.empty((nl,), dtype=torch.float64, device=device())
    ws, nb = _adjoint_workspace(nl, nf, nray, 1, launch)
    check(lib().rbx_integrate_adjoint(_ptr(a), aprec, _ptr(T), nl, nf, _ptr(hds), ld, _ptr(nsteps), nray, _ptr(od),
                                      prec, float(tau_cut), _ptr(g), _ptr(galpha), _ptr(gT),
                                      _lib.launch_ref(launch), _ptr(ws), nb, _stream()))
    return galpha, gT


def integrate_adjoint_batch(alpha_table, T, layer_off, ray_atm, hds, nsteps, gTb, order=None, precision=None,
                            tau_cut=0.0, launch=None):
    """integrate_adjoint() for rays through several atmospheres laid end to end: galpha
    (nl_total, nf) and gT (nl_total,).  order: the ray schedule (device int32), whose
    consecutive launch.adj_rays entries should share an atmosphere (RetrievalBatch's does)."""
    aprec = _lib.FP32 if alpha_table.dtype == torch.float32 else _lib.FP64
    prec = aprec if precision is None else _lib.PRECISIONS[precision]
    a = alpha_table.contiguous()
    T = as_device(T)
    nl, nf = a.shape
    nray, ld = hds.shape
    hds, nsteps = hds.contiguous(), nsteps.contiguous()
    g = _grad_input(gTb, nray, nf)
    off = layer_off if isinstance(layer_off, torch.Tensor) else _i32(layer_off)
    ra = ray_atm if isinstance(ray_atm, torch.Tensor) else _i32(ray_atm)
    od = None if order is None else (order if isinstance(order, torch.Tensor) else _i32(order))
    galpha = torch.empty((nl, nf), dtype=torch.float64, device=device())
    gT = torch.empty((nl,), dtype=torch.float64, device=device())
    ws, nb = _adjoint_workspace(nl, nf, nray, off.shape[0] - 1, launch)
    check(lib().rbx_integrate_adjoint_batch(_ptr(a), aprec, _ptr(T), nl, nf, _ptr(off), off.shape[0] - 1,
                                            _ptr(hds), ld, _ptr(nsteps), _ptr(ra), _ptr(od), nray, prec,
                                            float(tau_cut), _ptr(g), _ptr(galpha), _ptr(gT),
                                            _lib.launch_ref(launch), _ptr(ws), nb, _stream()))
    return galpha, gT


def alpha_vjp(dalpha, galpha, gT=None):
    """grad (3, nl) float64: dL/dT, dL/dq_NH3, dL/dq_H2O per layer from the adjoint's galpha
    (nl, nf) and gT (nl,) and the tangents dalpha (3, nl, nf) (rbx_alpha_vjp)."""
    dalpha = dalpha.contiguous()
    galpha = galpha.contiguous()
    nl, nf = galpha.shape
    if dalpha.shape != (3, nl, nf) or dalpha.dtype != torch.float64 or galpha.dtype != torch.float64:
        raise ValueError('dalpha (3, nl, nf) and galpha (nl, nf) float64 expected')
    if gT is not None:
        gT = gT.contiguous()
        if gT.shape != (nl,) or gT.dtype != torch.float64:
            raise ValueError('gT must be ({},) float64'.format(nl))
    grad = torch.empty((3, nl), dtype=torch.float64, device=device())
    check(lib().rbx_alpha_vjp(_ptr(dalpha), _ptr(galpha), _ptr(gT), nl, nf, _ptr(grad), _stream()))
    return grad


# ------------------------------------------------------------- Jupiter lines
def jupiter_catalog(raw, truncate_strength=None, truncate_freq=None, nh3_inversion=None, h2o=None, co=False,
                    orton=False):
    """Catalogue with the Jupiter line lists and the config's truncation.

    raw: lines.raw_catalog(); truncate_strength/freq: per-constituent dicts as
    config.truncate_* (defaults: h2s and ph3 at I0 > 1e-22); nh3_inversion /
    h2o: optional replacement lists (synthetic workloads); co / orton: also load
    the CO lines (co_ddb) / the Orton CIA table (h2_orton).
    """
    from . import lines as L
    ts = {'h2s': 1e-22, 'ph3': 1e-22} if truncate_strength is None else truncate_strength
    tf = truncate_freq or {}
    cat = Catalog()
    amm = dict(raw['ammonia'])
    if nh3_inversion is not None:
        for k in ('fo', 'Io', 'Eo', 'gammaNH3o'):
            amm[k] = nh3_inversion[k]
    cat.set_lines(_lib.LINES_NH3_INV, amm)
    cat.set_lines(_lib.LINES_NH3_ROT, amm)
    cat.set_lines(_lib.LINES_NH3_V2, amm)
    cat.set_lines(_lib.LINES_NH3_SJS, raw['nh3'])
    cat.set_lines(_lib.LINES_H2S, L.truncate(raw['h2s'], ('f0', 'I0', 'E', 'GH2S'),
                                             ts.get('h2s'), tf.get('h2s')))
    cat.set_lines(_lib.LINES_PH3, L.truncate_ph3(raw['ph3jh'], raw['PH3WGT'], ts.get('ph3'), tf.get('ph3')))
    h2o_lines = L.h2o_bk_lines() if h2o is None else h2o
    cat.set_lines(_lib.LINES_H2O, L.truncate_h2o(h2o_lines, ts.get('h2o'), tf.get('h2o')))
    if co:
        cat.set_lines(_lib.LINES_CO, L.load('co'))      # co_ddb reads no truncation (co_ddb.py:14-19)
    if orton:
        cat.set_lines(_lib.TABLE_H2_ORTON, L.orton_columns(L.load_orton()))
    return cat


# ------------------------------------------------------- many atmospheres
def _i32(x):
    return as_device(np.asarray(x, dtype=np.int32), dtype=torch.int32)


def raypaths_batch(planet, R, N, layer_off, ray_atm, b, ld, units='half_cm', off_dev=None, ray_atm_dev=None):
    """Ray paths through several atmospheres laid end to end (rbx_raypaths_batch).

    R, N: (nl_total,) concatenated layer radii / refractive indices; layer_off (natm+1,)
    host ints; ray_atm (nray,) host ints (checked here); off_dev / ray_atm_dev: their
    int32 device copies if the caller holds them; ld >= max layers - 1.
    Returns (ds, nsteps)."""
    off = np.asarray(layer_off, dtype=np.int64)
    ra = np.asarray(ray_atm, dtype=np.int64)
    _check_batch(off, ra, R.shape[0] if isinstance(R, torch.Tensor) else len(R), ld)
    R = as_device(R)
    N = as_device(N)
    b = _rays(b)
    nray = b.shape[0]
    off_d = off_dev if off_dev is not None else _i32(off)     # keep the device copies alive over the call
    ra_d = ray_atm_dev if ray_atm_dev is not None else _i32(ra)
    ds = torch.empty((nray, ld), dtype=torch.float64, device=device())
    ns = torch.empty((nray,), dtype=torch.int32, device=device())
    check(lib().rbx_raypaths_batch(ctypes.byref(planet), _ptr(R), _ptr(N), R.shape[0], _ptr(off_d), len(off) - 1,
                                   _ptr(b), nray, _ptr(ra_d), 1 if units == 'half_cm' else 0, _ptr(ds), ld,
                                   _ptr(ns), _stream()))
    return ds, ns


def _check_batch(off, ray_atm, nl_total, ld):
    if off.ndim != 1 or len(off) < 2 or off[0] != 0 or off[-1] != nl_total or np.any(np.diff(off) < 2):
        raise ValueError('layer offsets must run from 0 to nl_total with >= 2 layers per atmosphere')
    if len(ray_atm) and (ray_atm.min() < 0 or ray_atm.max() >= len(off) - 1):
        raise ValueError('ray atmosphere index out of range')
    if ld < int(np.max(np.diff(off))) - 1:
        raise ValueError('ld={} < largest atmosphere - 1'.format(ld))


def jacobian_batch(alpha_table, dalpha, T, layer_off, ray_atm, hds, nsteps, ldj, order=None, tau_cut=0.0,
                   precision=None):
    """Tb (nray, nf) and J (nray, 3, ldj, nf) float32 for rays through several atmospheres.

    alpha_table (nl_total, nf), dalpha (3, nl_total, nf) and T (nl_total,) concatenated;
    layer_off / ray_atm / order as device or host int arrays (order None: 0..nray-1)."""
    aprec = _lib.FP32 if alpha_table.dtype == torch.float32 else _lib.FP64
    prec = aprec if precision is None else _lib.PRECISIONS[precision]
    a = alpha_table.contiguous()
    T = as_device(T)
    nl, nf = a.shape
    nray, ld = hds.shape
    hds, nsteps = hds.contiguous(), nsteps.contiguous()   # held here: the C call reads their memory
    off = layer_off if isinstance(layer_off, torch.Tensor) else _i32(layer_off)
    ra = ray_atm if isinstance(ray_atm, torch.Tensor) else _i32(ray_atm)
    od = None if order is None else (order if isinstance(order, torch.Tensor) else _i32(order))
    dalpha = dalpha.contiguous()
    J = torch.empty((nray, 3, ldj, nf), dtype=torch.float32, device=device())
    Tb = torch.empty((nray, nf), dtype=torch.float64, device=device())
    check(lib().rbx_jacobian_batch(_ptr(a), aprec, _ptr(dalpha), _ptr(T), nl, nf, _ptr(off),
                                   off.shape[0] - 1, _ptr(hds), ld, _ptr(nsteps), _ptr(ra),
                                   _ptr(od), nray, prec, float(tau_cut), _ptr(J), ldj, _ptr(Tb), _stream()))
    return J, Tb


def integrate_batch(alpha_table, T, layer_off, ray_atm, hds, nsteps, order=None, disc=False, precision=None,
                    tau_cut=0.0, launch=None):
    """Tb (nray, nf) float64 for rays through several atmospheres laid end to end."""
    aprec = _lib.FP32 if alpha_table.dtype == torch.float32 else _lib.FP64
    prec = aprec if precision is None else _lib.PRECISIONS[precision]
    a = alpha_table.contiguous()
    T = as_device(T)
    nl, nf = a.shape
    nray, ld = hds.shape
    hds, nsteps = hds.contiguous(), nsteps.contiguous()   # held here: the C call reads their memory
    off = layer_off if isinstance(layer_off, torch.Tensor) else _i32(layer_off)
    ra = ray_atm if isinstance(ray_atm, torch.Tensor) else _i32(ray_atm)
    od = None if order is None else (order if isinstance(order, torch.Tensor) else _i32(order))
    Tb = torch.empty((nray, nf), dtype=torch.float64, device=device())
    check(lib().rbx_integrate_batch(_ptr(a), aprec, _ptr(T), nl, nf, _ptr(off), off.shape[0] - 1,
                                    _ptr(hds), ld, _ptr(nsteps), _ptr(ra), _ptr(od), nray,
                                    1 if disc else 0, prec, float(tau_cut), _ptr(Tb), None, _lib.launch_ref(launch),
                                    _stream()))
    return Tb


def atmosphere_props(P, T, mix, Z, layer_off, req, gm, p_ref=1.0, off_dev=None):
    """computeProp's property table (rbx_atmosphere_props): a dict of (nl_total,) float64
    device tensors keyed by _lib.PROP_ROWS (Z, R, P, GM, AMU, REFR, N, H, LAPSE, LAPSEP, g)
    of atmospheres laid end to end, from their P, T, mix and Z."""
    P, T, mix, Z = as_device(P), as_device(T), as_device(mix), as_device(Z)
    nlt = P.shape[0]
    if T.shape != (nlt,) or Z.shape != (nlt,) or mix.shape != (len(_lib.SPECIES), nlt):
        raise ValueError('shape mismatch: P {}, T {}, Z {}, mix {}'.format(tuple(P.shape), tuple(T.shape),
                                                                           tuple(Z.shape), tuple(mix.shape)))
    off = np.asarray(layer_off.cpu() if isinstance(layer_off, torch.Tensor) else layer_off, dtype=np.int64)
    if off[0] != 0 or off[-1] != nlt or np.any(np.diff(off) < 1):
        raise ValueError('layer offsets must run from 0 to nl_total')
    off_d = _i32(off) if off_dev is None else off_dev
    props = torch.empty((len(_lib.PROP_ROWS), nlt), dtype=torch.float64, device=device())
    check(lib().rbx_atmosphere_props(_ptr(P), _ptr(T), _ptr(mix), _ptr(Z), nlt, _ptr(off_d), len(off) - 1,
                                     float(req), float(gm), float(p_ref), _ptr(props), _stream()))
    return {k: props[i] for i, k in enumerate(_lib.PROP_ROWS)}


def atmosphere_batch(P, T, mix, layer_off, req, gm, p_ref=1.0, Z=None, off_dev=None):
    """Z, R, N (each (nl_total,) float64 on the device) of atmospheres laid end to end;
    Z None: hydrostatic z (rbx_atmosphere_batch).  layer_off: host offsets (checked here);
    off_dev: their int32 device copy if the caller holds one (then nothing is copied)."""
    P, T, mix = as_device(P), as_device(T), as_device(mix)
    nlt = P.shape[0]
    if T.shape != (nlt,) or mix.shape != (len(_lib.SPECIES), nlt):
        raise ValueError('shape mismatch: P {}, T {}, mix {}'.format(tuple(P.shape), tuple(T.shape),
                                                                     tuple(mix.shape)))
    off = np.asarray(layer_off.cpu() if isinstance(layer_off, torch.Tensor) else layer_off, dtype=np.int64)
    if off[0] != 0 or off[-1] != nlt or np.any(np.diff(off) < 1):
        raise ValueError('layer offsets must run from 0 to nl_total')
    off_d = _i32(off) if off_dev is None else off_dev
    Zin = None if Z is None else as_device(Z)
    Zo, R, N = (torch.empty((nlt,), dtype=torch.float64, device=device()) for _ in range(3))
    check(lib().rbx_atmosphere_batch(_ptr(P), _ptr(T), _ptr(mix), _ptr(Zin), nlt, _ptr(off_d), len(off) - 1,
                                     float(req), float(gm), float(p_ref), _ptr(Zo), _ptr(R), _ptr(N), _stream()))
    return Zo, R, N
