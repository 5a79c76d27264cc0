"""Float64 torch-autograd restatement of the forward model (TEST INFRASTRUCTURE ONLY).

The gradient reference for the tangent / adjoint kernels: the absorber
plug-ins and Brightness.single written once more, vectorised over layers, in
torch float64, so that torch.autograd differentiates Tb with respect to the
layer temperatures and mixing ratios exactly (no finite-difference step).

Restated from the reference formulas (the same ones oracle/absorbers.py and
oracle/rt.py follow, file:line cited per function), with the layer inputs as
tensors:
  T [nl] K, P [nl] bar, mix [7][nl] mixing ratios in the rows
  H2, HE, CH4, NH3, H2O, H2S, PH3 (the librbx mix-table order).
nh3_hs_sjs is restated for P < 400 bar only (the nh3_hs branch,
nh3_hs_sjs.py:14-15): deeper layers raise, they are not needed by the
gradient tests.  The forward values are pinned against oracle.layers /
oracle.rt in tests/test_oracle.py, so the autograd derivatives are those of
the pinned forward model.
"""
import math

import numpy as np
import torch

from . import absorbers as A

DT = torch.float64
ROWS = {'H2': 0, 'HE': 1, 'CH4': 2, 'NH3': 3, 'H2O': 4, 'H2S': 5, 'PH3': 6}


def _t(a):
    return torch.as_tensor(np.asarray(a, dtype=np.float64), dtype=DT)


def _pp(P, mix, name):
    return P * mix[ROWS[name]]


def h2_jj_ddb(f, T, P, mix, state='e'):
    """constituents/h2/h2_jj_ddb.py:7-38 -> (nl, nf) cm^-1."""
    P_h2, P_he, P_ch4 = _pp(P, mix, 'H2'), _pp(P, mix, 'HE'), _pp(P, mix, 'CH4')
    th = 273.0 / T
    if state == 'e':                                     # h2_jj_ddb.py:17-23
        pre = torch.clamp((T / 55.0) ** 2.7, max=1.0)
        pre = torch.clamp(pre * (T / 120.0) ** 0.55, max=1.0)
    else:                                                # h2_jj_ddb.py:24-27
        pre = torch.clamp((T / 40.0) ** 2.5, max=1.0)
    k = 3.9522E-14 * P_h2 * pre * (P_h2 * th ** 3.12 + 1.382 * P_he * th ** 2.24 + 9.322 * P_ch4 * th ** 3.34)
    return k[:, None] * (f ** 2)[None, :]


def h2o_bk(f, T, P, mix, data):
    """constituents/h2o/h2o_bk.py:65-189 (lines with base, continuum) -> (nl, nf) cm^-1."""
    P_h2, P_he, P_h2o = _pp(P, mix, 'H2'), _pp(P, mix, 'HE'), _pp(P, mix, 'H2O')
    d = {k: _t(v) for k, v in data.items()}
    Th = (300.0 / T)[:, None]                                             # (nl, 1)
    density = (A.H2O_M_AMU * P_h2o) / (8.314472e-5 * T)                   # h2o_bk.py:85
    density = A.H2O_ISO * (density / A.H2O_M_AMU) * A.H2O_NA * 1e-6       # h2o_bk.py:88
    S = d['I_o'] * Th ** 2.5 * torch.exp(d['E_o'] * (1.0 - Th))           # (nl, L)
    df = (d['w_s'] * P_h2o[:, None] * Th ** d['x_s'] + d['w_h2'] * P_h2[:, None] * Th ** d['x_h2']
          + d['w_he'] * P_he[:, None] * Th ** d['x_he'])
    fo = d['f_o'][None, :, None]
    sh = d['SR'][None, :, None]
    dfm = df[:, :, None]
    ff = f[None, None, :]
    base = (df / (562500.0 + df ** 2))[:, :, None]                        # h2o_bk.py:174-184
    Aa = (ff / fo) ** 2 / math.pi
    B = dfm / ((ff - fo - sh) ** 2 + dfm ** 2)
    C = dfm / ((ff + fo + sh) ** 2 + dfm ** 2)
    lines = (S[:, :, None] * Aa * (B - base + C - base)).sum(dim=1)
    lines = 4.342945e-4 * density[:, None] * lines                        # h2o_bk.py:99
    Cf_he = 1e6 * 1.03562010226e-10                                       # h2o_bk.py:102-119
    Cf_h2 = 1e6 * 5.07722009423e-11
    Th1 = Th[:, 0]
    foreign = (Cf_he * P_he * P_h2o + Cf_h2 * P_h2 * P_h2o) * Th1 ** 3.0
    self_c = 3.1e-07 * Th1 ** 12.0 * (P_h2o / 0.001) ** 2
    cont = 4.342945 * (foreign + self_c)
    return (lines + cont[:, None] * (f ** 2)[None, :]) / A.DB_PER_INVCM


def _br(f, f0, gamma, zeta, delta, ITG):
    """The Ben-Reuven sum shared by h2s_ddb / ph3_jh (h2s_ddb.py:76-79): (nl, nf)."""
    ff = f[None, None, :]
    g, z, dl = gamma[:, :, None], zeta[:, :, None], delta[:, :, None]
    fd2 = (f0[None, :, None] + dl) ** 2
    num = (g - z) * ff ** 2 + (g + z) * (fd2 + g ** 2 - z ** 2)
    den = (ff ** 2 - fd2 - g ** 2 + z ** 2) ** 2 + 4.0 * ff ** 2 * g ** 2
    shape = A.GHZ_PER_INVCM * 2.0 * (ff / f0[None, :, None]) ** 2 * num / (math.pi * den)
    return (shape * ITG[:, :, None]).sum(dim=1)


def h2s_ddb(f, T, P, mix, data):
    """constituents/h2s/h2s_ddb.py:42-87 -> (nl, nf) cm^-1."""
    T0 = 296.0
    P_h2, P_he, P_h2s = _pp(P, mix, 'H2'), _pp(P, mix, 'HE'), _pp(P, mix, 'H2S')
    f0, I0, E, G = (_t(data[k]) for k in ('f0', 'I0', 'E', 'GH2S'))
    gamma = (T0 / T)[:, None] ** 0.7 * (1.960 * P_h2[:, None] + 1.200 * P_he[:, None] + G * P_h2s[:, None])
    delta = (1.28 * P_h2s)[:, None].expand_as(gamma)
    ITG = I0 * torch.exp(-((1.0 / T)[:, None] - 1.0 / T0) * E * A.HCK)
    s = _br(f, f0, gamma, gamma, delta, ITG)
    return (A.GEISA_COEF * (P_h2s / T0) * (T0 / T) ** 3.5)[:, None] * s


def ph3_jh(f, T, P, mix, data):
    """constituents/ph3/ph3_jh.py:64-108 -> (nl, nf) cm^-1."""
    T0 = 300.0
    P_h2, P_he, P_ph3 = _pp(P, mix, 'H2'), _pp(P, mix, 'HE'), _pp(P, mix, 'PH3')
    f0, I0, E = (_t(data[k]) for k in ('f0', 'I0', 'E'))
    wI, wF, wS = (_t(data[k]) for k in ('WgtI0', 'WgtFGB', 'WgtSB'))
    gamma = ((T0 / T)[:, None] ** (2.0 / 3.0) * (3.2930 * P_h2 + 1.6803 * P_he)[:, None] * wF
             + (T0 / T)[:, None] * 4.2157 * P_ph3[:, None] * wS)
    zero = torch.zeros_like(gamma)
    ITG = I0 * wI * torch.exp(-((1.0 / T)[:, None] - 1.0 / T0) * E * A.HCK)
    s = _br(f, f0, gamma, zero, zero, ITG)
    return (A.GEISA_COEF * (P_ph3 / T0) * (T0 / T) ** 3.5)[:, None] * s


def _gross(f, fo, gamma):
    """nh3_hs.py:286-290: (nl, L, nf)."""
    ff = f[None, None, :]
    return (4.0 / math.pi) * ff ** 2 * gamma / ((fo ** 2 - ff ** 2) ** 2 + 4.0 * ff ** 2 * gamma ** 2)


def _nh3_hs_part(f, T, P, mix, d, low_set=A.HS_INV_LO):
    """nh3_hs.__alpha__ (constituents/nh3/nh3_hs.py:117-340) for one side of 30 GHz
    (nh3_dbs.__alpha__, nh3_dbs.py:92-313, with low_set=DBS_INV_LO)."""
    P_h2, P_he, P_nh3 = _pp(P, mix, 'H2'), _pp(P, mix, 'HE'), _pp(P, mix, 'NH3')
    cs = low_set if bool(torch.all(f <= A.HS_F_SPLIT)) else A.HS_INV_HI
    (gH2, gHe, gNH3, GH2, GHe, GNH3, zH2, zHe, zNH3, ZH2, ZHe, ZNH3, dd, Con) = cs
    Tdiv = (300.0 / T)[:, None]
    T295 = (295.0 / T)[:, None]
    go = d['gammaNH3o'][None, :]
    gamma = (gH2 * P_h2[:, None]) * Tdiv ** GH2 + (gHe * P_he[:, None]) * Tdiv ** GHe \
        + (gNH3 * P_nh3[:, None] * go) * T295 ** GNH3
    zeta = (zH2 * P_h2[:, None]) * Tdiv ** ZH2 + (zHe * P_he[:, None]) * Tdiv ** ZHe \
        + (zNH3 * P_nh3[:, None] * go) * T295 ** ZNH3
    pst = dd * gamma
    hck = A.HS_HC / A.HS_KB
    ST = d['Io'] * torch.exp(-(1.0 / T[:, None] - 1.0 / 300.0) * d['Eo'] * hck)
    noshape = Con * A.HS_COEF * (P_nh3 / 300.0)[:, None] * Tdiv ** 3.5 * ST
    ff = f[None, None, :]
    fo = d['fo'][None, :, None]
    g, z, p = gamma[:, :, None], zeta[:, :, None], pst[:, :, None]
    Aa = (2.0 / math.pi) * (ff / fo) ** 2
    Ff = Aa * ((g - z) * ff ** 2 + (g + z) * ((fo + p) ** 2 + g ** 2 - z ** 2)) / (
        (ff ** 2 - (fo + p) ** 2 - g ** 2 + z ** 2) ** 2 + 4.0 * ff ** 2 * g ** 2)
    a_inv = (noshape[:, :, None] * (Ff / A.HS_GHZ_TO_INVCM)).sum(dim=1)
    ST_rot = d['Io_rot'] * torch.exp((1.0 / 300.0 - 1.0 / T[:, None]) * d['Eo_rot'] * hck)
    g_rot = (0.2984 * P_h2[:, None] * d['gH2_rot']) * Tdiv ** 0.8730 \
        + (0.75 * P_he[:, None] * d['gHe_rot']) * Tdiv ** (2.0 / 3.0) \
        + (3.1789 * P_nh3[:, None] * d['gNH3_rot']) * Tdiv
    F_rot = _gross(f, d['fo_rot'][None, :, None], g_rot[:, :, None]) / A.HS_GHZ_TO_INVCM
    a_rot = ((2.4268 * A.HS_COEF * (P_nh3 / 300.0)[:, None] * Tdiv ** 3.5 * ST_rot)[:, :, None] * F_rot).sum(dim=1)
    ST_v2 = d['Io_v2'] * torch.exp((1.0 / 300.0 - 1.0 / T[:, None]) * d['Eo_v2'] * hck)
    g_v2 = P_h2 * 1.4 * (300.0 / T) ** 0.73 + P_he * 0.68 * (300.0 / T) ** 0.5716 + P_nh3 * 9.5 * (300.0 / T)
    F_v2 = _gross(f, d['fo_v2'][None, :, None], g_v2[:, None, None]) / A.HS_GHZ_TO_INVCM
    a_v2 = ((1.1206 * A.HS_COEF * (P_nh3 / 300.0)[:, None] * Tdiv ** 3.5 * ST_v2)[:, :, None] * F_v2).sum(dim=1)
    return a_inv + a_rot + a_v2


def nh3_hs(f, T, P, mix, data, low_set=A.HS_INV_LO):
    """nh3_hs.alpha (nh3_hs.py:96-114, 333-338): low half then high half, each clamped."""
    if bool(torch.any(P >= 400.0)):
        raise ValueError('oracle.jacobian restates nh3_hs_sjs for P < 400 bar only')
    d = {k: _t(v) for k, v in data.items() if k in ('fo', 'Io', 'Eo', 'gammaNH3o', 'fo_rot', 'Io_rot',
                                                     'Eo_rot', 'gNH3_rot', 'gH2_rot', 'gHe_rot',
                                                     'fo_v2', 'Io_v2', 'Eo_v2')}
    parts = []
    for sel in (f <= A.HS_F_SPLIT, f > A.HS_F_SPLIT):
        if bool(sel.any()):
            a = _nh3_hs_part(f[sel], T, P, mix, d, low_set)
            parts.append(torch.where(a < 0.0, torch.full_like(a, 1.0E-8), a))
    return torch.cat(parts, dim=1)


def alpha_layers(freqs, T, P, mix, lines, h2state='e', nh3='nh3_hs_sjs'):
    """Alpha.get_layers of the Jupiter absorber set, layer-major (nl, nf) cm^-1
    (alpha.py:151-305; sorted constituent order h2, h2o, h2s, nh3, ph3); nh3='nh3_dbs_sjs'
    gives the Saturn / Uranus / Neptune set.
    freqs must be ascending (nh3_hs returns its <= 30 GHz half first)."""
    low = {'nh3_hs_sjs': A.HS_INV_LO, 'nh3_dbs_sjs': A.DBS_INV_LO}[nh3]
    f = _t(freqs)
    if bool(torch.any(f[1:] < f[:-1])):
        raise ValueError('freqs must be ascending')
    return (h2_jj_ddb(f, T, P, mix, h2state) + h2o_bk(f, T, P, mix, lines['h2o'])
            + h2s_ddb(f, T, P, mix, lines['h2s']) + nh3_hs(f, T, P, mix, lines['nh3_hs'], low)
            + ph3_jh(f, T, P, mix, lines['ph3']))


def brightness(alpha, T, ds, layer4ds):
    """Brightness.single, point ray (brightness.py:30-126): Tb [nf] of one ray.
    alpha (nl, nf); ds km and layer4ds from oracle.geometry.compute_ds."""
    nf = alpha.shape[1]
    if ds is None:
        return torch.full((nf,), 2.725, dtype=DT)
    tau = torch.zeros(nf, dtype=DT)
    W = torch.zeros(nf, dtype=DT)
    intW = torch.zeros(nf, dtype=DT)
    Tb_lyr = torch.zeros(nf, dtype=DT)
    for i in range(len(ds) - 1):
        d = ds[i] * 1.0e3 / 1.0e-2                       # brightness.py:66
        i0, i1 = layer4ds[i], layer4ds[i + 1]
        a0, a1 = alpha[i0], alpha[i1]
        tau = tau + (a0 + a1) * d / 2.0
        W1 = a1 * torch.exp(-tau)
        intW = intW + (W1 + W) * d / 2.0
        Tb_lyr = Tb_lyr + (T[i1] * W1 + T[i0] * W) * d / 2.0
        W = W1
    return torch.where(Tb_lyr < 2.725, torch.full_like(Tb_lyr, 2.725), Tb_lyr / intW)


def jacobian(freqs, T, P, mix, lines, paths, planes=2):
    """dTb/dT[l], dTb/dq_NH3[l] (and, planes=3, dTb/dq_H2O[l]) by reverse-mode autograd.

    paths: list of (ds, layer4ds) per ray.  Returns (Tb [nray][nf],
    J [nray][planes][nl][nf]) as numpy float64.
    """
    T = _t(T).requires_grad_(True)
    mix = _t(mix).requires_grad_(True)
    P = _t(P)
    alpha = alpha_layers(freqs, T, P, mix, lines)
    nl, nf = alpha.shape
    Tb = torch.stack([brightness(alpha, T, ds, l4) for ds, l4 in paths])
    J = np.zeros((len(paths), planes, nl, nf))
    rows = (ROWS['NH3'], ROWS['H2O'])[:planes - 1]
    for r in range(len(paths)):
        for j in range(nf):
            gT, gm = torch.autograd.grad(Tb[r, j], (T, mix), retain_graph=True, allow_unused=True)
            if gT is not None:
                J[r, 0, :, j] = gT.numpy()
            if gm is not None:
                for p, row in enumerate(rows):
                    J[r, 1 + p, :, j] = gm[row].numpy()
    return Tb.detach().numpy(), J


def vjp(freqs, T, P, mix, lines, paths, gTb):
    """One reverse pass of L = sum_{ray,f} gTb[ray][f] Tb[ray][f] (rays whose gTb row is all
    zero are left out of L, so a NaN Tb there does not reach the gradient).  Returns
    (Tb [nray][nf], grad [3][nl]): dL/dT, dL/dq_NH3, dL/dq_H2O per layer, numpy float64."""
    T = _t(T).requires_grad_(True)
    mix = _t(mix).requires_grad_(True)
    P = _t(P)
    alpha = alpha_layers(freqs, T, P, mix, lines)
    g = np.asarray(gTb, dtype=np.float64)
    Tb = torch.stack([brightness(alpha, T, ds, l4) for ds, l4 in paths])
    keep = [r for r in range(len(paths)) if np.any(g[r] != 0.0)]
    L = sum((Tb[r] * _t(g[r])).sum() for r in keep) if keep else (T * 0.0).sum()
    gT, gm = torch.autograd.grad(L, (T, mix), allow_unused=True)
    nl = T.shape[0]
    grad = np.zeros((3, nl))
    if gT is not None:
        grad[0] = gT.numpy()
    if gm is not None:
        grad[1] = gm[ROWS['NH3']].numpy()
        grad[2] = gm[ROWS['H2O']].numpy()
    return Tb.detach().numpy(), grad


def alpha_tangents(freqs, T, P, mix, lines):
    """alpha (nl, nf) and d alpha / d(T, q_NH3, q_H2O) (3, nl, nf), numpy float64, by
    forward-mode automatic differentiation of alpha_layers: a layer's absorption depends
    on that layer's inputs only, so one Jacobian-vector product with a tangent of ones in
    one variable gives that variable's derivative for every layer at once (no graph is
    kept, so this scales to full line lists)."""
    from torch.func import jvp
    T = _t(T)
    P = _t(P)
    mix = _t(mix)
    nl = T.shape[0]

    def of_T(t):
        return alpha_layers(freqs, t, P, mix, lines)

    def of_row(row):
        def f(q):
            m = mix.clone()
            m[row] = q
            return alpha_layers(freqs, T, P, m, lines)
        return f

    ones = torch.ones(nl, dtype=DT)
    alpha, dT = jvp(of_T, (T,), (ones,))
    _, dq_nh3 = jvp(of_row(ROWS['NH3']), (mix[ROWS['NH3']],), (ones,))
    _, dq_h2o = jvp(of_row(ROWS['H2O']), (mix[ROWS['H2O']],), (ones,))
    return alpha.numpy(), torch.stack([dT, dq_nh3, dq_h2o]).numpy()


def rt_jacobian(alpha, T, paths, freq_index=None):
    """The integration's own derivatives by reverse-mode autograd of brightness():
    for each ray and frequency column, dTb/dalpha[l] and the direct dTb/dT[l] (T entering
    through T_l W_l only).  alpha (nl, nf), T (nl,); paths: list of (ds, layer4ds).
    Returns (Tb [nray][nc], Ja [nray][nl][nc], JT [nray][nl][nc]) over the columns
    freq_index (default all); chained with alpha_tangents these give the full Jacobian."""
    a = _t(alpha)
    nl, nf = a.shape
    cols = list(range(nf)) if freq_index is None else list(freq_index)
    a = a[:, cols].clone().requires_grad_(True)
    Tt = _t(T).requires_grad_(True)
    nc = len(cols)
    Tb = torch.stack([brightness(a, Tt, ds, l4) for ds, l4 in paths])
    Ja = np.zeros((len(paths), nl, nc))
    JT = np.zeros((len(paths), nl, nc))
    for r in range(len(paths)):
        for j in range(nc):
            ga, gt = torch.autograd.grad(Tb[r, j], (a, Tt), retain_graph=True, allow_unused=True)
            if ga is not None:
                Ja[r, :, j] = ga[:, j].numpy()
            if gt is not None:
                JT[r, :, j] = gt.numpy()
    return Tb.detach().numpy(), Ja, JT
