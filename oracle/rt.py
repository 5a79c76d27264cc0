"""FP64 restatement of Brightness.single (TEST INFRASTRUCTURE ONLY).

radiobear/brightness.py:30-126, vectorised over frequency: the trapezoid
tau prefix sum, W = alpha*exp(-tau) (point) or 2*alpha*E2(tau) (disc, via
scipy.special.expn as the reference uses), the integrated weight, the layer
brightness sum, its normalisation by the integrated weight and the T_cmb floor.
"""
import numpy as np
from scipy.special import expn

T_CMB = 2.725                      # utils.py:77
KM_TO_CM = (1.0E3, 1.0E-2)         # utils.Units['km'], utils.Units['cm']


def single(layers, T_layers, ds, layer4ds, disc=False, return_profiles=False):
    """Tb (list over frequency) for one ray.

    layers: alpha table (nf, nl) cm^-1; T_layers: (nl,) K;
    ds, layer4ds: the ray from compute_ds (km, layer indices); ds=None -> off planet.
    """
    layers = np.asarray(layers, dtype=np.float64)
    nf = layers.shape[0]
    if ds is None:                                   # brightness.py:46-51
        return [T_CMB] * nf
    tau = np.zeros(nf)
    W = np.zeros(nf)
    Tb_lyr = np.zeros(nf)
    intW = np.zeros(nf)
    prof_tau, prof_W = [tau.copy()], [W.copy()]
    with np.errstate(invalid='ignore', over='ignore'):
        for i in range(len(ds) - 1):
            d = ds[i] * KM_TO_CM[0] / KM_TO_CM[1]    # brightness.py:66
            ii = layer4ds[i]
            ii1 = layer4ds[i + 1]
            T1 = T_layers[ii1]
            T0 = T_layers[ii]
            a1 = layers[:, ii1]
            a0 = layers[:, ii]
            dtau = (a0 + a1) * d / 2.0
            tau1 = tau + dtau
            if disc:
                W1 = 2.0 * a1 * expn(2, tau1)
            else:
                W1 = a1 * np.exp(-tau1)
            intW = intW + (W1 + W) * d / 2.0
            Tb_lyr = Tb_lyr + (T1 * W1 + T0 * W) * d / 2.0
            tau, W = tau1, W1
            if return_profiles:
                prof_tau.append(tau.copy())
                prof_W.append(W.copy())
    Tb = []
    for j in range(nf):                              # brightness.py:107-117
        top = Tb_lyr[j]
        if top < T_CMB:
            top = T_CMB
        else:
            top = top / intW[j]
        Tb.append(top)
    if return_profiles:
        return Tb, dict(tau=np.array(prof_tau).T, W=np.array(prof_W).T, integrated_W=intW)
    return Tb
