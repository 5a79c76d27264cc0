"""FP64 restatement of the file-atmosphere pipeline (TEST INFRASTRUCTURE ONLY).

Used only to rebuild the reference benchmark atmosphere (scripts/benchmark.py,
Jupiter/config_benchmark.par) so the oracle can be checked against the
reference's golden T_rb vector.  Follows readGas (atm_base.py:39-78),
regrid (regrid.py:11-83, linear scipy interp1d in P), the JupiterNoTweak
tweak (Jupiter/JupiterNoTweak.py:1-25) and computeProp (atm_base.py:120-194):
R, N, P, Z, which the ray path consumes, and the rest of the property table (GM,
AMU, REFR, H, LAPSE, LAPSEP, g), pinned to the reference-written table of the
benchmark atmosphere (tests/golden/props.npz).
"""
import numpy as np
from scipy.interpolate import interp1d

# config.C of the Jupiter config (default_config.json 'constituents')
C_JUPITER = {'Z': 0, 'T': 1, 'P': 2, 'H2': 3, 'HE': 4, 'CH4': 5, 'NH3': 6, 'H2O': 7,
             'H2S': 8, 'SOLN': 9, 'OTHER': 10, 'PH3': 11, 'CO': 12, 'CO13': 13,
             'HCN': 14, 'DZ': 15}
COMMENT_CHARS = ['!', '#', '$', '%', '&', '*']

# chemistry.py:18-46 (amu g/mol, refractivity coefficient; H2O is T dependent)
AMU_OF = {'H2': 2.016, 'HE': 4.003, 'H2S': 34.076, 'NH3': 17.030, 'H2O': 18.015,
          'CH4': 16.04, 'PH3': 33.997, 'NH4SH': 51.110}
REFR = {'H2': 124.43, 'HE': 35.832, 'H2S': 2247.0, 'NH3': 2700.0, 'CH4': 413.0}


def refractivity(key, T):
    if key == 'H2O':
        return 245.0 + 1.28E6 / T        # chemistry.py:33-41
    return REFR.get(key, 0.0)


def _data_from(line):
    """utils.get_data_from (utils.py:152-160)."""
    if line[0] in COMMENT_CHARS or len(line) < 4:
        return None
    try:
        return [float(x) for x in line.split()]
    except ValueError:
        return None


def read_gas(text, C=C_JUPITER):
    """atm_base.readGas (atm_base.py:39-78) from the file's text."""
    lines = text.splitlines(keepends=True)
    counts = {}
    for ln in lines:
        cv = _data_from(ln)
        if cv is not None:
            counts[len(cv)] = counts.get(len(cv), 0) + 1
    inv = {v: k for k, v in counts.items()}
    sk = sorted(inv, reverse=True)
    if len(sk) > 1 and 2 * sum(sk[1:]) > sk[0]:
        raise ValueError("Not enough data lines in file: {}".format(sk))
    expected = inv[sk[0]]
    nC = len(C)
    gas = [[] for _ in range(nC)]
    for ln in lines:
        cv = _data_from(ln)
        if cv is None or len(cv) != expected:
            continue
        for n in range(nC):
            gas[n].append(cv[n] if n < len(cv) else 0.0)
    gas = np.array(gas)
    if not np.all(np.diff(gas[C['P']]) > 0.0):
        gas = np.fliplr(gas)
        if not np.all(np.diff(gas[C['P']]) > 0.0):
            raise ValueError("Pressure not monotonically increasing")
    renorm_z(gas, C)
    return gas


def renorm_z(gas, C):
    """atm_base._renorm_z (atm_base.py:206-219)."""
    gas[C['Z']] = gas[C['Z']] - gas[C['Z']][-1]
    dz = np.abs(np.diff(gas[C['Z']])) * 1.0E5
    gas[C['DZ']] = np.append(np.array([0.0]), dz)


def regrid_logp(gas, C, pmin, pmax, nsteps):
    """regrid.regrid with an integer regridType (regrid.py:11-83), interior only."""
    Pgrid = np.logspace(np.log10(pmin), np.log10(pmax), nsteps)
    out = np.zeros((gas.shape[0], len(Pgrid)))
    out[C['P']] = Pgrid
    fill = -999.9
    for key, idx in C.items():
        if key in ('P', 'DZ'):
            continue
        fv = interp1d(gas[C['P']], gas[idx], kind='linear', fill_value=fill, bounds_error=False)
        out[idx] = fv(Pgrid)
    if fill in out:
        raise NotImplementedError("regrid extrapolation is outside the oracle's scope")
    renorm_z(out, C)
    return out


def jupiter_no_tweak(gas, C):
    """Jupiter/JupiterNoTweak.py:1-25: CO, CO13, HCN, SOLN, PH3 set to zero."""
    gas = gas.copy()
    for k in ('CO', 'HCN', 'SOLN', 'PH3'):
        gas[C[k]] = 0.0
    gas[C['CO13']] = 1.0E-2 * gas[C['CO']]
    return gas


GRAV_CONST = 6.6738e-20   # km^3 / kg / s^2 (chemistry.py:7)
R_GAS = 8.314462          # J / K / mol (chemistry.py:8)
# the property table's rows (default_config.json "LP")
PROP_ROWS = ('Z', 'R', 'P', 'GM', 'AMU', 'REFR', 'N', 'H', 'LAPSE', 'LAPSEP', 'g')


def compute_prop(gas, C, Req, p_ref=1.0, GM_ref=None):
    """atm_base.computeProp (atm_base.py:120-194) -> dict of arrays: R, N, P, Z, and with
    GM_ref (km^3/s^2) also GM, AMU, REFR, H, LAPSE, LAPSEP and g."""
    P = gas[C['P']]
    T = gas[C['T']]
    Z = gas[C['Z']]
    psep = 1.0E6
    iOffset = 0
    for i in range(len(Z)):                          # atm_base.py:135-140
        if abs(P[i] - p_ref) < psep:
            psep = abs(P[i] - p_ref)
            iOffset = i
    R = Req + Z - Z[iOffset]                         # atm_base.py:148-150
    keys = [k for k in C if k not in ('P', 'T', 'Z', 'DZ')]
    n = len(P)
    N, REFR, AMU = np.zeros(n), np.zeros(n), np.zeros(n)
    GM, LAPSE, LAPSEP = np.zeros(n), np.zeros(n), np.zeros(n)
    for i in range(n):
        AMU[i] = sum(AMU_OF.get(k, 0.0) * gas[C[k]][i] for k in keys)          # atm_base.py:152-155
        if i:                                                                  # atm_base.py:161-173
            rho = (AMU[i] * P[i]) / (R_GAS * T[i])
            dr = abs(Z[i] - Z[i - 1])
            dM = 1.0e11 * rho * 4.0 * np.pi * R[i] ** 2 * dr
            GM[i] = GM[i - 1] + GRAV_CONST * dM
            LAPSE[i] = abs(T[i] - T[i - 1]) / dr
            LAPSEP[i] = abs(T[i] - T[i - 1]) / abs(P[i] - P[i - 1])
        refr = 0.0
        for k in keys:                                                         # atm_base.py:175-181
            refr += refractivity(k, T[i]) * gas[C[k]][i]
        REFR[i] = refr * P[i] * (293.0 / T[i])
        N[i] = REFR[i] / 1.0E6 + 1.0
    out = {'P': P.copy(), 'Z': Z.copy(), 'R': R, 'N': N}
    if GM_ref is not None:                                                     # atm_base.py:183-193
        gm = GM_ref - (GM - GM[iOffset])
        g = gm / R ** 2
        out.update({'GM': gm, 'AMU': AMU, 'REFR': REFR, 'LAPSE': LAPSE, 'LAPSEP': LAPSEP, 'g': g,
                    'H': (R_GAS * T) / (g * AMU) / 1000.0})
    return out
