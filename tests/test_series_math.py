"""The far-line series the FP32 absorption path adds for lines away from a tile
(radiobear_amd/csrc/far.hpp): for term(x) = (a x + c) / (x^2 - 2 Ds x + P0),

    below (|poles| <= x/3):  term = sum_k m_k x^-(k+1),  m_0 = a, m_1 = c + 2 Ds a,
                             m_k = 2 Ds m_{k-1} - P0 m_{k-2}
    above (|poles| >= 3 x):  term = sum_k n_k x^k,       n_0 = c / P0, n_1 = (a + 2 Ds n_0) / P0,
                             n_k = (2 Ds n_{k-1} - n_{k-2}) / P0

Pinned against the partial-fraction closed form (term = sum_i r_i / (x - p_i) with the poles
p_i from numpy.roots and residues r_i = (a p_i + c) / (p_i - p_j), so m_k = sum_i r_i p_i^k and
n_k = -sum_i r_i / p_i^(k+1)) and against direct evaluation at the ratio the kernels use, on
random lines of every canonical form: Ben-Reuven (G = 0, complex or real poles), Gross (a = 0)
and Van Vleck-Weisskopf (G > 0)."""
import numpy as np

import pytest

TERMS = 16        # the kernels' kFarTerms and far_ratio<FP32>() / far_ratio<FP64>()
RATIO = 1.0 / 3.0
RATIO_FP64 = 0.15


def below(a, c, Ds, P0):
    m = [a, c + 2 * Ds * a]
    for _ in range(2, TERMS):
        m.append(2 * Ds * m[-1] - P0 * m[-2])
    return np.array(m)


def above(a, c, Ds, P0):
    n = [c / P0]
    n.append((a + 2 * Ds * n[0]) / P0)
    for _ in range(2, TERMS):
        n.append((2 * Ds * n[-1] - n[-2]) / P0)
    return np.array(n)


def random_lines(rng, count):
    """(a, c, D, g, G) of random lines of the three forms, centres 0.5-1000 GHz, widths up to
    eight times the centre (so some Ben-Reuven lines have real poles)."""
    out = []
    for i in range(count):
        f0 = np.exp(rng.uniform(np.log(0.5), np.log(1000.0)))
        gam = f0 * np.exp(rng.uniform(np.log(1e-4), np.log(8.0)))
        W = rng.uniform(0.1, 2.0)
        kind = i % 3
        if kind == 0:     # Ben-Reuven with coupling zeta
            zet = gam * rng.uniform(0.0, 1.5)
            D = f0 ** 2 + gam ** 2 - zet ** 2
            out.append((W * (gam - zet), W * (gam + zet) * D, D, 4 * gam ** 2, 0.0))
        elif kind == 1:   # Gross
            out.append((0.0, W * gam, f0 ** 2, 4 * gam ** 2, 0.0))
        else:             # Van Vleck-Weisskopf pair over one denominator
            df = gam
            a = 2 * df * W
            out.append((a, a * (f0 ** 2 + df ** 2), f0 ** 2 - df ** 2, 0.0, 4 * f0 ** 2 * df ** 2))
    return out


def shifted(D, g, G):
    Ds = D - 0.5 * g
    h = g * (D - 0.25 * g) + G
    return Ds, h, Ds * Ds + h


def test_recurrences_match_partial_fractions():
    rng = np.random.default_rng(7)
    real_poles = 0
    for a, c, D, g, G in random_lines(rng, 300):
        Ds, h, P0 = shifted(D, g, G)
        real_poles += h < 0
        p = np.roots([1.0, -2.0 * Ds, P0]).astype(complex)
        if abs(p[0] - p[1]) < 1e-6 * abs(p[0]):
            continue   # (near-)double pole: residues ill-conditioned; the recurrence is not
        r = np.array([(a * p[0] + c) / (p[0] - p[1]), (a * p[1] + c) / (p[1] - p[0])])
        k = np.arange(TERMS)
        m_ref = np.real(np.sum(r[:, None] * p[:, None] ** k, axis=0))
        n_ref = np.real(-np.sum(r[:, None] / p[:, None] ** (k + 1), axis=0))
        pm = np.max(np.abs(p))
        pn = np.min(np.abs(p))
        # compare in the scaled bases in which the series are summed (x = |p|/RATIO and RATIO |p|)
        m = below(a, c, Ds, P0)
        n = above(a, c, Ds, P0)
        sm = (RATIO / pm) ** k
        sn = (RATIO * pn) ** k
        scale_m = np.max(np.abs(m_ref * sm)) + 1e-300
        scale_n = np.max(np.abs(n_ref * sn)) + 1e-300
        assert np.max(np.abs((m - m_ref) * sm)) < 1e-9 * scale_m
        assert np.max(np.abs((n - n_ref) * sn)) < 1e-9 * scale_n
    assert real_poles > 20   # both pole kinds exercised


@pytest.mark.parametrize('ratio', [RATIO, RATIO_FP64])
def test_series_sums_match_direct_evaluation(ratio):
    """Sixteen terms at the kernels' ratios, 1/3 (FP32 mode) and 0.15 (FP64 mode, where the bound
    below is 8.3e-12).  The tail is sum_i r_i p_i^16 / (x^16 (x - p_i));
    for two close poles (narrow lines, or real poles near a double root) the residues cancel and
    the combined coefficients grow like (k + 1) p^k, and a Gross line (a = 0, so m_0 = 0) keeps
    only 15 terms: the tail is then at most sum_{k>=15} (k + 1) (1/3)^k = 1.7e-6 of c / x^2, i.e.
    7.7e-7 of the line's own value at the ratio limit (4e-8 for two well-separated poles)."""
    rng = np.random.default_rng(11)
    bound = ratio ** (TERMS - 1) * (TERMS / (1 - ratio) + ratio / (1 - ratio) ** 2)   # 1.7e-6 at 1/3
    for a, c, D, g, G in random_lines(rng, 300):
        Ds, h, P0 = shifted(D, g, G)
        p = np.abs(np.roots([1.0, -2.0 * Ds, P0]))
        m = below(a, c, Ds, P0)
        n = above(a, c, Ds, P0)
        for x in (p.max() / ratio, 3.0 * p.max() / ratio):       # below: every |p| <= ratio x
            direct = (a * x + c) / (x * x - 2 * Ds * x + P0)
            series = np.sum(m * x ** -(np.arange(TERMS) + 1.0))
            size = (abs(a) * x + abs(c)) / abs(x * x - 2 * Ds * x + P0)   # the line's magnitude there
            assert abs(series - direct) <= bound * size + 1e-15 * abs(direct)
        if p.min() == 0.0:
            continue
        for x in (ratio * p.min(), ratio * p.min() / 5.0):        # above: every |p| >= x / ratio
            direct = (a * x + c) / (x * x - 2 * Ds * x + P0)
            series = np.sum(n * x ** np.arange(TERMS))
            size = (abs(a) * x + abs(c)) / abs(x * x - 2 * Ds * x + P0)
            assert abs(series - direct) <= bound * size + 1e-15 * abs(direct)
