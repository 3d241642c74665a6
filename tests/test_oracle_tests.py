"""Pins of the chi-square test (kernel1.cu:121-203, 420-431): hand-worked
tables, Yates' closed form for 2x2, scipy's Pearson statistic when no Yates
correction applies, the dof rule and the edge thresholds."""
import math
from fractions import Fraction

import numpy as np
import pytest
from scipy.stats import chi2_contingency

import oracle_ref as O


def reduced(t):
    t = np.asarray(t).reshape(3, 3)
    t = t[t.sum(1) > 0][:, t.sum(0) > 0]
    return t


def test_no_yates_matches_scipy_pearson():
    rng = np.random.default_rng(3)
    n = 0
    while n < 200:
        t = rng.integers(6, 60, 9)
        stat, p, d, exp = chi2_contingency(t.reshape(3, 3), correction=False)
        if exp.min() < 5:
            continue
        v, dof = O.chi2(t)
        assert dof == d == 4
        assert abs(v - stat) <= 1e-12 * stat
        n += 1


def test_yates_2x2_closed_form():
    # only two states per gene: Yates applies to every cell when some E < 5,
    # chi2 = N (|ad - bc| - N/2)^2 / (r1 r2 c1 c2)
    for a, b, c, d in [(3, 9, 12, 2), (1, 4, 7, 9), (5, 2, 2, 6), (0, 7, 8, 3)]:
        t = np.zeros((3, 3), int)
        t[0, 0], t[0, 2], t[2, 0], t[2, 2] = a, b, c, d
        N = a + b + c + d
        r1, r2, c1, c2 = a + b, c + d, a + c, b + d
        exp_min = min(r1 * c1, r1 * c2, r2 * c1, r2 * c2) / N
        v, dof = O.chi2(t)
        assert dof == 1
        if exp_min < 5:
            closed = N * (abs(a * d - b * c) - N / 2) ** 2 / (r1 * r2 * c1 * c2)
        else:
            closed = N * (a * d - b * c) ** 2 / (r1 * r2 * c1 * c2)
        assert abs(v - closed) <= 1e-12 * max(closed, 1)


def test_yates_3x3_hand_worked():
    # exact rational evaluation of sum (|O-E| - 1/2)^2 / E for a sparse table
    t = np.array([[1, 4, 0], [3, 2, 6], [0, 5, 9]])
    rows, cols, N = t.sum(1), t.sum(0), t.sum()
    tot = Fraction(0)
    for i in range(3):
        for j in range(3):
            E = Fraction(int(rows[i]) * int(cols[j]), int(N))
            dv = abs(Fraction(int(t[i, j])) - E) - Fraction(1, 2)
            tot += dv * dv / E
    v, dof = O.chi2(t)
    assert dof == 4
    assert abs(v - float(tot)) <= 1e-12 * float(tot)


def test_dof_rule_and_empty_table():
    assert O.chi2(np.array([[5, 0, 7], [0, 0, 0], [6, 0, 9]]))[1] == 1
    assert O.chi2(np.array([[5, 3, 7], [0, 0, 0], [6, 4, 9]]))[1] == 2
    assert O.chi2(np.array([[5, 3, 7], [0, 0, 0], [0, 0, 0]]))[1] == 0
    v, dof = O.chi2(np.zeros(9, int))
    assert v == 0.0 and dof == 1


def test_transpose_invariance():
    rng = np.random.default_rng(4)
    for _ in range(100):
        t = rng.integers(0, 30, (3, 3))
        a, da = O.chi2(t)
        b, db = O.chi2(t.T.copy())
        assert da == db and abs(a - b) <= 1e-12 * max(a, 1)


def test_edge_thresholds():
    prm = O.params()
    # strong but not extreme dependence: edge; independent table: no edge
    dep = np.array([[30, 5, 5], [5, 30, 5], [5, 5, 30]])
    ind = np.array([[12, 13, 12], [13, 12, 13], [12, 13, 12]])
    assert O.edge_decision(dep, 100, 0, prm) == 1
    assert O.edge_decision(ind, 100, 0, prm) == 0
    # extreme dependence underflows AS 147 -> p = 1 -> no edge
    ext = np.array([[200, 0, 0], [0, 200, 0], [0, 0, 200]])
    assert O.chi2(ext)[0] > 350
    assert O.edge_decision(ext, 100, 0, prm) == 0
    # Bonferroni: p*c capped at 1; -rawp turns it off
    t = np.array([[16, 8, 8], [8, 16, 8], [8, 8, 16]])
    chi, dof = O.chi2(t)
    p = 1 - O.gammds(chi / 2, dof / 2)
    assert 0.001 < p < 0.05
    assert O.edge_decision(t, 1, 0, prm) == 1
    assert O.edge_decision(t, 10000, 0, prm) == 0
    raw = O.params(p_adjust=0)
    assert O.edge_decision(t, 10000, 0, raw) == 1
    # prior lowers the bar to 1 - (1 - aE) pw^(1/lambda) = 0.328...
    c = int(0.2 / p)
    assert O.edge_decision(t, c, 0, prm) == 0
    assert O.edge_decision(t, c, 1, prm) == 1
    assert abs(prm.alphaEdgePrior - (1 - 0.95 * math.sqrt(0.5))) < 1e-15
