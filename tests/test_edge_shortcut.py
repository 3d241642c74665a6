"""The engine's edge decision (edgetest.h: chi-square + AS 147 with exact
certainty intervals that skip the series) against the oracle's restatement of
kernel1.cu's decision, on random tables and on sweeps that cross every
threshold (significance for each dof, with and without prior, with and without
Bonferroni, and AS 147's underflow cut)."""
import ctypes as C

import numpy as np
import pytest

import eddy_amd
import oracle_ref as O


def engine_decision(t, c, prior, prm):
    t = np.ascontiguousarray(np.asarray(t, np.int32).reshape(9))
    return eddy_amd.lib().eddy_edge_decision_host(t.ctypes.data_as(C.c_void_p), c, prior, C.byref(prm))


def params(p_adjust=1, alpha=0.05, pw=0.5):
    aep = 1 - (1 - alpha) * pw ** 0.5
    return (eddy_amd.Params(3, 100, p_adjust, alpha, aep, 0.05, 0, 1, 0),
            O.params(p_adjust=p_adjust, alpha_edge=alpha, prior_weight=pw))


@pytest.mark.parametrize("c", [1, 45, 780, 11175, 44850, 124750])
def test_random_tables(c):
    rng = np.random.default_rng(c)
    for p_adjust in (1, 0):
        ep, op = params(p_adjust)
        for _ in range(4000):
            n = int(rng.integers(10, 1300))
            kind = rng.integers(0, 3)
            if kind == 0:
                w = rng.dirichlet(np.ones(9))
            elif kind == 1:  # strong dependence
                w = np.eye(3).reshape(9) * rng.uniform(1, 20) + rng.dirichlet(np.ones(9))
                w /= w.sum()
            else:  # a missing state
                w = rng.dirichlet(np.ones(9))
                w.reshape(3, 3)[rng.integers(0, 3), :] = 0
                w /= w.sum()
            t = rng.multinomial(n, w)
            pr = int(rng.integers(0, 2))
            assert engine_decision(t, c, pr, ep) == O.edge_decision(t, c, pr, op), (t, c, pr, p_adjust)


@pytest.mark.parametrize("c", [1, 780, 44850, 124750])
def test_threshold_sweeps(c):
    """Tables whose chi-square grows in small steps through the significance
    threshold of each dof and through the underflow cut."""
    ep, op = params(1)
    ep0, op0 = params(0)
    n_checked = 0
    for n in (60, 300, 1200):
        for shape in ("dof1", "dof2", "dof4"):
            for shift in range(0, n // 3):
                k = n // 3
                if shape == "dof1":
                    t = np.array([[k + shift, 0, k - shift], [0, 0, 0], [k - shift, 0, k + shift]])
                elif shape == "dof2":
                    t = np.array([[k // 2 + shift, k // 2, k // 2 - shift], [0, 0, 0],
                                  [k // 2 - shift, k // 2, k // 2 + shift]])
                else:
                    a = k // 3
                    t = np.array([[a + shift, a, a - shift], [a, a + shift, a - shift], [a - shift, a, a + shift]])
                if t.min() < 0:
                    break
                for pr in (0, 1):
                    assert engine_decision(t, c, pr, ep) == O.edge_decision(t, c, pr, op), (t, c, pr)
                    assert engine_decision(t, c, pr, ep0) == O.edge_decision(t, c, pr, op0), (t, c, pr)
                    n_checked += 1
    assert n_checked > 1000


def test_unusual_thresholds_fall_back_to_exact():
    for alpha, pw in ((1.5, 0.5), (0.05, 0.0), (1e-320, 0.5)):
        ep, op = params(1, alpha, pw)
        rng = np.random.default_rng(7)
        for _ in range(500):
            t = rng.multinomial(int(rng.integers(10, 400)), rng.dirichlet(np.ones(9)))
            pr = int(rng.integers(0, 2))
            assert engine_decision(t, 100, pr, ep) == O.edge_decision(t, 100, pr, op)
