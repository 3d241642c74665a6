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


def engine_chi2(t, fast):
    t = np.ascontiguousarray(np.asarray(t, np.int32).reshape(9))
    dof = C.c_int()
    x = C.c_double()
    rc = eddy_amd.lib().eddy_chi2_host(t.ctypes.data_as(C.c_void_p), fast, C.byref(dof), C.byref(x))
    assert rc == 0
    return x.value, dof.value


def _random_tables(rng, count, nmax=2046):
    for _ in range(count):
        n = int(rng.integers(2, nmax + 1))
        kind = rng.integers(0, 4)
        if kind == 0:
            w = rng.dirichlet(np.ones(9))
        elif kind == 1:
            w = np.eye(3).reshape(9) * rng.uniform(1, 20) + rng.dirichlet(np.ones(9))
        elif kind == 2:
            w = rng.dirichlet(np.ones(9))
            w.reshape(3, 3)[rng.integers(0, 3), :] = 0
        else:  # very skewed margins: Yates' rule in play
            w = rng.dirichlet(np.full(9, 0.2))
        yield rng.multinomial(n, w / w.sum())


def test_fast_chi2_error_bound():
    """The reciprocal form of chi-square stays within 1e-10 (1 + x) of the
    reference formula (x = chi2 / 2), the guard the fast decisions assume
    with a factor 10 to spare; dof is identical."""
    rng = np.random.default_rng(17)
    worst = 0.0
    for t in _random_tables(rng, 20000):
        for k in [-1] + [int(i) for i in np.nonzero(t)[0]][:3]:
            u = t.copy()
            if k >= 0:
                u[k] -= 1
            if u.sum() == 0:
                continue
            xe, de = engine_chi2(u, 0)
            xf, df = engine_chi2(u, 1)
            assert de == df
            worst = max(worst, abs(xf - xe) / 2 / (1 + xe / 2))
    assert worst < 1e-10, worst


def test_loo_chi2_error_bound():
    """The leave-one-out form (the identity n (sum O^2 / (R K) - 1) with the
    column partial sums shared by a column's three cells, or the per-cell
    form under Yates' rule) stays within 1e-10 (1 + x) of the reference
    formula on the reduced tally; dof is identical."""
    rng = np.random.default_rng(23)
    worst = 0.0
    checked = 0
    for t in _random_tables(rng, 12000):
        if t.sum() < 2:
            continue
        for k in [int(i) for i in np.nonzero(t)[0]]:
            u = t.copy()
            u[k] -= 1
            xe, de = engine_chi2(u, 0)
            xl, dl = engine_chi2(t, 2 + k)
            assert de == dl, (t, k)
            worst = max(worst, abs(xl - xe) / 2 / (1 + xe / 2))
            checked += 1
    assert checked > 50000
    assert worst < 1e-10, worst


def test_loo_form_edge_cases():
    # the identity drops the empty rows and columns; a reduced tally whose
    # row empties out (R_a = 1) must match the reference with that row gone
    t = np.array([1, 0, 0, 30, 40, 35, 20, 33, 41])
    u = t.copy()
    u[0] -= 1
    xe, de = engine_chi2(u, 0)
    xl, dl = engine_chi2(t, 2)
    assert de == dl == 2 and abs(xe - xl) < 1e-10
    # a column with a small count: 1/(K_b - 1) rather than 1/K_b for the
    # reduced column matters far beyond the guard here
    v = np.array([30, 2, 25, 28, 1, 33, 35, 3, 27])
    w = v.copy()
    w[4] -= 1
    xe, _ = engine_chi2(w, 0)
    xl, _ = engine_chi2(v, 2 + 4)
    assert abs(xe - xl) < 1e-10 * (1 + xe)


def test_planted_yates_mistake_is_caught():
    # a table whose Yates decision sits at E = R K / n = 5 exactly: a float
    # comparison off by one rounding would flip Yates and change chi2 by O(1)
    t = np.array([5, 5, 5, 5, 5, 5, 5, 5, 5])  # E = 15 * 15 / 45 = 5
    xe, _ = engine_chi2(t, 0)
    xf, _ = engine_chi2(t, 1)
    assert xe == 0.0 and xf == 0.0
    t = np.array([6, 4, 5, 4, 6, 5, 5, 5, 5])
    xe, _ = engine_chi2(t, 0)
    xf, _ = engine_chi2(t, 1)
    assert abs(xe - xf) < 1e-12
    # Yates applies only below 5: chi2 without the 0.5 correction is larger
    d = np.array([[6, 4, 5], [4, 6, 5], [5, 5, 5]], float)
    E = 5.0
    assert abs(xe - ((np.abs(d - E)) ** 2 / E).sum()) < 1e-12


@pytest.mark.parametrize("c", [45, 11175, 124750])
def test_loo_mask_matches_oracle(c):
    """Every bit of the leave-one-out mask equals the oracle's decision of
    that leave-one-out tally (empty cells copy the full decision)."""
    rng = np.random.default_rng(c + 1)
    ep, op = params(1)
    L = eddy_amd.lib()
    checked = 0
    tables = list(_random_tables(rng, 1500, 700))
    # tallies just around the significance threshold, where the guard bands matter
    for n in (30, 120, 600):
        k = n // 3
        a = k // 3
        for shift in range(0, a):
            tables.append(np.array([a + shift, a, a - shift, a, a + shift, a - shift, a - shift, a, a + shift]))
    for t in tables:
        pr = int(rng.integers(0, 2))
        full = O.edge_decision(t, c, pr, op)
        t32 = np.ascontiguousarray(t, np.int32)
        m = L.eddy_edge_loo_mask_host(t32.ctypes.data_as(C.c_void_p), c, pr, full, C.byref(ep))
        for k in range(9):
            if t[k] > 0:
                u = t.copy()
                u[k] -= 1
                want = O.edge_decision(u, c, pr, op)
            else:
                want = full
            assert (m >> k) & 1 == want, (t, k, c, pr)
            checked += 1
    assert checked > 10000


def _dependent_tables(rng, count):
    """Tertile-like tables of 150..1200 samples, from independent to strongly
    dependent, so the statistic sweeps through both thresholds."""
    for _ in range(count):
        n = int(rng.integers(150, 1201))
        pr = rng.dirichlet([8, 8, 8])
        pc = rng.dirichlet([8, 8, 8])
        lam = float(rng.uniform(0.0, 0.35))
        perm = rng.permutation(3)
        dep = np.zeros((3, 3))
        for i in range(3):
            dep[i, perm[i]] = 1.0 / 3.0
        prob = (1 - lam) * np.outer(pr, pc) + lam * dep
        yield rng.multinomial(n, prob.ravel() / prob.sum()).astype(np.int64)


@pytest.mark.parametrize("c", [45, 11175, 124750])
def test_loo_all_certain_matches_oracle(c):
    """Where the all-at-once check answers, every leave-one-out tally of a
    nonzero cell gets that decision from the oracle (the reference formula);
    it answers for most tables and both ways."""
    rng = np.random.default_rng(7 * c + 3)
    ep, op = params(1)
    L = eddy_amd.lib()
    seen = {-1: 0, 0: 0, 1: 0}
    for t in _dependent_tables(rng, 2500):
        pr = int(rng.integers(0, 2))
        t32 = np.ascontiguousarray(t, np.int32)
        a = L.eddy_loo_all_certain_host(t32.ctypes.data_as(C.c_void_p), c, pr, C.byref(ep))
        seen[a] += 1
        if a < 0:
            continue
        for k in range(9):
            if t[k] > 0:
                u = t.copy()
                u[k] -= 1
                assert O.edge_decision(u, c, pr, op) == a, (t, k, c, pr, a)
    assert seen[0] > 150 and seen[1] > 150 and seen[0] + seen[1] > 1250, seen


def _boundary_tables(rng, walks, c, op):
    """Tables around a decision flip: from an independent table, move one
    sample at a time onto a fixed diagonal until the full decision turns to an
    edge, keeping the last 60 tables before the flip and the 20 after it."""
    out = []
    for _ in range(walks):
        n = int(rng.integers(150, 1201))
        pr = int(rng.integers(0, 2))
        p_r = rng.dirichlet([8, 8, 8])
        p_c = rng.dirichlet([8, 8, 8])
        t = rng.multinomial(n, np.outer(p_r, p_c).ravel()).astype(np.int64)
        diag = rng.permutation(3)
        hist, after = [], -1
        for _ in range(4 * n):
            hist.append(t.copy())
            if after < 0 and O.edge_decision(t, c, pr, op) == 1:
                after = 0
            if after >= 0:
                after += 1
                if after > 20:
                    break
            off = [k for k in range(9) if t[k] > 1 and k % 3 != diag[k // 3]]
            if not off:
                break
            k = off[int(rng.integers(0, len(off)))]
            t[k] -= 1
            t[(k // 3) * 3 + diag[k // 3]] += 1
        if after > 0:
            out += [(h, pr) for h in hist[-80:]]
    return out


@pytest.mark.parametrize("c", [45, 11175])
def test_loo_all_certain_at_decision_flips(c):
    """Near a flip some leave-one-out tallies decide differently from the full
    one; wherever the all-at-once check answers there, the oracle agrees on
    every leave-one-out tally."""
    rng = np.random.default_rng(11 * c + 5)
    ep, op = params(1)
    L = eddy_amd.lib()
    answered = 0
    for t, pr in _boundary_tables(rng, 120, c, op):
        t32 = np.ascontiguousarray(t, np.int32)
        a = L.eddy_loo_all_certain_host(t32.ctypes.data_as(C.c_void_p), c, pr, C.byref(ep))
        if a < 0:
            continue
        answered += 1
        for k in range(9):
            if t[k] > 0:
                u = t.copy()
                u[k] -= 1
                assert O.edge_decision(u, c, pr, op) == a, (t, k, c, pr, a)
    assert answered > 1000


def _screen(t, c, pr, ep):
    t32 = np.ascontiguousarray(np.asarray(t, np.int32).reshape(9))
    return eddy_amd.lib().eddy_screen32_host(t32.ctypes.data_as(C.c_void_p), c, pr, C.byref(ep))


def _exact_half_chi2(t):
    """x = chi2 / 2 without Yates' correction, in exact rationals:
    n (sum O^2 / (R K) - 1) / 2 over the nonempty rows and columns."""
    from fractions import Fraction

    m = np.asarray(t, np.int64).reshape(3, 3)
    R, K, n = m.sum(1), m.sum(0), int(m.sum())
    S = Fraction(0)
    for i in range(3):
        for j in range(3):
            if R[i] and K[j]:
                S += Fraction(int(m[i, j]) ** 2, int(R[i]) * int(K[j]))
    return Fraction(n, 2) * (S - 1)


def test_screen32_error_bound():
    """The single-precision x of the screens stays within the bound their
    guard doubles (1.4e-6 n + 6e-8 x) of the exact value, for tables of up to
    20000 samples; dof is the reference formula's."""
    rng = np.random.default_rng(23)
    worst = 0.0
    for t in _random_tables(rng, 3000, 20000):
        if t.sum() == 0:
            continue
        x32, d32 = engine_chi2(t, 11)
        _, de = engine_chi2(t, 0)
        assert d32 == de
        xe = _exact_half_chi2(t)
        err = abs(x32 / 2 - float(xe))
        bound = 1.4e-6 * t.sum() + 6e-8 * abs(float(xe))
        assert err <= bound, (t, x32 / 2, float(xe))
        worst = max(worst, err / bound)
    assert worst > 0.0


@pytest.mark.parametrize("c", [1, 780, 44850, 124750])
def test_screen32_matches_oracle(c):
    """Wherever the single-precision screen of the full test answers, the
    oracle's decision agrees; it answers for most tables, both ways."""
    rng = np.random.default_rng(5 * c + 1)
    seen = {-1: 0, 0: 0, 1: 0}
    for p_adjust in (1, 0):
        ep, op = params(p_adjust)
        tables = list(_dependent_tables(rng, 1500)) + list(_random_tables(rng, 1500, 20000))
        for t in tables:
            pr = int(rng.integers(0, 2))
            s = _screen(t, c, pr, ep)
            seen[s] += 1
            if s >= 0:
                assert s == O.edge_decision(t, c, pr, op), (t, c, pr, p_adjust)
    assert seen[0] > 1000 and seen[1] > 200 and seen[-1] < seen[0] + seen[1], seen


@pytest.mark.parametrize("c", [45, 11175])
def test_screen32_at_decision_flips(c):
    """Next to a decision flip the screen answers only where the oracle
    agrees, and leaves the tables closest to the flip to the double-precision
    path."""
    rng = np.random.default_rng(13 * c + 7)
    ep, op = params(1)
    answered = left = 0
    for t, pr in _boundary_tables(rng, 80, c, op):
        s = _screen(t, c, pr, ep)
        if s >= 0:
            answered += 1
            assert s == O.edge_decision(t, c, pr, op), (t, c, pr)
        else:
            left += 1
    assert answered > 1000 and left > 0, (answered, left)
