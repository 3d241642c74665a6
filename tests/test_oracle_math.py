"""Pins of the FP64 primitives (fpmath.h, shared by the kernels and the oracle)
against the C library, numpy and scipy; and of AS 147 (kernel1.cu:10-74)
against the closed forms of the regularised lower incomplete gamma."""
import math
import os

import numpy as np
import pytest
import scipy.special as sp

import oracle_ref as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def ulps(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.abs(a - b) / np.spacing(np.abs(b))


def test_exp_log_within_2ulp_of_libm():
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.uniform(-745, 709, 20000), rng.uniform(-1, 1, 5000), [0.0, 1.0, -1.0, 88.0]])
    e = np.array([O.lib().eo_exp(float(x)) for x in xs])
    assert ulps(e, np.exp(xs)).max() <= 2
    ys = np.exp(rng.uniform(-700, 700, 20000))
    l = np.array([O.lib().eo_log(float(y)) for y in ys])
    assert ulps(l, np.log(ys)).max() <= 2
    assert O.lib().eo_log(1.0) == 0.0
    assert O.lib().eo_log(2.0) == math.log(2.0)


def test_exp_log_special_values():
    L = O.lib()
    assert L.eo_exp(float("inf")) == float("inf")
    assert L.eo_exp(-float("inf")) == 0.0
    assert math.isnan(L.eo_exp(float("nan")))
    assert L.eo_exp(710.0) == float("inf")
    assert L.eo_log(0.0) == -float("inf")
    assert math.isnan(L.eo_log(-1.0))
    assert L.eo_log(5e-324) == math.log(5e-324)


def test_lgamma_pos_against_math():
    rng = np.random.default_rng(2)
    for x in np.concatenate([rng.uniform(1e-3, 10, 500), rng.uniform(10, 1e5, 500)]):
        assert abs(O.lib().eo_lgamma_pos(float(x)) - math.lgamma(x)) <= 1e-14 * max(1.0, abs(math.lgamma(x)))


@pytest.mark.parametrize("p,closed", [
    (1.0, lambda x: -math.expm1(-x)),
    (2.0, lambda x: 1.0 - math.exp(-x) * (1.0 + x)),
    (0.5, lambda x: math.erf(math.sqrt(x))),
])
def test_gammds_closed_forms(p, closed):
    # AS 147 stops when a term is <= 1e-9 of the sum, so agreement is to ~1e-9
    for x in np.concatenate([np.linspace(0.01, 5, 50), np.linspace(5, 80, 50)]):
        assert abs(O.gammds(float(x), p) - closed(x)) <= 2e-9 * closed(x) + 1e-300


def test_gammds_scipy_golden():
    g = np.load(os.path.join(GOLD, "gammainc.npz"))
    for p in (0.5, 1.0, 2.0):
        ref = g["p%g" % p]
        got = np.array([O.gammds(float(x), p) for x in g["x"]])
        live = np.array([p * math.log(x) - math.lgamma(p + 1) - x >= math.log(1e-37) for x in g["x"]])
        assert live.sum() > 80
        assert np.allclose(got[live], ref[live], rtol=2e-9, atol=0)
        assert np.all(got[~live] == 0.0)  # underflow cut (kernel1.cu:36-41)


def test_gammds_guards_and_underflow_cut():
    assert O.gammds(0.0, 1.0) == 0.0
    assert O.gammds(-1.0, 1.0) == 0.0
    assert O.gammds(3.0, 0.0) == 0.0
    # p ln x - lgamma(p+1) - x < ln 1e-37 returns 0: the p-value is then 1
    for p in (0.5, 1.0, 2.0):
        lo = [x for x in np.linspace(60, 120, 6001) if O.gammds(float(x), p) == 0.0]
        assert lo, "cut must trigger below x = 120"
        xc = lo[0]
        arg = p * math.log(xc) - math.lgamma(p + 1) - xc
        assert arg < math.log(1e-37) + 1e-9
        assert O.gammds(float(xc) - 0.02, p) > 0.99


def test_ibetac_scipy_golden_and_closed_forms():
    g = np.load(os.path.join(GOLD, "betaincc.npz"))
    for a, b, x, q in zip(g["a"], g["b"], g["x"], g["q"]):
        v, de = O.ibetac(float(a), float(b), float(x))
        assert de == 0
        assert abs(v - q) <= 1e-12 * max(q, 1e-300) + 1e-300 or abs(v - q) <= 1e-14
    for a in (0.3, 1.0, 2.5, 17.0):
        for x in (0.05, 0.3, 0.7):
            assert abs(O.ibetac(a, 1.0, x)[0] - (1 - x ** a)) < 1e-13
            assert abs(O.ibetac(1.0, a, x)[0] - (1 - x) ** a) < 1e-13
            # symmetry I_x(a,b) = 1 - I_{1-x}(b,a)
            assert abs(O.ibetac(a, 3.0, x)[0] - (1 - O.ibetac(3.0, a, 1 - x)[0])) < 1e-13


def test_ibetac_large_shapes_against_scipy():
    """Shape parameters up to 5000 (a beta fit to tightly spread JS values),
    one small and one large, results down to 1e-290: within 1e-12 relative
    of scipy's betaincc, computed live."""
    from scipy.special import betaincc
    rng = np.random.default_rng(11)
    n = 0
    for _ in range(4000):
        a, b = np.exp(rng.uniform(np.log(0.3), np.log(5000), 2))
        x = rng.uniform(0.0, 1.0)
        q = float(betaincc(a, b, x))
        if not q > 1e-290:
            continue
        v, de = O.ibetac(float(a), float(b), float(x))
        assert de == 0
        assert abs(v - q) <= 1e-12 * q, (a, b, x, q, v)
        n += 1
    assert n > 3000


def test_ibetac_domain():
    assert O.ibetac(-1.0, 2.0, 0.5)[1] == 1
    assert O.ibetac(0.0, 0.0, 0.5)[1] == 1
    assert O.ibetac(float("inf"), 2.0, 0.5)[1] == 1
    assert O.ibetac(2.0, 2.0, 1.5)[1] == 1
    assert O.ibetac(2.0, 3.0, 0.0) == (1.0, 0)
    assert O.ibetac(2.0, 3.0, 1.0) == (0.0, 0)
