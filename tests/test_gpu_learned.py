"""The learned proposal on the GPU: network kernels against the FP64 oracle,
the logged training data against the CPU oracle, and the learned job against
the method without it (identical results, fewer tests)."""
import ctypes as C
import os
import sys

import numpy as np
import pytest

import eddy_amd
import oracle_ref

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as M  # noqa: E402

pytestmark = pytest.mark.gpu


def _data(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(0, 1, (n, M.NIN)).astype(np.float32)
    Y = rng.uniform(0, 1, n).astype(np.float32)
    return X, Y


@pytest.mark.parametrize("n", [1, 64, 1000])
def test_forward_and_gradient_match_oracle(engine, n):
    print("case: mlp gradient n", n, flush=True)
    engine.mlp_init(7)
    P = engine.mlp_params()
    X, Y = _data(n, 11 + n)
    prob = engine.mlp_forward(X)
    assert np.allclose(prob, M.forward(P, X), rtol=0, atol=2e-6)
    g, L = engine.mlp_gradient(X, Y)
    gr, Lr = M.gradient(P, X, Y)
    assert abs(L - Lr) <= 2e-6 * max(1.0, abs(Lr))
    scale = np.abs(gr).max()
    assert np.abs(g - gr).max() <= 2e-5 * scale, np.abs(g - gr).max() / scale
    assert np.linalg.norm(g - gr) <= 1e-5 * np.linalg.norm(gr)


def test_adam_step_matches_oracle(engine):
    X, Y = _data(512, 3)
    engine.mlp_init(9)
    P0 = engine.mlp_params().astype(np.float64)
    lr = 1e-2
    engine.mlp_train(X, Y, epochs=1, batch=512, lr=lr, seed=5)  # one step over the whole batch
    P1 = engine.mlp_params()
    g, _ = M.gradient(P0, X, Y)
    Pr, _, _ = M.adam_step(P0, g, np.zeros_like(P0), np.zeros_like(P0), 1, lr)
    assert np.abs(P1 - Pr).max() <= 2e-3 * lr


def test_training_is_deterministic_and_learns(engine):
    rng = np.random.default_rng(21)
    X = rng.normal(0, 1, (20000, M.NIN)).astype(np.float32)
    Y = (1 / (1 + np.exp(-(2 * X[:, 0] - X[:, 1] + X[:, 2] * X[:, 3])))).astype(np.float32)
    out = []
    for _ in range(2):
        engine.mlp_init(3)
        P0 = engine.mlp_params()
        l0 = M.loss(P0, X, Y)
        lf = engine.mlp_train(X, Y, epochs=6, batch=1024, lr=3e-3, seed=8)
        out.append(engine.mlp_params())
    assert np.array_equal(out[0].view(np.uint32), out[1].view(np.uint32)), "training not reproducible"
    lr_ = M.loss(out[0], X, Y)
    assert abs(lf - M.loss(out[0], X[:65536], Y[:65536])) < 1e-4
    # entropy of the labels is the floor; the network must close most of the gap
    floor = float(np.mean(-(Y * np.log(Y) + (1 - Y) * np.log(1 - Y))))
    assert lr_ - floor < 0.25 * (l0 - floor), (l0, lr_, floor)


def _features_ref(xg, prior, c):
    """FP64 restatement of the 12 pair features (pooled 3x3 table of both classes)."""
    G = xg.shape[0]
    out = []
    for t in range(1, G):
        for col in range(t):
            tab = np.zeros((3, 3))
            np.add.at(tab, (xg[t], xg[col]), 1)
            n = tab.sum()
            r, k = tab.sum(1), tab.sum(0)
            e = np.outer(r, k) / n
            m = e > 0
            chi = float((((tab - e) ** 2)[m] / e[m]).sum())
            mm = (tab > 0) & m
            mi = float((tab[mm] / n * np.log(tab[mm] / e[mm])).sum())
            hr = float(-sum(v / n * np.log(v / n) for v in r if v > 0))
            hk = float(-sum(v / n * np.log(v / n) for v in k if v > 0))
            nr, nk = int((r > 0).sum()), int((k > 0).sum())
            mn = min(nr, nk)
            out.append([chi / n, np.log1p(chi), mi, hr, hk, (tab == 0).sum() / 9, e[m].min() / (n / 9), n / 1000,
                        (nr - 1) * (nk - 1) / 4, prior[t, col], np.log(max(c, 1)) / 12,
                        np.sqrt(chi / (n * (mn - 1))) if mn > 1 else 0.0])
    return np.array(out)


def test_logged_features_and_labels_match_oracle(engine):
    wl = eddy_amd.Workload(0)
    engine.upload(wl.X, wl.classes)
    genes, pairs = wl.sets[0]
    prm = wl.params(mode=eddy_amd.MODE_BATCHED, schedule=eddy_amd.TESTS_ALL)
    engine.reset_stream()
    X, Y = engine.learn_collect(wl.sets[:1], prm)
    G = len(genes)
    c = G * (G - 1) // 2
    assert X.shape == (c, 12) and Y.shape == (c,)
    prior = wl.prior_matrix(0)
    xg = wl.X[np.asarray(genes)].astype(np.int64)
    F = _features_ref(xg, np.maximum(prior, prior.T), c)
    assert np.allclose(X, F, rtol=2e-5, atol=2e-6), np.abs(X - F).max(0)
    # labels: full-test edges of both classes over permutations 1..r-1
    engine.reset_stream()
    data_ref = wl.X.astype(np.int32) - 1
    N, n1 = wl.N, wl.cfg.nA
    L = engine.L
    hits = np.zeros(c)
    lab = np.zeros(N, np.uint8)
    r = prm.perms
    for pm in range(r):
        assert L.eddy_stream_permutation(engine.h, N, n1, pm, lab.ctypes.data_as(C.c_void_p)) == 0
        if pm == 0:
            continue
        t1, t2, a1, a2 = oracle_ref.class_split(data_ref, genes, lab)
        o = oracle_ref.run_perm(t1, t2, a1, a2, prior, oracle_ref.params(mp=prm.mp, threads=8), want_bits=True)
        hits += o["bits"][0].astype(np.float64) + o["bits"][N + 1]
    assert np.array_equal(Y, (hits / (2 * (r - 1))).astype(np.float32))


def _same_results(run_a, run_b):
    (a_res, a_js), (b_res, b_js) = run_a, run_b
    assert len(a_res) == len(b_res)
    for s in range(len(a_res)):
        a, b = a_res[s], b_res[s]
        assert (a.status, a.perms_run, a.U0) == (b.status, b.perms_run, b.U0), s
        for f in ("js0", "pDDN", "alpha", "beta"):
            x, y = getattr(a, f), getattr(b, f)
            assert (np.isnan(x) and np.isnan(y)) or np.float64(x).view(np.uint64) == np.float64(y).view(np.uint64)
        assert np.array_equal(a_js[s].view(np.uint64), b_js[s].view(np.uint64)), s


def test_learned_windows_are_exact_and_save_tests(engine):
    """The trained network's windows against the method without it: identical
    results; fewer tests than running every full test in the count kernel;
    close to the fewest possible (every full test deferred to the plan); and
    on held-out sets the network predicts the edge rates better than their
    training mean."""
    wl = eddy_amd.Workload(3, nsets=48)
    engine.upload(wl.X, wl.classes)
    prm = wl.params(perms=40, mode=eddy_amd.MODE_BATCHED, schedule=eddy_amd.TESTS_CUTOFF, learned=0)
    print("case: learned windows, 48 sets of the 5000 x 600 workload at r=40", flush=True)
    engine.reset_stream()
    Xtr, Ytr = engine.learn_collect(wl.sets[:16], prm)
    Xte, Yte = engine.learn_collect(wl.sets[16:], prm)
    engine.mlp_init(1)
    engine.mlp_train(Xtr, Ytr, epochs=4, batch=2048, lr=3e-3, seed=1)
    P = engine.mlp_params()
    yte = np.clip(Yte.astype(np.float64), 0, 1)
    net_loss = M.loss(P, Xte, yte)
    m = float(np.clip(Ytr.mean(), 1e-6, 1 - 1e-6))
    const_loss = float(np.mean(-(yte * np.log(m) + (1 - yte) * np.log(1 - m))))
    print("held-out loss: network %.4f, training-mean predictor %.4f" % (net_loss, const_loss), flush=True)
    assert net_loss < 0.8 * const_loss
    runs = {}
    for name, learned in (("all", 0), ("deferred", 3), ("learned", 1)):
        engine.reset_stream()
        res = engine.run_sets(wl.sets, wl.params(perms=40, mode=eddy_amd.MODE_BATCHED,
                                                   schedule=eddy_amd.TESTS_CUTOFF, learned=learned))
        runs[name] = (res, [engine.js(s).copy() for s in range(len(wl.sets))])
    for name in ("deferred", "learned"):
        _same_results(runs[name], runs["all"])
    tests = {k: sum(r.tests_eval for r in v[0]) for k, v in runs.items()}
    print("tests: every full test %d, deferred %d, learned %d" % (tests["all"], tests["deferred"], tests["learned"]),
          flush=True)
    assert tests["learned"] < 0.97 * tests["all"]
    assert tests["learned"] <= 1.05 * tests["deferred"]


def test_learned_job_is_exact(engine):
    wl = eddy_amd.Workload(3, nsets=48)
    engine.upload(wl.X, wl.classes)
    prm = wl.params(perms=40, mode=eddy_amd.MODE_BATCHED, schedule=eddy_amd.TESTS_CUTOFF, learned=0)
    print("case: learned job, 48 sets of the 5000 x 600 workload at r=40", flush=True)
    engine.reset_stream()
    base = engine.run_sets(wl.sets, prm)
    base_js = [engine.js(s).copy() for s in range(len(wl.sets))]
    base_net = [engine.perm0_networks(s, base[s].U0, prm.mp) for s in (0, 20, 47)]
    engine.reset_stream()
    cfg = eddy_amd.learn_cfg(train_sets=16, epochs=4, batch=2048, lr=3e-3, margin=1.0, seed=1, max_rows=0,
                             label_perms=10)
    # the sets marshalled once (as the bench passes them), against the list
    res, rep = engine.run_learned(wl.batch, prm, cfg)
    assert rep.trained == 1 and rep.train_rows > 1024
    assert rep.loss_final < rep.loss_initial
    assert rep.first_windowed_set == -1 or rep.first_windowed_set >= 16
    _same_results((base, base_js), (res, [engine.js(s).copy() for s in range(len(wl.sets))]))
    for k, s in enumerate((0, 20, 47)):
        got = engine.perm0_networks(s, res[s].U0, prm.mp)
        for x, y in zip(base_net[k], got):
            assert np.array_equal(np.asarray(x).view(np.uint8), np.asarray(y).view(np.uint8))


def test_repermute_driver_matches_direct_run(engine):
    """The re-permutation driver reruns the most significant sets (smallest
    pDDN, status OK) with more permutations on a fresh stream, in set order:
    the same results as resetting the stream and running those sets directly."""
    wl = eddy_amd.Workload(3, nsets=30)
    engine.upload(wl.X, wl.classes)
    prm = wl.params(perms=20, mode=eddy_amd.MODE_BATCHED, schedule=eddy_amd.TESTS_CUTOFF)
    print("case: repermute top 6 of 30 sets at r=40", flush=True)
    engine.reset_stream()
    first = engine.run_sets(wl.sets, prm)
    ok = [q for q in range(len(first)) if first[q].status == eddy_amd.SET_OK and not np.isnan(first[q].pDDN)]
    want = sorted(sorted(ok, key=lambda q: (first[q].pDDN, q))[:6])
    chosen, res = engine.repermute(wl.sets, first, prm, top=6, perms=40)
    assert chosen == want
    got_js = [engine.js(k).copy() for k in range(len(chosen))]
    engine.reset_stream()
    direct = engine.run_sets([wl.sets[q] for q in chosen], wl.params(perms=40, mode=eddy_amd.MODE_BATCHED,
                                                                      schedule=eddy_amd.TESTS_CUTOFF))
    _same_results((res, got_js), (direct, [engine.js(k).copy() for k in range(len(chosen))]))
    assert all(r.perms_run in (1, 40) for r in res)
