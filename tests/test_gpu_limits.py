"""GPU parity against the oracle at the reference's own input limits.

* Sample counts above 2046, up to the reference's MAX_NUM_SAMPLES = 20000
  (definitions.cuh:17): 4000 samples (1000/3000 classes; the score kernels'
  network tables in global memory) and 20000 samples (5000/15000; the rows
  kernel's deltas and class lists in global memory as well), one permuted
  labelling in detail, and two sets through the permutation stream.
* Parent caps above 16 (the reference takes any positive -mp,
  main.cu:170-172): 17, 20 and 100 on the 40-gene set of the 200 x 200
  workload (100 caps nothing there: a row has at most 39 candidate parents),
  both schedules, and a whole set through the stream at -mp 20.
* Samples outside the three states (an expression value other than -1, 0,
  1 counts in no cell and no observed state, as in kernel1.cu:338-387 and
  kernel5.cu:197-218): the tallies then store their ninth cell instead of
  deriving it from the class size.
Every output is compared bit for bit.
"""
import ctypes as C
import os

import numpy as np
import pytest

import eddy_amd
import oracle_ref

pytestmark = pytest.mark.gpu


def _threads():
    return max(1, min(16, os.cpu_count() or 1))


def _labels(N, n1, seed):
    rng = np.random.default_rng(seed)
    lab = np.ones(N, np.uint8)
    lab[rng.permutation(N)[:n1]] = 0
    return lab


class _Wide:
    """200 genes x (nA + nB) samples from the seeded DAG generator, one set."""

    def __init__(self, nA, nB, G, seed=0xB16):
        self.nA, self.nB, self.N = nA, nB, nA + nB
        self.X = np.zeros((200, self.N), np.uint8)
        rc = eddy_amd.lib().eddy_synth_dataset(seed, 200, nA, nB, 0.1, self.X.ctypes.data_as(C.c_void_p))
        assert rc == eddy_amd.EDDY_OK
        self.classes = np.array([0] * nA + [1] * nB, np.uint8)
        rng = np.random.default_rng(seed + G)
        self.genes = rng.permutation(200)[:G].astype(np.int32)
        pr = []
        while len(pr) < G // 3:
            a, b = rng.integers(0, G, 2)
            if a != b:
                pr.append((int(a), int(b)))
        self.pairs = np.array(pr, np.int32).reshape(-1, 2)
        self.prior = np.zeros((G, G), np.int32)
        for a, b in pr:
            self.prior[a, b] = self.prior[b, a] = 1

    def params(self, schedule, perms=100, mp=3, mode=eddy_amd.MODE_BATCHED):
        aep = 1 - 0.95 * 0.5 ** 0.5
        return eddy_amd.Params(mp, perms, 1, 0.05, aep, 0.05, mode, schedule, 0)


def _compare(g, o, bits):
    if bits:
        assert np.array_equal(g["bits"], o["bits"]), "edge bits (out23) differ"
    assert g["U"] == o["U"]
    assert np.array_equal(g["uniq_ids"], o["uniq_ids"])
    assert np.array_equal(g["uni_nodes"], o["uni_nodes"])
    assert np.array_equal(g["uni_epn"], o["uni_epn"])
    assert np.array_equal(g["uni_edges"], o["uni_edges"])
    assert np.array_equal(g["out5"].view(np.uint64), o["out5"].view(np.uint64)), "BDeu per node differs"
    assert np.array_equal(g["lval"].view(np.uint64), o["lval"].view(np.uint64)), "likelihoods differ"
    assert (np.isnan(g["js"]) and np.isnan(o["js"])) or g["js"] == o["js"]


@pytest.mark.parametrize("nA,nB,G", [(1000, 3000, 20), (5000, 15000, 8)])
def test_sample_counts_above_2046(engine, nA, nB, G):
    w = _Wide(nA, nB, G)
    engine.upload(w.X, w.classes)
    lab = _labels(w.N, nA, 5)
    data_ref = w.X.astype(np.int32) - 1
    t1, t2, n1, n2 = oracle_ref.class_split(data_ref, w.genes, lab)
    bits = w.N <= 4000
    o = oracle_ref.run_perm(t1, t2, n1, n2, w.prior, oracle_ref.params(mp=3, threads=_threads()), want_bits=bits)
    schedules = [eddy_amd.TESTS_CUTOFF] + ([eddy_amd.TESTS_ALL] if bits else [])
    for sched in schedules:
        print("case: N %d (%d/%d) G %d schedule %d" % (w.N, nA, nB, G, sched), flush=True)
        g = engine.perm_detail(w.genes, w.pairs, lab, w.params(sched), bits=sched == eddy_amd.TESTS_ALL, out5=True)
        _compare(g, o, sched == eddy_amd.TESTS_ALL)


def test_4000_samples_through_stream(engine):
    w = _Wide(1000, 3000, 16)
    engine.upload(w.X, w.classes)
    sets = [(w.genes[:9], w.pairs[:0]), (w.genes, w.pairs)]
    priors = [np.zeros((9, 9), np.int32), w.prior]
    prm = w.params(eddy_amd.TESTS_CUTOFF, perms=3)
    print("case: N 4000, two sets through the stream at r = 3", flush=True)
    engine.reset_stream()
    res = engine.run_sets(sets, prm)
    data_ref = w.X.astype(np.int32) - 1
    oracle_ref.reset_stream()
    for k, ((g, _), pm) in enumerate(zip(sets, priors)):
        o = oracle_ref.run_set(data_ref, w.classes.astype(np.int32), g, pm,
                               oracle_ref.params(mp=3, perms=3, threads=_threads()))
        assert res[k].perms_run == o["ran"]
        n = o["ran"]
        assert np.array_equal(engine.js(k)[:n].view(np.uint64), o["js"][:n].view(np.uint64)), k
        if not o["failed"]:
            assert np.float64(res[k].pDDN).view(np.uint64) == np.float64(o["pDDN"]).view(np.uint64)
            assert res[k].U0 == o["U0"]


@pytest.mark.parametrize("mp", [17, 20, 100])
@pytest.mark.parametrize("schedule", [eddy_amd.TESTS_ALL, eddy_amd.TESTS_CUTOFF])
def test_parent_caps_above_16(engine, schedule, mp):
    wl = eddy_amd.Workload(0)
    engine.upload(wl.X, wl.classes)
    genes, pairs = wl.sets[0]
    data_ref = wl.X.astype(np.int32) - 1
    for seed in range(2):
        lab = wl.classes if seed == 0 else _labels(wl.N, wl.cfg.nA, 30 + seed)
        prm = wl.params(schedule=schedule, mode=eddy_amd.MODE_BATCHED, mp=mp)
        print("case: G %d mp %d schedule %d seed %d" % (len(genes), mp, schedule, seed), flush=True)
        g = engine.perm_detail(genes, pairs, lab, prm, bits=schedule == eddy_amd.TESTS_ALL, out5=True)
        t1, t2, n1, n2 = oracle_ref.class_split(data_ref, genes, lab)
        o = oracle_ref.run_perm(t1, t2, n1, n2, wl.prior_matrix(0), oracle_ref.params(mp=mp, threads=_threads()),
                                want_bits=schedule == eddy_amd.TESTS_ALL)
        _compare(g, o, schedule == eddy_amd.TESTS_ALL)


def test_parent_cap_20_through_stream(engine):
    wl = eddy_amd.Workload(0)
    engine.upload(wl.X, wl.classes)
    genes, pairs = wl.sets[0]
    sets = [(genes, pairs), (genes[:25], pairs[:0])]
    priors = [wl.prior_matrix(0), np.zeros((25, 25), np.int32)]
    data_ref = wl.X.astype(np.int32) - 1
    for mode, sched in ((eddy_amd.MODE_BATCHED, eddy_amd.TESTS_CUTOFF), (eddy_amd.MODE_PLAIN, eddy_amd.TESTS_ALL)):
        prm = wl.params(perms=6, mode=mode, schedule=sched, mp=20)
        print("case: mp 20 through the stream, mode %d" % mode, flush=True)
        engine.reset_stream()
        res = engine.run_sets(sets, prm)
        oracle_ref.reset_stream()
        for k, ((g, _), pm) in enumerate(zip(sets, priors)):
            o = oracle_ref.run_set(data_ref, wl.classes.astype(np.int32), g, pm,
                                   oracle_ref.params(mp=20, perms=6, threads=_threads()))
            n = o["ran"]
            assert res[k].perms_run == n
            assert np.array_equal(engine.js(k)[:n].view(np.uint64), o["js"][:n].view(np.uint64)), k
            assert res[k].U0 == o["U0"]
            if not o["failed"]:
                assert np.float64(res[k].pDDN).view(np.uint64) == np.float64(o["pDDN"]).view(np.uint64)
        # permutation-0 networks of the first set through the capture (wide event records)
        U = res[0].U0
        ids, nodes, epn, edges, lval = engine.perm0_networks(0, U, 20)
        t1, t2, n1, n2 = oracle_ref.class_split(data_ref, genes, wl.classes)
        o = oracle_ref.run_perm(t1, t2, n1, n2, priors[0], oracle_ref.params(mp=20, threads=_threads()),
                                want_bits=False)
        assert np.array_equal(ids, o["uniq_ids"]) and np.array_equal(edges, o["uni_edges"])
        assert np.array_equal(nodes, o["uni_nodes"]) and np.array_equal(epn, o["uni_epn"])
        assert np.array_equal(lval.view(np.uint64), o["lval"].view(np.uint64))


@pytest.mark.parametrize("schedule", [eddy_amd.TESTS_ALL, eddy_amd.TESTS_CUTOFF])
def test_samples_outside_the_three_states(engine, schedule):
    wl = eddy_amd.Workload(0)
    X = wl.X.copy()
    rng = np.random.default_rng(0x5EED)
    X[rng.random(X.shape) < 0.03] = 3  # about 3% of the values out of range
    engine.upload(X, wl.classes)
    genes, pairs = wl.sets[0]
    data_ref = X.astype(np.int32) - 1
    for seed in range(2):
        lab = wl.classes if seed == 0 else _labels(wl.N, wl.cfg.nA, 50 + seed)
        prm = wl.params(schedule=schedule, mode=eddy_amd.MODE_BATCHED)
        print("case: out-of-range states, schedule %d seed %d" % (schedule, seed), flush=True)
        g = engine.perm_detail(genes, pairs, lab, prm, bits=schedule == eddy_amd.TESTS_ALL, out5=True)
        t1, t2, n1, n2 = oracle_ref.class_split(data_ref, genes, lab)
        o = oracle_ref.run_perm(t1, t2, n1, n2, wl.prior_matrix(0), oracle_ref.params(mp=3, threads=_threads()),
                                want_bits=schedule == eddy_amd.TESTS_ALL)
        _compare(g, o, schedule == eddy_amd.TESTS_ALL)
    if schedule == eddy_amd.TESTS_CUTOFF:
        sets = [(genes, pairs), (genes[:17], pairs[:0])]
        priors = [wl.prior_matrix(0), np.zeros((17, 17), np.int32)]
        print("case: out-of-range states, two sets through the stream at r = 4", flush=True)
        engine.reset_stream()
        res = engine.run_sets(sets, wl.params(perms=4))
        oracle_ref.reset_stream()
        for k in range(2):
            o = oracle_ref.run_set(data_ref, wl.classes.astype(np.int32), sets[k][0], priors[k],
                                   oracle_ref.params(mp=3, perms=4, threads=_threads()))
            n = o["ran"]
            assert res[k].perms_run == n
            assert np.array_equal(engine.js(k)[:n].view(np.uint64), o["js"][:n].view(np.uint64)), k
            assert res[k].U0 == o["U0"]
