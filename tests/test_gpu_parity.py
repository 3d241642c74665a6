"""GPU parity: the HIP engine against the CPU oracle on the same inputs.

Integer outputs (edge bits, unique networks, CSR) must be identical; FP64
outputs (BDeu per node, likelihoods, JS) are compared bit for bit as well,
since both sides round through the same FP64 routines (tolerance 1e-12
relative is the stated bar; the assertions use exact equality, which implies it).
"""
import os

import numpy as np
import pytest

import eddy_amd
import oracle_ref

pytestmark = pytest.mark.gpu


def _labels(N, n1, seed):
    rng = np.random.default_rng(seed)
    lab = np.ones(N, np.uint8)
    lab[rng.permutation(N)[:n1]] = 0
    return lab


def _compare_perm(engine, wl, s, labels, schedule, check_bits, mp=None):
    genes, pairs = wl.sets[s]
    prm = wl.params(schedule=schedule, mode=eddy_amd.MODE_PLAIN, mp=mp)
    print("case: set", s, "G", len(genes), "schedule", schedule, "mp", prm.mp, flush=True)
    g = engine.perm_detail(genes, pairs, labels, prm, bits=check_bits, out5=True)
    data_ref = wl.X.astype(np.int32) - 1
    t1, t2, n1, n2 = oracle_ref.class_split(data_ref, genes, labels)
    o = oracle_ref.run_perm(t1, t2, n1, n2, wl.prior_matrix(s), oracle_ref.params(mp=prm.mp), want_bits=check_bits)
    if check_bits:
        assert np.array_equal(g["bits"], o["bits"]), "edge bits (out23) differ"
    assert g["U"] == o["U"]
    assert np.array_equal(g["uniq_ids"], o["uniq_ids"])
    assert np.array_equal(g["uni_nodes"], o["uni_nodes"])
    assert np.array_equal(g["uni_epn"], o["uni_epn"])
    assert np.array_equal(g["uni_edges"], o["uni_edges"])
    assert np.array_equal(g["out5"].view(np.uint64), o["out5"].view(np.uint64)), "BDeu per node differs"
    assert np.array_equal(g["lval"].view(np.uint64), o["lval"].view(np.uint64)), "likelihoods differ"
    assert (np.isnan(g["js"]) and np.isnan(o["js"])) or g["js"] == o["js"]


@pytest.mark.parametrize("schedule", [eddy_amd.TESTS_ALL, eddy_amd.TESTS_CUTOFF])
def test_perm_parity_config0(engine, schedule):
    wl = eddy_amd.Workload(0)
    engine.upload(wl.X, wl.classes)
    N, n1 = wl.N, wl.cfg.nA
    for seed in range(3):
        lab = wl.classes if seed == 0 else _labels(N, n1, seed)
        _compare_perm(engine, wl, 0, lab, schedule, check_bits=(schedule == eddy_amd.TESTS_ALL))


@pytest.mark.parametrize("mp", [1, 2, 5, 8])
@pytest.mark.parametrize("schedule", [eddy_amd.TESTS_ALL, eddy_amd.TESTS_CUTOFF])
def test_perm_parity_other_parent_caps(engine, schedule, mp):
    """Parent caps other than 3: row lists, events (n and mp parents per
    event) and families with more parents than BDeu scores (at most 3,
    kernel5.cu:119-130) against the oracle."""
    wl = eddy_amd.Workload(0)
    engine.upload(wl.X, wl.classes)
    for seed in range(2):
        lab = wl.classes if seed == 0 else _labels(wl.N, wl.cfg.nA, 10 + seed)
        _compare_perm(engine, wl, 0, lab, schedule, check_bits=False, mp=mp)


def _oracle_threads():
    return max(1, min(8, os.cpu_count() or 1))


def test_sets_stream_parity(engine):
    """Whole gene sets through the permutation stream: several sets in a row,
    one of which is NaN at permutation 0 (a single network) and so consumes
    only one permutation's draws (main.cu:1431-1435); JS values, pDDN and the
    failed status must match the oracle run through glibc rand()."""
    wl = eddy_amd.Workload(0)
    engine.upload(wl.X, wl.classes)
    genes, pairs = wl.sets[0]
    sets = [(genes[:12], pairs[:0]), (genes[12:14], pairs[:0]), (genes[14:30], pairs[:0]), (genes[:1], pairs[:0]),
            (genes[30:40], pairs[:0])]
    data_ref = wl.X.astype(np.int32) - 1
    runs = [(eddy_amd.MODE_BATCHED, eddy_amd.TESTS_CUTOFF, None), (eddy_amd.MODE_PLAIN, eddy_amd.TESTS_ALL, None),
            (eddy_amd.MODE_BATCHED, eddy_amd.TESTS_CUTOFF, 2), (eddy_amd.MODE_BATCHED, eddy_amd.TESTS_CUTOFF, 3)]
    for mode, sched, learn in runs:
        prm = wl.params(perms=8, mode=mode, schedule=sched)
        print("case: stream parity mode", mode, "learned job" if learn else "", learn, flush=True)
        engine.reset_stream()
        if learn is None:
            res = engine.run_sets(sets, prm)
        else:  # the learned job: the NaN set falls in the logged part (2) or right after it (3)
            res, rep = engine.run_learned(sets, prm, eddy_amd.learn_cfg(train_sets=learn, epochs=2, batch=64))
            assert rep.trained == 1
        oracle_ref.reset_stream()
        for s, (g, _) in enumerate(sets):
            prior = np.zeros((len(g), len(g)), np.int32)
            o = oracle_ref.run_set(data_ref, wl.classes.astype(np.int32), g, prior,
                                   oracle_ref.params(mp=3, perms=8, threads=_oracle_threads()))
            r = res[s]
            assert r.perms_run == o["ran"]
            assert (r.status == eddy_amd.SET_FAILED_NAN) == bool(o["failed"])
            js = engine.js(s)
            n = o["ran"]
            assert np.array_equal(js[:n].view(np.uint64), o["js"][:n].view(np.uint64)), s
            if not o["failed"]:
                assert r.pDDN == o["pDDN"] or (np.isnan(r.pDDN) and np.isnan(o["pDDN"]))
                assert r.U0 == o["U0"]


def test_perm_parity_config1_cutoff(engine):
    """The 1000 x 400 workload's shape (a 150-gene set, 400 samples): networks, BDeu, likelihoods, JS."""
    wl = eddy_amd.Workload(1)
    engine.upload(wl.X, wl.classes)
    genes, pairs = wl.sets[0]
    prm = wl.params(schedule=eddy_amd.TESTS_CUTOFF, mode=eddy_amd.MODE_BATCHED)
    print("case: config1 cutoff", flush=True)
    g = engine.perm_detail(genes, pairs, wl.classes, prm, bits=False, out5=True)
    data_ref = wl.X.astype(np.int32) - 1
    t1, t2, n1, n2 = oracle_ref.class_split(data_ref, genes, wl.classes)
    o = oracle_ref.run_perm(t1, t2, n1, n2, wl.prior_matrix(0), oracle_ref.params(mp=3, threads=_oracle_threads()),
                            want_bits=False)
    assert g["U"] == o["U"]
    assert np.array_equal(g["uniq_ids"], o["uniq_ids"])
    assert np.array_equal(g["uni_nodes"], o["uni_nodes"])
    assert np.array_equal(g["uni_edges"], o["uni_edges"])
    assert np.array_equal(g["out5"].view(np.uint64), o["out5"].view(np.uint64))
    assert np.array_equal(g["lval"].view(np.uint64), o["lval"].view(np.uint64))
    assert g["js"] == o["js"]


@pytest.mark.parametrize("nA,nB", [(1000, 1000), (300, 1700)])
def test_perm_parity_large_sample_counts(engine, nA, nB):
    """Near the engine's sample limit (2046): 16 sample chunks in the count
    kernel, 16-27 mask words per class, the tail kernel's largest LDS use;
    the 40-gene set of the 200 x 200 workload on a seeded dataset of nA + nB samples."""
    wl = eddy_amd.Workload(0)
    L = eddy_amd.lib()
    X = np.zeros((wl.cfg.genes, nA + nB), np.uint8)
    assert L.eddy_synth_dataset(wl.cfg.seed + 7, wl.cfg.genes, nA, nB, wl.cfg.rewire,
                                X.ctypes.data_as(eddy_amd.C.c_void_p)) == eddy_amd.EDDY_OK
    wl.X, wl.N = X, nA + nB
    wl.classes = np.array([0] * nA + [1] * nB, np.uint8)
    wl.cfg.nA, wl.cfg.nB = nA, nB
    engine.upload(wl.X, wl.classes)
    lab = _labels(wl.N, nA, 21)
    _compare_perm(engine, wl, 0, lab, eddy_amd.TESTS_CUTOFF, check_bits=False)
