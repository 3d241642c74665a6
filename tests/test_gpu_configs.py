"""GPU parity against the oracle at the benchmark workloads' own shapes.

The seeded workloads (eddy_synth_config) in their real sizes:
  5000 genes x 600 samples (300/300), 500 sets of 10..300 genes - the batched
      form and the learned job (sets through the permutation stream, JS, pDDN
      and U0 bit for bit), and one permutation of the largest set in detail;
  20000 genes x 1200 samples (300/900), sets of 5..500 genes - one permuted
      labelling of sets of about 5, 150 and 500 genes in detail (every
      network, per-node BDeu, likelihoods, JS), and two small sets through
      the stream;
  1000 genes x 400 samples (200/200), the 150-gene set - a permuted labelling
      with every edge bit (out23) of every network.
The oracle (oracle/) is the checker; it runs the reference algorithm on the
CPU (full leave-one-out tallies, brute-force BDeu), threaded over networks.
Every FP64 output is compared bit for bit, which implies the 1e-12 relative
bound.
"""
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


def _spread(wl, k):
    """k sets spanning the size range: smallest, largest and quantiles between."""
    sizes = np.array([len(g) for g, _ in wl.sets])
    order = np.argsort(sizes, kind="stable")
    pick = sorted({int(order[int(round(q * (len(order) - 1)))]) for q in np.linspace(0, 1, k)})
    return pick


def _near(wl, G):
    sizes = np.array([len(g) for g, _ in wl.sets])
    return int(np.argmin(np.abs(sizes - G)))


def _eq64(x, y):
    return (np.isnan(x) and np.isnan(y)) or np.float64(x).view(np.uint64) == np.float64(y).view(np.uint64)


def _detail_vs_oracle(engine, wl, s, labels, schedule, bits, oracle_out=None):
    genes, pairs = wl.sets[s]
    prm = wl.params(schedule=schedule, mode=eddy_amd.MODE_BATCHED)
    print("case: set %d G %d N %d schedule %d" % (s, len(genes), wl.N, schedule), flush=True)
    g = engine.perm_detail(genes, pairs, labels, prm, bits=bits, out5=True)
    o = oracle_out
    if o is None:
        data_ref = wl.X.astype(np.int32) - 1
        t1, t2, n1, n2 = oracle_ref.class_split(data_ref, genes, labels)
        o = oracle_ref.run_perm(t1, t2, n1, n2, wl.prior_matrix(s), oracle_ref.params(mp=prm.mp, threads=_threads()),
                                want_bits=bits)
    if bits:
        assert np.array_equal(g["bits"], o["bits"]), "edge bits (out23) differ"
    assert g["U"] == o["U"]
    assert np.array_equal(g["uniq_ids"], o["uniq_ids"])
    assert np.array_equal(g["uni_nodes"], o["uni_nodes"])
    assert np.array_equal(g["uni_epn"], o["uni_epn"])
    assert np.array_equal(g["uni_edges"], o["uni_edges"])
    assert np.array_equal(g["out5"].view(np.uint64), o["out5"].view(np.uint64)), "BDeu per node differs"
    assert np.array_equal(g["lval"].view(np.uint64), o["lval"].view(np.uint64)), "likelihoods differ"
    assert _eq64(g["js"], o["js"])
    return o


def _oracle_stream(wl, idx, perms):
    """The oracle's run of sets idx (in order) through a reset rand() stream."""
    data_ref = wl.X.astype(np.int32) - 1
    oracle_ref.reset_stream()
    out = []
    for s in idx:
        genes, _ = wl.sets[s]
        out.append(oracle_ref.run_set(data_ref, wl.classes.astype(np.int32), genes, wl.prior_matrix(s),
                                      oracle_ref.params(mp=wl.cfg.mp, perms=perms, threads=_threads())))
    return out


def _check_stream(engine, res, ref, perms):
    for k, (r, o) in enumerate(zip(res, ref)):
        assert r.perms_run == o["ran"], k
        assert (r.status == eddy_amd.SET_FAILED_NAN) == bool(o["failed"]), k
        n = o["ran"]
        js = engine.js(k)
        assert np.array_equal(js[:n].view(np.uint64), o["js"][:n].view(np.uint64)), k
        if not o["failed"]:
            assert _eq64(r.pDDN, o["pDDN"]), k
            assert r.U0 == o["U0"], k


# ---------------------------------------------------------------- 5000 x 600
@pytest.fixture(scope="module")
def wl3():
    return eddy_amd.Workload(3)


@pytest.fixture(scope="module")
def oracle3(wl3):
    idx = _spread(wl3, 6)
    return idx, _oracle_stream(wl3, idx, 4)


@pytest.mark.parametrize("form", ["batched", "learned"])
def test_5000x600_sets_through_stream(engine, wl3, oracle3, form):
    """Six of the 500 sets spanning 10..300 genes, r = 4, through the stream:
    the batched form (cutoff schedule) and the learned job (network trained
    on the first two sets, its windows on the other four)."""
    idx, ref = oracle3
    engine.upload(wl3.X, wl3.classes)
    sets = [wl3.sets[s] for s in idx]
    prm = wl3.params(perms=4, mode=eddy_amd.MODE_BATCHED, schedule=eddy_amd.TESTS_CUTOFF)
    print("case: 5000x600 sets", idx, "sizes", [len(g) for g, _ in sets], form, flush=True)
    engine.reset_stream()
    if form == "batched":
        res = engine.run_sets(sets, prm)
    else:
        res, rep = engine.run_learned(sets, prm, eddy_amd.learn_cfg(train_sets=2, epochs=2, batch=256,
                                                                      label_perms=3))
        assert rep.trained == 1
    _check_stream(engine, res, ref, 4)


def test_5000x600_largest_set_detail(engine, wl3):
    """The largest set (about 300 genes) at one permuted labelling: every edge
    bit in the all-tests schedule, and the cutoff schedule's networks."""
    s = _spread(wl3, 2)[-1]
    engine.upload(wl3.X, wl3.classes)
    lab = _labels(wl3.N, wl3.cfg.nA, 31)
    o = _detail_vs_oracle(engine, wl3, s, lab, eddy_amd.TESTS_ALL, bits=True)
    _detail_vs_oracle(engine, wl3, s, lab, eddy_amd.TESTS_CUTOFF, bits=False, oracle_out=o)


# ---------------------------------------------------------------- 20000 x 1200
@pytest.fixture(scope="module")
def wl4():
    return eddy_amd.Workload(4)


@pytest.mark.parametrize("G", [5, 150, 500])
def test_20000x1200_detail(engine, wl4, G):
    """Uneven classes (300/900) at one permuted labelling, sets of about 5,
    150 and 500 genes: 19 mask words per class, events and families of the
    largest sets, the tail kernel's 1202-network likelihoods."""
    s = _near(wl4, G)
    engine.upload(wl4.X, wl4.classes)
    _detail_vs_oracle(engine, wl4, s, _labels(wl4.N, wl4.cfg.nA, 40 + G), eddy_amd.TESTS_CUTOFF, bits=False)


def test_20000x1200_sets_through_stream(engine, wl4):
    """Two small sets at r = 3 through the stream (1200-sample permutations)."""
    idx = [_near(wl4, 8), _near(wl4, 20)]
    ref = _oracle_stream(wl4, idx, 3)
    engine.upload(wl4.X, wl4.classes)
    engine.reset_stream()
    print("case: 20000x1200 sets", idx, flush=True)
    res = engine.run_sets([wl4.sets[s] for s in idx],
                          wl4.params(perms=3, mode=eddy_amd.MODE_BATCHED, schedule=eddy_amd.TESTS_CUTOFF))
    _check_stream(engine, res, ref, 3)


# ---------------------------------------------------------------- 1000 x 400
def test_1000x400_permuted_all_bits(engine):
    """The 150-gene set at a permuted labelling: every edge bit of every
    network (all-tests schedule) and the cutoff schedule's networks."""
    wl = eddy_amd.Workload(1)
    engine.upload(wl.X, wl.classes)
    lab = _labels(wl.N, wl.cfg.nA, 77)
    o = _detail_vs_oracle(engine, wl, 0, lab, eddy_amd.TESTS_ALL, bits=True)
    _detail_vs_oracle(engine, wl, 0, lab, eddy_amd.TESTS_CUTOFF, bits=False, oracle_out=o)


# ---------------------------------------------------------------- re-permutation
def test_repermute_against_oracle(engine):
    """The re-permutation driver against the oracle: 12 sets at r = 6, one of
    them a single gene (NaN at permutation 0, excluded from the selection);
    the four most significant run again at r = 8 on a fresh stream, and each
    one's JS and pDDN equal the oracle's run of those sets from a reset stream."""
    wl = eddy_amd.Workload(3, nsets=12)
    wl.sets[3] = (wl.sets[3][0][:1], wl.sets[3][1][:0])
    engine.upload(wl.X, wl.classes)
    prm = wl.params(perms=6, mode=eddy_amd.MODE_BATCHED, schedule=eddy_amd.TESTS_CUTOFF)
    print("case: repermute 4 of 12 sets, r 6 -> 8", flush=True)
    engine.reset_stream()
    first = engine.run_sets(wl.sets, prm)
    assert first[3].status == eddy_amd.SET_FAILED_NAN
    first_ref = _oracle_stream(wl, list(range(12)), 6)
    _check_stream(engine, first, first_ref, 6)
    chosen, res = engine.repermute(wl.sets, first, prm, top=4, perms=8)
    assert 3 not in chosen and len(chosen) == 4
    ok = [q for q in range(12) if first_ref[q]["failed"] == 0 and not np.isnan(first_ref[q]["pDDN"])]
    assert chosen == sorted(sorted(ok, key=lambda q: (first_ref[q]["pDDN"], q))[:4])
    _check_stream(engine, res, _oracle_stream(wl, chosen, 8), 8)
