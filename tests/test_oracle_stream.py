"""Pins of the oracle's permutation stream, class split and uncapped EdgeList
pass, independent of the engine.

* eo_perm_labels (main.cu:538-573) against a statement-by-statement Python
  restatement of the reference loop (N discarded draws `rand() % N`, the
  linear-scan rejection loop, the relabelling of permutations > 0) run over
  the committed glibc rand() fixture: labels of consecutive permutations and
  the stream position after them (the next live rand() value).  A variant
  that drops the discarded draws is shown to fail the same comparison.
* eo_run_set's stream consumption: a set that is NaN at permutation 0 takes
  one permutation's draws (main.cu:1431-1435), any other set r of them.
* eo_class_split (main.cu:575-605) against numpy's stable partition of the
  samples by label, under permuted labels.
* eo_uncapped_pair (main.cu:1530-1612, run22 with PARENTS_LIMIT = INT_MAX)
  against a direct construction from the edge bits and against the oracle's
  own capped construction with the cap at G (no row can hold more).
"""
import ctypes as C
import os

import numpy as np
import pytest

import oracle_ref as O

FIX = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "glibc_rand_1e4.npy"))
LIBC = C.CDLL(None)


class FixtureRand:
    """rand() replayed from the committed fixture (glibc, no srand call)."""

    def __init__(self):
        self.i = 0

    def __call__(self):
        v = int(FIX[self.i])
        self.i += 1
        return v


def reference_permutation(rand, N, n1, perm, classIDs, discard=True):
    """main.cu:541-570 as the reference writes it."""
    randNums = [0] * N
    if discard:
        for c in range(N):
            randNums[c] = rand() % N
    n = 0
    while n < N:
        r = rand() % N
        x = 0
        while x < n:
            if randNums[x] == r:
                break
            x += 1
        if x == n:
            randNums[n] = r
            n += 1
    if perm > 0:
        for c in range(n1):
            classIDs[randNums[c]] = 0
        for c in range(n1, N):
            classIDs[randNums[c]] = 1
    return classIDs


def _oracle_labels(N, n1, perm, cls):
    arr = np.ascontiguousarray(cls, np.int32)
    O.lib().eo_perm_labels(N, n1, perm, arr.ctypes.data_as(C.c_void_p))
    return arr


@pytest.mark.parametrize("N,n1", [(7, 3), (30, 14), (64, 32)])
def test_perm_labels_against_reference_loop(N, n1):
    # stop while one more permutation (about N (1 + ln N) draws) surely fits the fixture
    budget = len(FIX) - int(12 * N * (1 + np.log(N)))
    rand = FixtureRand()
    O.reset_stream()
    cls_ref = [0] * n1 + [1] * (N - n1)
    cls_or = np.array(cls_ref, np.int32)
    perm = 0
    while rand.i < budget:
        cls_ref = reference_permutation(rand, N, n1, perm, list(cls_ref))
        cls_or = _oracle_labels(N, n1, perm, cls_or)
        assert list(cls_or) == cls_ref, (N, perm)
        assert sum(1 for v in cls_ref if v == 0) == n1
        perm += 1
    assert perm >= 10
    assert LIBC.rand() == int(FIX[rand.i])  # both consumed the same draws


def test_dropped_discard_draws_are_caught():
    """Sensitivity of the comparison above: without the N discarded draws of
    main.cu:542-544 the labels differ from the oracle's."""
    N, n1 = 30, 14
    rand = FixtureRand()
    O.reset_stream()
    cls = np.array([0] * n1 + [1] * (N - n1), np.int32)
    differ = False
    for perm in range(6):
        bad = reference_permutation(rand, N, n1, perm, list(cls), discard=False)
        cls = _oracle_labels(N, n1, perm, cls)
        differ |= list(cls) != bad
    assert differ


def _small_data(G, N, seed):
    rng = np.random.default_rng(seed)
    return rng.integers(-1, 2, size=(G, N)).astype(np.int32)


def test_run_set_stream_consumption():
    N, n1 = 24, 11
    data = _small_data(8, N, 3)
    lab = np.array([0] * n1 + [1] * (N - n1), np.int32)
    rand = FixtureRand()
    O.reset_stream()
    # one gene: a single network per class, U = 1, NaN at permutation 0
    r1 = O.run_set(data, lab, np.array([2], np.int32), np.zeros((1, 1), np.int32), O.params(perms=5))
    assert r1["ran"] == 1 and r1["failed"] == 1
    reference_permutation(rand, N, n1, 0, list(lab))
    # eight genes, r = 4: four permutations' draws
    r2 = O.run_set(data, lab, np.arange(8, dtype=np.int32), np.zeros((8, 8), np.int32),
                   O.params(perms=4, p_adjust=0, alpha_edge=0.3))
    assert r2["ran"] == 4
    cls = list(lab)
    for perm in range(4):
        cls = reference_permutation(rand, N, n1, perm, cls)
    assert LIBC.rand() == int(FIX[rand.i])


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_class_split_under_permuted_labels(seed):
    N, n1, Gall = 40, 17, 12
    data = _small_data(Gall, N, seed)
    rng = np.random.default_rng(seed)
    lab = np.ones(N, np.int32)
    lab[rng.permutation(N)[:n1]] = 0
    idx = rng.permutation(Gall)[:7].astype(np.int32)
    t1, t2, m1, m2 = O.class_split(data, idx, lab)
    assert (m1, m2) == (n1, N - n1)
    order = np.argsort(lab, kind="stable")  # class 0 first, ascending samples within each class
    assert np.array_equal(t1, data[idx][:, order[:n1]])
    assert np.array_equal(t2, data[idx][:, order[n1:]])


def _uncapped(b0, bL, G):
    nodes = np.zeros(2 * G, np.int32)
    edges = np.zeros(max(1, G * G), np.int32)
    epn = np.zeros(3, np.int32)
    n = O.lib().eo_uncapped_pair(b0.ctypes.data_as(C.c_void_p), bL.ctypes.data_as(C.c_void_p), G,
                                 nodes.ctypes.data_as(C.c_void_p), edges.ctypes.data_as(C.c_void_p), len(edges),
                                 epn.ctypes.data_as(C.c_void_p))
    assert n == epn[2]
    return nodes, edges[:n], epn


@pytest.mark.parametrize("G,seed", [(6, 1), (11, 2), (17, 3)])
def test_uncapped_pair_from_bits(G, seed):
    c = G * (G - 1) // 2
    rng = np.random.default_rng(seed)
    b0 = np.ascontiguousarray((rng.random(c) < 0.5).astype(np.uint8))
    bL = np.ascontiguousarray((rng.random(c) < 0.3).astype(np.uint8))
    nodes, edges, epn = _uncapped(b0, bL, G)
    want_nodes, want_edges, ends = [], [], [0]
    for b in (b0, bL):
        acc = 0
        for t in range(G):
            want_nodes.append(acc)
            cols = [col for col in range(t) if b[t * (t - 1) // 2 + col]]
            want_edges += cols
            acc += len(cols)
        ends.append(ends[-1] + acc)
    assert list(nodes) == want_nodes
    assert list(edges) == want_edges
    assert list(epn) == ends


def test_uncapped_pair_equals_capped_construction_at_cap_G():
    """The EdgeList pass is the network construction of H8-H9 without the cap:
    networks 0 and S-1 of a permutation built with mp = G (a row holds at
    most G - 1 parents) give the same CSR."""
    G, n1, n2 = 9, 14, 12
    data = _small_data(G, n1 + n2, 9)
    lab = np.array([0] * n1 + [1] * n2, np.int32)
    t1, t2, _, _ = O.class_split(data, np.arange(G, dtype=np.int32), lab)
    prm = O.params(mp=G, p_adjust=0, alpha_edge=0.4)
    S, c = n1 + n2 + 2, G * (G - 1) // 2
    out = O.EoPermOut()
    bits = np.zeros(S * c, np.uint8)
    nodes = np.zeros(S * G, np.int32)
    edges = np.zeros(S * G * G, np.int32)
    epn = np.zeros(S + 1, np.int32)
    keep = [np.zeros(n, t) for n, t in ((2 * G, np.int32), (6 * G, np.int32), (S, np.int32), (S, np.int32),
                                         (S * G, np.int32), (S * G * G, np.int32), (S + 1, np.int32),
                                         (2 * S * G, np.float64), (2 * S, np.float64))]
    out.edgebits, out.nodes, out.edges, out.edgesPN = (a.ctypes.data for a in (bits, nodes, edges, epn))
    (out.ppn, out.stf, out.uniqueN, out.uni_ids, out.uni_nodes, out.uni_edges, out.uni_epn, out.out5,
     out.lval) = (a.ctypes.data for a in keep)
    O.lib().eo_run_perm(t1.ctypes.data_as(C.c_void_p), t2.ctypes.data_as(C.c_void_p), G, n1, n2,
                        np.zeros((G, G), np.int32).ctypes.data_as(C.c_void_p), C.byref(prm), C.byref(out))
    bits = bits.reshape(S, c)
    u_nodes, u_edges, u_epn = _uncapped(np.ascontiguousarray(bits[0]), np.ascontiguousarray(bits[S - 1]), G)
    for k, s in enumerate((0, S - 1)):
        assert list(u_nodes[k * G:(k + 1) * G]) == list(nodes[s * G:(s + 1) * G])
        assert list(u_edges[u_epn[k]:u_epn[k + 1]]) == list(edges[epn[s]:epn[s + 1]])
    assert u_epn[2] > 0
