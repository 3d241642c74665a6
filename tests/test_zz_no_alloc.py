"""The entry point that allocates nothing (eddy_run_perm_detail_noalloc)
against eddy_run_perm_detail: identical outputs, bit for bit, on calls in a
row whose second input is smaller (the scratch array then holds the first
call's contents), on both schedules, at a sample count whose score tables
live in global memory, and a clear error for a set that does not fit."""
import ctypes as C

import numpy as np
import pytest

import eddy_amd

pytestmark = pytest.mark.gpu


def _labels(N, n1, seed):
    rng = np.random.default_rng(seed)
    lab = np.ones(N, np.uint8)
    lab[rng.permutation(N)[:n1]] = 0
    return lab


def _same(a, b):
    assert a["U"] == b["U"] and a["edSum"] == b["edSum"]
    assert np.float64(a["js"]).view(np.uint64) == np.float64(b["js"]).view(np.uint64) or (
        np.isnan(a["js"]) and np.isnan(b["js"]))
    for k in ("bits", "uniq_ids", "uni_nodes", "uni_epn", "uni_edges"):
        if a[k] is not None:
            assert np.array_equal(a[k], b[k]), k
    for k in ("out5", "lval"):
        assert np.array_equal(a[k].view(np.uint64), b[k].view(np.uint64)), k


def _case(engine, wl, genes, pairs, lab, sched):
    prm = wl.params(schedule=sched, mode=eddy_amd.MODE_BATCHED)
    bits = sched == eddy_amd.TESTS_ALL
    print("case: noalloc G %d N %d schedule %d" % (len(genes), wl.N, sched), flush=True)
    return (engine.perm_detail(genes, pairs, lab, prm, bits=bits, out5=True, noalloc=True),
            (genes, pairs, lab, prm, bits))


def test_noalloc_matches_perm_detail(engine):
    wl = eddy_amd.Workload(0)
    engine.upload(wl.X, wl.classes)
    genes, pairs = wl.sets[0]
    small = pairs[(pairs[:, 0] < 20) & (pairs[:, 1] < 20)]
    lab = _labels(wl.N, wl.cfg.nA, 3)
    for sched in (eddy_amd.TESTS_ALL, eddy_amd.TESTS_CUTOFF):
        # two no-allocation calls in a row, the second smaller, then the
        # allocating entry point on the same inputs
        first = _case(engine, wl, genes, pairs, lab, sched)
        second = _case(engine, wl, genes[:20], small, wl.classes, sched)
        for got, (g, p, lb, prm, bits) in (first, second):
            want = engine.perm_detail(g, p, lb, prm, bits=bits, out5=True, noalloc=False)
            _same(got, want)


def test_noalloc_at_4000_samples_and_capacity_error(engine):
    L = eddy_amd.lib()
    nA, nB = 1000, 3000
    wl = eddy_amd.Workload(0)
    X = np.zeros((wl.cfg.genes, nA + nB), np.uint8)
    assert L.eddy_synth_dataset(wl.cfg.seed + 9, wl.cfg.genes, nA, nB, 0.1, X.ctypes.data_as(C.c_void_p)) == 0
    wl.X, wl.N = X, nA + nB
    wl.classes = np.array([0] * nA + [1] * nB, np.uint8)
    wl.cfg.nA, wl.cfg.nB = nA, nB
    engine.upload(wl.X, wl.classes)
    genes, pairs = wl.sets[0]
    lab = _labels(wl.N, nA, 4)
    got, (g, p, lb, prm, bits) = _case(engine, wl, genes[:24], pairs[(pairs[:, 0] < 24) & (pairs[:, 1] < 24)], lab,
                                       eddy_amd.TESTS_CUTOFF)
    _same(got, engine.perm_detail(g, p, lb, prm, bits=bits, out5=True, noalloc=False))
    # 1024 genes at 20000 samples need more than the scratch array holds
    big = np.zeros((1030, 20000), np.uint8)
    assert L.eddy_synth_dataset(77, 1030, 5000, 15000, 0.1, big.ctypes.data_as(C.c_void_p)) == 0
    engine.upload(big, np.array([0] * 5000 + [1] * 15000, np.uint8))
    prm2 = wl.params(schedule=eddy_amd.TESTS_CUTOFF, mode=eddy_amd.MODE_BATCHED)
    print("case: noalloc capacity error, 1024 genes x 20000 samples", flush=True)
    with pytest.raises(RuntimeError, match="scratch array"):
        engine.perm_detail(np.arange(1024, dtype=np.int32), np.zeros((0, 2), np.int32), _labels(20000, 5000, 5),
                           prm2, bits=False, out5=False, noalloc=True)
    # the context and the scratch array still work afterwards
    engine.upload(wl.X, wl.classes)
    got, (g, p, lb, prm, bits) = _case(engine, wl, genes[:12], pairs[:0], lab, eddy_amd.TESTS_CUTOFF)
    _same(got, engine.perm_detail(g, p, lb, prm, bits=bits, out5=True, noalloc=False))
