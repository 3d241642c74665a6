"""Launch shapes, graph launches and the autotuner on the GPU.

Every launch configuration must give bit-identical results: the graph launch
(one hipGraph per slot, updated in place per chunk) against one-by-one
launches, the XCD-grouped tile order, both item orders, both loo kernel
instances, every BDeu lane-group count and persistent grid.  Runs with many
chunks (plain mode: one chunk per set and permutation) exercise the graph's
in-place updates.  A 550/550-sample dataset (9 mask words per class) checks
the BDeu lane groups against the oracle where the group count used to be
derived as 3.
"""
import os

import numpy as np
import pytest

import eddy_amd
import oracle_ref

pytestmark = pytest.mark.gpu

SHAPES = [
    dict(graph=1),
    dict(tile_order=1),
    dict(item_order=1),
    dict(loo_waves=3),
    dict(bdeu_groups=1),
    dict(bdeu_groups=2),
    dict(bdeu_blocks=64),
    dict(graph=1, tile_order=1, item_order=1, loo_waves=3, bdeu_groups=4, bdeu_blocks=4096),
]


@pytest.fixture
def shapes(engine):
    keep = engine.launch_cfg()
    yield keep
    engine.set_launch_cfg(**keep)


def _run(engine, wl, prm, probe, learned=None):
    engine.reset_stream()
    if learned is None:
        res = engine.run_sets(wl.sets, prm)
    else:
        res, _ = engine.run_learned(wl.sets, prm, learned)
    js = [engine.js(s).copy() for s in range(len(wl.sets))]
    nets = [engine.perm0_networks(s, res[s].U0, prm.mp) if res[s].U0 > 0 else None for s in probe]
    return res, js, nets


def _same(a, b, what):
    (ra, ja, na), (rb, jb, nb) = a, b
    for s in range(len(ra)):
        x, y = ra[s], rb[s]
        assert (x.status, x.perms_run, x.U0, x.tests_ref) == (y.status, y.perms_run, y.U0, y.tests_ref), (what, s)
        for f in ("js0", "pDDN", "alpha", "beta"):
            u, v = getattr(x, f), getattr(y, f)
            assert (np.isnan(u) and np.isnan(v)) or np.float64(u).view(np.uint64) == np.float64(v).view(np.uint64), \
                (what, s, f)
        assert np.array_equal(ja[s].view(np.uint64), jb[s].view(np.uint64)), (what, s)
    for p, q in zip(na, nb):
        assert (p is None) == (q is None), what
        if p is not None:
            for x, y in zip(p, q):
                assert np.array_equal(np.asarray(x).view(np.uint8), np.asarray(y).view(np.uint8)), what


def test_launch_config_validation(engine, shapes):
    for bad in (dict(graph=2), dict(tile_order=-1), dict(loo_waves=5), dict(bdeu_groups=3), dict(bdeu_blocks=0),
                dict(bdeu_blocks=20000), dict(item_order=2)):
        with pytest.raises(RuntimeError):
            engine.set_launch_cfg(**bad)
    assert engine.launch_cfg() == shapes


def test_batched_results_do_not_depend_on_launch_shapes(engine, shapes):
    wl = eddy_amd.Workload(3, nsets=40)
    engine.upload(wl.X, wl.classes)
    prm = wl.params(perms=12, mode=eddy_amd.MODE_BATCHED, schedule=eddy_amd.TESTS_CUTOFF)
    probe = (0, 17, 39)
    engine.set_launch_cfg(**dict(shapes, graph=0))
    print("case: batched, 40 sets of the 5000 x 600 workload at r=12, one-by-one launches", flush=True)
    ref = _run(engine, wl, prm, probe)
    for sh in SHAPES:
        cfg = dict(shapes, graph=0)
        cfg.update(sh)
        engine.set_launch_cfg(**cfg)
        print("case: batched,", sh, flush=True)
        _same(ref, _run(engine, wl, prm, probe), sh)
    # a second job on the same graphs (their nodes updated in place again)
    engine.set_launch_cfg(**dict(shapes, graph=1))
    print("case: batched, graph, second job", flush=True)
    _same(ref, _run(engine, wl, prm, probe), "graph again")


def test_learned_job_and_plain_mode_under_graphs(engine, shapes):
    wl = eddy_amd.Workload(3, nsets=24)
    engine.upload(wl.X, wl.classes)
    prm = wl.params(perms=16, mode=eddy_amd.MODE_BATCHED, schedule=eddy_amd.TESTS_CUTOFF)
    cfg = eddy_amd.learn_cfg(train_sets=8, epochs=2, batch=2048, lr=3e-3, margin=1.0, seed=1, max_rows=0,
                             label_perms=8)
    engine.set_launch_cfg(**dict(shapes, graph=0))
    print("case: learned job, 24 sets at r=16, one-by-one launches", flush=True)
    ref = _run(engine, wl, prm, (0, 23), learned=cfg)
    engine.set_launch_cfg(**dict(shapes, graph=1))
    print("case: learned job, graph", flush=True)
    _same(ref, _run(engine, wl, prm, (0, 23), learned=cfg), "learned graph")
    # plain mode, every test: one chunk (graph update) per set and permutation
    small = eddy_amd.Workload(3, nsets=4)
    prm = small.params(perms=6, mode=eddy_amd.MODE_PLAIN, schedule=eddy_amd.TESTS_ALL)
    engine.set_launch_cfg(**dict(shapes, graph=0))
    print("case: plain mode, all tests, 4 sets at r=6, one-by-one launches", flush=True)
    ref = _run(engine, small, prm, (0, 3))
    engine.set_launch_cfg(**dict(shapes, graph=1))
    print("case: plain mode, all tests, graph", flush=True)
    _same(ref, _run(engine, small, prm, (0, 3)), "plain graph")


def _wide_workload():
    """The 200 x 200 workload's genes and set on a 550/550-sample dataset (9 words per class mask)."""
    wl = eddy_amd.Workload(0)
    L = eddy_amd.lib()
    nA = nB = 550
    X = np.zeros((wl.cfg.genes, nA + nB), np.uint8)
    assert L.eddy_synth_dataset(wl.cfg.seed + 1, wl.cfg.genes, nA, nB, wl.cfg.rewire,
                                X.ctypes.data_as(eddy_amd.C.c_void_p)) == eddy_amd.EDDY_OK
    wl.X = X
    wl.N = nA + nB
    wl.classes = np.array([0] * nA + [1] * nB, np.uint8)
    wl.cfg.nA, wl.cfg.nB = nA, nB
    return wl


def test_bdeu_lane_groups_match_oracle_on_wide_classes(engine, shapes):
    wl = _wide_workload()
    engine.upload(wl.X, wl.classes)
    genes, pairs = wl.sets[0]
    prm = wl.params(mode=eddy_amd.MODE_PLAIN, schedule=eddy_amd.TESTS_CUTOFF)
    rng = np.random.default_rng(5)
    lab = np.ones(wl.N, np.uint8)
    lab[rng.permutation(wl.N)[:550]] = 0
    data_ref = wl.X.astype(np.int32) - 1
    t1, t2, n1, n2 = oracle_ref.class_split(data_ref, genes, lab)
    print("case: oracle, the 40-gene set on 550/550 samples", flush=True)
    o = oracle_ref.run_perm(t1, t2, n1, n2, wl.prior_matrix(0), oracle_ref.params(mp=prm.mp), want_bits=False)
    for groups in (0, 1, 2, 4):
        engine.set_launch_cfg(**dict(shapes, bdeu_groups=groups))
        print("case: 550/550 samples, BDeu lane groups", groups, flush=True)
        g = engine.perm_detail(genes, pairs, lab, prm, bits=False, out5=True)
        assert g["U"] == o["U"]
        assert np.array_equal(g["uni_edges"], o["uni_edges"])
        assert np.array_equal(g["out5"].view(np.uint64), o["out5"].view(np.uint64)), groups
        assert np.array_equal(g["lval"].view(np.uint64), o["lval"].view(np.uint64)), groups


def test_autotune_picks_a_valid_shape_and_caches_it(engine, shapes, tmp_path):
    wl = eddy_amd.Workload(3, nsets=60)
    engine.upload(wl.X, wl.classes)
    prm = wl.params(perms=10, mode=eddy_amd.MODE_BATCHED, schedule=eddy_amd.TESTS_CUTOFF)
    engine.reset_stream()
    ref = _run(engine, wl, prm, (0, 59))
    cache = str(tmp_path / "tune.txt")
    print("case: autotune, 60 sets of the 5000 x 600 workload", flush=True)
    c1, r1 = engine.autotune(wl.sets, prm, cache=cache)
    assert r1["from_cache"] == 0 and r1["trials"] >= 10 and r1["pairs"] > 0
    assert r1["ms_best"] <= r1["ms_default"] * 1.15
    assert engine.launch_cfg() == c1
    assert os.path.exists(cache) and len(open(cache).read().splitlines()) == 1
    c2, r2 = engine.autotune(wl.sets, prm, cache=cache)
    assert r2["from_cache"] == 1 and r2["trials"] == 0 and c2 == c1
    # the stream and the results are untouched by calibration
    print("case: run with the tuned shapes", flush=True)
    _same(ref, _run(engine, wl, prm, (0, 59)), c1)
