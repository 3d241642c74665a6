"""CPU checks of libeddy's C ABI: the library loads, exports every function
include/eddy.h declares, replays the glibc rand() stream (main.cu:538-573)
without a GPU, and generates seeded synthetic workloads deterministically."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import eddy_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def header_functions():
    src = open(os.path.join(ROOT, "include", "eddy.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(eddy_[a-z0-9_]+)\s*\(", src)))


def test_exports_match_header():
    L = eddy_amd.lib()
    names = header_functions()
    assert set(names) == set(eddy_amd.EXPORTS)
    for n in names:
        assert hasattr(L, n), n


def host_ctx():
    L = eddy_amd.lib()
    h = C.c_void_p()
    assert L.eddy_ctx_create(-1, C.byref(h)) == 0
    return L, h


def test_stream_matches_glibc_rand():
    L, h = host_ctx()
    ref = np.load(os.path.join(GOLD, "glibc_rand_1e4.npy"))
    out = np.zeros(len(ref), np.int32)
    assert L.eddy_stream_draw(h, len(out), out.ctypes.data_as(C.c_void_p)) == 0
    assert np.array_equal(out, ref)
    # the live C library agrees with the committed fixture on this machine
    libc = C.CDLL(None)
    libc.srand(1)
    assert [libc.rand() for _ in range(100)] == list(ref[:100])
    assert L.eddy_stream_reset(h) == 0
    assert L.eddy_stream_draw(h, 5, out.ctypes.data_as(C.c_void_p)) == 0
    assert list(out[:5]) == list(ref[:5])
    # a host-only context refuses GPU work
    x = np.zeros((2, 4), np.uint8)
    cl = np.array([0, 0, 1, 1], np.uint8)
    assert L.eddy_upload_dataset(h, x.ctypes.data_as(C.c_void_p), 2, 4, cl.ctypes.data_as(C.c_void_p)) == \
        eddy_amd.EDDY_ERR_NODEV
    L.eddy_ctx_destroy(h)


def test_synth_deterministic_and_tertiles():
    a = eddy_amd.Workload(0)
    b = eddy_amd.Workload(0)
    assert np.array_equal(a.X, b.X)
    N = a.N
    counts = np.stack([(a.X == v).sum(1) for v in range(3)], 1)
    assert np.all(counts[:, 0] == N // 3)
    assert np.all(counts[:, 1] == 2 * N // 3 - N // 3)
    genes, pairs = a.sets[0]
    assert len(genes) == 40 and len(set(genes.tolist())) == 40 and len(pairs) == 20
    w2 = eddy_amd.Workload(2, nsets=50)
    sizes = [len(g) for g, _ in w2.sets]
    assert min(sizes) >= 10 and max(sizes) <= 300
    assert all(len(p) == int(0.3 * len(g)) for g, p in w2.sets)


def test_no_gpu_engine_fails_loudly():
    # Without a device the engine refuses to run rather than computing on the CPU
    L = eddy_amd.lib()
    h = C.c_void_p()
    rc = L.eddy_ctx_create(0, C.byref(h))
    if rc == 0:
        L.eddy_ctx_destroy(h)
        pytest.skip("a GPU is present")
    assert rc == eddy_amd.EDDY_ERR_NODEV
    with pytest.raises(RuntimeError):
        eddy_amd.Engine(0)


def test_stream_permutations_match_oracle():
    """The engine's permutation labels (main.cu:538-573) equal the oracle's,
    which calls the C library's rand(), over consecutive permutations."""
    import oracle_ref

    L, h = host_ctx()
    for N, n1 in ((200, 100), (601, 300), (1200, 300), (4000, 1000), (20000, 5000)):
        L.eddy_stream_reset(h)
        oracle_ref.reset_stream()
        lab_e = np.array([0] * n1 + [1] * (N - n1), np.uint8)
        lab_o = lab_e.astype(np.int32)
        for perm in range(12):
            assert L.eddy_stream_permutation(h, N, n1, perm, lab_e.ctypes.data_as(C.c_void_p)) == 0
            oracle_ref.lib().eo_perm_labels(N, n1, perm, lab_o.ctypes.data_as(C.c_void_p))
            assert np.array_equal(lab_e, lab_o.astype(np.uint8)), (N, perm)
            assert (lab_e == 0).sum() == n1
    L.eddy_ctx_destroy(h)


def test_stream_index_matches_stream():
    """The stream index (worker threads, block-wise scan) gives the same
    class-2 bits and leaves the stream at the same place as one permutation at
    a time, across many superblocks and for odd, even, small and maximal n."""
    L = eddy_amd.lib()
    for N, n1, count in ((600, 300, 3000), (601, 17, 200), (2046, 1023, 40), (64, 63, 500), (65, 1, 500),
                         (2, 1, 2000), (1200, 300, 50), (4000, 1000, 12), (20000, 7001, 3)):
        ha, hb = C.c_void_p(), C.c_void_p()
        assert L.eddy_ctx_create(-1, C.byref(ha)) == 0 and L.eddy_ctx_create(-1, C.byref(hb)) == 0
        words = (N + 31) // 32
        # start both a little way into the stream
        pre = np.zeros(777, np.int32)
        for h in (ha, hb):
            assert L.eddy_stream_draw(h, len(pre), pre.ctypes.data_as(C.c_void_p)) == 0
        got = np.zeros((count, words), np.uint32)
        assert L.eddy_stream_class2_batch(ha, N, n1, count, got.ctypes.data_as(C.c_void_p)) == 0
        lab = np.zeros(N, np.uint8)
        for k in range(count):
            assert L.eddy_stream_permutation(hb, N, n1, 1, lab.ctypes.data_as(C.c_void_p)) == 0
            bits = (got[k][np.arange(N) >> 5] >> (np.arange(N) & 31)) & 1
            assert np.array_equal(bits.astype(np.uint8), lab), (N, n1, k)
        # same position afterwards
        a, b = np.zeros(100, np.int32), np.zeros(100, np.int32)
        assert L.eddy_stream_draw(ha, 100, a.ctypes.data_as(C.c_void_p)) == 0
        assert L.eddy_stream_draw(hb, 100, b.ctypes.data_as(C.c_void_p)) == 0
        assert np.array_equal(a, b), (N, n1)
        L.eddy_ctx_destroy(ha)
        L.eddy_ctx_destroy(hb)


def test_mlp_training_order_is_a_sample_and_permutations():
    """Each epoch visits the same m distinct sampled rows, each exactly once, in
    a different order; the sample and the orders depend on the seed; with
    m = n every row is visited."""
    L = eddy_amd.lib()
    for n, m, epochs in ((614123, 131072, 2), (1000, 1000, 3), (5, 3, 4), (1, 1, 1), (70000, 65536, 1)):
        out = np.zeros(m * epochs, np.int32)
        assert L.eddy_mlp_order_host(1, n, m, epochs, out.ctypes.data_as(C.c_void_p)) == 0
        ep = out.reshape(epochs, m)
        sample = np.sort(ep[0])
        assert sample[0] >= 0 and sample[-1] < n and len(np.unique(sample)) == m
        for e in range(1, epochs):
            assert np.array_equal(np.sort(ep[e]), sample)
            if m > 8:
                assert not np.array_equal(ep[e], ep[0])
        if m == n:
            assert np.array_equal(sample, np.arange(n))
        other = np.zeros_like(out)
        assert L.eddy_mlp_order_host(2, n, m, epochs, other.ctypes.data_as(C.c_void_p)) == 0
        if m > 8:
            assert not np.array_equal(other, out)
    # the sample of 131072 of 614123 rows is spread evenly over the rows
    out = np.zeros(131072, np.int32)
    assert L.eddy_mlp_order_host(7, 614123, 131072, 1, out.ctypes.data_as(C.c_void_p)) == 0
    hist = np.bincount(out // 61413, minlength=10)[:10]
    assert hist.min() > 0.9 * 131072 / 10
    # refused arguments
    assert L.eddy_mlp_order_host(1, 10, 11, 1, out.ctypes.data_as(C.c_void_p)) != 0
    assert L.eddy_mlp_order_host(1, 10, 5, 65, out.ctypes.data_as(C.c_void_p)) != 0


def test_launch_config_is_validated_without_a_gpu():
    """eddy_set_launch_cfg accepts every documented value and rejects the rest
    (leaving the configuration as it was); eddy_autotune needs a device."""
    L, h = host_ctx()
    try:
        c = eddy_amd.LaunchCfg()
        assert L.eddy_get_launch_cfg(h, C.byref(c)) == 0
        default = c.as_dict()
        assert default == {"graph": 1, "tile_order": 0, "item_order": 0, "loo_waves": 4, "bdeu_groups": 0,
                           "bdeu_blocks": 2048}
        good = [dict(graph=0), dict(tile_order=1), dict(item_order=1), dict(loo_waves=3), dict(bdeu_groups=1),
                dict(bdeu_groups=2), dict(bdeu_groups=4), dict(bdeu_blocks=64), dict(bdeu_blocks=16384)]
        bad = [dict(graph=2), dict(tile_order=2), dict(item_order=-1), dict(loo_waves=2), dict(loo_waves=5),
               dict(bdeu_groups=3), dict(bdeu_groups=8), dict(bdeu_blocks=63), dict(bdeu_blocks=16385)]
        for kw in good + bad:
            x = eddy_amd.LaunchCfg(**default)
            for k, v in kw.items():
                setattr(x, k, v)
            rc = L.eddy_set_launch_cfg(h, C.byref(x))
            got = eddy_amd.LaunchCfg()
            assert L.eddy_get_launch_cfg(h, C.byref(got)) == 0
            if kw in good:
                assert rc == eddy_amd.EDDY_OK and got.as_dict() == dict(default, **kw), kw
            else:
                assert rc == eddy_amd.EDDY_ERR_ARG and got.as_dict() == default, kw
            assert L.eddy_set_launch_cfg(h, C.byref(eddy_amd.LaunchCfg(**default))) == 0
        prm = eddy_amd.Params(3, 10, 1, 0.05, 0.05, 0.05, eddy_amd.MODE_BATCHED, eddy_amd.TESTS_CUTOFF, 0)
        rep = eddy_amd.TuneReport()
        assert L.eddy_autotune(h, None, 0, C.byref(prm), None, C.byref(c), C.byref(rep)) == eddy_amd.EDDY_ERR_NODEV
    finally:
        L.eddy_ctx_destroy(h)


def test_set_batch_marshals_like_a_list():
    """A SetBatch holds the same C descriptors a call builds from the list of
    (genes, pairs) itself, and keeps the arrays they point into alive."""
    sets = [(np.arange(5) * 3, np.array([[0, 1], [2, 4]])), ([7, 8, 9], []), (np.array([1, 2]), [[1, 0]])]
    b = eddy_amd.SetBatch(sets)
    arr, keep = eddy_amd.Engine._descs(sets)
    assert len(b) == 3 and [len(g) for g, _ in b] == [5, 3, 2]
    for i in range(3):
        assert b.arr[i].n_genes == arr[i].n_genes and b.arr[i].n_prior == arr[i].n_prior
        assert [b.arr[i].genes[k] for k in range(arr[i].n_genes)] == [arr[i].genes[k] for k in range(arr[i].n_genes)]
        assert [b.arr[i].prior[k] for k in range(2 * arr[i].n_prior)] == \
               [arr[i].prior[k] for k in range(2 * arr[i].n_prior)]
    assert eddy_amd.Engine._descs(b)[0] is b.arr
