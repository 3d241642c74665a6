"""GPU parity: the HIP engine against the CPU oracle on the same inputs.

Integer outputs (edge bits, unique networks, CSR) must be identical; FP64
outputs (BDeu per node, likelihoods, JS) are compared bit for bit as well,
since both sides round through the same FP64 routines (tolerance 1e-12
relative is the stated bar; the assertions use exact equality, which implies it).
"""
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


def _compare_perm(engine, wl, s, labels, schedule, check_bits):
    genes, pairs = wl.sets[s]
    prm = wl.params(schedule=schedule, mode=eddy_amd.MODE_PLAIN)
    print("case: set", s, "G", len(genes), "schedule", schedule, flush=True)
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
