"""eddy_amd - Python access to libeddy (the gfx950 EDDY engine) through its C ABI.

The engine itself is C++/HIP (``csrc/``) behind ``include/eddy.h``; this module
only loads ``libeddy.so`` with ctypes, mirrors the header's structs and wraps
the calls used by the tests and ``bench.py``.  It never computes anything
itself: when the library is missing, or no GPU is present, the calls raise.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libeddy.so")

EDDY_OK, EDDY_ERR_ARG, EDDY_ERR_HIP, EDDY_ERR_CAPACITY, EDDY_ERR_STATE, EDDY_ERR_NOMEM, EDDY_ERR_NODEV = range(7)
SET_OK, SET_FAILED_NAN, SET_DOMAIN = 0, 1, 2
MODE_PLAIN, MODE_BATCHED = 0, 1
TESTS_ALL, TESTS_CUTOFF = 0, 1

# Every symbol include/eddy.h declares (checked by the CPU test suite).
EXPORTS = [
    "eddy_ctx_create", "eddy_ctx_destroy", "eddy_last_error", "eddy_version",
    "eddy_upload_dataset", "eddy_stream_reset", "eddy_stream_draw", "eddy_run_sets",
    "eddy_get_js", "eddy_get_perm0_networks", "eddy_get_edgelist", "eddy_run_perm_detail",
    "eddy_get_timing", "eddy_synth_config", "eddy_synth_dataset", "eddy_synth_sets",
    "eddy_edge_decision_host", "eddy_stream_permutation",
    "eddy_mlp_init", "eddy_mlp_get_params", "eddy_mlp_set_params", "eddy_mlp_set_margin", "eddy_mlp_forward",
    "eddy_mlp_gradient", "eddy_mlp_train", "eddy_learn_collect", "eddy_run_learned",
    "eddy_edge_loo_mask_host", "eddy_chi2_host", "eddy_screen32_host", "eddy_stream_class2_batch",
    "eddy_mlp_order_host", "eddy_run_repermute", "eddy_get_launch_cfg", "eddy_set_launch_cfg", "eddy_autotune",
    "eddy_loo_all_certain_host", "eddy_run_perm_detail_noalloc",
]
MLP_FEATURES = 12
MLP_NPARAMS = 5057


class Params(C.Structure):
    _fields_ = [("mp", C.c_int), ("perms", C.c_int), ("p_adjust", C.c_int),
                ("alpha_edge", C.c_double), ("alpha_edge_prior", C.c_double),
                ("alpha_ddn", C.c_double), ("mode", C.c_int), ("schedule", C.c_int),
                ("learned", C.c_int)]


class SetDesc(C.Structure):
    _fields_ = [("n_genes", C.c_int), ("genes", C.POINTER(C.c_int)),
                ("n_prior", C.c_int), ("prior", C.POINTER(C.c_int))]


class SetResult(C.Structure):
    _fields_ = [("status", C.c_int), ("perms_run", C.c_int), ("U0", C.c_int),
                ("js0", C.c_double), ("pDDN", C.c_double), ("alpha", C.c_double),
                ("beta", C.c_double), ("tests_ref", C.c_longlong), ("tests_eval", C.c_longlong)]


class PermDetail(C.Structure):
    _fields_ = [("bits", C.c_void_p), ("uniq_ids", C.c_void_p), ("uni_nodes", C.c_void_p),
                ("uni_epn", C.c_void_p), ("uni_edges", C.c_void_p), ("out5", C.c_void_p),
                ("lval", C.c_void_p), ("U", C.c_int), ("edSum", C.c_int), ("js", C.c_double)]


KERNELS = ["count_test", "loo", "rows", "score", "pvalue", "gather"]
# the batched launch sequence's kernels one by one (eddy_timing.split_ms)
SPLIT = ["k_count_test", "k_loo", "k_rows", "k_score", "k_bdeu", "k_score_tail", "k_gather_class1"]


class Timing(C.Structure):
    _fields_ = [("total_ms", C.c_double), ("kernel_ms", C.c_double * 6), ("launches", C.c_longlong * 6),
                ("items", C.c_longlong), ("pairs", C.c_longlong), ("tests", C.c_longlong),
                ("host_stream_ms", C.c_double), ("sched", C.c_longlong * 4), ("stats", C.c_longlong * 6),
                ("phase1_ms", C.c_double), ("phase_ms", C.c_double * 16), ("split_ms", C.c_double * 7),
                ("split_launches", C.c_longlong * 7)]


class LaunchCfg(C.Structure):
    """Launch shapes (include/eddy.h eddy_launch_cfg); results never depend on them."""
    _fields_ = [("graph", C.c_int), ("tile_order", C.c_int), ("item_order", C.c_int), ("loo_waves", C.c_int),
                ("bdeu_groups", C.c_int), ("bdeu_blocks", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class TuneReport(C.Structure):
    _fields_ = [("from_cache", C.c_int), ("trials", C.c_int), ("tune_ms", C.c_double), ("ms_default", C.c_double),
                ("ms_best", C.c_double), ("items", C.c_longlong), ("pairs", C.c_longlong),
                ("kernel_ms", C.c_double * 6)]


class LearnCfg(C.Structure):
    _fields_ = [("train_sets", C.c_int), ("epochs", C.c_int), ("batch", C.c_int), ("lr", C.c_float),
                ("margin", C.c_float), ("seed", C.c_uint64), ("max_rows", C.c_longlong), ("label_perms", C.c_int)]


class LearnReport(C.Structure):
    _fields_ = [("collect_ms", C.c_double), ("train_ms", C.c_double), ("apply_ms", C.c_double),
                ("train_rows", C.c_longlong), ("loss_initial", C.c_float), ("loss_final", C.c_float),
                ("tests_collect", C.c_longlong), ("tests_apply", C.c_longlong), ("trained", C.c_int),
                ("first_windowed_set", C.c_int)]


def learn_cfg(train_sets=100, epochs=1, batch=16384, lr=3e-3, margin=3.0, seed=1, max_rows=65536, label_perms=10):
    return LearnCfg(train_sets, epochs, batch, lr, margin, seed, max_rows, label_perms)


class SynthCfg(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("genes", C.c_int), ("nA", C.c_int), ("nB", C.c_int),
                ("rewire", C.c_double), ("nsets", C.c_int), ("gmin", C.c_int), ("gmax", C.c_int),
                ("prior_fixed", C.c_int), ("prior_per_gene", C.c_double), ("mp", C.c_int),
                ("perms", C.c_int), ("prior_weight", C.c_double)]


_lib = None


def lib():
    """Load libeddy.so (built in-tree by __graft_entry__.build())."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libeddy.so is not built: run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        L.eddy_last_error.restype = C.c_char_p
        L.eddy_version.restype = C.c_char_p
        L.eddy_ctx_create.argtypes = [C.c_int, C.POINTER(C.c_void_p)]
        for name in ("eddy_ctx_destroy", "eddy_stream_reset"):
            getattr(L, name).argtypes = [C.c_void_p]
        L.eddy_last_error.argtypes = [C.c_void_p]
        L.eddy_upload_dataset.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p]
        L.eddy_stream_draw.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.eddy_run_sets.argtypes = [C.c_void_p, C.POINTER(SetDesc), C.c_int, C.POINTER(Params),
                                    C.POINTER(SetResult)]
        L.eddy_get_js.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int]
        L.eddy_get_perm0_networks.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int,
                                              C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
        L.eddy_get_edgelist.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
        L.eddy_run_perm_detail.argtypes = [C.c_void_p, C.POINTER(SetDesc), C.c_void_p, C.POINTER(Params),
                                           C.POINTER(PermDetail)]
        L.eddy_run_perm_detail_noalloc.argtypes = [C.c_void_p, C.POINTER(SetDesc), C.c_void_p, C.POINTER(Params),
                                                   C.POINTER(PermDetail)]
        L.eddy_get_timing.argtypes = [C.c_void_p, C.POINTER(Timing)]
        L.eddy_get_launch_cfg.argtypes = [C.c_void_p, C.POINTER(LaunchCfg)]
        L.eddy_set_launch_cfg.argtypes = [C.c_void_p, C.POINTER(LaunchCfg)]
        L.eddy_autotune.argtypes = [C.c_void_p, C.POINTER(SetDesc), C.c_int, C.POINTER(Params), C.c_char_p,
                                    C.POINTER(LaunchCfg), C.POINTER(TuneReport)]
        L.eddy_synth_config.argtypes = [C.c_int, C.POINTER(SynthCfg)]
        L.eddy_synth_dataset.argtypes = [C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_double, C.c_void_p]
        L.eddy_synth_sets.argtypes = [C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                      C.c_void_p, C.c_void_p, C.c_longlong, C.c_void_p, C.c_void_p,
                                      C.c_longlong]
        L.eddy_stream_permutation.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p]
        L.eddy_stream_class2_batch.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p]
        L.eddy_mlp_order_host.argtypes = [C.c_uint64, C.c_longlong, C.c_longlong, C.c_int, C.c_void_p]
        L.eddy_edge_decision_host.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(Params)]
        L.eddy_edge_loo_mask_host.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.POINTER(Params)]
        L.eddy_loo_all_certain_host.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(Params)]
        L.eddy_screen32_host.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(Params)]
        L.eddy_chi2_host.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double)]
        L.eddy_mlp_init.argtypes = [C.c_void_p, C.c_uint64]
        L.eddy_mlp_get_params.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
        L.eddy_mlp_set_params.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
        L.eddy_mlp_set_margin.argtypes = [C.c_void_p, C.c_float]
        L.eddy_mlp_forward.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]
        L.eddy_mlp_gradient.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
        L.eddy_mlp_train.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_longlong, C.c_int, C.c_int,
                                     C.c_float, C.c_uint64, C.c_void_p]
        L.eddy_learn_collect.argtypes = [C.c_void_p, C.POINTER(SetDesc), C.c_int, C.POINTER(Params), C.c_void_p,
                                         C.c_void_p, C.c_longlong, C.POINTER(C.c_longlong)]
        L.eddy_run_learned.argtypes = [C.c_void_p, C.POINTER(SetDesc), C.c_int, C.POINTER(Params),
                                       C.POINTER(LearnCfg), C.POINTER(SetResult), C.POINTER(LearnReport)]
        L.eddy_run_repermute.argtypes = [C.c_void_p, C.POINTER(SetDesc), C.c_int, C.POINTER(Params),
                                         C.POINTER(SetResult), C.c_int, C.c_int, C.c_void_p, C.POINTER(SetResult),
                                         C.POINTER(C.c_int)]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


# ---------------------------------------------------------------- synthetic data
class Workload:
    """A seeded synthetic workload in one of the benchmark shapes."""

    def __init__(self, cfg_index, nsets=None):
        L = lib()
        cfg = SynthCfg()
        if L.eddy_synth_config(cfg_index, C.byref(cfg)) != EDDY_OK:
            raise ValueError("unknown config %d" % cfg_index)
        self.cfg = cfg
        self.N = cfg.nA + cfg.nB
        self.X = np.zeros((cfg.genes, self.N), np.uint8)
        rc = L.eddy_synth_dataset(cfg.seed, cfg.genes, cfg.nA, cfg.nB, cfg.rewire, _ptr(self.X))
        if rc != EDDY_OK:
            raise RuntimeError("synth dataset failed")
        self.classes = np.array([0] * cfg.nA + [1] * cfg.nB, np.uint8)
        n = cfg.nsets if nsets is None else min(nsets, cfg.nsets)
        gcap = cfg.nsets * cfg.gmax
        sizes = np.zeros(cfg.nsets, np.int32)
        genes = np.zeros(gcap, np.int32)
        pcount = np.zeros(cfg.nsets, np.int32)
        pmax = cfg.gmax * (cfg.gmax - 1) // 2
        pcap = cfg.nsets * (cfg.prior_fixed if cfg.prior_fixed >= 0 else int(cfg.prior_per_gene * cfg.gmax) + 1)
        pcap = max(1, min(pcap, cfg.nsets * pmax))
        pairs = np.zeros(2 * pcap, np.int32)
        rc = L.eddy_synth_sets(cfg.seed, cfg.genes, cfg.nsets, cfg.gmin, cfg.gmax, cfg.prior_fixed,
                               cfg.prior_per_gene, _ptr(sizes), _ptr(genes), gcap, _ptr(pcount), _ptr(pairs), pcap)
        if rc != EDDY_OK:
            raise RuntimeError("synth sets failed")
        self.sets = []
        go = po = 0
        for s in range(cfg.nsets):
            G = int(sizes[s])
            P = int(pcount[s])
            if s < n:
                self.sets.append((genes[go:go + G].copy(), pairs[2 * po:2 * (po + P)].reshape(P, 2).copy()))
            go += G
            po += P

    @property
    def batch(self):
        """The sets as a SetBatch (marshalled on first use)."""
        if getattr(self, "_batch", None) is None:
            self._batch = SetBatch(self.sets)
        return self._batch

    def params(self, perms=None, mode=MODE_BATCHED, schedule=TESTS_CUTOFF, learned=0, p_adjust=1,
               alpha_edge=0.05, prior_weight=None, lam=2.0, alpha_ddn=0.05, mp=None):
        pw = self.cfg.prior_weight if prior_weight is None else prior_weight
        aep = 1 - (1 - alpha_edge) * pw ** (1 / lam)
        return Params(mp or self.cfg.mp, perms or self.cfg.perms, p_adjust, alpha_edge, aep, alpha_ddn,
                      mode, schedule, learned)

    def prior_matrix(self, s):
        genes, pairs = self.sets[s]
        G = len(genes)
        P = np.zeros((G, G), np.int32)
        for a, b in pairs:
            P[a, b] = P[b, a] = 1
        return P


# ---------------------------------------------------------------- engine
class SetBatch:
    """Gene sets marshalled once into the C descriptors (eddy_set_desc) and
    the arrays they point into, for repeated calls on the same sets; the
    engine's calls take it wherever they take a list of (genes, pairs)."""

    def __init__(self, sets):
        self.sets = list(sets)
        self.arr, self.keep = Engine._descs(self.sets)

    def __len__(self):
        return len(self.sets)

    def __iter__(self):
        return iter(self.sets)

    def __getitem__(self, i):
        return self.sets[i]


class Engine:
    def __init__(self, device=0):
        L = lib()
        h = C.c_void_p()
        rc = L.eddy_ctx_create(device, C.byref(h))
        if rc != EDDY_OK:
            raise RuntimeError("eddy_ctx_create failed (%d): no usable HIP device" % rc)
        self.h = h
        self.L = L

    def close(self):
        if self.h:
            self.L.eddy_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc != EDDY_OK:
            raise RuntimeError("%s failed (%d): %s" % (what, rc, self.L.eddy_last_error(self.h).decode()))

    def upload(self, X, classes):
        X = np.ascontiguousarray(X, np.uint8)
        classes = np.ascontiguousarray(classes, np.uint8)
        self._chk(self.L.eddy_upload_dataset(self.h, _ptr(X), X.shape[0], X.shape[1], _ptr(classes)), "upload")
        self.N = X.shape[1]
        self.n1 = int((classes == 0).sum())

    def reset_stream(self):
        self._chk(self.L.eddy_stream_reset(self.h), "stream_reset")

    def draw(self, n):
        out = np.zeros(n, np.int32)
        self._chk(self.L.eddy_stream_draw(self.h, n, _ptr(out)), "stream_draw")
        return out

    @staticmethod
    def _descs(sets):
        if isinstance(sets, SetBatch):
            return sets.arr, sets.keep
        keep = []
        arr = (SetDesc * max(1, len(sets)))()
        for i, (genes, pairs) in enumerate(sets):
            g = np.ascontiguousarray(genes, np.int32)
            p = np.ascontiguousarray(np.asarray(pairs, np.int32).reshape(-1))
            keep += [g, p]
            arr[i].n_genes = len(g)
            arr[i].genes = g.ctypes.data_as(C.POINTER(C.c_int))
            arr[i].n_prior = len(p) // 2
            arr[i].prior = p.ctypes.data_as(C.POINTER(C.c_int))
        return arr, keep

    def run_sets(self, sets, params):
        arr, keep = self._descs(sets)
        res = (SetResult * max(1, len(sets)))()
        self._chk(self.L.eddy_run_sets(self.h, arr, len(sets), C.byref(params), res), "run_sets")
        self.last_perms = params.perms
        self.last_G = [len(g) for g, _ in sets]
        return [res[i] for i in range(len(sets))]

    def run_learned(self, sets, params, cfg):
        """The learned job (eddy_run_learned): returns (results, report)."""
        arr, keep = self._descs(sets)
        res = (SetResult * max(1, len(sets)))()
        rep = LearnReport()
        self._chk(self.L.eddy_run_learned(self.h, arr, len(sets), C.byref(params), C.byref(cfg), res,
                                          C.byref(rep)), "run_learned")
        self.last_perms = params.perms
        self.last_G = [len(g) for g, _ in sets]
        return [res[i] for i in range(len(sets))], rep

    def repermute(self, sets, first, params, top, perms):
        """eddy_run_repermute: the `top` most significant sets of `first` run
        again with `perms` permutations on a fresh stream.  Returns
        (chosen set indices, their results)."""
        arr, keep = self._descs(sets)
        fr = (SetResult * max(1, len(sets)))(*first) if len(first) else (SetResult * 1)()
        chosen = np.zeros(max(1, top), np.int32)
        res = (SetResult * max(1, top))()
        n = C.c_int()
        self._chk(self.L.eddy_run_repermute(self.h, arr, len(sets), C.byref(params), fr, top, perms, _ptr(chosen),
                                            res, C.byref(n)), "run_repermute")
        self.last_perms = perms
        self.last_G = [len(sets[i][0]) for i in chosen[:n.value]]
        return chosen[:n.value].tolist(), [res[i] for i in range(n.value)]

    # ---- learned proposal network
    def mlp_init(self, seed):
        self._chk(self.L.eddy_mlp_init(self.h, seed), "mlp_init")

    def mlp_params(self):
        out = np.zeros(MLP_NPARAMS, np.float32)
        self._chk(self.L.eddy_mlp_get_params(self.h, _ptr(out), MLP_NPARAMS), "mlp_get_params")
        return out

    def mlp_set_params(self, P):
        P = np.ascontiguousarray(P, np.float32)
        self._chk(self.L.eddy_mlp_set_params(self.h, _ptr(P), len(P)), "mlp_set_params")

    def mlp_set_margin(self, kappa):
        self._chk(self.L.eddy_mlp_set_margin(self.h, kappa), "mlp_set_margin")

    def mlp_forward(self, X):
        X = np.ascontiguousarray(X, np.float32)
        out = np.zeros(len(X), np.float32)
        self._chk(self.L.eddy_mlp_forward(self.h, _ptr(X), len(X), _ptr(out)), "mlp_forward")
        return out

    def mlp_gradient(self, X, Y):
        X = np.ascontiguousarray(X, np.float32)
        Y = np.ascontiguousarray(Y, np.float32)
        g = np.zeros(MLP_NPARAMS, np.float32)
        loss = np.zeros(1, np.float32)
        self._chk(self.L.eddy_mlp_gradient(self.h, _ptr(X), _ptr(Y), len(X), _ptr(g), _ptr(loss)), "mlp_gradient")
        return g, float(loss[0])

    def mlp_train(self, X, Y, epochs, batch, lr, seed=1):
        X = np.ascontiguousarray(X, np.float32)
        Y = np.ascontiguousarray(Y, np.float32)
        loss = np.zeros(1, np.float32)
        self._chk(self.L.eddy_mlp_train(self.h, _ptr(X), _ptr(Y), len(X), epochs, batch, lr, seed, _ptr(loss)),
                  "mlp_train")
        return float(loss[0])

    def learn_collect(self, sets, params):
        """Features and soft labels of every pair of `sets`, from the plain method."""
        arr, keep = self._descs(sets)
        n = sum(len(g) * (len(g) - 1) // 2 for g, _ in sets)
        X = np.zeros((max(1, n), MLP_FEATURES), np.float32)
        Y = np.zeros(max(1, n), np.float32)
        got = C.c_longlong()
        self._chk(self.L.eddy_learn_collect(self.h, arr, len(sets), C.byref(params), _ptr(X), _ptr(Y), n,
                                            C.byref(got)), "learn_collect")
        self.last_perms = params.perms
        self.last_G = [len(g) for g, _ in sets]
        return X[:got.value], Y[:got.value]

    def js(self, s):
        out = np.zeros(self.last_perms, np.float64)
        self._chk(self.L.eddy_get_js(self.h, s, _ptr(out), len(out)), "get_js")
        return out

    def perm0_networks(self, s, U, mp):
        G = self.last_G[s]
        ids = np.zeros(U, np.int32)
        nodes = np.zeros(max(1, U * G), np.int32)
        epn = np.zeros(U + 1, np.int32)
        edges = np.zeros(max(1, U * G * min(mp, max(1, G - 1))), np.int32)
        lval = np.zeros(2 * U, np.float64)
        ne = C.c_int()
        self._chk(self.L.eddy_get_perm0_networks(self.h, s, _ptr(ids), _ptr(nodes), len(nodes), _ptr(epn),
                                                 _ptr(edges), len(edges), _ptr(lval), C.byref(ne)), "perm0_networks")
        return ids, nodes[:U * G].reshape(U, G) if G else nodes[:0], epn, edges[:ne.value], lval

    def edgelist(self, s):
        G = self.last_G[s]
        nodes = np.zeros(max(1, 2 * G), np.int32)
        epn = np.zeros(3, np.int32)
        edges = np.zeros(max(1, G * G), np.int32)
        self._chk(self.L.eddy_get_edgelist(self.h, s, _ptr(nodes), _ptr(epn), _ptr(edges), len(edges)), "edgelist")
        return nodes[:2 * G], epn, edges[:epn[2]]

    def perm_detail(self, genes, pairs, labels, params, bits=True, out5=True, noalloc=False):
        G = len(genes)
        S = self.N + 2
        c = G * (G - 1) // 2
        d = PermDetail()
        bufs = {
            "bits": np.zeros(max(1, S * c), np.uint8) if bits else None,
            "uniq_ids": np.zeros(S, np.int32),
            "uni_nodes": np.zeros(S * G, np.int32),
            "uni_epn": np.zeros(S + 1, np.int32),
            "uni_edges": np.zeros(max(1, S * G * min(params.mp, max(1, G - 1))), np.int32),
            "out5": np.zeros(2 * S * G, np.float64) if out5 else None,
            "lval": np.zeros(2 * S, np.float64),
        }
        for k, v in bufs.items():
            setattr(d, k, _ptr(v) if v is not None else None)
        arr, keep = self._descs([(genes, pairs)])
        lab = np.ascontiguousarray(labels, np.uint8)
        fn = self.L.eddy_run_perm_detail_noalloc if noalloc else self.L.eddy_run_perm_detail
        self._chk(fn(self.h, arr, _ptr(lab), C.byref(params), C.byref(d)), "perm_detail")
        U = d.U
        return {
            "U": U, "js": d.js, "edSum": d.edSum,
            "bits": bufs["bits"].reshape(S, c) if bits else None,
            "uniq_ids": bufs["uniq_ids"][:U], "uni_nodes": bufs["uni_nodes"][:U * G].reshape(U, G),
            "uni_epn": bufs["uni_epn"][:U + 1], "uni_edges": bufs["uni_edges"][:d.edSum],
            "out5": bufs["out5"][:2 * U * G].reshape(2, U, G) if out5 else None,
            "lval": bufs["lval"][:2 * U],
        }

    def timing(self):
        t = Timing()
        self._chk(self.L.eddy_get_timing(self.h, C.byref(t)), "timing")
        return t

    # ---- launch shapes and the autotuner
    def launch_cfg(self):
        c = LaunchCfg()
        self._chk(self.L.eddy_get_launch_cfg(self.h, C.byref(c)), "get_launch_cfg")
        return c.as_dict()

    def set_launch_cfg(self, **kw):
        """Changes the named fields of the launch configuration."""
        c = LaunchCfg()
        self._chk(self.L.eddy_get_launch_cfg(self.h, C.byref(c)), "get_launch_cfg")
        for k, v in kw.items():
            if k not in dict(LaunchCfg._fields_):
                raise KeyError(k)
            setattr(c, k, int(v))
        self._chk(self.L.eddy_set_launch_cfg(self.h, C.byref(c)), "set_launch_cfg")

    def autotune(self, sets, params, cache=None):
        """eddy_autotune: returns (chosen configuration, report dict)."""
        arr, keep = self._descs(sets)
        c, rep = LaunchCfg(), TuneReport()
        path = None if cache is None else os.fsencode(cache)
        self._chk(self.L.eddy_autotune(self.h, arr, len(sets), C.byref(params), path, C.byref(c), C.byref(rep)),
                  "autotune")
        out = {k: getattr(rep, k) for k, _ in TuneReport._fields_}
        out["kernel_ms"] = dict(zip(["count", "loo", "rows", "score", "bdeu", "tail"], list(rep.kernel_ms)))
        return c.as_dict(), out
