"""ctypes access to the CPU oracle (oracle/build/liboracle.so) for the tests.

The oracle is the checker: a plain-C restatement of the reference's
permutation loop.  Only the test suite and bench.py's cpu_baseline leg use it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB = os.path.join(ORACLE_DIR, "build", "liboracle.so")
CLI = os.path.join(ORACLE_DIR, "build", "eddy_oracle")


class EoParams(C.Structure):
    _fields_ = [("mp", C.c_int), ("perms", C.c_int), ("pAdjust", C.c_int), ("alphaEdge", C.c_double),
                ("alphaEdgePrior", C.c_double), ("alphaDDN", C.c_double), ("threads", C.c_int)]


class EoPermOut(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("edgebits", "ppn", "stf", "nodes", "edges", "edgesPN", "uniqueN",
                                          "uni_ids", "uni_nodes", "uni_edges", "uni_epn", "out5", "lval")] + \
               [("S", C.c_int), ("c", C.c_int), ("E", C.c_int), ("U", C.c_int), ("edSum", C.c_int),
                ("js", C.c_double)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.eo_gammds.restype = C.c_double
        L.eo_gammds.argtypes = [C.c_double, C.c_double]
        L.eo_chi2.restype = C.c_double
        L.eo_chi2.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
        L.eo_edge_decision.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(EoParams)]
        L.eo_bdeu_family.restype = C.c_double
        L.eo_bdeu_family.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int]
        L.eo_pvalue.restype = C.c_double
        L.eo_pvalue.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                C.POINTER(C.c_int)]
        L.eo_run_perm.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                  C.POINTER(EoParams), C.POINTER(EoPermOut)]
        L.eo_class_split.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int,
                                     C.c_void_p, C.c_void_p]
        L.eo_perm_labels.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p]
        L.eo_run_set.restype = C.c_int
        L.eo_run_set.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int,
                                 C.c_void_p, C.POINTER(EoParams), C.c_void_p, C.POINTER(C.c_int),
                                 C.POINTER(C.c_double), C.POINTER(C.c_int)]
        L.eo_uncapped_pair.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int,
                                       C.c_void_p]
        for f in ("eo_exp", "eo_log", "eo_lgamma_pos"):
            getattr(L, f).restype = C.c_double
            getattr(L, f).argtypes = [C.c_double]
        L.eo_ibetac.restype = C.c_double
        L.eo_ibetac.argtypes = [C.c_double, C.c_double, C.c_double, C.POINTER(C.c_int)]
        _lib = L
    return _lib


def ibetac(a, b, x):
    de = C.c_int()
    v = lib().eo_ibetac(a, b, x, C.byref(de))
    return v, de.value


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def params(mp=3, perms=100, p_adjust=1, alpha_edge=0.05, prior_weight=0.5, lam=2.0, alpha_ddn=0.05, threads=1):
    aep = 1 - (1 - alpha_edge) * prior_weight ** (1 / lam)
    return EoParams(mp, perms, p_adjust, alpha_edge, aep, alpha_ddn, threads)


def gammds(x, p):
    return lib().eo_gammds(x, p)


def chi2(tally):
    t = np.ascontiguousarray(np.asarray(tally, np.int32).reshape(9))
    dof = C.c_int()
    v = lib().eo_chi2(_p(t), C.byref(dof))
    return v, dof.value


def edge_decision(tally, c, prior, prm):
    t = np.ascontiguousarray(np.asarray(tally, np.int32).reshape(9))
    return lib().eo_edge_decision(_p(t), c, prior, C.byref(prm))


def pvalue(js):
    js = np.ascontiguousarray(js, np.float64)
    a, b, de = C.c_double(), C.c_double(), C.c_int()
    p = lib().eo_pvalue(_p(js), len(js), C.byref(a), C.byref(b), C.byref(de))
    return p, a.value, b.value, de.value


def class_split(data_ref, set_idx, labels):
    """data_ref: [genes x N] int32 reference values (-1/0/1); labels 0/1."""
    N = data_ref.shape[1]
    G = len(set_idx)
    n1 = int((np.asarray(labels) == 0).sum())
    n2 = N - n1
    t1 = np.zeros((G, max(1, n1)), np.int32)
    t2 = np.zeros((G, max(1, n2)), np.int32)
    idx = np.ascontiguousarray(set_idx, np.int32)
    lab = np.ascontiguousarray(labels, np.int32)
    lib().eo_class_split(_p(np.ascontiguousarray(data_ref, np.int32)), N, _p(idx), G, _p(lab), n1, n2, _p(t1), _p(t2))
    return t1, t2, n1, n2


def run_perm(t1, t2, n1, n2, prior, prm, want_bits=True):
    G = t1.shape[0]
    S = n1 + n2 + 2
    c = G * (G - 1) // 2
    o = EoPermOut()
    mp = prm.mp
    bufs = {
        "edgebits": np.zeros(max(1, S * c), np.uint8) if want_bits else None,
        "ppn": np.zeros(2 * G, np.int32), "stf": np.zeros(6 * G, np.int32),
        "uniqueN": np.zeros(S, np.int32), "uni_ids": np.zeros(S, np.int32),
        "uni_nodes": np.zeros(max(1, S * G), np.int32), "uni_edges": np.zeros(max(1, S * G * min(mp, max(G, 1))), np.int32),
        "uni_epn": np.zeros(S + 1, np.int32), "out5": np.zeros(max(1, 2 * S * G), np.float64),
        "lval": np.zeros(2 * S, np.float64),
    }
    for k, v in bufs.items():
        setattr(o, k, _p(v))
    prior = np.ascontiguousarray(prior, np.int32)
    lib().eo_run_perm(_p(np.ascontiguousarray(t1)), _p(np.ascontiguousarray(t2)), G, n1, n2, _p(prior),
                      C.byref(prm), C.byref(o))
    U = o.U
    return {
        "U": U, "js": o.js, "E": o.E, "edSum": o.edSum,
        "bits": bufs["edgebits"][:S * c].reshape(S, c) if want_bits else None,
        "uniq_ids": bufs["uni_ids"][:U], "uni_nodes": bufs["uni_nodes"][:U * G].reshape(U, G),
        "uni_epn": bufs["uni_epn"][:U + 1], "uni_edges": bufs["uni_edges"][:o.edSum],
        "out5": bufs["out5"][:2 * U * G].reshape(2, U, G), "lval": bufs["lval"][:2 * U],
        "ppn": bufs["ppn"], "stf": bufs["stf"],
    }


def run_set(data_ref, true_labels, set_idx, prior, prm):
    """Whole gene set through the process-wide rand() stream (caller resets it)."""
    N = data_ref.shape[1]
    lab = np.ascontiguousarray(true_labels, np.int32)
    n1 = int((lab == 0).sum())
    js = np.zeros(prm.perms, np.float64)
    failed, pddn, u0 = C.c_int(), C.c_double(), C.c_int()
    idx = np.ascontiguousarray(set_idx, np.int32)
    ran = lib().eo_run_set(_p(np.ascontiguousarray(data_ref, np.int32)), N, _p(lab), n1, N - n1, _p(idx), len(idx),
                           _p(np.ascontiguousarray(prior, np.int32)), C.byref(prm), _p(js), C.byref(failed),
                           C.byref(pddn), C.byref(u0))
    return {"js": js, "ran": ran, "failed": failed.value, "pDDN": pddn.value, "U0": u0.value}


def reset_stream():
    C.CDLL(None).srand(1)
