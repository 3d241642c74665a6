"""Writes the committed fixtures under tests/golden/ (run in the build container).

* glibc_rand_1e4.npy   the first 10,000 values of the C library's rand() with
                       no srand call (the reference's permutation stream)
* betaincc.npz         scipy.special.betaincc(a, b, x) on a grid of shapes a
                       beta fit to JS values produces, plus edge cases
* gammainc.npz         scipy.special.gammainc(p, x) for p in {1/2, 1, 2}
* cli_small/           a tiny synthetic run (30 genes x 60 samples, 3 sets)
                       of the oracle CLI: its inputs and the four output files
* reference_shapes/    the reference's own example inputs, as data: its class
                       file (575 labels, 303 GBApPD then 272 GBApUnaffected)
                       and its two GMT files (14 sets of 11..50 genes), copied
                       from resources/ of the reference tree; its expression
                       matrix is absent there, so the tests pair them with a
                       seeded synthetic matrix over the GMT gene names
"""
import ctypes
import os
import shutil
import subprocess
import sys

import numpy as np
import scipy
import scipy.special as sp

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def rand_fixture():
    libc = ctypes.CDLL(None)
    libc.srand(1)
    v = np.array([libc.rand() for _ in range(10000)], np.int32)
    np.save(os.path.join(HERE, "glibc_rand_1e4.npy"), v)


def betaincc_fixture():
    rng = np.random.default_rng(20261203)
    a = np.concatenate([rng.uniform(0.5, 50, 300), rng.uniform(50, 2000, 100), [1.0, 2.5, 0.7, 8.9, 1e-3]])
    b = np.concatenate([rng.uniform(0.5, 400, 300), rng.uniform(100, 5000, 100), [3.0, 30.0, 1.0, 80.0, 2.0]])
    x = np.concatenate([rng.uniform(0.0, 0.6, 400), [0.2, 0.2, 0.5, 0.15, 0.9]])
    q = sp.betaincc(a, b, x)
    np.savez(os.path.join(HERE, "betaincc.npz"), a=a, b=b, x=x, q=q, scipy=scipy.__version__)


def gammainc_fixture():
    xs = np.concatenate([np.linspace(0.01, 5, 60), np.linspace(5, 95, 60)])
    out = {}
    for p in (0.5, 1.0, 2.0):
        out["p%g" % p] = sp.gammainc(p, xs)
    np.savez(os.path.join(HERE, "gammainc.npz"), x=xs, **out)


def cli_fixture():
    """Small end-to-end case: inputs written here, expected outputs from the oracle CLI."""
    d = os.path.join(HERE, "cli_small")
    if os.path.exists(d):
        shutil.rmtree(d)
    os.makedirs(os.path.join(d, "PRIORS"))
    rng = np.random.default_rng(7)
    G, nA, nB = 30, 30, 30
    N = nA + nB
    # two-class linear chains so that edges exist and differ between classes
    x = np.zeros((G, N))
    for s in range(N):
        for g in range(G):
            par = [g - 1] if g % 5 else []
            if s >= nA and g % 7 == 3:
                par = [max(0, g - 3)] if g >= 3 else []
            x[g, s] = sum(x[p, s] for p in par) * 0.9 + rng.normal()
    t = np.zeros((G, N), np.int32)
    for g in range(G):
        r = np.argsort(np.argsort(x[g], kind="stable"), kind="stable")
        t[g] = np.where(r < N // 3, -1, np.where(r < 2 * N // 3, 0, 1))
    with open(os.path.join(d, "expr.txt"), "w") as f:
        f.write("gene_id" + "".join("\tS%03d" % (i + 1) for i in range(N)) + "\n")
        for g in range(G):
            f.write("GENE%02d" % g + "".join("\t%d" % v for v in t[g]) + "\n")
        f.write("PAD" + "\t0" * N + "\n")
    with open(os.path.join(d, "class.txt"), "w") as f:
        f.write("\t".join(["ctrl"] * nA + ["case"] * nB) + "\n")
    sets = {
        "SET_A": ["GENE%02d" % g for g in range(0, 12)],
        "SET_B": ["GENE%02d" % g for g in (14, 3, 9, 21, 5, 27, 11, 18, 2, 24, 7, 16, 29, 0)] + ["MISSING1"],
        "SET_C": ["GENE%02d" % g for g in range(10, 30)],
    }
    with open(os.path.join(d, "sets.gmt"), "w") as f:
        for k, v in sets.items():
            f.write(k + "\thttp://example/" + k + "\t" + "\t".join(v) + "\n")
    with open(os.path.join(d, "PRIORS", "SET_A.prior"), "w") as f:
        f.write("GENE01\tinteracts-with\tGENE05\nGENE02\tneighbor-of\tGENE03\nGENE09\tcontrols\tGENE08\n")
    with open(os.path.join(d, "PRIORS", "SET_C.prior"), "w") as f:
        f.write("GENE12\tinteracts-with\tGENE20\nGENE25\tinteracts-with\tGENE28\n")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    exe = os.path.join(ROOT, "oracle", "build", "eddy_oracle")
    args = ["-d", "expr.txt", "-c", "class.txt", "-g", "sets.gmt", "-r", "20", "-pD", "0.5"]
    with open(os.path.join(d, "args.txt"), "w") as f:
        f.write(" ".join(args) + "\n")
    subprocess.run([exe] + args, cwd=d, check=True, stdout=subprocess.DEVNULL)
    os.makedirs(os.path.join(d, "expected"), exist_ok=True)
    for fn in os.listdir(d):
        if fn.endswith(".txt") and (fn == "results.txt" or fn.startswith("SET_")):
            shutil.move(os.path.join(d, fn), os.path.join(d, "expected", fn))


def reference_shapes_fixture():
    src = "/root/reference/resources"
    d = os.path.join(HERE, "reference_shapes")
    os.makedirs(d, exist_ok=True)
    for rel, dst in (("data/GBApPD_vs_GBApUnaffected_ClassFile.txt", "class.txt"), ("gmt/test.gmt", "test.gmt"),
                     ("gmt/failed.gmt", "failed.gmt")):
        shutil.copyfile(os.path.join(src, rel), os.path.join(d, dst))


if __name__ == "__main__":
    what = sys.argv[1:] or ["rand", "betaincc", "gammainc", "cli", "reference_shapes"]
    if "rand" in what:
        rand_fixture()
    if "betaincc" in what:
        betaincc_fixture()
    if "gammainc" in what:
        gammainc_fixture()
    if "cli" in what:
        cli_fixture()
    if "reference_shapes" in what:
        reference_shapes_fixture()
