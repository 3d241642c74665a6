"""Inputs in the shape of the reference's own example run (resources/ of the
reference tree): its class file and GMT files (tests/golden/reference_shapes,
committed as data) with a seeded synthetic expression matrix over the GMT gene
names, since the reference's expression matrix is not part of its tree.

The matrix holds every GMT gene except every ninth one (those are dropped by
the gene lookup, main.cu:400-430), 40 genes in no set, and the padding row
the parser drops (main.cu:283).  Values are the seeded DAG generator's
tertiles written as -1/0/1, 303 class-1 then 272 class-2 samples, matching
the class file's order.
"""
import os
import shutil

import numpy as np

import eddy_amd

HERE = os.path.dirname(os.path.abspath(__file__))
CASE = os.path.join(HERE, "golden", "reference_shapes")
N1, N2 = 303, 272


def gene_names():
    names, seen = [], set()
    for fn in ("test.gmt", "failed.gmt"):
        with open(os.path.join(CASE, fn)) as f:
            for line in f:
                for g in line.rstrip("\n").split("\t")[2:]:
                    if g and g not in seen:
                        seen.add(g)
                        names.append(g)
    return names


def write_inputs(d, seed=0xEDD15EED):
    """Writes expr.txt, class.txt, test.gmt and failed.gmt into directory d."""
    names = gene_names()
    rows = [g for i, g in enumerate(names) if i % 9 != 4] + ["FILL%03d" % i for i in range(40)]
    X = np.zeros((len(rows), N1 + N2), np.uint8)
    rc = eddy_amd.lib().eddy_synth_dataset(seed, len(rows), N1, N2, 0.15, X.ctypes.data_as(eddy_amd.C.c_void_p))
    assert rc == eddy_amd.EDDY_OK
    # rows in a seeded order, so a set's genes are not in the generator's DAG order
    perm = np.random.default_rng(seed).permutation(len(rows))
    with open(os.path.join(d, "expr.txt"), "w") as f:
        f.write("gene_id" + "".join("\tGSM%05d" % (i + 1) for i in range(N1 + N2)) + "\n")
        for k, r in enumerate(perm):
            f.write(rows[k] + "\t" + "\t".join(str(int(v) - 1) for v in X[r]) + "\n")
        f.write("PAD" + "\t0" * (N1 + N2) + "\n")
    for fn in ("class.txt", "test.gmt", "failed.gmt"):
        shutil.copyfile(os.path.join(CASE, fn), os.path.join(d, fn))
