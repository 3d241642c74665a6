"""End-to-end drop-in check: the `eddy` CLI on the GPU writes the same
results.txt, _Networks, _BDEU_SCORES and _EdgeList files, byte for byte, as the
oracle CLI: committed golden files of a small synthetic run, and the
reference's own example inputs (its 575-label class file and GMT files with a
seeded expression matrix), run by both CLIs with the reference's example
arguments (resources/running_parameters.txt) at 20 permutations."""
import os
import shutil
import subprocess

import pytest

import refshapes

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CASE = os.path.join(ROOT, "tests", "golden", "cli_small")
EXE = os.path.join(ROOT, "eddy-gpu_amd", "eddy")
ORACLE = os.path.join(ROOT, "oracle", "build", "eddy_oracle")


@pytest.mark.parametrize("extra", [[], ["--mode", "plain", "--tests", "all"], ["--no-learned"], ["--learn-sets", "2"],
                                   ["--no-autotune"], ["--tune-cache", "shapes.txt"]])
def test_cli_golden(tmp_path, extra):
    for fn in ("expr.txt", "class.txt", "sets.gmt"):
        shutil.copy(os.path.join(CASE, fn), tmp_path / fn)
    shutil.copytree(os.path.join(CASE, "PRIORS"), tmp_path / "PRIORS")
    args = open(os.path.join(CASE, "args.txt")).read().split()
    print("case: cli_small", extra, flush=True)
    r = subprocess.run([EXE] + args + extra, cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    exp_dir = os.path.join(CASE, "expected")
    for fn in sorted(os.listdir(exp_dir)):
        want = open(os.path.join(exp_dir, fn)).read()
        got = open(tmp_path / fn).read()
        assert got == want, fn
    produced = {f for f in os.listdir(tmp_path) if f.startswith("SET_")}
    assert produced == {f for f in os.listdir(exp_dir) if f.startswith("SET_")}


@pytest.mark.parametrize("gmt,extra", [("failed.gmt", []), ("failed.gmt", ["-rawp"]), ("test.gmt", []),
                                       ("test.gmt", ["-mp", "20"])])
def test_cli_reference_shapes(tmp_path, gmt, extra):
    """The reference's class file (303/272 samples) and GMT files (sets of
    11..50 genes, some genes absent from the matrix), also with a parent cap
    above 16: every output file of the GPU CLI equals the oracle CLI's, byte
    for byte."""
    args = ["-d", "expr.txt", "-c", "class.txt", "-g", gmt, "-pE", "0.05", "-pw", "0.5", "-pD", "0.1", "-r",
            "20"] + extra
    runs = {}
    for who, exe in (("oracle", ORACLE), ("gpu", EXE)):
        d = tmp_path / who
        d.mkdir()
        refshapes.write_inputs(str(d))
        env = dict(os.environ, EDDY_ORACLE_THREADS=str(max(1, min(16, os.cpu_count() or 1))))
        print("case: reference shapes", gmt, extra, who, flush=True)
        r = subprocess.run([exe] + args, cwd=d, capture_output=True, text=True, timeout=110, env=env)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        runs[who] = {f: open(d / f).read() for f in os.listdir(d)
                     if f == "results.txt" or f.endswith(("_Networks.txt", "_BDEU_SCORES.txt", "_EdgeList.txt"))}
    assert runs["gpu"].keys() == runs["oracle"].keys()
    for fn in sorted(runs["oracle"]):
        assert runs["gpu"][fn] == runs["oracle"][fn], fn
    rows = runs["oracle"]["results.txt"].splitlines()
    assert len(rows) == 1 + (9 if gmt == "failed.gmt" else 5)
