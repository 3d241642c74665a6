"""End-to-end drop-in check: the `eddy` CLI on the GPU writes the same
results.txt, _Networks, _BDEU_SCORES and _EdgeList files, byte for byte, as the
oracle CLI (committed golden files of a small synthetic run)."""
import os
import shutil
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CASE = os.path.join(ROOT, "tests", "golden", "cli_small")
EXE = os.path.join(ROOT, "eddy-gpu_amd", "eddy")


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
