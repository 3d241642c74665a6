"""CPU checks of the `eddy` CLI's input front end (main.cu:206-503 quirks) and
of the oracle CLI against the committed golden run."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "eddy-gpu_amd", "eddy")
CASE = os.path.join(ROOT, "tests", "golden", "cli_small")


def parse(tmp_path, expr, cls, gmt, priors=None, extra=()):
    (tmp_path / "e.txt").write_text(expr)
    (tmp_path / "c.txt").write_text(cls)
    (tmp_path / "g.gmt").write_text(gmt)
    os.makedirs(tmp_path / "PRIORS", exist_ok=True)
    for k, v in (priors or {}).items():
        (tmp_path / "PRIORS" / (k + ".prior")).write_text(v)
    r = subprocess.run([EXE, "-d", "e.txt", "-c", "c.txt", "-g", "g.gmt", "--parse-only", *extra],
                       cwd=tmp_path, capture_output=True, text=True, timeout=60)
    return r


def sets_of(out):
    res = {}
    for line in out.splitlines():
        if line.startswith("set "):
            head, rows = line.split(":")
            w = head.split()
            res[w[1]] = (int(w[3]), int(w[5]), [int(x) for x in rows.split()])
    return res


def test_quirks(tmp_path):
    # header: trailing tab -> a "\n" token of length 1 is dropped; the last
    # data row (G5) is dropped; G2 appears twice (first row wins)
    expr = "id\tS1\tS2\tS3\tS4\t\n" + \
           "G0\t1\t0\t-1\t1\nG1\t0\t0\t1\t-1\nG2\t1\t1\t1\t0\nG2\t-1\t-1\t-1\t-1\nG3\t0\t1\t0\t1\nG5\t1\t1\t1\t1\n"
    cls = "A\tB\tA\tB\n"
    gmt = "\n" + "SETX\turl\tG3\tG9\tG2\tG0\tG5\n" + "SETY\turl\tG1\n"
    pri = "G3\tinteracts-with\tG0\nG2\tneighbor-of\tG0\nG0\tinteracts-with\tG7\nG9\tx\tG3\n" + \
          "G2\tinteracts-with\tG3" + "\t" * 120 + "\n"
    r = parse(tmp_path, expr, cls, gmt, {"SETX": pri})
    assert r.returncode == 0, r.stdout
    assert "4 samples" in r.stdout and "5 genes" in r.stdout
    assert "Classes: 2 A, 2 B" in r.stdout
    s = sets_of(r.stdout)
    # G9 and G5 (dropped last row) are not in the data; G2 -> first row 2
    assert s["SETX"][2] == [4, 2, 0]
    # prior: only G3-G0 (neighbor-of skipped, unknown genes skipped, the long
    # line's third token is cut at the 100-byte read)
    assert s["SETX"][1] == 1
    assert s["SETY"][2] == [1]


def test_bad_inputs_exit_failure(tmp_path):
    expr = "id\tS1\tS2\tS3\nG0\t1\t0\t-1\nG1\t0\t0\t1\n"
    r = parse(tmp_path, expr, "A\tA\tA\n", "S\tu\tG0\n")
    assert r.returncode != 0 and "two class" in r.stdout
    r = parse(tmp_path, expr, "A\tB\tC\n", "S\tu\tG0\n")
    assert r.returncode != 0
    r = parse(tmp_path, expr, "A\tB\n", "S\tu\tG0\n")
    assert r.returncode != 0
    r = subprocess.run([EXE, "-d", "e.txt"], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode != 0 and "Invalid class file" in r.stdout


def test_oracle_cli_golden(tmp_path):
    """The oracle CLI still reproduces the committed golden files."""
    for fn in ("expr.txt", "class.txt", "sets.gmt"):
        shutil.copy(os.path.join(CASE, fn), tmp_path / fn)
    shutil.copytree(os.path.join(CASE, "PRIORS"), tmp_path / "PRIORS")
    exe = os.path.join(ROOT, "oracle", "build", "eddy_oracle")
    args = open(os.path.join(CASE, "args.txt")).read().split()
    r = subprocess.run([exe] + args, cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0
    exp = os.path.join(CASE, "expected")
    for fn in os.listdir(exp):
        assert open(tmp_path / fn).read() == open(os.path.join(exp, fn)).read(), fn
