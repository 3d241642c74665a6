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


def _ref_parse(text):
    """The reference's expression reading (main.cu:255-301) restated: lines
    keep their newline, tokens are strtok runs of non-tabs, values are
    sscanf("%d") read as a long and stored into an int, 1-byte header tokens
    do not count as samples, the last data row is dropped."""
    import re

    parts = text.split("\n")  # only '\n' ends a line (a '\r' stays in its token)
    lines = [x + "\n" for x in parts[:-1]] + ([parts[-1]] if parts[-1] else [])
    head = [t for t in lines[0].split("\t") if t]
    ns = max(0, len(head) - 1) - sum(1 for t in head[1:] if len(t) == 1)
    rows = lines[1:]
    genes, states = [], []
    for r in rows[:-1]:
        toks = [t for t in r.split("\t") if t]
        if len(toks) < ns + 1:
            return None
        genes.append(toks[0])
        for t in toks[1:ns + 1]:
            m = re.match(r"[ \t\n\r\v\f]*([+-]?)(\d+)", t)
            if not m:
                return None
            v = int(m.group(2)) * (-1 if m.group(1) == "-" else 1)
            v = max(-(1 << 63), min((1 << 63) - 1, v))
            v = (v & 0xFFFFFFFF) - (1 << 32 if v & 0x80000000 else 0)
            states.append(v + 1 if -1 <= v <= 1 else 3)
    h = 1469598103934665603
    for s in states:
        h = ((h ^ s) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    for g in genes:
        for ch in g.encode():
            h = ((h ^ ch) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
        h = ((h ^ 0xFF) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return ns, len(genes), h


def _hash_of(out):
    for line in out.splitlines():
        if line.startswith("expression hash "):
            return int(line.split()[2], 16)
    return None


def test_expression_reader_matches_reference_semantics(tmp_path):
    """Random files with the awkward cases (runs of tabs, trailing tabs, CRLF,
    signs, leading blanks, decimals, out-of-range and huge values, extra
    columns, a last line without newline) read into the same states and gene
    names as the restated reference reader."""
    import random

    rnd = random.Random(5)
    vals = ["1", "0", "-1", "+1", " -1", "1.7", "2", "-3", "007", "4294967296", "99999999999999999999999", "1\r"]
    for case in range(25):
        ns = rnd.randint(2, 9)
        head = "id" + "".join("\t" * rnd.randint(1, 2) + "S%d" % i for i in range(ns))
        head += rnd.choice(["", "\t", "\t\t"]) + "\n"
        body = []
        for g in range(rnd.randint(2, 12)):
            toks = ["G%d" % rnd.randint(0, 5)] + [rnd.choice(vals) for _ in range(ns + rnd.randint(0, 2))]
            body.append("\t".join(t if rnd.random() > 0.1 else t + "\t" for t in toks))
        text = head + "\n".join(body) + rnd.choice(["", "\n"])
        want = _ref_parse(text)
        r = parse(tmp_path, text, "A\tB" + "\tA" * 20 + "\n", "S\tu\tG0\n")
        if want is None:
            assert r.returncode != 0, (case, text)
            continue
        assert f"{want[0]} samples" in r.stdout and f"{want[1]} genes" in r.stdout, (case, r.stdout)
        if r.returncode == 0:
            assert _hash_of(r.stdout) == want[2], (case, text)


def test_expression_cache(tmp_path):
    expr = "id\tS1\tS2\tS3\tS4\nG0\t1\t0\t-1\t1\nG1\t0\t0\t1\t-1\nG2\t1\t1\t1\t0\nG3\t0\t1\t0\t1\n"
    r1 = parse(tmp_path, expr, "A\tB\tA\tB\n", "S\tu\tG0\tG1\n", extra=("--cache",))
    assert r1.returncode == 0 and "expression cache" not in r1.stdout
    assert (tmp_path / "e.txt.eddybin").exists()
    r2 = subprocess.run([EXE, "-d", "e.txt", "-c", "c.txt", "-g", "g.gmt", "--parse-only", "--cache"],
                        cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r2.returncode == 0 and "expression cache e.txt.eddybin" in r2.stdout
    assert _hash_of(r2.stdout) == _hash_of(r1.stdout) and _hash_of(r1.stdout) is not None
    # a changed source is read again
    (tmp_path / "e.txt").write_text(expr.replace("G3\t0\t1\t0\t1", "G3\t1\t1\t1\t1") + "G4\t0\t0\t0\t0\n")
    r3 = subprocess.run([EXE, "-d", "e.txt", "-c", "c.txt", "-g", "g.gmt", "--parse-only", "--cache"],
                        cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r3.returncode == 0 and "expression cache" not in r3.stdout
    assert _hash_of(r3.stdout) != _hash_of(r1.stdout)
