#!/usr/bin/env python3
"""bench.py - EDDY permutation loop on one MI355X.

One step = eddy_run_sets over the whole batched workload (500 gene sets of the
5000-gene x 600-sample synthetic dataset, sizes log-uniform in 10..300, r = 100
permutations each, mp = 3, pw = 0.5), with the batched engine and the learned
cutoff schedule.  The dataset is resident in HBM before the timed region; each
step replays the same permutation stream (stream reset) so every step is the
same work.  Prints one JSON line:

  value          gene sets/s (whole job, r = 100), batched + learned
  plain          the plain GPU form (one set and one permutation per launch
                 sequence, every leave-one-out test) measured on every set
  ci_tests       reference-equivalent and evaluated chi-square tests/s
  roofline       the kernel with the most device time against MI355X's
                 published peak; rooflines: the count kernel (int8 MFMA),
                 the leave-one-out kernel (HBM) and BDeu (LDS)
  cpu_baseline   the CPU oracle (reference algorithm) on one item of each
                 set-size decile, on all of the GPU's host cores and on one
                 thread, extrapolated decile by decile to the same workload
"""
import argparse
import faulthandler
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "eddy-gpu_amd")
sys.path.insert(0, PKG)

PEAK_INT8_OPS = 5.0e15      # dense MFMA int8, OP/s
PEAK_FP64 = 78.6e12         # FP64, FLOP/s
PEAK_HBM = 8.0e12           # B/s
PEAK_LDS = 256 * 128 * 2.4e9  # B/s: 256 CUs x 128 B/clk at 2.4 GHz
HOST_CORE_SHARE = 16        # host cores one GPU's jobs may use on the GPU box


def ensure_built():
    if not os.path.exists(os.path.join(PKG, "libeddy.so")):
        subprocess.run([sys.executable, os.path.join(ROOT, "__graft_entry__.py")], check=True)


def algorithmic_work(sets, N, nA, perms_run):
    """Per-step algorithmic quantities summed over all (set, permutation) items."""
    # pairwise counting as a dense 0/1 contraction: 2 * 9 * c * n_A int8 ops per
    # item (class A's nine one-hot products per pair; class B's table is the
    # pair total minus class A's)
    ops = 0.0
    ref_tests = 0  # reference-equivalent CI tests: S c per item
    for (genes, _), r in zip(sets, perms_run):
        G = len(genes)
        c = G * (G - 1) // 2
        ops += r * 18.0 * c * nA
        ref_tests += r * (N + 2) * c
    return ops, ref_tests


def host_description():
    """nproc and the CPU model of this host (lscpu), for the baseline's record."""
    model = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
                break
    except (OSError, subprocess.SubprocessError):
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model}


def cpu_baseline(wl, perms_run, seconds_1t, threads):
    """The oracle (reference algorithm: a full leave-one-out tally per network
    and pair, AS 147 for every test, brute-force BDeu) on one permuted item of
    each set-size decile of the workload, on `threads` threads and on one
    thread.  Each decile's measured seconds per unit of the reference's cost
    S*c*n extrapolate that decile's sets at their permutations; the one-thread
    run stops after `seconds_1t` and takes the remaining deciles from the
    threaded run scaled by the one-thread/threaded ratio measured before."""
    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ref

    data_ref = wl.X.astype(np.int32) - 1
    N = wl.N
    cost = lambda G: (N + 2) * (G * (G - 1) // 2) * N
    sizes = np.array([len(g) for g, _ in wl.sets])
    order = np.argsort(sizes, kind="stable")
    deciles = np.array_split(order, 10)
    pick = [int(d[len(d) // 2]) for d in deciles]
    rng = np.random.default_rng(0)
    labs = [rng.permutation(wl.classes) for _ in pick]

    def run(th, budget=None):
        secs = []
        t_start = time.perf_counter()
        for s, lab in zip(pick, labs):
            if budget is not None and time.perf_counter() - t_start > budget:
                break
            genes, _ = wl.sets[s]
            t1, t2, n1, n2 = oracle_ref.class_split(data_ref, genes, lab)
            t0 = time.perf_counter()
            oracle_ref.run_perm(t1, t2, n1, n2, wl.prior_matrix(s), oracle_ref.params(mp=wl.cfg.mp, threads=th),
                                want_bits=False)
            secs.append(time.perf_counter() - t0)
        return secs

    def extrapolate(secs):
        total = 0.0
        for d, s, t in zip(deciles, pick, secs):
            rate = t / max(1, cost(sizes[s]))
            total += rate * sum(perms_run[q] * cost(sizes[q]) for q in d)
        return total

    sec_n = run(threads)
    sec_1 = run(1, budget=seconds_1t)
    measured_1 = len(sec_1)
    if measured_1 < len(pick):
        ratio = sum(sec_1) / max(1e-9, sum(sec_n[:measured_1]))
        sec_1 = sec_1 + [t * ratio for t in sec_n[measured_1:]]
    t_n, t_1 = extrapolate(sec_n), extrapolate(sec_1)
    host = host_description()
    sample = ("one permuted item of each set-size decile (G %s) of the same workload: %.1f s on %d threads, "
              "%.1f s on one thread%s; each decile's seconds per S*c*n extrapolate its %d sets at r=%d"
              % ("/".join(str(int(sizes[s])) for s in pick), sum(sec_n), threads, sum(sec_1[:measured_1]),
                 "" if measured_1 == len(pick) else
                 " (deciles %d..9 from the threaded times x the one-thread/threaded ratio %.1f of the others)"
                 % (measured_1, sum(sec_1[:measured_1]) / max(1e-9, sum(sec_n[:measured_1]))),
                 len(wl.sets) // 10, wl.cfg.perms))
    return {
        "value": len(wl.sets) / t_n, "unit": "gene sets/s", "cores": threads, "kind": "port",
        "sample": sample, "host": host,
        "single_thread": {"value": len(wl.sets) / t_1, "unit": "gene sets/s", "cores": 1,
                          "deciles_measured": measured_1},
        "decile_item_seconds": {"threads": [round(x, 3) for x in sec_n], "one_thread": [round(x, 3) for x in sec_1]},
    }


def workflow4(args):
    """The 20000-gene x 1200-sample workflow (uneven classes 300/900, 2000 sets of 5..500
    genes): the learned job at r = 100 over every set, then the 200 most
    significant sets again at r = 1000 on a fresh stream.  One warm-up run of
    the first phase, then both phases timed once.  Prints one JSON line."""
    import eddy_amd

    t_setup = time.perf_counter()
    wl = eddy_amd.Workload(4)
    eng = eddy_amd.Engine(0)
    eng.upload(wl.X, wl.classes)
    setup_s = time.perf_counter() - t_setup
    prm = wl.params(mode=eddy_amd.MODE_BATCHED, schedule=eddy_amd.TESTS_CUTOFF, learned=1)
    cfg = eddy_amd.learn_cfg(train_sets=100, epochs=args.epochs, batch=16384, lr=3e-3, margin=args.margin, seed=1,
                             max_rows=args.train_rows, label_perms=args.label_perms)
    print("bench: 20000 x 1200 workflow, %d sets, N=%d (%d/%d), setup %.1fs" % (len(wl.sets), wl.N, wl.cfg.nA, wl.cfg.nB,
                                                                   setup_s), file=sys.stderr, flush=True)
    eng.reset_stream()
    eng.run_learned(wl.batch, prm, cfg)
    eng.reset_stream()
    t0 = time.perf_counter()
    res, rep = eng.run_learned(wl.batch, prm, cfg)
    t1 = time.perf_counter()
    chosen, res2 = eng.repermute(wl.batch, res, prm, top=200, perms=1000)
    t2 = time.perf_counter()
    print("bench: r=100 phase %.2fs, r=1000 phase %.2fs" % (t1 - t0, t2 - t1), file=sys.stderr, flush=True)
    # the r = 100 phase without the network (same stream, same results), the
    # two forms interleaved over a few rounds (medians)
    uprm = wl.params(mode=eddy_amd.MODE_BATCHED, schedule=eddy_amd.TESTS_CUTOFF, learned=0)
    tl, tu, same = [t1 - t0], [], True
    k = min(100, len(wl.sets))
    tests_l = rep.tests_apply
    for r in range(args.compare_rounds4):
        eng.reset_stream()
        ta = time.perf_counter()
        ures = eng.run_sets(wl.batch, uprm)
        tu.append(time.perf_counter() - ta)
        same = same and all((a.status, a.perms_run, a.U0) == (b.status, b.perms_run, b.U0) and
                            (a.status != eddy_amd.SET_OK or a.pDDN == b.pDDN) for a, b in zip(ures, res))
        if r + 1 < args.compare_rounds4:
            eng.reset_stream()
            ta = time.perf_counter()
            eng.run_learned(wl.batch, prm, cfg)
            tl.append(time.perf_counter() - ta)
    med = lambda v: sorted(v)[len(v) // 2]
    tests_u = sum(r.tests_eval for r in ures[k:])
    G = [len(wl.sets[q][0]) for q in chosen]
    line = {
        "metric": "20000 x 1200 workflow: gene sets/s at r=100 (batched+learned), then r=1000 on the top 200",
        "value": len(wl.sets) / (t1 - t0), "unit": "gene sets/s", "n_gpus": 1, "higher_is_better": True,
        "data": "synthetic (seeded linear-Gaussian DAG, tertiles)",
        "config": {"workload": "20000 genes x 1200 samples (300/900), 2000 gene sets (G 5..500 log-uniform), "
                               "r=100, mp=3, pw=0.5; then the 200 smallest-pDDN sets at r=1000"},
        "phase1_ms": (t1 - t0) * 1e3, "phase2_ms": (t2 - t1) * 1e3,
        "phase2_sets": len(chosen), "phase2_sets_per_s": len(chosen) / (t2 - t1),
        "phase2_genes_min_median_max": [min(G), sorted(G)[len(G) // 2], max(G)] if G else None,
        "significant_at_0.05": int(sum(1 for r in res2 if r.status == eddy_amd.SET_OK and r.pDDN < 0.05)),
        "failed_sets_phase1": int(sum(1 for r in res if r.status != eddy_amd.SET_OK)),
        "learned": {"collect_ms": rep.collect_ms, "train_gpu_ms": rep.train_ms, "first_windowed_set":
                    rep.first_windowed_set, "loss_final": rep.loss_final},
        "vs_unlearned": {"rounds": len(tu), "learned_ms_median": med(tl) * 1e3, "unlearned_ms_median": med(tu) * 1e3,
                         "learned_speedup_vs_unlearned": med(tu) / med(tl),
                         "tests_applied_sets_learned": tests_l, "tests_applied_sets_unlearned": tests_u,
                         "results_identical": same},
    }
    print(json.dumps(line))


def run_step(eng, wl, prm, cfg):
    eng.reset_stream()
    res, rep = eng.run_learned(wl.batch, prm, cfg)
    return res, rep, eng.timing()


# ---------------------------------------------------------------- profile summaries
# rocprofv3 writes one SQLite database per run (-d DIR -o run: DIR/run_results.db).
# The summaries below read its `kernels` and `pmc_events` views.  Every job
# starts with the set-row gather (`k_gather`); the autotuner's calibration
# runs its own first, so the dispatches from the (skip + 1)-th `k_gather` on
# are the timed jobs (and the warm-up ones).

def _kname(n):
    return n.split("(")[0]


def _dispatches(db):
    import sqlite3
    c = sqlite3.connect(db)
    return c, c.execute("select dispatch_id, name, start, end from kernels order by start").fetchall()


def _job_start(rows, skip):
    g = [i for i, r in enumerate(rows) if _kname(r[1]) == "k_gather"]
    if len(g) <= skip:
        raise SystemExit("profile summary: fewer than %d k_gather dispatches" % (skip + 1))
    return g[skip], len(g) - skip


def summarize_trace(db, skip):
    _, rows = _dispatches(db)
    i0, jobs = _job_start(rows, skip)
    rows = rows[i0:]
    agg = {}
    for _, n, s, e in rows:
        a = agg.setdefault(_kname(n), [0, 0.0, 1e30, 0.0])
        d = (e - s) / 1e6
        a[0] += 1
        a[1] += d
        a[2] = min(a[2], d)
        a[3] = max(a[3], d)
    tot = sum(a[1] for a in agg.values())
    out = ["| kernel | calls | total ms | ms per job | avg ms | min ms | max ms | % |", "|---|---|---|---|---|---|---|---|"]
    for k, a in sorted(agg.items(), key=lambda x: -x[1][1]):
        out.append("| %s | %d | %.2f | %.2f | %.4f | %.4f | %.4f | %.1f |"
                   % (k, a[0], a[1], a[1] / jobs, a[1] / a[0], a[2], a[3], 100 * a[1] / tot))
    out.append("")
    out.append("jobs: %d, span of their dispatches %.1f ms" % (jobs, (rows[-1][3] - rows[0][2]) / 1e6))
    return "\n".join(out)


def _counters(db, skip):
    """{kernel: {counter: total}}, dispatch counts, jobs, and the dispatches before the jobs."""
    c, rows = _dispatches(db)
    i0, jobs = _job_start(rows, skip)
    after = {r[0] for r in rows[i0:]}
    vals = {}
    for d, n, cn, v in c.execute("select dispatch_id, name, counter_name, counter_value from pmc_events"):
        vals.setdefault(d, [n, {}])[1][cn] = vals.setdefault(d, [n, {}])[1].get(cn, 0.0) + v
    agg, disp, before = {}, {}, []
    for d, (n, cv) in vals.items():
        if d not in after:
            before.append((d, _kname(n), cv))
            continue
        k = _kname(n)
        disp[k] = disp.get(k, 0) + 1
        for cn, v in cv.items():
            agg.setdefault(k, {})[cn] = agg.setdefault(k, {}).get(cn, 0.0) + v
    return agg, disp, jobs, sorted(before)


def summarize_counters(db, skip):
    agg, disp, jobs, _ = _counters(db, skip)
    names = sorted({cn for v in agg.values() for cn in v})
    out = ["per job (%d jobs):" % jobs, "",
           "| kernel | dispatches | " + " | ".join(names) + " |", "|---|---|" + "---|" * len(names)]
    for k in sorted(agg, key=lambda k: -agg[k].get(names[0], 0.0)):
        out.append("| %s | %d | " % (k, disp[k] // jobs) +
                   " | ".join("%.4g" % (agg[k].get(cn, 0.0) / jobs) for cn in names) + " |")
    return "\n".join(out)


def summarize_traffic(db_fetch, db_write, skip, copy_mib, json_path):
    """FETCH_SIZE / WRITE_SIZE (KiB per dispatch) scaled by what they read for
    the calibration copy of copy_mib MiB (bench.py --calibrate-copy: the run's
    first dispatch), as bytes per dispatch and per job."""
    res = {}
    factors = {}
    for key, db, cn in (("read", db_fetch, "FETCH_SIZE"), ("write", db_write, "WRITE_SIZE")):
        agg, disp, jobs, before = _counters(db, skip)
        cal = before[0][2].get(cn, 0.0) if before and "copy" in before[0][1].lower() else 0.0
        if cal <= 0:
            raise SystemExit("profile summary: no calibration copy before the jobs in %s" % db)
        factors[key] = copy_mib * 1024.0 / cal
        for k, cv in agg.items():
            r = res.setdefault(k, {"dispatches_per_job": disp[k] / jobs})
            r[key + "_bytes_per_job"] = cv.get(cn, 0.0) * 1024.0 * factors[key] / jobs
    kern = {}
    out = ["scale factors: fetch %.3f, write %.3f (copied KiB per counted KiB, from the %d MiB copy)"
           % (factors["read"], factors["write"], copy_mib), "",
           "| kernel | dispatches per job | read MB per dispatch | written MB per dispatch | total MB per job |",
           "|---|---|---|---|---|"]
    for k, r in sorted(res.items(), key=lambda x: -(x[1].get("read_bytes_per_job", 0) + x[1].get("write_bytes_per_job", 0))):
        rd, wr, n = r.get("read_bytes_per_job", 0.0), r.get("write_bytes_per_job", 0.0), r["dispatches_per_job"]
        kern[k] = {"dispatches_per_step": n, "read_bytes_per_dispatch": rd / n, "write_bytes_per_dispatch": wr / n,
                   "bytes_per_step": rd + wr}
        out.append("| %s | %.1f | %.1f | %.1f | %.1f |" % (k, n, rd / n / 1e6, wr / n / 1e6, (rd + wr) / 1e6))
    if json_path:
        with open(json_path, "w") as f:
            json.dump({"fetch_factor": factors["read"], "write_factor": factors["write"], "kernels": kern}, f,
                      indent=1)
    return "\n".join(out)


def summarize_main(argv):
    ap = argparse.ArgumentParser(prog="bench.py --summarize")
    ap.add_argument("kind", choices=["trace", "counters", "traffic"])
    ap.add_argument("db", nargs="+", help="run_results.db (traffic: the FETCH_SIZE run's, then the WRITE_SIZE run's)")
    ap.add_argument("--skip", type=int, default=1, help="k_gather dispatches before the first job (the autotuner's)")
    ap.add_argument("--copy-mib", type=int, default=1024)
    ap.add_argument("--json", default=None, help="traffic: also write bytes per kernel as JSON here")
    a = ap.parse_args(argv)
    if a.kind == "trace":
        print(summarize_trace(a.db[0], a.skip))
    elif a.kind == "counters":
        print(summarize_counters(a.db[0], a.skip))
    else:
        print(summarize_traffic(a.db[0], a.db[1], a.skip, a.copy_mib, a.json))


def main():
    faulthandler.enable()  # a crash inside the library prints the Python call it came from
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=3)
    ap.add_argument("--nsets", type=int, default=None)
    ap.add_argument("--train-sets", type=int, default=100)
    ap.add_argument("--epochs", type=int, default=1)
    ap.add_argument("--train-rows", type=int, default=65536)
    ap.add_argument("--margin", type=float, default=3.0)
    ap.add_argument("--label-perms", type=int, default=10)
    ap.add_argument("--plain-sets", type=int, default=None,
                    help="sets the plain GPU form runs (default: all of them); 0 skips it")
    ap.add_argument("--cpu-seconds", type=float, default=40.0,
                    help="time limit of the one-thread CPU baseline run (the deciles past it are scaled)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-compare", action="store_true")
    ap.add_argument("--compare-rounds", type=int, default=5,
                    help="interleaved rounds of the learned job, the job without the network and the deferred form")
    ap.add_argument("--compare-rounds4", type=int, default=3,
                    help="rounds of the 20000 x 1200 workflow's learned / unlearned comparison")
    ap.add_argument("--workflow4", action="store_true",
                    help="run the 20000-gene x 1200-sample workflow (r=100 over 2000 sets, then r=1000 on the top 200) instead")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch shapes instead of running the autotuner in setup")
    ap.add_argument("--launch", default="",
                    help="fixed launch shapes after tuning, e.g. graph=1,loo_waves=3 (results never change)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="first run one device-to-device copy of MIB MiB (a kernel of known traffic for "
                         "scaling hardware byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this engine runs on one GPU; multi-GPU sharding is not built",
              file=sys.stderr)
        sys.exit(2)
    ensure_built()
    if args.workflow4:
        workflow4(args)
        return
    import eddy_amd

    if args.calibrate_copy > 0:
        import torch

        src = torch.empty(args.calibrate_copy << 20, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print("bench: calibration copy of %d MiB done" % args.calibrate_copy, file=sys.stderr, flush=True)

    t_setup = time.perf_counter()
    wl = eddy_amd.Workload(args.config, nsets=args.nsets)
    eng = eddy_amd.Engine(0)
    eng.upload(wl.X, wl.classes)
    setup_s = time.perf_counter() - t_setup
    nsets = len(wl.sets)
    prm = wl.params(mode=eddy_amd.MODE_BATCHED, schedule=eddy_amd.TESTS_CUTOFF, learned=1)
    cfg = eddy_amd.learn_cfg(train_sets=args.train_sets, epochs=args.epochs, batch=16384, lr=3e-3,
                             margin=args.margin, seed=1, max_rows=args.train_rows,
                             label_perms=args.label_perms)

    print("bench: setup %.1fs, %d sets, N=%d" % (setup_s, nsets, wl.N), file=sys.stderr, flush=True)
    tune = None
    if not args.no_autotune:
        chosen, tune = eng.autotune(wl.sets, prm)
        print("bench: autotune %.0f ms, %d trials: %s" % (tune["tune_ms"], tune["trials"], chosen), file=sys.stderr,
              flush=True)
    if args.launch:
        eng.set_launch_cfg(**{k: int(v) for k, v in (kv.split("=") for kv in args.launch.split(","))})
    launch = eng.launch_cfg()
    for w in range(args.warmup):
        tw = time.perf_counter()
        run_step(eng, wl, prm, cfg)
        print("bench: warmup %d %.2fs" % (w, time.perf_counter() - tw), file=sys.stderr, flush=True)
    split_ms = [0.0] * 7
    split_n = [0] * 7
    tests_eval = 0
    items = 0
    host_stream_ms = 0.0
    sched = [0, 0, 0, 0]
    stats = [0] * 6
    phase = [0.0] * 16
    reps = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res, rep, tm = run_step(eng, wl, prm, cfg)
        reps.append(rep)
        for k in range(7):
            split_ms[k] += tm.split_ms[k]
            split_n[k] += tm.split_launches[k]
        tests_eval += tm.tests
        items += tm.items
        host_stream_ms += tm.host_stream_ms
        for k in range(4):
            sched[k] += tm.sched[k]
        for k in range(6):
            stats[k] += tm.stats[k]
        for k in range(16):
            phase[k] += tm.phase_ms[k]
        print("bench: step done, total %.2fs (collect %.0f ms, then %.0f ms; training %.1f ms on a second stream, "
              "windows from set %d)" % (time.perf_counter() - t0, rep.collect_ms, rep.apply_ms, rep.train_ms,
                                       rep.first_windowed_set), file=sys.stderr, flush=True)
    elapsed = time.perf_counter() - t0
    ms_per_step = 1000.0 * elapsed / args.steps
    value = nsets * args.steps / elapsed
    perms_run = [r.perms_run for r in res]
    ops, ref_tests = algorithmic_work(wl.sets, wl.N, wl.cfg.nA, perms_run)
    js_head = [eng.js(s).copy() for s in range(nsets)]
    rep = reps[-1]
    learned = {
        "train_sets": min(args.train_sets, nsets), "logged_rows": rep.train_rows,
        "train_rows": min(rep.train_rows, args.train_rows) if args.train_rows > 0 else rep.train_rows,
        "epochs": args.epochs,
        "loss_initial": rep.loss_initial, "loss_final": rep.loss_final, "margin": args.margin,
        "collect_ms": rep.collect_ms, "train_gpu_ms_overlapped": rep.train_ms, "apply_ms": rep.apply_ms,
        "first_windowed_set": rep.first_windowed_set, "label_perms": args.label_perms,
        "tests_applied_sets": rep.tests_apply,
    }

    compare = None
    if not args.no_compare:
        # the learned job beside the same job without the network and with every
        # full test deferred to the plan (no window at all): identical results;
        # the three forms interleaved over several rounds (medians), so drift
        # over time does not fall on one form
        uprm = wl.params(mode=eddy_amd.MODE_BATCHED, schedule=eddy_amd.TESTS_CUTOFF, learned=0)
        zprm = wl.params(mode=eddy_amd.MODE_BATCHED, schedule=eddy_amd.TESTS_CUTOFF, learned=3)
        k = min(args.train_sets, nsets)
        same = True

        def same_as_learned(out):
            ok = all((a.status, a.perms_run, a.U0) == (b.status, b.perms_run, b.U0) for a, b in zip(out, res))
            return ok and all(bool((eng.js(s).view("u8") == js_head[s].view("u8")).all()) for s in range(nsets))

        tl, tu, tz = [], [], []
        for _ in range(args.compare_rounds):
            t0 = time.perf_counter()
            lres, _, _ = run_step(eng, wl, prm, cfg)
            tl.append(time.perf_counter() - t0)
            same = same and same_as_learned(lres)
            eng.reset_stream()
            t0 = time.perf_counter()
            ures = eng.run_sets(wl.batch, uprm)
            tu.append(time.perf_counter() - t0)
            same = same and same_as_learned(ures)
            eng.reset_stream()
            t0 = time.perf_counter()
            zres = eng.run_sets(wl.batch, zprm)
            tz.append(time.perf_counter() - t0)
            same = same and same_as_learned(zres)
        med = lambda v: sorted(v)[len(v) // 2]
        tests_u = sum(r.tests_eval for r in ures[k:])
        tests_z = sum(r.tests_eval for r in zres[k:])
        compare = {"rounds": args.compare_rounds,
                   "learned_ms_median": med(tl) * 1000.0, "learned_sets_per_s_median": nsets / med(tl),
                   "unlearned_ms_median": med(tu) * 1000.0, "unlearned_sets_per_s_median": nsets / med(tu),
                   "deferred_full_tests_ms_median": med(tz) * 1000.0,
                   "deferred_full_tests_sets_per_s_median": nsets / med(tz),
                   "learned_speedup_vs_unlearned": med(tu) / med(tl),
                   "tests_applied_sets_deferred": tests_z,
                   "tests_applied_sets_unlearned": tests_u, "tests_applied_sets_learned": rep.tests_apply,
                   "test_reduction_applied_sets": 1.0 - rep.tests_apply / max(1, tests_u),
                   "results_bit_identical": same}
        learned["vs_unlearned"] = compare

    # plain GPU form (one set and one permutation per launch sequence, every
    # test), measured on the same sets (all of them by default)
    plain = None
    nplain = nsets if args.plain_sets is None else min(args.plain_sets, nsets)
    if nplain > 0:
        sub = wl.sets[:nplain]
        pprm = wl.params(mode=eddy_amd.MODE_PLAIN, schedule=eddy_amd.TESTS_ALL, learned=0)
        eng.reset_stream()
        tp = time.perf_counter()
        pres = eng.run_sets(sub, pprm)
        dtp = time.perf_counter() - tp
        same = all((r.status, r.perms_run, r.U0) == (q.status, q.perms_run, q.U0) for r, q in zip(pres, res))
        for s in range(len(sub)):
            same = same and bool((eng.js(s).view("u8") == js_head[s].view("u8")).all())
        plain = {"sets": len(sub), "seconds": dtp, "plain_sets_per_s": len(sub) / dtp,
                 "measured": "every set" if nplain == nsets else "the first %d sets" % nplain,
                 "speedup_batched_learned_vs_plain": value / (len(sub) / dtp), "results_bit_identical": same}
        print("bench: plain form %d sets in %.1fs" % (len(sub), dtp), file=sys.stderr, flush=True)

    # Rooflines.  Algorithmic work per launch over the kernel's average launch
    # time (HIP events on the slot's stream, summed over the timed steps):
    #   k_count_test  2*9*c*n_A int8 ops per item (class A's nine one-hot
    #                 products per pair; class B's table is the pooled table
    #                 minus class A's) against dense int8 MFMA
    #   k_loo         22 B per planned (pair, class) entry (its 9 u16 cells
    #                 and 4 B of info) against HBM
    #   BDeu          (1 + |pa|) state bytes per sample and class plus 4 q r
    #                 bytes of bins, per distinct family and scoring class,
    #                 over k_score + k_bdeu (which score them) against LDS
    per_step = lambda x: x / args.steps
    avg_s = [split_ms[k] / max(1, split_n[k]) / 1000.0 for k in range(7)]
    lps = [split_n[k] / args.steps for k in range(7)]
    roofs = {}
    if split_n[0]:
        a_cnt = per_step(ops * args.steps) / lps[0] / avg_s[0]
        roofs["k_count_test"] = {"bound": "mfma", "achieved": a_cnt / 1e12, "peak": PEAK_INT8_OPS / 1e12,
                                 "unit": "TOP/s", "frac": a_cnt / PEAK_INT8_OPS,
                                 "work_per_launch": per_step(ops * args.steps) / lps[0],
                                 "avg_launch_ms": avg_s[0] * 1e3, "launches_per_step": lps[0]}
    if split_n[1]:
        a_loo = per_step(sched[2] * 22) / lps[1] / avg_s[1]
        roofs["k_loo"] = {"bound": "hbm", "achieved": a_loo / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                          "frac": a_loo / PEAK_HBM, "work_per_launch": per_step(sched[2] * 22) / lps[1],
                          "avg_launch_ms": avg_s[1] * 1e3, "launches_per_step": lps[1]}
    if split_n[3] and split_n[4]:
        t_bd = (split_ms[3] + split_ms[4]) / split_n[3] / 1000.0
        a_bd = per_step(stats[0] + stats[4]) / lps[3] / t_bd
        roofs["bdeu"] = {"kernels": "k_score+k_bdeu", "bound": "lds", "achieved": a_bd / 1e9,
                         "peak": PEAK_LDS / 1e9, "unit": "GB/s", "frac": a_bd / PEAK_LDS,
                         "work_per_launch": per_step(stats[0] + stats[4]) / lps[3],
                         "avg_launch_ms": t_bd * 1e3, "launches_per_step": lps[3]}
    dom = max(range(7), key=lambda k: split_ms[k])
    dname = eddy_amd.SPLIT[dom]
    if dname in roofs:
        roof = {"kernel": dname, **{k: roofs[dname][k] for k in ("bound", "achieved", "peak", "unit", "frac")}}
    elif dom in (3, 4):
        roof = {"kernel": dname, **{k: roofs["bdeu"][k] for k in ("bound", "achieved", "peak", "unit", "frac")},
                "note": "BDeu bytes over k_score + k_bdeu"}
    else:
        roof = {"kernel": dname, "bound": "hbm", "achieved": None, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                "frac": None}
    roof["avg_launch_ms"] = avg_s[dom] * 1e3
    roof["launches_per_step"] = lps[dom]
    roof["traffic"] = None
    # HBM bytes per launch from the committed byte-counter profile of this
    # workload (FETCH_SIZE + WRITE_SIZE, copy-calibrated): a kernel's bytes per
    # step over this run's launches per step.  The loo kernel runs under the
    # name of the instance the launch configuration picked.
    tpath = os.path.join(ROOT, "profiles", "hbm_traffic_learned_job_5000x600.json")
    if args.config == 3 and args.nsets is None and os.path.exists(tpath):
        kern = json.load(open(tpath))["kernels"]
        prof_name = lambda n: "k_loo_w3" if n == "k_loo" and launch.get("loo_waves") == 3 else n
        for k, n in enumerate(eddy_amd.SPLIT):
            if n in roofs and prof_name(n) in kern:
                roofs[n]["traffic"] = kern[prof_name(n)]["bytes_per_step"] / max(1e-9, lps[k])
        if prof_name(dname) in kern:
            roof["traffic"] = kern[prof_name(dname)]["bytes_per_step"] / max(1e-9, lps[dom])
            roof["traffic_source"] = os.path.relpath(tpath, ROOT)

    cpu = None
    if not args.no_cpu:
        # every host core this GPU's jobs may use (the box's share), and one
        threads = max(1, min(HOST_CORE_SHARE, os.cpu_count() or 1))
        tc = time.perf_counter()
        cpu = cpu_baseline(wl, perms_run, args.cpu_seconds, threads)
        print("bench: cpu baseline %.1fs" % (time.perf_counter() - tc), file=sys.stderr, flush=True)

    line = {
        "metric": "gene sets/s at r=100 perms (batched+learned); CI tests/s",
        "value": value,
        "unit": "gene sets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8 counts / fp64 tests",
        "data": "synthetic (seeded linear-Gaussian DAG, tertiles)",
        "config": {"workload": "%d genes x %d samples (%d/%d), %d gene sets (G %d..%d log-uniform), r=%d, mp=%d, "
                               "pw=%.1f; network trained on the first %d sets, applied to the rest"
                               % (wl.cfg.genes, wl.N, wl.cfg.nA, wl.cfg.nB, nsets, wl.cfg.gmin, wl.cfg.gmax,
                                  wl.cfg.perms, wl.cfg.mp, wl.cfg.prior_weight, min(args.train_sets, nsets)),
                   "global_batch": nsets, "parallelism": "single GPU"},
        "ci_tests_per_s_ref_equiv": ref_tests * args.steps / elapsed,
        "ci_tests_per_s_evaluated": tests_eval / elapsed,
        "items_per_step": items // args.steps,
        "kernel_ms_per_step": {eddy_amd.SPLIT[k]: split_ms[k] / args.steps for k in range(7)},
        "kernel_ms_per_step_note": "HIP events between the kernels of each chunk's sequence on its slot's stream; "
                                   "two chunks are in flight, so intervals overlap",
        "host_stream_ms_per_step": host_stream_ms / args.steps,
        "schedule": {"rows_past_plan": sched[0] // args.steps, "cols_tested_past_plan": sched[1] // args.steps,
                     "planned_entries": sched[2] // args.steps, "rows": sched[3] // args.steps},
        "score_stats": {"hist_bytes": stats[0] // args.steps, "bin_bytes": stats[4] // args.steps,
                        "families": stats[1] // args.steps, "unique_networks": stats[2] // args.steps,
                        "events": stats[3] // args.steps},
        "phase_wg_ms": [round(x / args.steps, 1) for x in phase],
        "setup_s": setup_s,
        "launch": launch,
        "autotune": tune,
        "learned": learned,
        "plain": plain,
        "roofline": roof,
        "rooflines": roofs,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "--summarize":
        summarize_main(sys.argv[2:])
    else:
        main()
