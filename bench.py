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
                 sequence, every leave-one-out test) on a prefix of the sets
  ci_tests       reference-equivalent and evaluated chi-square tests/s
  roofline       the dominant kernel against MI355X's published peak
  cpu_baseline   the CPU oracle (reference algorithm) on a bounded sample,
                 extrapolated to the same workload
"""
import argparse
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


def cpu_baseline(wl, perms_run, seconds, threads):
    """Oracle (reference algorithm, full LOO tally per network and pair) on the
    first (set, permutation) items until `seconds` of CPU work, extrapolated by
    the reference's cost S*c*n per item to the whole step."""
    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ref

    data_ref = wl.X.astype(np.int32) - 1
    N = wl.N
    rng = np.random.default_rng(0)
    done_cost, t0, items = 0.0, time.perf_counter(), 0
    for s, (genes, _) in enumerate(wl.sets):
        if len(genes) < 2:
            continue
        prior = wl.prior_matrix(s)
        lab = wl.classes if items == 0 else rng.permutation(wl.classes)
        t1, t2, n1, n2 = oracle_ref.class_split(data_ref, genes, lab)
        prm = oracle_ref.params(mp=wl.cfg.mp, threads=threads)
        oracle_ref.run_perm(t1, t2, n1, n2, prior, prm, want_bits=False)
        G = len(genes)
        done_cost += (N + 2) * (G * (G - 1) // 2) * N
        items += 1
        if time.perf_counter() - t0 > seconds and items >= 2:
            break
    dt = time.perf_counter() - t0
    total_cost = sum(r * (N + 2) * (len(g) * (len(g) - 1) // 2) * N for (g, _), r in zip(wl.sets, perms_run))
    t_total = dt * total_cost / done_cost
    return {
        "value": len(wl.sets) / t_total,
        "unit": "gene sets/s",
        "cores": threads,
        "kind": "port",
        "sample": "%d (set, permutation) items of the same workload in %.1f s with %d OpenMP threads, "
                  "extrapolated by S*c*n to %d sets at r=%d" % (items, dt, threads, len(wl.sets), wl.cfg.perms),
    }


def workflow4(args):
    """Config 4's workflow (uneven classes 300/900, 20000 genes, 2000 sets of 5..500
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
    print("bench: config 4, %d sets, N=%d (%d/%d), setup %.1fs" % (len(wl.sets), wl.N, wl.cfg.nA, wl.cfg.nB,
                                                                   setup_s), file=sys.stderr, flush=True)
    eng.reset_stream()
    eng.run_learned(wl.sets, prm, cfg)
    eng.reset_stream()
    t0 = time.perf_counter()
    res, rep = eng.run_learned(wl.sets, prm, cfg)
    t1 = time.perf_counter()
    chosen, res2 = eng.repermute(wl.sets, res, prm, top=200, perms=1000)
    t2 = time.perf_counter()
    print("bench: r=100 phase %.2fs, r=1000 phase %.2fs" % (t1 - t0, t2 - t1), file=sys.stderr, flush=True)
    G = [len(wl.sets[q][0]) for q in chosen]
    line = {
        "metric": "config-4 workflow: gene sets/s at r=100 (batched+learned), then r=1000 on the top 200",
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
    }
    print(json.dumps(line))


def run_step(eng, wl, prm, cfg):
    eng.reset_stream()
    res, rep = eng.run_learned(wl.sets, prm, cfg)
    return res, rep, eng.timing()


def main():
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
    ap.add_argument("--plain-sets", type=int, default=10)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-compare", action="store_true")
    ap.add_argument("--compare-rounds", type=int, default=5,
                    help="interleaved rounds of the learned job, the job without the network and the deferred form")
    ap.add_argument("--workflow4", action="store_true",
                    help="run config 4's workflow (r=100 over 2000 sets, then r=1000 on the top 200) instead")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch shapes instead of running the autotuner in setup")
    ap.add_argument("--launch", default="",
                    help="fixed launch shapes after tuning, e.g. graph=1,loo_waves=3 (results never change)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="first run one device-to-device copy of MIB MiB (a kernel of known traffic for "
                         "scaling hardware byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this engine runs on one GPU; multi-GPU sharding is not built (see DESIGN.md)",
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
    kern_ms = [0.0] * 6
    launches = [0] * 6
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
        for k in range(6):
            kern_ms[k] += tm.kernel_ms[k]
            launches[k] += tm.launches[k]
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
            ures = eng.run_sets(wl.sets, uprm)
            tu.append(time.perf_counter() - t0)
            same = same and same_as_learned(ures)
            eng.reset_stream()
            t0 = time.perf_counter()
            zres = eng.run_sets(wl.sets, zprm)
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
    # test) on a prefix of the sets; its time is extrapolated to the whole
    # workload by the reference's cost S*c*r per set
    plain = None
    if args.plain_sets > 0:
        sub = wl.sets[:args.plain_sets]
        pprm = wl.params(mode=eddy_amd.MODE_PLAIN, schedule=eddy_amd.TESTS_ALL, learned=0)
        eng.reset_stream()
        tp = time.perf_counter()
        pres = eng.run_sets(sub, pprm)
        dtp = time.perf_counter() - tp
        same = all((r.status, r.perms_run, r.U0) == (q.status, q.perms_run, q.U0) for r, q in zip(pres, res))
        for s in range(len(sub)):
            same = same and bool((eng.js(s).view("u8") == js_head[s].view("u8")).all())
        cost = [(wl.N + 2) * len(g) * (len(g) - 1) // 2 * r for (g, _), r in zip(wl.sets, perms_run)]
        frac = sum(cost[:len(sub)]) / max(1, sum(cost))
        plain_rate = nsets / (dtp / max(frac, 1e-12))
        plain = {"sets": len(sub), "seconds": dtp, "cost_fraction_of_workload": frac,
                 "plain_sets_per_s_extrapolated": plain_rate, "speedup_batched_learned_vs_plain": value / plain_rate,
                 "results_bit_identical": same}

    # roofline of the dominant kernel
    dom = max(range(6), key=lambda k: kern_ms[k])
    dom_avg_s = kern_ms[dom] / max(1, launches[dom]) / 1000.0
    per_launch = lambda total: total / max(1, launches[dom])
    if dom == 0:
        achieved = per_launch(ops * args.steps) / dom_avg_s
        roof = {"kernel": "k_count_test", "bound": "mfma", "achieved": achieved / 1e12, "peak": PEAK_INT8_OPS / 1e12,
                "unit": "TOP/s", "frac": achieved / PEAK_INT8_OPS, "traffic": None}
    elif dom == 3:
        # BDeu histograms: (1 + |pa|) state bytes per sample and class, plus
        # 4 q r bytes of bins, per distinct family and scoring class
        achieved = per_launch(stats[0] + stats[4]) / dom_avg_s
        # the score interval spans k_score, k_bdeu and k_score_tail
        roof = {"kernel": "k_score+k_bdeu+k_score_tail", "bound": "hbm", "achieved": achieved / 1e9,
                "peak": PEAK_HBM / 1e9,
                "unit": "GB/s", "frac": achieved / PEAK_HBM, "traffic": None}
    elif dom == 1:
        # leave-one-out tests: each planned (pair, class) entry reads its 9 u16 cells and 4 B of info
        achieved = per_launch(sched[2] * 22) / dom_avg_s
        roof = {"kernel": "k_loo", "bound": "hbm", "achieved": achieved / 1e9, "peak": PEAK_HBM / 1e9,
                "unit": "GB/s", "frac": achieved / PEAK_HBM, "traffic": None}
    else:
        roof = {"kernel": eddy_amd.KERNELS[dom], "bound": "hbm", "achieved": None, "peak": PEAK_HBM / 1e9,
                "unit": "GB/s", "frac": None, "traffic": None}
    roof["avg_launch_ms"] = dom_avg_s * 1000.0
    roof["launches_per_step"] = launches[dom] / args.steps
    # HBM bytes of the same kernel from the committed byte-counter profile of
    # this workload (FETCH_SIZE + WRITE_SIZE, copy-calibrated): its bytes per
    # step over this run's launches per step (chunk boundaries follow the
    # stream replay's timing, the work per step does not)
    tpath = os.path.join(ROOT, "profiles", "hbm_traffic_learned_job_5000x600.json")
    if args.config == 3 and args.nsets is None and os.path.exists(tpath):
        kern = json.load(open(tpath))["kernels"]
        parts = [kern.get(k) for k in roof["kernel"].split("+")]
        if all(parts):
            roof["traffic"] = sum(q["bytes_per_step"] for q in parts) / max(1e-9, roof["launches_per_step"])
            roof["traffic_source"] = os.path.relpath(tpath, ROOT)

    cpu = None
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        threads = max(1, min(threads, 16))
        cpu = cpu_baseline(wl, perms_run, args.cpu_seconds, threads)

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
        "kernel_ms_per_step": {eddy_amd.KERNELS[k]: kern_ms[k] / args.steps for k in range(6)},
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
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
