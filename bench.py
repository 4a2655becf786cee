#!/usr/bin/env python3
"""Benchmark: kernel-PC dcc.gamma skeleton (CI tests/s) on one MI355X.

One step = one full skeleton run (levels, sepsets, deletion trace) of the
dcc.gamma kernel-PC path over the synthetic nonlinear SEM n=351, p=48
(edge probability 0.08, per-node noise variance U[0.25, 1], seed 2),
alpha = 0.1, numCol = 35, index = 1, m.max = Inf.  Each timed step starts
from the data on the host: the upload and the per-call arenas are inside the
step; the skeleton call is synchronous (it returns after the last level's
p-values are back on the host).  The result caches are off (every step runs
every test).

Reported beside the batched form:
  * plain form -- one launch sequence per CI test, measured on the first
    --plain-per-level tests of every level and extrapolated with the level's
    test count;
  * roofline of the dominant kernel from HIP-event kernel times;
  * cpu_baseline -- the C oracle (a port of the reference CPU path) running the
    whole skeleton of the same workload on this host's cores (OpenMP), with a
    parity line against the GPU run, and a 1-thread per-level sample.
Prints ONE JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6  # MI355X dense FP64 (vector and matrix)
HBM_PEAK_GBS = 8000.0


def design_dims(level):
    """(q, penalties, rank per penalty) of the regrXonS design at |S| = level."""
    if level == 1:
        return 10, 1, 8
    if level == 2:
        return 30, 1, 27
    return 1 + 9 * level, level, 8


def gcv_flops(n, level, fits, evals):
    """Algorithmic FLOPs of the GCV stage: per target Q'y (4nq) and the final fit
    (2nq), per objective evaluation q^3/3 (pivoted Cholesky of the aggregate
    penalty) + 4(q+r)q^2 + 8q^3 (SVD of the (q+r) x q augmented system) +
    q^2 r m (derivatives)."""
    if level == 0:
        return 0.0
    q, m, rk = design_dims(level)
    r = min(q, m * rk)
    per_eval = q ** 3 / 3.0 + 4.0 * (q + r) * q * q + 8.0 * q ** 3 + q * q * r * m
    return fits * (6.0 * n * q) + evals * per_eval


def component_flops(n, k, comps):
    """Algorithmic FLOPs per dcov component: (4/3)n^3 (tridiagonal reduction)
    + 2n^2 k (eigenvector back-transform) + 3n^2 (distances, centring)."""
    return comps * (4.0 / 3.0 * n ** 3 + 2.0 * n * n * k + 3.0 * n * n)


def pair_flops(n, k, pairs):
    return pairs * (2.0 * k * k * n + 3.0 * k * k)


def read_traffic(kernel, workload_tag):
    """Calibrated HBM bytes per launch of `kernel` from the committed counter
    profile, when that profile was taken on this workload (else None)."""
    path = os.path.join(ROOT, "profiles", "traffic_%s.json" % kernel.replace("+", "_"))
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        d = json.load(fh)
    if workload_tag not in d.get("workload", ""):
        return None
    return d.get("hbm_bytes_per_launch")


def host_info():
    """CPU model, core counts and the threads this process may use."""
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    affinity = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS")
    threads = min(affinity, int(env)) if env and env.isdigit() and int(env) > 0 else affinity
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity_cpus": affinity,
            "threads": threads}


def cpu_baseline_full(data, res, k, threads):
    """The whole skeleton on the C oracle (a port of the reference CPU path:
    kpcalg/mgcv/pcalg restated in C), OpenMP over each level's residual fits
    and dcov components, as BASELINE.md's CPU plan asks; plus its parity line
    against the GPU run."""
    from oracle import orc
    n, p = data.shape
    t0 = time.perf_counter()
    ref = orc.skeleton(data, alpha=0.1, m_max=p - 2, num_col=k, threads=threads, method="dcc.gamma")
    wall = time.perf_counter() - t0
    tests = int(sum(ref["n.edgetests"]))
    same = bool(np.array_equal(ref["adjacency"], res["adjacency"]) and
                ref["n.edgetests"] == res["n.edgetests"])
    mine = dict(zip(zip(res["tasks"]["level"].tolist(), res["tasks"]["x"].tolist(),
                        res["tasks"]["y"].tolist(), res["tasks"]["S_key"]),
                    res["tasks"]["p_used"].tolist()))
    worst = 0.0
    for key, pv in zip(zip(ref["tasks"]["level"].tolist(), ref["tasks"]["x"].tolist(),
                           ref["tasks"]["y"].tolist(), ref["tasks"]["S_key"]),
                       ref["tasks"]["p_used"].tolist()):
        if key in mine:
            worst = max(worst, abs(mine[key] - pv) / max(abs(pv), 1e-300))
    return tests / wall, wall, same, worst


def cpu_baseline(data, res, n, k, seconds_budget):
    """Time the C oracle on a per-level sample of the workload's residual fits
    and dcov components (each level gets an equal share of the time budget);
    extrapolate with the GPU run's per-level fit/component counts."""
    from oracle import orc
    s = res["summary"]
    tasks = res["tasks"]
    levels = res["levels"]
    active = [lev for lev in range(levels) if np.any(tasks["level"] == lev)]
    share = seconds_budget / max(1, len(active))
    total = 0.0
    sampled = []
    t_start = time.time()
    for lev in active:
        idx = np.nonzero(tasks["level"] == lev)[0]
        fit_t, comp_t, nf, nc = 0.0, 0.0, 0, 0
        t_lev = time.time()
        for i in idx:
            x = int(tasks["x"][i]) - 1
            S = [int(v) - 1 for v in tasks["S_key"][i].split("|")] if tasks["S_key"][i] else []
            if S:
                t0 = time.perf_counter()
                r, _ = orc.regrxons(data, x, S)
                fit_t += time.perf_counter() - t0
                nf += 1
            else:
                r = np.ascontiguousarray(data[:, x])
            t0 = time.perf_counter()
            orc.dcov_component(r, k)
            comp_t += time.perf_counter() - t0
            nc += 1
            if time.time() - t_lev > share:
                break
        est = (fit_t / nf * s["level_fits"][lev] if nf else 0.0) + \
            comp_t / max(1, nc) * s["level_components"][lev]
        total += est
        sampled.append("L%d:%dfit/%dcomp" % (lev, nf, nc))
    tests = int(sum(res["n.edgetests"]))
    return {
        "value": tests / total if total > 0 else None,
        "unit": "CI tests/s",
        "cores": 1,
        "kind": "port",
        "sample": "C oracle (kpcalg/mgcv/pcalg restatement), 1 thread, per-level sample %s of the "
                  "run's residual fits and dcov components, extrapolated with the run's per-level "
                  "fit/component counts (%.1f s of CPU work; est. skeleton %.1f s)"
                  % (",".join(sampled), time.time() - t_start, total),
    }


def bench_regression(args):
    """The standalone regrXonS batch: 1024 fits, n = 500, |S| ~ U{1..6},
    x ~ U[-2, 2], y = sum_k sin(2 x_k) + 0.3 x_k^2 + N(0, 0.3^2); lambda by the
    batched in-kernel Newton on the GCV score with the hand-written
    gradient/Hessian.  One step = one kpc_regrxons_mixed call (data upload,
    designs, one GCV launch per design width, residual download)."""
    import fastkpc_amd as fk
    from fastkpc_amd.sem import regression_batch
    data, targets, sets = regression_batch()
    n = data.shape[0]
    for _ in range(args.warmup):
        fk.regrxons_mixed(data, targets, sets)
    walls, setup_s, gcv_s, launches = [], 0.0, 0.0, 0
    info = None
    for _ in range(args.steps):
        t0 = time.perf_counter()
        _, info = fk.regrxons_mixed(data, targets, sets)
        walls.append(time.perf_counter() - t0)
        t = fk.last_fit_timing()
        setup_s += t["setup"]
        gcv_s += t["gcv"]
        launches += t["gcv_launches"]
    wall = float(np.mean(walls))
    fits = len(targets)
    lens = np.array([len(x) for x in sets])
    flops = sum(gcv_flops(n, int(lens[i]), 1, info[i, 5]) for i in range(fits))
    ktime = gcv_s / args.steps
    achieved = flops / ktime / 1e12
    # plain form: one call per fit on that fit's own columns
    plain = None
    if not args.no_plain:
        pick = [i for L in range(1, 7) for i in np.nonzero(lens == L)[0][:5]]
        t_plain = {}
        for i in pick:
            cols = sets[i] + [int(targets[i])]
            sub = np.asfortranarray(data[:, cols])
            t0 = time.perf_counter()
            fk.regrxons_batch(sub, [len(cols) - 1], [list(range(len(cols) - 1))])
            t_plain.setdefault(int(lens[i]), []).append(time.perf_counter() - t0)
        est = sum(np.mean(t_plain[L]) * np.sum(lens == L) for L in t_plain)
        plain = {"fits_per_s": round(fits / est, 2), "batch_wall_s_est": round(est, 3),
                 "sampled_fits": len(pick), "batched_over_plain": round(est / wall, 2)}
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import orc
        share = args.cpu_baseline_seconds / 6.0
        est, sampled = 0.0, []
        t_start = time.time()
        for L in range(1, 7):
            idx = np.nonzero(lens == L)[0]
            tot, k = 0.0, 0
            for i in idx:
                t0 = time.perf_counter()
                orc.regrxons(data, int(targets[i]), sets[i])
                tot += time.perf_counter() - t0
                k += 1
                if tot > share:
                    break
            est += tot / k * len(idx)
            sampled.append("|S|=%d:%d" % (L, k))
        cpu = {"value": round(fits / est, 2), "unit": "fits/s", "cores": 1, "kind": "port",
               "sample": "C oracle (mgcv magic restatement), 1 thread, %s fits timed (%.1f s), "
                         "extrapolated per |S| (est. %.1f s for the batch)"
                         % (",".join(sampled), time.time() - t_start, est)}
    line = {
        "metric": "regrXonS fits/s, 1024 fits n=500 |S|~U{1..6}: batched adjoint-Newton GCV",
        "value": round(fits / wall, 2),
        "unit": "fits/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (fastkpc_amd/sem.py regression_batch, seed 4)",
        "config": {"workload": "regrXonS batch: 1024 fits, n=500, |S|~U{1..6}, x~U[-2,2], "
                               "y=sum sin(2x)+0.3x^2+N(0,0.3^2)",
                   "fits": fits, "n": n, "fits_per_S": np.bincount(lens)[1:].tolist(),
                   "evaluations": int(info[:, 5].sum()),
                   "setup_s_per_step": round(setup_s / args.steps, 4),
                   "gcv_s_per_step": round(ktime, 4)},
        "plain_gpu": plain,
        "roofline": {"bound": "mfma", "kernel": "magic_kernel+magic_slots_kernel", "achieved": round(achieved, 4),
                     "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / FP64_PEAK_TFLOPS, 6),
                     "traffic": None, "launches_per_step": launches / args.steps,
                     "avg_launch_us": round(ktime / max(1, launches / args.steps) * 1e6, 1),
                     "algorithmic_flops_per_launch": flops / max(1, launches / args.steps)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


def copy_calibration(nbytes):
    """Device-to-device copy of nbytes through the HIP runtime: the reference
    kernel for scaling FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a = ctypes.c_void_p()
    b = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0  # DeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_calibration_bytes": nbytes, "copies": 3}))


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="sem351x48",
                    help="sem351x48 (headline), sem200x8, sem500x50, sem2000x64, or regr1024 "
                         "(the standalone 1024-fit regrXonS batch)")
    ap.add_argument("--method", default="dcc.gamma",
                    choices=["dcc.gamma", "hsic.gamma", "dcc.perm", "hsic.perm"],
                    help="CI method (the headline metric is dcc.gamma)")
    ap.add_argument("--replicates", type=int, default=100,
                    help="permutation replicates for dcc.perm / hsic.perm")
    ap.add_argument("--plain-per-level", type=int, default=30)
    ap.add_argument("--cpu-baseline-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="only run a device copy of this many bytes (counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print("bench.py measures the single-GPU path only: this build has no multi-GPU execution "
              "(one MI355X per box); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.config == "regr1024":
        bench_regression(args)
        return

    import fastkpc_amd as fk
    from fastkpc_amd.sem import config_data

    data, mask, cfg = config_data(args.config)
    n, p = data.shape
    k = 35
    opts = {"numCol": k, "ci_method": args.method, "index": 1}
    if args.method.endswith("perm"):
        opts["permutation_params"] = {"replicates": args.replicates, "seed": 2024}
    for _ in range(args.warmup):
        fk.fastkpc_compatible_cuda_skeleton(data, 0.1, options=opts)
    walls = []
    res = None
    stage_tot = {}
    launch_tot = {}
    level_evals = None
    for _ in range(args.steps):
        t0 = time.perf_counter()
        res = fk.fastkpc_compatible_cuda_skeleton(data, 0.1, options=opts)
        walls.append(time.perf_counter() - t0)
        s = res["summary"]
        for key, v in s["stage_seconds"].items():
            stage_tot[key] = stage_tot.get(key, 0.0) + v
        for key, v in s["stage_launches"].items():
            launch_tot[key] = launch_tot.get(key, 0) + v
    s = res["summary"]
    tests = int(sum(res["n.edgetests"]))
    wall = float(np.mean(walls))
    value = tests / wall

    # ---- roofline of the dominant kernel (HIP events inside the library)
    kernels = {
        # the GCV stage: magic_slots_kernel<m> (additive designs, |S| = 3, 4)
        # and magic_kernel (every other design)
        "gcv": ("magic_kernel+magic_slots_kernel",
                sum(gcv_flops(n, l, s["level_fits"][l], s["level_evaluations"][l])
                    for l in range(res["levels"]))),
    }
    if args.method == "dcc.gamma":
        kernels["dcov_component"] = ("dcov_lanczos_kernel",
                                     component_flops(n, k, sum(s["level_components"])))
        kernels["dcov_pair"] = ("dcov_pair_kernel", pair_flops(n, k, sum(s["level_physical"])))
    elif args.method == "dcc.perm":
        # per replicate (and the observed sum): the upper triangle of 64 x 64
        # tiles, 10 flops per (i, j) (two centred distances and a product)
        ns = (n + 63) // 64
        per = (args.replicates + 1) * ns * (ns + 1) / 2 * 64 * 64 * 10.0
        kernels["dcov_pair"] = ("dcc_perm_kernel", per * sum(s["level_physical"]))
    dom = max(kernels, key=lambda key: stage_tot.get(key, 0.0))
    kname, flops_per_run = kernels[dom]
    ktime = stage_tot[dom] / args.steps
    launches = launch_tot[dom] / args.steps
    if ktime <= 0.0:
        # KPC_GRAPH=1: each wave is one graph and only the whole pipeline is
        # timed, so no kernel's own time is available for the roofline
        print(json.dumps({"metric": "CI tests/s (KPC_GRAPH=1 wave graphs)", "value": round(value, 2),
                          "unit": "CI tests/s", "ms_per_step": round(wall * 1e3, 3),
                          "graph_seconds_per_step": s["graph_seconds"],
                          "graph_launches_per_step": s["graph_launches"]}))
        return
    achieved = flops_per_run / ktime / 1e12
    roofline = {
        # FP64 compute roof (vector and matrix FP64 share the 78.6 TFLOP/s
        # peak; the path's HBM bytes are far below the HBM roof).  The
        # kernels' limiter is the latency of their sequential steps: the
        # occupancy / stall counters are in profiles/counters_sem351x48.md
        "bound": "mfma",
        "limiter": "latency (sequential optimiser and Lanczos steps; waves wait in 60-68% of "
                   "their resident cycles)",
        "kernel": kname,
        "achieved": round(achieved, 4),
        "peak": FP64_PEAK_TFLOPS,
        "unit": "TFLOP/s",
        "frac": round(achieved / FP64_PEAK_TFLOPS, 6),
        "traffic": read_traffic(kname, "%s %s" % (args.method, args.config)),
        "launches_per_step": launches,
        "avg_launch_us": round(ktime / max(1, launches) * 1e6, 1),
        "algorithmic_flops_per_launch": flops_per_run / max(1, launches),
    }

    # the path's algorithmic FLOPs per skeleton (GCV + components + pairs),
    # against the wall time of each form
    path_flops = sum(v[1] for v in kernels.values())
    path_roofline = {"bound": "mfma", "unit": "TFLOP/s", "peak": FP64_PEAK_TFLOPS,
                     "algorithmic_flops": path_flops,
                     "achieved": round(path_flops / wall / 1e12, 5),
                     "frac": round(path_flops / wall / 1e12 / FP64_PEAK_TFLOPS, 7)}

    # ---- plain form (one launch sequence per test), stratified per level
    plain = None
    if not args.no_plain:
        popts = dict(opts, mode="plain", plain_test_limit=args.plain_per_level)
        rp = fk.fastkpc_compatible_cuda_skeleton(data, 0.1, options=popts)
        sp = rp["summary"]
        est = 0.0
        for lev in range(rp["levels"]):
            if sp["plain_tests"][lev] > 0:
                est += rp["n.edgetests"][lev] * sp["plain_seconds"][lev] / sp["plain_tests"][lev]
        same = bool(np.array_equal(rp["adjacency"], res["adjacency"]) and
                    rp["n.edgetests"] == res["n.edgetests"])
        plain = {
            "tests_per_s": round(tests / est, 2) if est > 0 else None,
            "skeleton_wall_s_est": round(est, 3),
            "sampled_tests": int(sum(sp["plain_tests"])),
            "same_skeleton": same,
            "batched_over_plain": round(value / (tests / est), 2) if est > 0 else None,
            "roofline": {"bound": "mfma", "unit": "TFLOP/s", "peak": FP64_PEAK_TFLOPS,
                         "algorithmic_flops": path_flops,
                         "achieved": round(path_flops / est / 1e12, 6) if est > 0 else None,
                         "frac": round(path_flops / est / 1e12 / FP64_PEAK_TFLOPS, 8)
                         if est > 0 else None,
                         "note": "whole-path FP64 work of the skeleton over the plain form's "
                                 "estimated wall time"},
        }

    cpu = None
    if not args.no_cpu_baseline and args.method == "dcc.gamma":
        host = host_info()
        rate, cwall, same, worst = cpu_baseline_full(data, res, k, host["threads"])
        single = cpu_baseline(data, res, n, k, args.cpu_baseline_seconds)
        cpu = {
            "value": round(rate, 2),
            "unit": "CI tests/s",
            "cores": host["threads"],
            "kind": "port",
            "sample": "whole skeleton of the same workload on the C oracle (kpcalg/mgcv/pcalg "
                      "restatement), OpenMP over each level's fits and components, %d threads on "
                      "%s; %.1f s wall; baseline only" % (host["threads"], host["cpu_model"], cwall),
            "skeleton_wall_s": round(cwall, 3),
            "same_skeleton_as_gpu": same,
            "max_rel_p_diff_vs_gpu": worst,
            "host": host,
            "single_thread": {"value": round(single["value"], 2), "cores": 1,
                              "sample": single["sample"]},
        }

    line = {
        "metric": ("CI tests/s + skeleton wall-s, 351x48 dcc.gamma: batched+adjoint vs plain GPU"
                   if args.method == "dcc.gamma" and args.config == "sem351x48" else
                   "CI tests/s, %s %s" % (args.config, args.method)),
        "value": round(value, 2),
        "unit": "CI tests/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic nonlinear SEM (generator fastkpc_amd/sem.py, seed %d)" % cfg["seed"],
        "config": {
            "workload": "kernel-PC skeleton, %s, %s: n=%d p=%d alpha=0.1 numCol=35 index=1 "
                        "m.max=Inf%s" % (args.method, args.config, n, p,
                                         " replicates=%d" % args.replicates
                                         if args.method.endswith("perm") else ""),
            "n": n, "p": p, "ci_method": args.method, "alpha": 0.1, "numCol": k,
            "m_max_resolved": p - 2, "form": "batched (frontier-speculative level waves)",
            "logical_tests": tests, "n_edgetests": res["n.edgetests"],
            "physical_tests": int(s["physical_tests"]), "levels": res["levels"],
            "skeleton_wall_s": round(wall, 4),
            "stage_seconds_per_step": {kk: round(v / args.steps, 4) for kk, v in stage_tot.items()},
        },
        "plain_gpu": plain,
        "roofline": roofline,
        "path_roofline": path_roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
