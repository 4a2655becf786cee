"""Python mirror of the reference's R-facing API for the kernel-PC CI path.

`fastkpc_compatible_cuda_skeleton` mirrors fastkpc/R/fast_kpc.R:56-130 (same
option names, m.max resolution fastkpc/R/cuda_native.R:7-27, result fields
fastkpc/src/full_cuda_ci_one_call.cpp:5597-5640, 6008-6030); `kernelCItest`
mirrors kpcalg/R/kernelCItest.R:68-129 for one test; `regrXonS` mirrors
kpcalg/R/regrXonS.R:33-61.  All computation runs in libkpcmi.so on the GPU.
"""
import math

import numpy as np

from . import _lib
from ._lib import KpcOptions, KpcResult, check, dptr, iptr

CI_METHODS = {"dcc.gamma": 0, "hsic.gamma": 1, "dcc.perm": 2, "hsic.perm": 3}
ALLOWED_OPTIONS = ("max_conditioning_size", "index", "numCol", "trace_level", "dcov_batch",
                   "mgcv_residual_backend", "mgcv_residual_backend_native_s_size_limit",
                   "mgcv_residual_backend_condition_threshold", "route",
                   "compatible_cuda_strict", "ci_method", "hsic_params", "permutation_params",
                   "mode", "plain_test_limit", "cache")


def resolve_max_conditioning_size(max_conditioning_size, p):
    """fastkpc/R/cuda_native.R:7-27: Inf -> p - 2, else min(floor(m), p - 2)."""
    m = max_conditioning_size
    if (not isinstance(m, (int, float)) or isinstance(m, bool) or (isinstance(m, float) and math.isnan(m))
            or m < 0 or not isinstance(p, int) or p < 2):
        raise ValueError("max_conditioning_size must be one non-negative number and p must be an integer >= 2")
    if math.isinf(m):
        return p - 2
    return int(min(math.floor(m), p - 2))


def _options(n, p, alpha, options):
    unknown = [k for k in options if k not in ALLOWED_OPTIONS]
    if unknown:
        raise ValueError("unknown compatible CUDA option(s): " + ",".join(sorted(unknown)))
    o = KpcOptions()
    _lib.lib().kpc_default_options(o)
    o.alpha = float(alpha)
    o.max_conditioning_size = resolve_max_conditioning_size(
        options.get("max_conditioning_size", float("inf")), p)
    o.index = float(options.get("index", 1))
    o.num_col = int(options.get("numCol", n // 10))
    method = options.get("ci_method", "dcc.gamma")
    if method not in CI_METHODS:
        raise ValueError("ci_method must be one of " + ", ".join(CI_METHODS))
    o.ci_method = CI_METHODS[method]
    o.hsic_sig = float(options.get("hsic_params", {}).get("sig", 1.0))
    perm = options.get("permutation_params", {})
    o.replicates = int(perm.get("replicates", 100))
    o.include_observed = 1 if perm.get("include_observed", True) else 0
    if perm.get("seed") is not None:
        o.has_seed = 1
        o.seed = int(perm["seed"])
    mode = options.get("mode", "batched")
    if mode not in ("batched", "plain"):
        raise ValueError("mode must be 'batched' or 'plain'")
    o.mode = 0 if mode == "batched" else 1
    o.plain_test_limit = int(options.get("plain_test_limit", 0))
    # permutation indices: "philox" -- the counter-based generator keyed by
    # each test's seed (default); "R" -- R's global Mersenne-Twister stream
    # after set.seed(seed) (or from r_state, .Random.seed[2:626]), drawn test
    # by test in pcalg's execution order, as kpcalg and the reference's strict
    # route draw them (fastkpc/src/full_cuda_ci_one_call.cpp:464-567)
    rng = perm.get("rng", "philox")
    if rng not in ("philox", "R"):
        raise ValueError("permutation_params$rng must be 'philox' or 'R'")
    o.perm_stream = 1 if rng == "R" else 0
    # the process-wide result and target-state caches (the reference's
    # one-call route always uses them); off unless asked for
    o.use_cache = 1 if options.get("cache", False) else 0
    if rng == "R" and perm.get("r_state") is not None:
        st = np.ascontiguousarray(np.asarray(perm["r_state"], dtype=np.int64).astype(np.uint32)
                                  .view(np.int32))
        if st.shape != (625,):
            raise ValueError("r_state must hold 625 integers (.Random.seed[2:626])")
        o._r_state_keep = st  # keeps the array alive while the options are in use
        o.r_state = iptr(st)
    return o


def _as_data(data):
    d = np.asarray(data, dtype=np.float64)
    if d.ndim != 2:
        raise ValueError("data must be a 2-D matrix")
    return np.asfortranarray(d)


def fastkpc_compatible_cuda_skeleton(data, alpha, labels=None, options=None):
    options = dict(options or {})
    d = _as_data(data)
    n, p = d.shape
    if labels is None:
        labels = ["V%d" % (i + 1) for i in range(p)]
    if len(labels) != p:
        raise ValueError("labels length must match number of columns")
    route = options.get("route", "full_cuda")
    if route != "full_cuda":
        raise ValueError("only route='full_cuda' is provided by this library")
    o = _options(n, p, alpha, options)
    res = _lib.ctypes.POINTER(KpcResult)()
    check(_lib.lib().kpc_skeleton_run(dptr(d), n, p, o, _lib.ctypes.byref(res)))
    try:
        r = res.contents
        adj = np.ctypeslib.as_array(r.adjacency, (p * p,)).reshape(p, p).astype(bool).copy()
        pmax = np.ctypeslib.as_array(r.pmax, (p * p,)).reshape(p, p).copy()
        slen = np.ctypeslib.as_array(r.sepset_len, (p * p,)).reshape(p, p).copy()
        sep = np.ctypeslib.as_array(r.sepset, (p * p * 64,)).reshape(p, p, 64)
        smax = max(int(slen.max()), 0)
        sp1 = (sep[:, :, :smax] + 1).tolist()
        sepsets = [[None if k < 0 else s[:k] for s, k in zip(srow, krow)]
                   for srow, krow in zip(sp1, slen.tolist())]
        L = r.n_tasks
        if L > 0:
            ts = np.ctypeslib.as_array(r.task_slen, (L,)).copy()
            tasks = {
                "level": np.ctypeslib.as_array(r.task_level, (L,)).copy(),
                "x": np.ctypeslib.as_array(r.task_x, (L,)).copy() + 1,
                "y": np.ctypeslib.as_array(r.task_y, (L,)).copy() + 1,
                "S_key": _lib.ctypes.string_at(r.task_keys).decode().split("\n"),
                "conditioning_size": ts,
                "p_used": np.ctypeslib.as_array(r.task_p, (L,)).copy(),
                "native_edge_deleted": np.ctypeslib.as_array(r.task_deleted, (L,)).astype(bool).copy(),
                "decision_margin": np.ctypeslib.as_array(r.task_margin, (L,)).copy(),
            }
            # trace rows in plan order over the consumed tests, numbered across
            # levels, and each row's 1-based position in its level's plan
            # (fastkpc/src/full_cuda_ci_one_call.cpp:5535-5537)
            tasks["canonical_test_order_id"] = np.arange(1, L + 1)
            tasks["task_index"] = np.ctypeslib.as_array(r.task_index, (L,)).copy()
            stat = np.ctypeslib.as_array(r.task_stat, (L,)).copy()
            mean = np.ctypeslib.as_array(r.task_mean, (L,)).copy()
            var = np.ctypeslib.as_array(r.task_var, (L,)).copy()
            seeds = np.ctypeslib.as_array(r.task_perm_seed, (L,)).astype(np.uint64)
        else:
            tasks = {"level": np.zeros(0, int), "x": np.zeros(0, int), "y": np.zeros(0, int),
                     "S_key": [], "conditioning_size": np.zeros(0, int), "p_used": np.zeros(0),
                     "native_edge_deleted": np.zeros(0, bool), "decision_margin": np.zeros(0)}
            tasks["canonical_test_order_id"] = np.zeros(0, int)
            tasks["task_index"] = np.zeros(0, int)
            stat = mean = var = np.zeros(0)
            seeds = np.zeros(0, np.uint64)
        # the tested edge as (smaller, larger) 1-based index, whichever
        # orientation the test conditioned from (the reference trace's
        # edge_x / edge_y, fastkpc/src/full_cuda_ci_one_call.cpp:5601-5602)
        # every trace row is a consumed test, so none is ignored (the reference
        # writes a constant FALSE column, fastkpc/src/full_cuda_ci_one_call.cpp:5610-5611)
        tasks["native_edge_ignored"] = np.zeros(len(tasks["x"]), bool)
        tasks["edge_x"] = np.minimum(tasks["x"], tasks["y"])
        tasks["edge_y"] = np.maximum(tasks["x"], tasks["y"])
        # the statistic and its null mean / variance of every row: the
        # reference trace carries them for hsic.gamma
        # (fastkpc/src/full_cuda_ci_one_call.cpp:5620-5627); here every method
        # has them (dcc.gamma: nV^2 and its gamma moments; *.perm: the observed
        # statistic and the replicates' mean / variance)
        tasks["method_statistic"] = stat
        tasks["method_mean"] = mean
        tasks["method_variance"] = var
        method = options.get("ci_method", "dcc.gamma")
        if method in ("dcc.perm", "hsic.perm"):
            # fastkpc/src/full_cuda_ci_one_call.cpp:5613-5616 (uint64 here:
            # the seed of the counter-based index generator)
            tasks["permutation_seed"] = seeds
        levels = r.levels
        # the reference's levels data frame (fastkpc/src/full_cuda_ci_one_call.cpp:5632-5640)
        lv, dl = tasks["level"], tasks["native_edge_deleted"]
        level_rows = {
            "level": list(range(levels)),
            "tasks_planned": [int(v) for v in list(r.level_tasks_planned)[:levels]],
            "tests_replayed": [int(v) for v in list(r.n_edgetests)[:levels]],
            "tasks_ignored_after_delete": [int(v) for v in list(r.level_tasks_ignored)[:levels]],
            "deletions": [int(np.count_nonzero(dl[lv == l])) for l in range(levels)],
            "rounds": [int(v) for v in list(r.level_rounds)[:levels]],
            "elapsed_ms": [1e3 * v for v in list(r.level_seconds)[:levels]],
        }
        # the reference's summary list, core fields first
        # (fastkpc/src/full_cuda_ci_one_call.cpp:6017-6060), then this
        # library's own counters
        summary = {
            "run_status": "ok",
            "entrypoint": "kpc_skeleton_run",
            "compatible_cuda_route": "compatible.cuda",
            "compatible_cuda_strict": True,
            "native_call_count": 1,
            "p": p, "n": n, "alpha": float(alpha),
            "index": float(o.index),
            "numCol": int(o.num_col),
            "ci_method": method,
            "ci_backend": "kpcmi-gfx950",
            "hsic_sig": float(o.hsic_sig),
            "permutation_replicates": int(o.replicates) if method in ("dcc.perm", "hsic.perm") else 0,
            "permutation_include_observed": bool(o.include_observed),
            "permutation_has_seed": bool(o.has_seed),
            "permutation_seed": float(o.seed) if o.has_seed else float("nan"),
            "scheduler": "kpcmi-frontier-waves",
            "frontier_batch_count": int(sum(level_rows["rounds"])),
            "levels": levels,
            "tasks_planned": int(sum(level_rows["tasks_planned"])),
            "logical_tests_consumed": int(sum(level_rows["tests_replayed"])),
            "physical_tests_evaluated": int(r.physical_tests),
            "speculative_tests_ignored": int(r.physical_tests) - int(sum(level_rows["tests_replayed"])),
            "tasks_ignored_after_delete": int(sum(level_rows["tasks_ignored_after_delete"])),
            "deletions": int(sum(level_rows["deletions"])),
            "r_callback_count": 0,
            "result_cache_hit_count": int(r.result_cache_hits),
            "target_cache_hit_count": int(r.target_cache_hits),
            "mode": options.get("mode", "batched"),
            "max_conditioning_size_resolved": o.max_conditioning_size,
            "wall_seconds": r.wall_seconds,
            "stage_seconds": dict(zip(["smooth_setup", "geometry", "gcv", "dcov_component",
                                       "dcov_pair", "host"], list(r.stage_seconds)[:6])),
            "physical_tests": r.physical_tests,
            "residual_fits": r.residual_fits,
            "components": r.components,
            "setups": r.setups,
            "magic_evaluations": r.magic_evaluations,
            "level_tests": list(r.level_tests)[:levels],
            "level_physical": list(r.level_physical)[:levels],
            "level_fits": list(r.level_fits)[:levels],
            "level_components": list(r.level_components)[:levels],
            "level_evaluations": list(r.level_evaluations)[:levels],
            "level_lanczos_steps": list(r.level_lanczos_steps)[:levels],
            "level_seconds": list(r.level_seconds)[:levels],
            "level_stage_seconds": [dict(zip(["smooth_setup", "geometry", "gcv", "dcov_component",
                                              "dcov_pair"], list(r.level_stage_seconds[l])))
                                    for l in range(levels)],
            "plain_tests": list(r.plain_tests)[:levels],
            "plain_seconds": list(r.plain_seconds)[:levels],
            "stage_launches": dict(zip(["smooth_setup", "geometry", "gcv", "dcov_component",
                                        "dcov_pair"], list(r.stage_launches)[:5])),
            "graph_seconds": r.stage_seconds[7],
            "graph_launches": r.stage_launches[7],
            "labels": list(labels),
        }
        if o.perm_stream == 1 and method in ("dcc.perm", "hsic.perm"):
            # the stream after the run (.Random.seed[2:626]), as unsigned words
            summary["r_state_out"] = [int(v) & 0xFFFFFFFF for v in list(r.r_state_out)]
        return {
            "adjacency": adj,
            "sepsets": sepsets,
            "pMax": pmax,
            "n.edgetests": [int(v) for v in list(r.n_edgetests)[:levels]],
            "tasks": tasks,
            "levels": levels,
            "level_rows": level_rows,
            "summary": summary,
        }
    finally:
        _lib.lib().kpc_result_free(res)


def ci_batch(data, x, y, S, num_col=35, ci_method="dcc.gamma", alpha=0.1, sig=1.0,
             replicates=100, seed=None, index=1.0, include_observed=True, rng="philox",
             r_state=None, return_state=False):
    """Batch of CI tests (x_i, y_i | S_i) sharing |S| (0-based indices):
    returns (count, 4) {p, statistic, mean, variance}.  rng="R": the
    permutation tests draw from R's stream after set.seed(seed) (or r_state)
    in batch order, as consecutive kernelCItest calls in R; return_state also
    returns the stream afterwards (625 unsigned words)."""
    d = _as_data(data)
    n, p = d.shape
    x = np.ascontiguousarray(x, dtype=np.int32)
    y = np.ascontiguousarray(y, dtype=np.int32)
    S = np.ascontiguousarray(np.asarray(S, dtype=np.int32).reshape(len(x), -1))
    perm = {"replicates": replicates, "include_observed": include_observed, "rng": rng,
            "r_state": r_state}
    if seed is not None:
        perm["seed"] = seed
    o = _options(n, p, alpha, {"numCol": num_col, "ci_method": ci_method, "index": index,
                               "hsic_params": {"sig": sig}, "permutation_params": perm})
    out = np.zeros((len(x), 4))
    st = np.zeros(625, dtype=np.int32)
    check(_lib.lib().kpc_ci_batch_rs(dptr(d), n, p, o, len(x), iptr(x), iptr(y), S.shape[1],
                                     iptr(S), dptr(out), iptr(st)))
    if return_state:
        return out, [int(v) & 0xFFFFFFFF for v in st]
    return out


def kernelCItest(x, y, S, suffStat):
    """kpcalg/R/kernelCItest.R:68-129 with 1-based x, y, S; suffStat holds data,
    ic.method, index, numCol, sig, p (permutations) and, for the permutation
    methods, an optional integer seed of the counter-based index generator."""
    data = suffStat["data"]
    n = np.asarray(data).shape[0]
    method = suffStat.get("ic.method", "dcc.perm")
    num_col = suffStat.get("numCol", n // 10)
    S0 = sorted(int(s) - 1 for s in (S or []))
    out = ci_batch(data, [int(x) - 1], [int(y) - 1], [S0], num_col=num_col, ci_method=method,
                   sig=suffStat.get("sig", 1.0), replicates=suffStat.get("p", 100),
                   seed=suffStat.get("seed"), index=suffStat.get("index", 1.0))
    return float(out[0, 0])


def task_seed(user_seed, ci_method, x, y, S):
    """Seed of the permutation test ({x, y}, S) (0-based)."""
    Sa = np.ascontiguousarray(sorted(S), dtype=np.int32)
    out = _lib.ctypes.c_ulonglong()
    check(_lib.lib().kpc_task_seed(int(user_seed), CI_METHODS[ci_method], int(x), int(y),
                                   iptr(Sa), len(Sa), _lib.ctypes.byref(out)))
    return int(out.value)


def permutation_table(n, replicates, seeds):
    """Device-generated permutation indices: (len(seeds), replicates, n)."""
    sd = (_lib.ctypes.c_ulonglong * len(seeds))(*[int(s) for s in seeds])
    out = np.zeros((len(seeds), replicates, n), dtype=np.int32)
    check(_lib.lib().kpc_permutation_table(int(n), int(replicates), len(seeds), sd, iptr(out)))
    return out


def regrXonS(data, targets, S):
    """Residuals of each target (0-based) on conditioning set S (0-based, shared)."""
    d = _as_data(data)
    n, p = d.shape
    t = np.ascontiguousarray(targets, dtype=np.int32)
    Sm = np.ascontiguousarray(np.tile(np.asarray(sorted(S), dtype=np.int32), (len(t), 1)))
    resid = np.zeros((len(t), n))
    info = np.zeros((len(t), 32))
    check(_lib.lib().kpc_regrxons_batch(dptr(d), n, p, len(t), iptr(t), Sm.shape[1], iptr(Sm),
                                        dptr(resid), dptr(info)))
    return resid, info


def regrxons_batch(data, targets, S):
    """Residuals for per-fit conditioning sets S (count x s_len, 0-based)."""
    d = _as_data(data)
    n, p = d.shape
    t = np.ascontiguousarray(targets, dtype=np.int32)
    Sm = np.ascontiguousarray(np.asarray(S, dtype=np.int32).reshape(len(t), -1))
    resid = np.zeros((len(t), n))
    info = np.zeros((len(t), 32))
    check(_lib.lib().kpc_regrxons_batch(dptr(d), n, p, len(t), iptr(t), Sm.shape[1], iptr(Sm),
                                        dptr(resid), dptr(info)))
    return resid, info


def regrxons_mixed(data, targets, S_list):
    """Residuals and fit info of fits with per-fit conditioning sets of any
    size (0-based; each S sorted ascending): (count, n), (count, 32)."""
    d = _as_data(data)
    n, p = d.shape
    t = np.ascontiguousarray(targets, dtype=np.int32)
    if len(S_list) != len(t):
        raise ValueError("one conditioning set per target")
    lens = np.ascontiguousarray([len(s) for s in S_list], dtype=np.int32)
    offs = np.ascontiguousarray(np.concatenate([[0], np.cumsum(lens)[:-1]]), dtype=np.int32)
    flat = np.ascontiguousarray(np.concatenate([np.asarray(s, dtype=np.int32) for s in S_list]),
                                dtype=np.int32)
    resid = np.zeros((len(t), n))
    info = np.zeros((len(t), 32))
    check(_lib.lib().kpc_regrxons_mixed(dptr(d), n, p, len(t), iptr(t), iptr(lens), iptr(offs),
                                        iptr(flat), dptr(resid), dptr(info)))
    return resid, info


def tuning_report():
    """The launch-shape autotuner's table: {key: (chosen value or None,
    {candidate: seconds per replicate or None})}."""
    buf = _lib.ctypes.create_string_buffer(1 << 16)
    _lib.lib().kpc_tuning_report(buf, len(buf))
    out = {}
    for line in buf.value.decode().splitlines():
        parts = line.split()
        if len(parts) < 2:
            continue
        v = parts[1].split("=")[1]
        cands = {}
        for c in parts[2:]:
            a, b = c.split(":")
            cands[int(a)] = None if float(b) < 0 else float(b)
        out[parts[0]] = (None if v == "?" else int(v), cands)
    return out


def empty_cache():
    """Free the device blocks the library caches between calls."""
    check(_lib.lib().kpc_empty_cache())


def last_fit_timing():
    """Device seconds of the last regrXonS call: {"setup", "gcv", "gcv_launches"}."""
    out = np.zeros(3)
    check(_lib.lib().kpc_last_fit_timing(dptr(out)))
    return {"setup": float(out[0]), "gcv": float(out[1]), "gcv_launches": int(out[2])}


def gcv_grad_hess(data, targets, S, log_sp):
    """GCV score, gradient and Hessian in log(sp) (hand-written derivative)."""
    d = _as_data(data)
    n, p = d.shape
    t = np.ascontiguousarray(targets, dtype=np.int32)
    Sm = np.ascontiguousarray(np.asarray(S, dtype=np.int32).reshape(len(t), -1))
    lsp = np.ascontiguousarray(np.asarray(log_sp, dtype=np.float64).reshape(len(t), -1))
    m = lsp.shape[1]
    score = np.zeros(len(t))
    grad = np.zeros((len(t), m))
    hess = np.zeros((len(t), m, m))
    check(_lib.lib().kpc_gcv_grad_hess(dptr(d), n, p, len(t), iptr(t), Sm.shape[1], iptr(Sm),
                                       dptr(lsp), dptr(score), dptr(grad), dptr(hess)))
    return score, grad, hess


def dcov_components(x, k):
    """dcov.gamma low-rank components of the columns of x (n x count)."""
    xx = np.asfortranarray(np.asarray(x, dtype=np.float64))
    if xx.ndim == 1:
        xx = xx[:, None]
    n, c = xx.shape
    vals = np.zeros((c, k))
    vecs = np.zeros((c, k, n))
    dsum = np.zeros(c)
    mom = np.zeros(c)
    check(_lib.lib().kpc_dcov_components(dptr(xx), n, c, k, dptr(vals), dptr(vecs), dptr(dsum),
                                         dptr(mom)))
    return vals, vecs.transpose(0, 2, 1), dsum, mom


CACHE_ACTIONS = {"info": 0, "reset": 1, "configure": 2, "configure_target": 3}


def _cache_dict(ci, action=None):
    out = {"schema_version": "kpcmi-cache-control-v1"}
    if action is not None:
        out["action"] = action
    for name, _ in _lib.KpcCacheInfo._fields_:
        if name == "dataset_key":
            continue
        key = "cache_epoch" if name == "epoch" else name
        key = key if not key.startswith("target_") else key
        out[key] = int(getattr(ci, name))
    return out


def cache_control(action="info", capacity=None):
    """The result / target-state caches' statistics, reset or capacity
    (fastkpc/src/full_cuda_ci_one_call.cpp:6812-6866: actions info, reset,
    configure, configure_target)."""
    if action not in CACHE_ACTIONS:
        raise ValueError("action must be one of " + ", ".join(CACHE_ACTIONS))
    ci = _lib.KpcCacheInfo()
    check(_lib.lib().kpc_cache_control(CACHE_ACTIONS[action],
                                       -1 if capacity is None else int(capacity),
                                       _lib.ctypes.byref(ci)))
    return _cache_dict(ci, action)


def cache_state(data):
    """The caches' statistics with the entries that belong to `data`
    (fastkpc/src/full_cuda_ci_one_call.cpp:6869-6900)."""
    d = _as_data(data)
    n, p = d.shape
    ci = _lib.KpcCacheInfo()
    check(_lib.lib().kpc_cache_state(dptr(d), n, p, _lib.ctypes.byref(ci)))
    out = _cache_dict(ci)
    out["dataset_key"] = ci.dataset_key.decode()
    return out


def _sepset_arrays(sepsets, p):
    slen = np.full(p * p, -1, dtype=np.int32)
    sep = np.zeros(p * p * 64, dtype=np.int32)
    for i in range(p):
        for j in range(p):
            ss = sepsets[i][j]
            if ss is not None:
                slen[i * p + j] = len(ss)
                sep[(i * p + j) * 64:(i * p + j) * 64 + len(ss)] = [int(v) - 1 for v in ss]
    return slen, sep


def fastkpc_orient_wanpdag(skeleton, data, alpha=0.1, index=1.0, numCol=None,
                           ci_method="dcc.gamma", hsic_params=None, permutation_params=None,
                           orient_collider=True, solve_confl=False, rules=(True, True, True),
                           rng_state=None):
    """kpcalg-authority WAN-PDAG orientation of a skeleton result
    (fastkpc/R/fast_kpc.R:263-445: kpcalg udag2wanpdag + regrVonPS) on the GPU.
    skeleton: the dict fastkpc_compatible_cuda_skeleton returns (adjacency,
    sepsets with 1-based members).  Permutation methods use the counter-based
    indices unless permutation_params$rng == "R" (R's stream from rng_state,
    .Random.seed[2:626] -- e.g. the skeleton's summary r_state_out -- or from
    set.seed(seed)).  Returns pdag (p x p ints), the regrVonPS call trace
    (0-based) and the residual CI trace."""
    d = _as_data(data)
    n, p = d.shape
    adj = np.ascontiguousarray(np.asarray(skeleton["adjacency"], dtype=np.uint8).reshape(p, p))
    slen, sep = _sepset_arrays(skeleton["sepsets"], p)
    o = _lib.KpcOrientOptions()
    _lib.lib().kpc_default_orient_options(o)
    if ci_method not in CI_METHODS:
        raise ValueError("ci_method must be one of " + ", ".join(CI_METHODS))
    o.ci_method = CI_METHODS[ci_method]
    o.alpha = float(alpha)
    o.num_col = int(n // 10 if numCol is None else numCol)
    o.hsic_sig = float((hsic_params or {}).get("sig", 1.0))
    o.index = float(index)
    perm = dict(permutation_params or {})
    o.replicates = int(perm.get("replicates", 100))
    o.include_observed = 1 if perm.get("include_observed", True) else 0
    if perm.get("seed") is not None:
        o.has_seed = 1
        o.seed = int(perm["seed"])
    rng = perm.get("rng", "philox")
    if rng not in ("philox", "R"):
        raise ValueError("permutation_params$rng must be 'philox' or 'R'")
    o.perm_stream = 1 if rng == "R" else 0
    keep = None
    if rng == "R" and rng_state is not None:
        keep = np.ascontiguousarray(np.asarray(rng_state, dtype=np.int64).astype(np.uint32)
                                    .view(np.int32))
        if keep.shape != (625,):
            raise ValueError("rng_state must hold 625 integers (.Random.seed[2:626])")
        o.r_state = iptr(keep)
    o.orient_collider = 1 if orient_collider else 0
    o.solve_confl = 1 if solve_confl else 0
    for i in range(3):
        o.rules[i] = 1 if rules[i] else 0
    res = _lib.ctypes.POINTER(_lib.KpcOrientResult)()
    check(_lib.lib().kpc_wanpdag_run(dptr(d), n, p, adj.ctypes.data_as(_lib._U8), iptr(slen),
                                     iptr(sep), o, _lib.ctypes.byref(res)))
    try:
        r = res.contents
        pdag = np.ctypeslib.as_array(r.pdag, (p * p,)).reshape(p, p).copy()
        calls = [{"target": r.call_v[k],
                  "subset": [r.call_S[k * 64 + i] for i in range(r.call_slen[k])],
                  "parents": [r.call_parents[k * 64 + i] for i in range(r.call_plen[k])],
                  "reject_count": r.call_reject[k]} for k in range(r.n_calls)]
        tests = [{"call": r.test_call[k], "other": r.test_other[k],
                  "p.value": r.test_out[4 * k], "statistic": r.test_out[4 * k + 1],
                  "mean": r.test_out[4 * k + 2], "variance": r.test_out[4 * k + 3],
                  "rejected": r.test_out[4 * k] < alpha} for k in range(r.n_tests)]
        out = {"pdag": pdag, "events": calls, "ci_trace": tests,
               "counts": {"regrvonps_calls": r.n_calls, "ci_test_count": r.n_tests},
               "orientation_elapsed_sec": r.wall_seconds,
               "authority": "kpcalg::udag2wanpdag (kpcmi GPU regrVonPS)"}
        if rng == "R" and ci_method in ("dcc.perm", "hsic.perm"):
            out["rng_end_state"] = [int(v) & 0xFFFFFFFF for v in list(r.r_state_out)]
        return out
    finally:
        _lib.lib().kpc_orient_result_free(res)


def fastkpc_compatible_cuda_wanpdag(data, alpha=0.1, labels=None, options=None):
    """Skeleton then kpcalg-authority orientation
    (fastkpc/R/fast_kpc.R:447-604): options as fastkpc_compatible_cuda_skeleton
    plus orient_collider, solve_confl and rules."""
    options = dict(options or {})
    orient = {k: options.pop(k) for k in ("orient_collider", "solve_confl", "rules")
              if k in options}
    d = _as_data(data)
    n, p = d.shape
    sk = fastkpc_compatible_cuda_skeleton(d, alpha, labels=labels, options=options)
    method = options.get("ci_method", "dcc.gamma")
    perm = options.get("permutation_params", {})
    orientation = fastkpc_orient_wanpdag(
        sk, d, alpha=alpha, index=options.get("index", 1.0),
        numCol=options.get("numCol", n // 10), ci_method=method,
        hsic_params=options.get("hsic_params"), permutation_params=perm,
        rng_state=sk["summary"].get("r_state_out"), **orient)
    pd = orientation["pdag"]
    directed = int(np.count_nonzero((pd == 1) & (pd.T == 0)))
    undirected = int(np.count_nonzero((pd == 1) & (pd.T == 1)) // 2)
    bidirected = int(np.count_nonzero((pd == 2) & (pd.T == 2)) // 2)
    return {"schema_version": "fastkpc-compatible-cuda-wanpdag-v1",
            "skeleton": sk, "orientation": orientation,
            "metrics": {"skeleton_edge_count": int(np.count_nonzero(np.triu(sk["adjacency"]))),
                        "directed_edge_count": directed,
                        "undirected_edge_count": undirected,
                        "bidirected_edge_count": bidirected,
                        "skeleton_ci_test_count": int(sum(sk["n.edgetests"])),
                        "orientation_ci_test_count": len(orientation["ci_trace"])}}
