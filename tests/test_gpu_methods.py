"""hsic.gamma, hsic.perm and dcc.perm on MI355X vs the oracle.

Tolerances: permutation indices identical; permutation p-values equal as
exact ratios; statistics, means and variances within 1e-9 relative;
hsic.gamma p-values within 1e-9 relative (1e-12 absolute); skeleton
decisions, sepsets and deletion traces exact.  Tests with a discrete decision
within 1e-12 of its threshold (see conftest.compare_skeleton) are held to
1e-6 relative.
"""
import os

import numpy as np
import pytest

from conftest import (GOLDEN, TIE_MARGIN, check_level_rows, compare_skeleton, p_close,
                      replay_at_gpu_sp)

pytestmark = pytest.mark.gpu


def fk():
    import fastkpc_amd
    from fastkpc_amd import _lib
    assert _lib.lib().kpc_device_count() >= 1, "no GPU visible"
    return fastkpc_amd


def orc():
    from oracle import orc as o
    return o


def rel_close(a, b, rtol=1e-9):
    return abs(a - b) <= rtol * abs(b) + 1e-300


@pytest.mark.parametrize("n", [2, 3, 64, 200, 351, 2048])
def test_device_permutations_equal_oracle_indices(n):
    f, o = fk(), orc()
    seeds = [f.task_seed(2024, "dcc.perm", 0, 5, [1, 2]), 0, 2 ** 64 - 1, 123456789]
    R = 7
    table = f.permutation_table(n, R, seeds)
    for i, s in enumerate(seeds):
        for r in range(R):
            assert np.array_equal(table[i, r], o.permutation(s, r, n)), (n, s, r)


def mixed_data(n=200, p=10, seed=14):
    from fastkpc_amd.sem import generate
    return generate(n=n, p=p, edge_prob=0.3, seed=seed, link="mixed")[0]


@pytest.mark.parametrize("method", ["hsic.gamma", "hsic.perm", "dcc.perm"])
@pytest.mark.parametrize("s_len", [0, 1, 2, 3])
def test_ci_batch_vs_oracle(method, s_len):
    f, o = fk(), orc()
    data = mixed_data()
    p = data.shape[1]
    rng = np.random.default_rng(100 + s_len)
    xs, ys, Ss = [], [], []
    for _ in range(4):
        perm = rng.permutation(p)
        xs.append(int(perm[0]))
        ys.append(int(perm[1]))
        Ss.append(sorted(int(v) for v in perm[2:2 + s_len]))
    S = np.array(Ss, dtype=np.int32).reshape(4, s_len)
    out = f.ci_batch(data, xs, ys, S, num_col=35, ci_method=method, replicates=60, seed=77)
    for i in range(4):
        ref = o.ci_test(data, xs[i], ys[i], Ss[i], num_col=35, method=method, replicates=60,
                        seed=77)
        tie = ref["margin"] < TIE_MARGIN
        if s_len:
            for t in (xs[i], ys[i]):
                _, gi = f.regrxons_batch(data, [t], [Ss[i]])
                tie |= gi[0, 30] < TIE_MARGIN
        ctx = (method, s_len, xs[i], ys[i], Ss[i], out[i], ref)
        if tie:
            assert abs(out[i, 1] - ref["statistic"]) <= 1e-6 * abs(ref["statistic"]), ctx
            continue
        assert rel_close(out[i, 1], ref["statistic"]), ctx
        if method.endswith("perm"):
            assert out[i, 0] == ref["p.value"], ctx
            assert rel_close(out[i, 2], ref["mean"], 1e-9), ctx
            assert rel_close(out[i, 3], ref["variance"], 1e-8), ctx
        else:
            assert p_close(out[i, 0], ref["p.value"]), ctx
            assert rel_close(out[i, 2], ref["mean"]) and rel_close(out[i, 3], ref["variance"]), ctx


def test_orientation_free_identity_and_twin_reuse():
    f = fk()
    data = mixed_data(seed=3)
    for method in ("hsic.perm", "dcc.perm", "hsic.gamma", "dcc.gamma"):
        out = f.ci_batch(data, [1, 6, 1], [6, 1, 6], np.array([[2, 4], [2, 4], [2, 4]], np.int32),
                         num_col=20, ci_method=method, replicates=40, seed=5)
        assert np.array_equal(out[0], out[1]) and np.array_equal(out[0], out[2]), method
        alone = f.ci_batch(data, [6], [1], np.array([[2, 4]], np.int32), num_col=20,
                           ci_method=method, replicates=40, seed=5)
        assert np.array_equal(alone[0], out[0]), method


@pytest.mark.parametrize("n,num_col,sig", [(6, 3, 1.0), (9, 8, 0.3), (120, 5, 1.0), (700, 35, 1.0),
                                           (2048, 64, 2.0)])
def test_hsic_gamma_edge_sizes_vs_oracle(n, num_col, sig):
    f, o = fk(), orc()
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n)
    y = np.sin(x) + rng.standard_normal(n)
    data = np.column_stack([x, y])
    out = f.ci_batch(data, [0], [1], np.zeros((1, 0), np.int32), num_col=num_col,
                     ci_method="hsic.gamma", sig=sig)
    ref = o.ci_test(data, 0, 1, [], num_col=num_col, method="hsic.gamma", sig=sig)
    # these inputs take no pivot or stopping decision near a rounding tie
    # (smallest oracle margin 2e-9), so every value is held to 1e-9
    assert ref["margin"] >= TIE_MARGIN, ref
    assert rel_close(out[0, 1], ref["statistic"]), (out[0], ref)
    assert rel_close(out[0, 2], ref["mean"]) and rel_close(out[0, 3], ref["variance"])
    assert p_close(out[0, 0], ref["p.value"])


@pytest.mark.parametrize("n,index,incl,R", [(6, 1.0, True, 10), (150, 0.5, True, 50),
                                             (150, 1.5, False, 50), (1000, 1.0, True, 20),
                                             (2048, 1.0, True, 3), (80, 1.0, True, 0)])
def test_dcc_perm_edge_cases_vs_oracle(n, index, incl, R):
    f, o = fk(), orc()
    rng = np.random.default_rng(n + R)
    x = rng.standard_normal(n)
    y = 0.5 * x * x + rng.standard_normal(n)
    data = np.column_stack([x, y])
    out = f.ci_batch(data, [0], [1], np.zeros((1, 0), np.int32), num_col=5, ci_method="dcc.perm",
                     replicates=R, seed=9, index=index, include_observed=incl)
    ref = o.ci_test(data, 0, 1, [], num_col=5, method="dcc.perm", replicates=R, seed=9,
                    index=index, include_observed=incl)
    assert rel_close(out[0, 1], ref["statistic"]), (out[0], ref)
    assert out[0, 0] == ref["p.value"], (out[0], ref)


@pytest.mark.parametrize("R,incl", [(0, True), (0, False), (1, True), (257, True)])
def test_hsic_perm_replicate_counts_vs_oracle(R, incl):
    f, o = fk(), orc()
    data = mixed_data(n=160, seed=8)
    out = f.ci_batch(data, [0, 2], [3, 5], np.zeros((2, 0), np.int32), num_col=35,
                     ci_method="hsic.perm", replicates=R, seed=31, include_observed=incl)
    for i, (x, y) in enumerate([(0, 3), (2, 5)]):
        ref = o.ci_test(data, x, y, [], num_col=35, method="hsic.perm", replicates=R, seed=31,
                        include_observed=incl)
        assert out[i, 0] == ref["p.value"], (R, incl, out[i], ref)
        assert rel_close(out[i, 1], ref["statistic"])


@pytest.mark.parametrize("name", ["sem200x8_hsic_gamma", "sem351x48_hsic_gamma",
                                  "sem200x8_hsic_perm", "sem200x8_dcc_perm",
                                  "mixed200x14_hsic_perm", "mixed200x14_dcc_perm"])
def test_skeleton_methods_vs_oracle(name):
    f = fk()
    import sys
    sys.path.insert(0, GOLDEN)
    from make_oracle_goldens import PERM_SEED, WORKLOADS, workload_data
    src, method, reps = WORKLOADS[name][:3]
    data = workload_data(src)
    gold = np.load(os.path.join(GOLDEN, "oracle_%s.npz" % name))
    opts = {"numCol": 35, "ci_method": method}
    if reps:
        opts["permutation_params"] = {"replicates": reps, "seed": PERM_SEED}
    res = f.fastkpc_compatible_cuda_skeleton(data, 0.1, options=opts)
    exact = method.endswith("perm")
    rep = {}
    ties = compare_skeleton(res, gold, (lambda a, b: a == b) if exact else p_close, rep)
    if exact:
        # no permutation row sits on a rounding tie in these workloads
        assert not ties, list(ties.items())[:3]
    else:
        rep["tie_score_excess"] = replay_at_gpu_sp(f, orc(), data, method, ties)
    print(name, "vs oracle:", rep, flush=True)
    p = data.shape[1]
    for i in range(p):
        for j in range(p):
            ln = gold["sepsets"][i, j, 0]
            want = None if ln < 0 else [int(v) for v in gold["sepsets"][i, j, 1:1 + ln]]
            assert res["sepsets"][i][j] == want, (i, j)


def test_plain_form_equals_batched_form_permutation():
    f = fk()
    data = mixed_data(n=120, p=8, seed=21)
    opts = {"numCol": 15, "ci_method": "dcc.perm",
            "permutation_params": {"replicates": 30, "seed": 4}}
    rb = f.fastkpc_compatible_cuda_skeleton(data, 0.1, options=opts)
    rp = f.fastkpc_compatible_cuda_skeleton(data, 0.1, options=dict(opts, mode="plain"))
    assert rb["n.edgetests"] == rp["n.edgetests"]
    assert np.array_equal(rb["adjacency"], rp["adjacency"])
    assert rb["tasks"]["S_key"] == rp["tasks"]["S_key"]
    assert np.array_equal(rb["tasks"]["p_used"], rp["tasks"]["p_used"])


@pytest.mark.parametrize("method", ["hsic.gamma", "dcc.gamma", "dcc.perm"])
def test_skeleton_trace_statistic_and_edge_columns(method):
    """The skeleton trace's edge_x / edge_y are the 1-based (min, max) of each
    tested pair; method_statistic / method_mean / method_variance (and p_used)
    of every row -- every level, rows filled from a twin test included -- equal
    bit for bit ci_batch on the same tests (whose values are checked against
    the oracle above); permutation rows carry their task seed."""
    import fastkpc_amd as fk
    from fastkpc_amd.sem import config_data
    data, _, _ = config_data("sem200x8")
    opts = {"numCol": 35, "ci_method": method}
    if method == "dcc.perm":
        opts["permutation_params"] = {"replicates": 50, "seed": 11}
    res = fk.fastkpc_compatible_cuda_skeleton(data, 0.1, options=opts)
    t = res["tasks"]
    assert np.array_equal(t["edge_x"], np.minimum(t["x"], t["y"]))
    assert np.array_equal(t["edge_y"], np.maximum(t["x"], t["y"]))
    assert np.all(t["edge_x"] < t["edge_y"]) and t["edge_x"].min() >= 1
    for size in sorted(set(t["conditioning_size"].tolist())):
        rows = [i for i in range(len(t["x"])) if t["conditioning_size"][i] == size]
        xs = [int(t["x"][i]) - 1 for i in rows]
        ys = [int(t["y"][i]) - 1 for i in rows]
        S = np.array([[int(v) - 1 for v in t["S_key"][i].split("|")] if size else []
                      for i in rows], np.int32).reshape(len(rows), size)
        out = fk.ci_batch(data, xs, ys, S, num_col=35, ci_method=method, replicates=50, seed=11)
        for j, i in enumerate(rows):
            got = (t["p_used"][i], t["method_statistic"][i], t["method_mean"][i],
                   t["method_variance"][i])
            assert tuple(out[j]) == got, (size, i, out[j], got)
            if method == "dcc.perm":
                assert int(t["permutation_seed"][i]) == fk.task_seed(11, method, xs[j], ys[j],
                                                                     S[j].tolist())
    assert ("permutation_seed" in t) == (method == "dcc.perm")


@pytest.mark.parametrize("cfg", ["sem200x8", "sem351x48"])
def test_level_rows_match_the_level_plan(cfg):
    """level_rows (the reference's levels data frame) against the plan
    rebuilt from the trace alone (conftest.check_level_rows): tasks planned,
    tasks ignored after each deleting row, deletions and the final
    adjacency."""
    import fastkpc_amd as fk
    from fastkpc_amd.sem import config_data
    data, _, _ = config_data(cfg)
    res = fk.fastkpc_compatible_cuda_skeleton(data, 0.1, options={"numCol": 35})
    check_level_rows(res, data.shape[1])
    assert all(r >= 1 for r in res["level_rows"]["rounds"])


def test_trace_task_index_is_the_position_in_the_level_plan():
    """task_index is each trace row's 1-based position in its level's plan,
    enumerated here independently in the documented order (edges x < y
    row-major, the x orientation's conditioning sets, then the y
    orientation's, each in lexicographic order, over the adjacency at the
    level's start); canonical_test_order_id numbers the rows 1..N."""
    from itertools import combinations
    import fastkpc_amd as fk
    from fastkpc_amd.sem import config_data
    data, _, _ = config_data("sem200x8")
    res = fk.fastkpc_compatible_cuda_skeleton(data, 0.1, options={"numCol": 35})
    t = res["tasks"]
    p = data.shape[1]
    assert np.array_equal(t["canonical_test_order_id"], np.arange(1, len(t["x"]) + 1))
    assert t["native_edge_ignored"].dtype == bool and not t["native_edge_ignored"].any()
    adj = np.ones((p, p), bool)
    np.fill_diagonal(adj, False)
    for l in range(res["levels"]):
        plan = {}
        for a in range(p):
            for b in range(a + 1, p):
                if not adj[a, b]:
                    continue
                for x, y in ((a, b), (b, a)):
                    nb = [i for i in range(p) if i != y and adj[x, i]]
                    for S in combinations(nb, l):
                        plan[(x + 1, y + 1, "|".join(str(s + 1) for s in S))] = len(plan) + 1
        assert len(plan) == res["level_rows"]["tasks_planned"][l]
        rows = np.nonzero(t["level"] == l)[0]
        got = [int(t["task_index"][i]) for i in rows]
        want = [plan[(int(t["x"][i]), int(t["y"][i]), t["S_key"][i])] for i in rows]
        assert got == want and got == sorted(got)
        sel = rows[t["native_edge_deleted"][rows]]
        for i in sel:
            adj[t["edge_x"][i] - 1, t["edge_y"][i] - 1] = adj[t["edge_y"][i] - 1, t["edge_x"][i] - 1] = False
