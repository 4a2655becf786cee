"""The stored oracle traces used by the GPU parity tests are the oracle's
current output (re-run here on the CPU)."""
import os

import numpy as np
import pytest

from conftest import GOLDEN, trace_map
from oracle import orc


@pytest.mark.slow
def test_stored_sem351x48_trace_is_current_oracle_output():
    from fastkpc_amd.sem import config_data
    data, _, _ = config_data("sem351x48")
    gold = np.load(os.path.join(GOLDEN, "oracle_sem351x48.npz"))
    res = orc.skeleton(data, alpha=0.1, m_max=46, num_col=35, threads=8)
    assert res["n.edgetests"] == [int(v) for v in gold["n_edgetests"]]
    assert np.array_equal(res["adjacency"], gold["adjacency"])
    t = res["tasks"]
    mine = trace_map(t["level"], t["x"], t["y"], t["S_key"], t["p_used"])
    ref = trace_map(gold["level"], gold["x"], gold["y"], gold["S"], gold["p"])
    assert mine == ref


def test_reference_trace_edge_columns_are_min_max_of_the_tested_pair():
    """The reference's own decision trace (dev500x50 fixture, 1-based) names
    the tested edge as edge_x = min(x, y), edge_y = max(x, y) of the tested
    pair, whichever orientation the test conditioned from; the library's
    trace derives its edge_x / edge_y columns the same way."""
    z = np.load(os.path.join(GOLDEN, "dev500x50_v2.npz"))
    tx, ty = z["dt_tested_x"], z["dt_tested_y"]
    assert len(tx) > 1000 and np.any(tx > ty) and np.any(tx < ty)
    assert np.array_equal(z["dt_edge_x"], np.minimum(tx, ty))
    assert np.array_equal(z["dt_edge_y"], np.maximum(tx, ty))
    assert tx.min() >= 1 and ty.max() <= z["data"].shape[1]


def test_reference_trace_task_index_and_ignored_counts_follow_the_level_plan():
    """The reference's logical trace (dev500x50 fixture) numbers each row by
    its position in the level plan -- edges x < y row-major, the x
    orientation's lexicographic conditioning sets, then the y orientation's,
    over the adjacency at the level's start -- and its levels frame counts
    the planned tasks left on an edge after the row that deleted it.  Both are
    rebuilt here from the plan order alone and must equal the fixture; the
    GPU test then compares the library's columns with the fixture directly."""
    from itertools import combinations
    z = np.load(os.path.join(GOLDEN, "dev500x50_v2.npz"))
    p = z["data"].shape[1]
    adj = np.ones((p, p), bool)
    np.fill_diagonal(adj, False)
    lv, lx, ly, lS, ld = z["lt_level"], z["lt_x"], z["lt_y"], z["lt_S"], z["lt_del"]
    planned = ignored = 0
    for l in range(int(z["bs_levels"])):
        plan, edge_end = {}, {}
        for a in range(p):
            for b in range(a + 1, p):
                if not adj[a, b]:
                    continue
                for x, y in ((a, b), (b, a)):
                    nb = [i for i in range(p) if i != y and adj[x, i]]
                    for S in combinations(nb, l):
                        plan[(x + 1, y + 1, "|".join(str(s + 1) for s in S))] = len(plan) + 1
                edge_end[(a + 1, b + 1)] = len(plan)
        planned += len(plan)
        rows = np.nonzero(lv == l)[0]
        want = [plan[(int(lx[i]), int(ly[i]), str(lS[i]))] for i in rows]
        assert want == z["lt_task_index"][rows].tolist()
        for i in rows[ld[rows]]:
            e = (min(lx[i], ly[i]), max(lx[i], ly[i]))
            ignored += edge_end[e] - plan[(int(lx[i]), int(ly[i]), str(lS[i]))]
            adj[e[0] - 1, e[1] - 1] = adj[e[1] - 1, e[0] - 1] = False
    assert planned == int(z["bs_tasks_planned"])
    assert ignored == int(z["bs_tasks_ignored_after_delete"])
    assert np.array_equal(z["lt_seq"], np.arange(1, len(lv) + 1))
    # the deletion trace is the logical trace's deleting rows, in order
    sel = np.nonzero(ld)[0]
    assert np.array_equal(lv[sel], z["dt_level"]) and np.array_equal(lx[sel], z["dt_tested_x"])
    assert np.array_equal(ly[sel], z["dt_tested_y"]) and np.array_equal(z["lt_p"][sel], z["dt_p"])
    assert [str(v) for v in lS[sel]] == [str(v) for v in z["dt_S"]]


@pytest.mark.slow
def test_oracle_deletion_order_and_statistics_of_the_500x50_fixture():
    """The oracle's ordered deletion trace equals the reference's (level,
    tested orientation, S_key; p within 2.4e-10 + 6e-9 relative), and the
    stored oracle trace carries the oracle's current statistics."""
    z = np.load(os.path.join(GOLDEN, "dev500x50_v2.npz"))
    gold = np.load(os.path.join(GOLDEN, "oracle_dev500x50.npz"))
    res = orc.skeleton(z["data"], alpha=0.1, m_max=48, num_col=35, threads=8)
    t = res["tasks"]
    sel = np.nonzero(t["native_edge_deleted"])[0]
    assert np.array_equal(t["level"][sel], z["dt_level"])
    assert np.array_equal(t["x"][sel], z["dt_tested_x"]) and np.array_equal(t["y"][sel], z["dt_tested_y"])
    assert [t["S_key"][i] for i in sel] == [str(v) for v in z["dt_S"]]
    dp = np.abs(t["p_used"][sel] - z["dt_p"])
    assert np.all(dp <= 2.4e-10 + 6.0e-9 * np.abs(z["dt_p"])), dp.max()
    for a, b in (("statistic", "stat"), ("mean", "mean"), ("variance", "var"), ("p_used", "p")):
        assert np.array_equal(t[a], gold[b]), a
