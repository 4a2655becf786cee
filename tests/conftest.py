import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# the launch-shape autotuner measures afresh in every test process
os.environ.setdefault("KPC_TUNE_CACHE", "off")
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (runs libkpcmi kernels)")
    config.addinivalue_line("markers", "slow: CPU test that takes tens of seconds")


@pytest.fixture(scope="session")
def dcov_fixture():
    with open(os.path.join(GOLDEN, "dcov_gamma_oracle_v1.json")) as fh:
        fx = json.load(fh)
    cases = []
    for c in fx["cases"]:
        c = dict(c)
        c["rx"] = np.array([float.fromhex(v) for v in c["rx"]])
        c["ry"] = np.array([float.fromhex(v) for v in c["ry"]])
        cases.append(c)
    return cases


@pytest.fixture(scope="session")
def dev500():
    return np.load(os.path.join(GOLDEN, "dev500x50_v2.npz"))


def trace_map(level, x, y, S, p):
    return {(int(a), int(b), int(c), str(d)): float(e) for a, b, c, d, e in zip(level, x, y, S, p)}


# Decision margins below this (relative) can flip under rounding differences.
TIE_MARGIN = 1e-12


def p_close(a, b):
    return abs(a - b) <= 1e-9 * abs(b) + 1e-12


def rel_diff(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def compare_skeleton(res, gold, tol_p, report=None):
    """Decisions exact; p-values within tol_p and the statistic, null mean and
    null variance of every trace row within 1e-9 relative of the oracle
    (permutation variances 1e-8: a sum of squares over the replicates).

    A row whose two fits (or incomplete-Cholesky factors, or replicate
    comparisons) took a discrete decision within TIE_MARGIN (relative) of its
    threshold on either implementation -- an optimiser step accepted on a
    ~1e-16 score change -- may legitimately end elsewhere.  Such rows are
    returned (row key -> GPU values) when they exceed 1e-9; the caller replays
    them on the oracle at the GPU's selected smoothing parameters
    (replay_at_gpu_sp), which holds them to 1e-9 as well."""
    assert res["n.edgetests"] == [int(v) for v in gold["n_edgetests"]]
    assert np.array_equal(res["adjacency"], gold["adjacency"])
    t = res["tasks"]
    keys = [(int(a), int(b), int(c), str(d)) for a, b, c, d in
            zip(t["level"], t["x"], t["y"], t["S_key"])]
    gkeys = [(int(a), int(b), int(c), str(d)) for a, b, c, d in
             zip(gold["level"], gold["x"], gold["y"], gold["S"])]
    # same rows in the same canonical order
    assert keys == gkeys
    cols = [("p", t["p_used"], gold["p"]), ("stat", t["method_statistic"], gold["stat"]),
            ("mean", t["method_mean"], gold["mean"]), ("var", t["method_variance"], gold["var"])]
    perm = "permutation_seed" in t
    tie_over = {}
    worst = {c: 0.0 for c, _, _ in cols}
    n_tie = 0
    bad = []
    for i, k in enumerate(keys):
        tie = min(t["decision_margin"][i], gold["margin"][i]) < TIE_MARGIN
        n_tie += tie
        ok = tol_p(t["p_used"][i], gold["p"][i])
        for c, mine, ref in cols[1:]:
            r = rel_diff(mine[i], ref[i])
            ok &= r <= (1e-8 if (perm and c == "var") else 1e-9)
        for c, mine, ref in cols:
            worst[c] = max(worst[c], rel_diff(mine[i], ref[i])) if not tie else worst[c]
        if ok:
            continue
        if tie:
            tie_over[k] = tuple(float(m[i]) for _, m, _ in cols)
        else:
            bad.append((k, [(c, m[i], r[i]) for c, m, r in cols]))
    assert not bad, bad[:5]
    if report is not None:
        report.update(rows=len(keys), tie_rows=n_tie, tie_rows_over_1e9=len(tie_over),
                      max_rel_non_tie=worst)
    dels = [k for k, e in zip(keys, t["native_edge_deleted"]) if e]
    gdels = [k for k, e in zip(gkeys, gold["deleted"]) if e]
    assert dels == gdels
    return tie_over


def replay_at_gpu_sp(fk, orc, data, method, rows, replicates=100, seed=0):
    """Hold rounding-tie rows to 1e-9: for each (level, x, y, S_key) -> GPU
    (p, statistic, mean, variance), refit both variables on the oracle at the
    log smoothing parameters the GPU selected, run the oracle's CI test on
    those residuals and compare every value at 1e-9 relative (p: 1e-9
    relative + 1e-12 absolute).  Also checks the GPU's GCV score at its own
    selection against the oracle's evaluation there (1e-9), and returns the
    largest relative excess of that score over the oracle's own optimum."""
    worst_excess = 0.0
    for (level, x1, y1, skey), (p, st, mn, vr) in rows.items():
        S = sorted(int(v) - 1 for v in skey.split("|"))
        assert S, "a rounding tie needs a fitted residual"
        x, y = x1 - 1, y1 - 1
        m = 1 if len(S) <= 2 else len(S)
        assert m <= 22, "fit info carries 22 log smoothing parameters"
        _, gi = fk.regrxons_batch(data, [x, y], [S, S])
        res = []
        for j, tgt in enumerate((x, y)):
            r, info = orc.regrxons_at_sp(data, tgt, S, gi[j, 8:8 + m])
            assert rel_diff(gi[j, 1], info["score"]) <= 1e-9, (level, x1, y1, skey, gi[j, 1], info)
            _, own = orc.regrxons(data, tgt, S)
            worst_excess = max(worst_excess, (gi[j, 1] - own["score"]) / own["score"])
            res.append(r)
        lo, hi = (res[0], res[1]) if x < y else (res[1], res[0])
        if method == "dcc.gamma":
            ref = orc.dcov_gamma(res[0], res[1])
            ref = (ref["p.value"], ref["nV2"], ref["mean"], ref["variance"])
        elif method == "hsic.gamma":
            ref = orc.hsic_gamma(res[0], res[1])
            ref = (ref["p.value"], ref["statistic"], ref["mean"], ref["variance"])
        else:
            # the permuted variable is the larger index, the indices those of
            # the test's task seed
            ts = orc.task_seed(seed, method, x, y, S)
            ref = (orc.dcc_perm(lo, hi, replicates=replicates, seed=ts) if method == "dcc.perm"
                   else orc.hsic_perm(lo, hi, replicates=replicates, seed=ts))
            ref = (ref["p.value"], ref["statistic"], ref["mean"], ref["variance"])
        ctx = (level, x1, y1, skey, (p, st, mn, vr), ref)
        assert p == ref[0] if method.endswith("perm") else p_close(p, ref[0]), ctx
        for a, b, tol in zip((st, mn, vr), ref[1:], (1e-9, 1e-9, 1e-8)):
            assert rel_diff(a, b) <= tol, ctx
    return worst_excess


def check_level_rows(res, p):
    """The levels frame against the plan rebuilt from the trace alone: per
    level, tasks_planned is the sum over the edges present at the level's
    start of C(|adj(x) \\ {y}|, l) + C(|adj(y) \\ {x}|, l); an edge still
    present replays every planned task; a deleted edge ignores the planned
    tasks after its deleting row (tasks_ignored_after_delete); deletions
    apply at the level's end and rebuild the final adjacency."""
    from math import comb
    rows, t = res["level_rows"], res["tasks"]
    adj = np.ones((p, p), bool)
    np.fill_diagonal(adj, False)
    assert rows["tests_replayed"] == res["n.edgetests"]
    for l in range(res["levels"]):
        lv = t["level"] == l
        planned = ignored = 0
        for a in range(p):
            for b in range(a + 1, p):
                if not adj[a, b]:
                    continue
                n_plan = comb(int(adj[a].sum()) - 1, l) + comb(int(adj[b].sum()) - 1, l)
                planned += n_plan
                er = np.nonzero(lv & (t["edge_x"] == a + 1) & (t["edge_y"] == b + 1))[0]
                dl = er[t["native_edge_deleted"][er]]
                if len(dl) == 0:
                    assert len(er) == n_plan, (l, a, b, len(er), n_plan)
                else:
                    assert len(dl) == 1 and dl[0] == er[-1]
                    ignored += n_plan - len(er)
        assert rows["tasks_planned"][l] == planned, (l, rows["tasks_planned"][l], planned)
        assert rows["tasks_ignored_after_delete"][l] == ignored, (l, ignored)
        assert rows["tasks_planned"][l] == res["n.edgetests"][l] + ignored
        sel = lv & t["native_edge_deleted"]
        assert rows["deletions"][l] == int(sel.sum())
        for x, y in zip(t["edge_x"][sel], t["edge_y"][sel]):
            adj[x - 1, y - 1] = adj[y - 1, x - 1] = False
    assert np.array_equal(adj, res["adjacency"])
