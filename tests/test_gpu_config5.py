"""BASELINE config 5 on MI355X: the sem2000x64 skeleton (n=2000, p=64,
levels 0-7) under dcc.perm and hsic.perm with 1000 permutations per test.
A full oracle skeleton is out of reach on a CPU, so:
* full size, every row: p is an exact permutation ratio k / (R + 1); the
  levels frame equals the plan rebuilt from the trace (tasks planned,
  ignored after deletion, deletions, final adjacency);
* a stratified sample of rows (tests/golden/oracle_sem2000x64_sample.npz,
  up to 6 per level and method, deleting rows included, computed by the
  oracle from the same seeded indices): each is in the GPU trace with an
  identical p, and statistic, mean and variance within 1e-9 (1e-8); a row
  whose fits took a rounding-decided optimiser step is replayed on the oracle
  at the GPU's smoothing parameters and held to the same bar;
* ci_batch replays the sampled rows bit for bit."""
import os

import numpy as np
import pytest

from conftest import GOLDEN, TIE_MARGIN, check_level_rows, rel_diff, replay_at_gpu_sp

pytestmark = [
    pytest.mark.gpu,
    pytest.mark.skip(reason="off the GPU: a run of this test (dcc.perm) left the GPU with a memory "
                            "fault after changes to the thin-plate set-up's Gram-Schmidt pass and "
                            "X / penalty stage; those changes were taken out, but the fault's cause "
                            "is not shown yet, so the n = 2000 skeletons stay off the GPU until it "
                            "is"),
]


@pytest.mark.parametrize("method", ["dcc.perm", "hsic.perm"])
def test_sem2000x64_permutation_skeleton(method):
    import fastkpc_amd as fk
    from fastkpc_amd import _lib
    from fastkpc_amd.sem import config_data
    assert _lib.lib().kpc_device_count() >= 1, "no GPU visible"
    data, _, _ = config_data("sem2000x64")
    gold = np.load(os.path.join(GOLDEN, "oracle_sem2000x64_sample.npz"))
    R = int(gold["replicates"])
    opts = {"numCol": 35, "ci_method": method,
            "permutation_params": {"replicates": R, "seed": int(gold["seed"])}}
    res = fk.fastkpc_compatible_cuda_skeleton(data, 0.1, options=opts)
    t = res["tasks"]
    print(method, "n.edgetests", res["n.edgetests"], flush=True)
    k = np.rint(t["p_used"] * (R + 1))
    assert np.all((k >= 1) & (k <= R + 1))
    assert np.array_equal(k / (R + 1), t["p_used"])
    check_level_rows(res, data.shape[1])
    rows = {(int(a), int(b), int(c), str(d)): i for i, (a, b, c, d) in
            enumerate(zip(t["level"], t["x"], t["y"], t["S_key"]))}
    sel = np.nonzero(gold["method"] == method)[0]
    assert len(sel) >= 30
    idx, ties = [], {}
    for g in sel:
        key = (int(gold["level"][g]), int(gold["x"][g]), int(gold["y"][g]), str(gold["S"][g]))
        assert key in rows, key
        i = rows[key]
        idx.append(i)
        tie = min(gold["margin"][g], t["decision_margin"][i]) < TIE_MARGIN
        ctx = (method, key, t["p_used"][i], gold["p"][g], tie)
        ok = (t["p_used"][i] == gold["p"][g] and
              rel_diff(t["method_statistic"][i], gold["stat"][g]) <= 1e-9 and
              rel_diff(t["method_mean"][i], gold["mean"][g]) <= 1e-9 and
              rel_diff(t["method_variance"][i], gold["var"][g]) <= 1e-8)
        if not tie:
            assert ok, ctx
        elif not ok:
            # a fit that took a rounding-decided optimiser step: replayed on the
            # oracle at the GPU's smoothing parameters below
            ties[key] = (t["p_used"][i], t["method_statistic"][i], t["method_mean"][i],
                         t["method_variance"][i])
    from oracle import orc
    replay_at_gpu_sp(fk, orc, data, method, ties, replicates=R, seed=int(gold["seed"]))
    print(method, "sampled rows", len(sel), "tie rows replayed", len(ties), flush=True)
    # the per-test entry point gives the trace's values bit for bit
    for size in sorted(set(int(t["conditioning_size"][i]) for i in idx)):
        part = [i for i in idx if int(t["conditioning_size"][i]) == size]
        S = np.array([[int(v) - 1 for v in t["S_key"][i].split("|")] if size else []
                      for i in part], np.int32).reshape(len(part), size)
        out = fk.ci_batch(data, [int(t["x"][i]) - 1 for i in part], [int(t["y"][i]) - 1 for i in part],
                          S, ci_method=method, replicates=R, seed=int(gold["seed"]))
        for j, i in enumerate(part):
            assert tuple(out[j]) == (t["p_used"][i], t["method_statistic"][i], t["method_mean"][i],
                                     t["method_variance"][i])
