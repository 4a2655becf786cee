"""R-stream permutation tests on MI355X (KPC_PERM_R_STREAM) against the
oracle: the permutation indices come from R's global Mersenne-Twister stream
after set.seed(seed), drawn test by test -- in batch order for ci_batch, in
pcalg's execution order for the skeleton -- so the p-values are the ones
kpcalg's dcov.test / hsic.perm give on the same residuals.

Tolerances: p-values equal (exact ratios), statistics, means and variances
within 1e-9 relative (variances 1e-8), the stream afterwards identical.
"""
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN, compare_skeleton

pytestmark = pytest.mark.gpu


def fk():
    import fastkpc_amd
    from fastkpc_amd import _lib
    assert _lib.lib().kpc_device_count() >= 1, "no GPU visible"
    return fastkpc_amd


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def mixed_data(n=200, p=10, seed=14):
    from fastkpc_amd.sem import generate
    return generate(n=n, p=p, edge_prob=0.3, seed=seed, link="mixed")[0]


@pytest.mark.parametrize("method", ["dcc.perm", "hsic.perm"])
@pytest.mark.parametrize("s_len", [0, 1, 3])
def test_ci_batch_r_stream_vs_oracle(method, s_len):
    from oracle import orc
    f = fk()
    data = mixed_data()
    n, p = data.shape
    rng = np.random.default_rng(40 + s_len)
    xs, ys, Ss = [], [], []
    for _ in range(5):
        perm = rng.permutation(p)
        xs.append(int(perm[0]))
        ys.append(int(perm[1]))
        Ss.append(sorted(int(v) for v in perm[2:2 + s_len]))
    S = np.array(Ss, dtype=np.int32).reshape(5, s_len)
    R = 57
    out, state = f.ci_batch(data, xs, ys, S, ci_method=method, replicates=R, seed=31, rng="R",
                            return_state=True)
    rs = orc.RStream(31)
    for i in range(5):
        rx = orc.regrxons(data, xs[i], Ss[i])[0] if s_len else data[:, xs[i]]
        ry = orc.regrxons(data, ys[i], Ss[i])[0] if s_len else data[:, ys[i]]
        tab = rs.perm_table(method, n, R)
        ref = (orc.dcc_perm_table(rx, ry, tab) if method == "dcc.perm"
               else orc.hsic_perm_table(rx, ry, tab))
        ctx = (method, s_len, i, out[i], ref)
        assert out[i, 0] == ref["p.value"], ctx
        assert rel(out[i, 1], ref["statistic"]) <= 1e-9, ctx
        assert rel(out[i, 2], ref["mean"]) <= 1e-9, ctx
        assert rel(out[i, 3], ref["variance"]) <= 1e-8, ctx
    assert state == rs.state()


def test_ci_batch_r_stream_continues_from_a_given_state():
    """Two batches, the second started from the first one's returned stream,
    give the results of one batch holding both (as consecutive R calls do)."""
    f = fk()
    data = mixed_data(seed=5)
    xs, ys = [0, 2, 4, 1], [3, 5, 7, 6]
    S = np.zeros((4, 0), np.int32)
    both, st = f.ci_batch(data, xs, ys, S, ci_method="dcc.perm", replicates=30, seed=8, rng="R",
                          return_state=True)
    a, st1 = f.ci_batch(data, xs[:2], ys[:2], S[:2], ci_method="dcc.perm", replicates=30, seed=8,
                        rng="R", return_state=True)
    b, st2 = f.ci_batch(data, xs[2:], ys[2:], S[2:], ci_method="dcc.perm", replicates=30,
                        rng="R", r_state=st1, return_state=True)
    assert np.array_equal(np.vstack([a, b]), both)
    assert st2 == st


@pytest.mark.parametrize("name", ["sem200x8_dcc_perm_R", "sem200x8_hsic_perm_R",
                                  "mixed200x14_dcc_perm_R", "mixed200x14_hsic_perm_R"])
def test_skeleton_r_stream_vs_oracle(name):
    f = fk()
    sys.path.insert(0, GOLDEN)
    from make_oracle_goldens import PERM_SEED, WORKLOADS, workload_data
    src, method, reps = WORKLOADS[name][:3]
    data = workload_data(src)
    gold = np.load(os.path.join(GOLDEN, "oracle_%s.npz" % name))
    opts = {"numCol": 35, "ci_method": method,
            "permutation_params": {"replicates": reps, "seed": PERM_SEED, "rng": "R"}}
    res = f.fastkpc_compatible_cuda_skeleton(data, 0.1, options=opts)
    rep = {}
    ties = compare_skeleton(res, gold, lambda a, b: a == b, rep)
    assert not ties, list(ties.items())[:3]
    assert res["summary"]["r_state_out"] == [int(v) for v in gold["r_state"]]
    print(name, "vs oracle:", rep, "rounds", res["level_rows"]["rounds"], flush=True)
    p = data.shape[1]
    for i in range(p):
        for j in range(p):
            ln = gold["sepsets"][i, j, 0]
            want = None if ln < 0 else [int(v) for v in gold["sepsets"][i, j, 1:1 + ln]]
            assert res["sepsets"][i][j] == want, (i, j)
