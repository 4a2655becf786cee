"""kpcalg-authority WAN-PDAG orientation on MI355X (kpc_wanpdag_run: GPU
regrVonPS fits and residual CI tests under the host graph algorithm) against
the oracle's orientation on the same skeleton: identical PDAG, identical
regrVonPS call sequence (target, subset, parents, reject count) and residual
test p-values within 1e-9 (exact for the permutation methods), including the
R stream continued from the skeleton's final state."""
import numpy as np
import pytest

from conftest import rel_diff

pytestmark = pytest.mark.gpu


def fk():
    import fastkpc_amd
    from fastkpc_amd import _lib
    assert _lib.lib().kpc_device_count() >= 1, "no GPU visible"
    return fastkpc_amd


def compare(mine, ref, exact):
    assert np.array_equal(mine["pdag"], ref["pdag"])
    got = [(c["target"], c["subset"], c["parents"], c["reject_count"]) for c in mine["events"]]
    want = [(c["V"], c["S"], c["parents"], c["reject"]) for c in ref["calls"]]
    assert got == want
    assert len(mine["ci_trace"]) == len(ref["tests"])
    for a, b in zip(mine["ci_trace"], ref["tests"]):
        assert (a["call"], a["other"]) == (b["call"], b["other"])
        if exact:
            assert a["p.value"] == b["out"][0], (a, b)
        else:
            assert abs(a["p.value"] - b["out"][0]) <= 1e-9 * b["out"][0] + 1e-12, (a, b)
        assert rel_diff(a["statistic"], b["out"][1]) <= 1e-9, (a, b)


@pytest.mark.parametrize("cfg,method", [("sem200x8", "dcc.gamma"), ("sem200x8", "hsic.gamma"),
                                        ("sem351x48", "dcc.gamma")])
def test_orientation_vs_oracle(cfg, method):
    from oracle import orc
    f = fk()
    from fastkpc_amd.sem import config_data
    data, _, _ = config_data(cfg)
    res = f.fastkpc_compatible_cuda_wanpdag(data, 0.1, options={"numCol": 35, "ci_method": method})
    sk = res["skeleton"]
    ref = orc.wanpdag(data, sk["adjacency"], sk["sepsets"], method=method, alpha=0.1)
    compare(res["orientation"], ref, exact=False)
    m = res["metrics"]
    print(cfg, method, m, flush=True)
    assert m["directed_edge_count"] + 2 * m["undirected_edge_count"] == \
        int(np.count_nonzero(res["orientation"]["pdag"]))


@pytest.mark.parametrize("rng", ["philox", "R"])
def test_orientation_permutation_methods_vs_oracle(rng):
    from oracle import orc
    f = fk()
    from fastkpc_amd.sem import generate
    data = generate(n=200, p=10, edge_prob=0.3, seed=14, link="mixed")[0]
    opts = {"numCol": 35, "ci_method": "dcc.perm",
            "permutation_params": {"replicates": 60, "seed": 77, "rng": rng}}
    res = f.fastkpc_compatible_cuda_wanpdag(data, 0.1, options=opts)
    sk = res["skeleton"]
    ref = orc.wanpdag(data, sk["adjacency"], sk["sepsets"], method="dcc.perm", alpha=0.1,
                      replicates=60, seed=77, r_stream=rng == "R",
                      r_state=sk["summary"].get("r_state_out"))
    compare(res["orientation"], ref, exact=True)
    if rng == "R":
        assert res["orientation"]["rng_end_state"] == ref["r_state"]


def test_orientation_on_a_hand_built_skeleton():
    """A skeleton with no collider (chain 0 - 1 - 2, triangle 1 - 2 - 3,
    3 - 4, and a separate 5 - 6) on nonlinear non-Gaussian data: every edge
    enters the generalised transitive step, so regrVonPS acceptances,
    rejections and rules 1-3 all run."""
    from oracle import orc
    f = fk()
    rng = np.random.default_rng(3)
    n, p = 200, 7
    x = rng.uniform(-1.5, 1.5, (n, p))
    x[:, 1] += np.tanh(2 * x[:, 0])
    x[:, 2] += 0.8 * x[:, 1] ** 2
    x[:, 3] += np.sin(2 * x[:, 1]) + 0.5 * x[:, 2]
    x[:, 4] += np.tanh(x[:, 3]) ** 2
    x[:, 6] += np.sin(2 * x[:, 5])
    adj = np.zeros((p, p), bool)
    for a, b in [(0, 1), (1, 2), (1, 3), (2, 3), (3, 4), (5, 6)]:
        adj[a, b] = adj[b, a] = True
    sep = [[None if (a == b or adj[a, b]) else [] for b in range(p)] for a in range(p)]
    # 1-based separating sets: no unshielded triple is a collider
    for (a, b), s1 in {(0, 2): [2], (0, 3): [2], (0, 4): [4], (1, 4): [4], (2, 4): [4]}.items():
        sep[a][b] = sep[b][a] = s1
    sk = {"adjacency": adj, "sepsets": sep}
    for method in ("dcc.gamma", "hsic.gamma"):
        mine = f.fastkpc_orient_wanpdag(sk, x, alpha=0.1, numCol=20, ci_method=method)
        ref = orc.wanpdag(x, adj, sep, method=method, alpha=0.1, num_col=20)
        compare(mine, ref, exact=False)
        assert len(mine["events"]) > 3, mine["events"]
        print(method, "calls", len(mine["events"]), "pdag", mine["pdag"].tolist(), flush=True)
