"""The oracle's WAN-PDAG orientation (oracle/orc_orient.c) against a literal
Python transcription of kpcalg/R/udag2wanpdag.R and kpcalg/R/regrvonps.R,
statement by statement with R's 1-based indexing, which(arr.ind = TRUE)
column order, combn order and loop structure; regrVonPS's regression and CI
tests use the oracle's own primitives, so the check isolates the graph
algorithm: the same regrVonPS calls in the same order, the same p-values and
the same final PDAG.  Also checks the collider invariant on the result."""
import itertools

import numpy as np
import pytest

from oracle import orc


def which_arr(M):
    """which(M, arr.ind = TRUE): 1-based (row, col) pairs, column by column."""
    p = M.shape[0]
    return [(i + 1, j + 1) for j in range(p) for i in range(p) if M[i, j]]


def combn_cols(x, m):
    return [list(c) for c in itertools.combinations(x, m)]


class RUdag2Wanpdag:
    """Transcription of udag2wanpdag(gInput, suffStat, kernelCItest, alpha,
    verbose = FALSE, unfVect = NULL, solve.confl, orientCollider, rules)."""

    def __init__(self, data, method, alpha):
        self.data = data
        self.method = method
        self.alpha = alpha
        self.calls = []

    # regrVonPS(G, V, S, suffStat, indepTest, alpha), 1-based V and S
    def regrVonPS(self, G, V, S):
        p = G.shape[0]
        parentsV = [i + 1 for i in range(p) if G[i, V - 1] == 1 and G[V - 1, i] == 0]
        cols = [v - 1 for v in list(S) + parentsV]
        residV, _ = orc.regrxons(self.data, V - 1, cols, ordered=True)
        pvals = []
        for w in S:
            y = self.data[:, w - 1]
            if self.method == "dcc.gamma":
                pv = orc.dcov_gamma(residV, y)["p.value"]
            else:
                pv = orc.hsic_gamma(residV, y)["p.value"]
            pvals.append(pv)
        self.calls.append((V - 1, [s - 1 for s in S], [q - 1 for q in parentsV], pvals))
        return sum(1 for v in pvals if v < self.alpha)

    @staticmethod
    def checkImmor(pdag, V, S):
        udag = np.minimum(pdag + pdag.T, 1)
        output = True
        parentsV = [i + 1 for i in range(pdag.shape[0]) if pdag[i, V - 1] == 1 and pdag[V - 1, i] == 0]
        if len(S) > 1:
            t1 = udag[np.ix_([s - 1 for s in S], [s - 1 for s in S])] + np.eye(len(S), dtype=int)
            if np.any((t1 == 0) & (t1.T == 0)):
                output = False
        t2 = udag[np.ix_([q - 1 for q in parentsV], [s - 1 for s in S])]
        if t2.sum() < len(S) * len(parentsV):
            output = False
        return output

    @staticmethod
    def rule1(pdag):
        for a, b in which_arr((pdag == 1) & (pdag.T == 0)):
            isC = [c + 1 for c in range(pdag.shape[0])
                   if pdag[b - 1, c] == 1 and pdag[c, b - 1] == 1 and pdag[a - 1, c] == 0
                   and pdag[c, a - 1] == 0]
            for c in isC:
                pdag[b - 1, c - 1] = 1
                pdag[c - 1, b - 1] = 0
        return pdag

    @staticmethod
    def rule2(pdag):
        for a, b in which_arr((pdag == 1) & (pdag.T == 1)):
            isC = [c + 1 for c in range(pdag.shape[0])
                   if pdag[a - 1, c] == 1 and pdag[c, a - 1] == 0 and pdag[c, b - 1] == 1
                   and pdag[b - 1, c] == 0]
            for _ in isC:
                pdag[a - 1, b - 1] = 1
                pdag[b - 1, a - 1] = 0
        return pdag

    @staticmethod
    def rule3(pdag):
        for a, b in which_arr((pdag == 1) & (pdag.T == 1)):
            c = [k + 1 for k in range(pdag.shape[0])
                 if pdag[a - 1, k] == 1 and pdag[k, a - 1] == 1 and pdag[k, b - 1] == 1
                 and pdag[b - 1, k] == 0]
            if len(c) >= 2:
                for c1, c2 in combn_cols(c, 2):
                    if pdag[c1 - 1, c2 - 1] == 0 and pdag[c2 - 1, c1 - 1] == 0:
                        pdag[a - 1, b - 1] = 1
                        pdag[b - 1, a - 1] = 0
                        break
        return pdag

    @staticmethod
    def nbhd(pdag, n):
        ind = which_arr((pdag == 1) & (pdag.T == 1))
        und = [[j for (i, j) in ind if i == v] for v in range(1, n + 1)]
        return und, [len(u) for u in und]

    def run(self, g, sepset):
        p = g.shape[0]
        pdag = g.copy()
        if g.sum() == 0:
            return pdag
        for x, y in which_arr(g == 1):
            allZ = [z for z in range(1, p + 1) if g[y - 1, z - 1] == 1 and z != x]
            for z in allZ:
                if g[x - 1, z - 1] == 0 and not (y in sepset[x - 1][z - 1] or y in sepset[z - 1][x - 1]):
                    pdag[x - 1, y - 1] = pdag[z - 1, y - 1] = 1
                    pdag[y - 1, x - 1] = pdag[y - 1, z - 1] = 0
        s = 1
        n = p
        undi, size = self.nbhd(pdag, n)
        updt = [False] * n
        while max(size) >= s:
            for i in range(1, n + 1):
                V = i
                s2 = s
                if size[V - 1] == s2 or (size[V - 1] < s2 and updt[V - 1]):
                    s2 = size[V - 1]
                    while s2 > 0:
                        if len(undi[V - 1]) == 1:
                            sub = [[undi[V - 1][0]]]
                        else:
                            sub = combn_cols(undi[V - 1], s2)
                        for S in sub:
                            to_update = True
                            if self.checkImmor(pdag, V, S):
                                pval1 = self.regrVonPS(pdag, V, S)
                                if pval1 > 0:
                                    to_update = False
                                else:
                                    for W in S:
                                        s3 = size[W - 1]
                                        to_continue = True
                                        while to_continue and s3 > 0:
                                            if len(undi[W - 1]) == 1:
                                                subW = [[undi[W - 1][0]]]
                                            else:
                                                subW = combn_cols(undi[W - 1], s3)
                                            for S2 in subW:
                                                if self.checkImmor(pdag, W, S2):
                                                    if sum(1 for v in S2 if v == V) == 1:
                                                        pval2 = self.regrVonPS(pdag, W, S2)
                                                        if pval2 == 0:
                                                            to_update = False
                                                            to_continue = False
                                            s3 -= 1
                                if to_update:
                                    pdag[V - 1, [v - 1 for v in S]] = 0
                                    pdag[[v - 1 for v in S], V - 1] = 1
                                    rest = [v for v in undi[V - 1] if v not in S]
                                    if rest:
                                        pdag[[v - 1 for v in rest], V - 1] = 0
                                        pdag[V - 1, [v - 1 for v in rest]] = 1
                                    s2 = 0
                                    while True:
                                        old = pdag.copy()
                                        pdag = self.rule1(pdag)
                                        pdag = self.rule2(pdag)
                                        pdag = self.rule3(pdag)
                                        if np.array_equal(pdag, old):
                                            break
                                    undi, size = self.nbhd(pdag, n)
                                    updt[V - 1] = True
                        s2 -= 1
            s += 1
        return pdag


def synthetic_skeleton(cfg):
    from fastkpc_amd.sem import config_data
    data, _, _ = config_data(cfg)
    sk = orc.skeleton(data, method="dcc.gamma", threads=8)
    return data, sk


@pytest.mark.slow
@pytest.mark.parametrize("cfg,method", [("sem200x8", "dcc.gamma"), ("sem200x8", "hsic.gamma")])
def test_oracle_orientation_equals_r_transcription(cfg, method):
    data, sk = synthetic_skeleton(cfg)
    p = data.shape[1]
    g = sk["adjacency"].astype(int)
    sepset = [[(s if s is not None else []) for s in row] for row in sk["sepsets"]]
    r = RUdag2Wanpdag(data, method, 0.1)
    want = r.run(g, sepset)
    got = orc.wanpdag(data, sk["adjacency"], sk["sepsets"], method=method, alpha=0.1)
    assert np.array_equal(got["pdag"], want)
    calls = [(c["V"], c["S"], c["parents"]) for c in got["calls"]]
    assert calls == [(v, s, q) for v, s, q, _ in r.calls]
    mine = [t["out"][0] for t in got["tests"]]
    assert mine == [pv for *_, pvs in r.calls for pv in pvs]


def test_orientation_on_a_hand_built_skeleton_matches_transcription():
    """A skeleton with no collider (chain 0 - 1 - 2, triangle 1 - 2 - 3,
    3 - 4, and a separate 5 - 6) on nonlinear non-Gaussian data: every edge
    enters the generalised transitive step, so regrVonPS acceptances and
    rejections and rules 1-3 run; the oracle against the transcription."""
    rng = np.random.default_rng(3)
    n, p = 200, 7
    x = rng.uniform(-1.5, 1.5, (n, p))
    x[:, 1] += np.tanh(2 * x[:, 0])
    x[:, 2] += 0.8 * x[:, 1] ** 2
    x[:, 3] += np.sin(2 * x[:, 1]) + 0.5 * x[:, 2]
    x[:, 4] += np.tanh(x[:, 3]) ** 2
    x[:, 6] += np.sin(2 * x[:, 5])
    g = np.zeros((p, p), int)
    for a, b in [(0, 1), (1, 2), (1, 3), (2, 3), (3, 4), (5, 6)]:
        g[a, b] = g[b, a] = 1
    sep = [[None if (a == b or g[a, b]) else [] for b in range(p)] for a in range(p)]
    for (a, b), s1 in {(0, 2): [2], (0, 3): [2], (0, 4): [4], (1, 4): [4], (2, 4): [4]}.items():
        sep[a][b] = sep[b][a] = s1
    sepset = [[(s if s is not None else []) for s in row] for row in sep]
    for method in ("dcc.gamma", "hsic.gamma"):
        r = RUdag2Wanpdag(x, method, 0.1)
        # the transcription's regrVonPS uses numCol = 35 (the oracle default)
        want = r.run(g, sepset)
        got = orc.wanpdag(x, g.astype(bool), sep, method=method, alpha=0.1)
        assert np.array_equal(got["pdag"], want), method
        calls = [(c["V"], c["S"], c["parents"]) for c in got["calls"]]
        assert calls == [(v, s, q) for v, s, q, _ in r.calls]
        assert len(calls) > 3
        assert [t["out"][0] for t in got["tests"]] == [pv for *_, pvs in r.calls for pv in pvs]
