"""Pins the CPU oracle against the reference's golden vectors, textbook /
library routines, brute-force restatements and invariants."""
import numpy as np
import pytest
from scipy import special

from oracle import orc


def kpcalg_dcov_gamma_numpy(x, y, num_col):
    """kpcalg/R/dcovgamma.R:56-82 transcribed literally (explicit H matrices,
    dense eigh, eigs() ordering by |lambda|) for small n."""
    n = len(x)
    H = np.eye(n) - np.full((n, n), 1.0 / n)
    matx = np.abs(x[:, None] - x[None, :])
    maty = np.abs(y[:, None] - y[None, :])

    def eigs(M):
        w, V = np.linalg.eigh(M)
        idx = sorted(range(n), key=lambda i: (-abs(w[i]), -i))[:num_col]
        return V[:, idx], np.diag(w[idx])

    Ux, Sx = eigs(matx)
    Uy, Sy = eigs(maty)
    nV2 = np.trace((H @ Ux) @ Sx @ ((Ux.T @ H) @ (H @ Uy)) @ Sy @ (Uy.T @ H)) / n
    mean = matx.mean() * maty.mean()
    vx = np.trace((H @ Ux) @ Sx @ ((Ux.T @ H) @ (H @ Ux)) @ Sx @ (Ux.T @ H))
    vy = np.trace((H @ Uy) @ Sy @ ((Uy.T @ H) @ (H @ Uy)) @ Sy @ (Uy.T @ H))
    var = 2 * (n - 4) * (n - 5) / n / (n - 1) / (n - 2) / (n - 3) * vx * vy / n ** 4 * n ** 2
    alpha = mean ** 2 / var
    beta = var / mean
    p = 1 - special.gammainc(alpha, nV2 / beta)
    return p, nV2, mean, var


def test_dcov_gamma_matches_reference_fixture(dcov_fixture):
    # fastkpc/tests/fixtures/legacy_dcov_gamma_oracle_v1.rds: kpcalg outputs at n=351
    for c in dcov_fixture:
        r = orc.dcov_gamma(c["rx"], c["ry"], 35)
        for key, ref in (("p.value", c["p_value"]), ("nV2", c["nV2"]), ("mean", c["nV2Mean"]),
                         ("variance", c["nV2Variance"])):
            assert abs(r[key] - ref) <= 1e-9 * max(abs(ref), 1e-12), (c["case_id"], key, r[key], ref)
        assert (r["p.value"] >= 0.1) == c["delete_edge"]


@pytest.mark.parametrize("a", [0.5, 1.0, 2.7, 3.8, 5.6, 12.0, 40.0])
def test_pgamma_matches_scipy(a):
    for x in np.concatenate([np.linspace(1e-3, a * 4, 37), [a, a + 1, a + 1 + 1e-9]]):
        ref = special.gammainc(a, x)
        got = orc.pgamma_lower(x, a)
        assert abs(got - ref) <= 2e-14 * max(ref, 1e-300) + 1e-16, (a, x, got, ref)


def test_gamma_pvalue_special_cases():
    # a = 1: exponential law, P(1, x) = 1 - exp(-x)
    for x in (0.1, 1.0, 3.0):
        assert abs(orc.pgamma_lower(x, 1.0) - (1 - np.exp(-x))) < 1e-15
    assert orc.pgamma_lower(0.0, 2.0) == 0.0


def test_dcov_component_vs_eigh():
    rng = np.random.default_rng(5)
    x = rng.standard_normal(120)
    vals, vecs, dsum, mom = orc.dcov_component(x, 20)
    D = np.abs(x[:, None] - x[None, :])
    w, V = np.linalg.eigh(D)
    idx = sorted(range(len(x)), key=lambda i: (-abs(w[i]), -i))[:20]
    np.testing.assert_allclose(vals, w[idx], rtol=1e-12, atol=1e-12 * abs(w).max())
    for c, i in enumerate(idx):
        v = V[:, i] - V[:, i].mean()
        s = np.sign(v @ vecs[:, c])
        np.testing.assert_allclose(vecs[:, c], s * v, atol=1e-10)
    assert abs(dsum - D.sum()) <= 1e-12 * D.sum()
    C = vecs
    ref_mom = np.sum(np.outer(vals, vals) * (C.T @ C) ** 2)
    assert abs(mom - ref_mom) <= 1e-12 * abs(ref_mom)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_dcov_gamma_vs_literal_kpcalg(seed):
    rng = np.random.default_rng(seed)
    n = 48
    x = rng.standard_normal(n)
    y = 0.5 * x ** 2 + rng.standard_normal(n) * (0.5 + seed * 0.3)
    p, nV2, mean, var = kpcalg_dcov_gamma_numpy(x, y, 9)
    r = orc.dcov_gamma(x, y, 9)
    assert abs(r["nV2"] - nV2) <= 1e-11 * abs(nV2)
    assert abs(r["mean"] - mean) <= 1e-13 * abs(mean)
    assert abs(r["variance"] - var) <= 1e-11 * abs(var)
    assert abs(r["p.value"] - p) <= 1e-11 * max(p, 1e-6)


def test_dcov_gamma_invariances():
    rng = np.random.default_rng(9)
    x = rng.standard_normal(80)
    y = np.tanh(x) + rng.standard_normal(80)
    base = orc.dcov_gamma(x, y, 12)
    # translation and scaling leave the test unchanged; swapping x and y too
    for xx, yy in ((x + 3.0, y), (2.5 * x, y), (y, x), (x, -y)):
        r = orc.dcov_gamma(xx, yy, 12)
        assert abs(r["p.value"] - base["p.value"]) <= 1e-10


@pytest.mark.parametrize("S", [[2], [3, 7], [1, 5, 9], [2, 4, 6, 8], [1, 3, 5, 7, 9, 11],
                               [2, 6, 10, 14, 18, 22, 26, 30]])
def test_gcv_derivatives_match_central_differences(dev500, S):
    data = dev500["data"]
    m = 1 if len(S) <= 2 else len(S)
    lsp = np.linspace(-1.0, 1.5, m)  # interior points with non-trivial gradients
    score, grad, hess = orc.gcv_eval(data, 0, S, lsp)
    assert np.abs(grad).max() > 1e-4 * score
    h = 1e-5
    for i in range(m):
        e = np.zeros(m)
        e[i] = h
        sp, gp, _ = orc.gcv_eval(data, 0, S, lsp + e)
        sm, gm, _ = orc.gcv_eval(data, 0, S, lsp - e)
        fd = (sp - sm) / (2 * h)
        assert abs(fd - grad[i]) <= 1e-6 * abs(grad[i]) + 1e-9 * score, (i, fd, grad[i])
        fdh = (gp - gm) / (2 * h)
        np.testing.assert_allclose(fdh, hess[i], rtol=1e-4, atol=1e-6 * (np.abs(hess).max() + score))


@pytest.mark.parametrize("S", [[2], [3, 7], [1, 5, 9], [2, 4, 6, 8], [1, 3, 5, 7, 9, 11],
                               [2, 6, 10, 14, 18, 22, 26, 30]])
def test_gcv_selection_meets_the_stopping_rule_and_fits_unpenalised_terms(dev500, S):
    data = dev500["data"]
    resid, info = orc.regrxons(data, 0, S)
    lsp = info["log_sp"]
    s0, g, _ = orc.gcv_eval(data, 0, S, lsp)
    assert abs(s0 - info["score"]) <= 1e-12 * s0
    # magic's stopping rule on its own gradient (mgcv_multi_penalty_cpp.cpp:1225-1244)
    assert np.linalg.norm(g) <= 1e-7 ** (1.0 / 3.0) * (1.0 + abs(s0))
    if len(S) <= 3:
        # full-rank end points here: a local minimum along every coordinate.
        # (With m >= 4 some log sp run to ~24, the numerical rank drops and
        # magic's truncated-basis gradient is no longer the score's exact
        # derivative, so mgcv stops where the exact slope is ~1e-3.)
        for i in range(len(lsp)):
            for d in (-0.05, 0.05):
                e = np.zeros(len(lsp))
                e[i] = d
                s1, _, _ = orc.gcv_eval(data, 0, S, lsp + e)
                assert s1 >= s0 * (1 - 1e-9)
    # intercept and the linear null-space terms are unpenalised: residual is
    # orthogonal to them
    y = data[:, 0]
    assert abs(resid.sum()) <= 1e-9 * np.abs(y).sum()
    for j in S:
        xs = data[:, j] - data[:, j].mean()
        assert abs(resid @ xs) <= 1e-8 * np.linalg.norm(y) * np.linalg.norm(xs)


def test_residual_scale_equivariance(dev500):
    data = np.array(dev500["data"])
    r1, _ = orc.regrxons(data, 4, [6])
    d2 = data.copy()
    d2[:, 4] *= 3.0
    r2, _ = orc.regrxons(d2, 4, [6])
    np.testing.assert_allclose(r2, 3.0 * r1, rtol=1e-8, atol=1e-9 * np.abs(d2[:, 4]).max())


@pytest.mark.slow
def test_oracle_skeleton_reproduces_500x50_fixture(dev500):
    # fastkpc/tests/fixtures/full_cuda_ci_development_500x50_v2.rds: kpcalg/mgcv
    # dcc.gamma skeleton (adjacency, sepsets, n.edgetests, 2722 p-values)
    d = dev500
    res = orc.skeleton(d["data"], alpha=0.1, m_max=48, num_col=35, threads=8)
    assert res["n.edgetests"] == [int(v) for v in d["n_edgetests"]]
    assert np.array_equal(res["adjacency"], d["adjacency"])
    p = d["adjacency"].shape[0]
    for row in d["sepsets"]:
        i, j, ln = row[0], row[1], row[2]
        if d["adjacency"][i, j] or i == j:
            continue
        assert res["sepsets"][i][j] == [int(v) for v in row[3:3 + ln]], (i, j)
    t = res["tasks"]
    from conftest import trace_map
    mine = trace_map(t["level"], t["x"], t["y"], t["S_key"], t["p_used"])
    ref = trace_map(d["lt_level"], d["lt_x"], d["lt_y"], d["lt_S"], d["lt_p"])
    assert set(mine) == set(ref)
    worst = max(abs(mine[k] - ref[k]) for k in ref)
    # the fixture's own eigenpairs carry RSpectra's tol = 1e-10 noise
    assert worst <= 1e-8, worst


@pytest.mark.parametrize("S", [[2, 4, 6, 8], [1, 3, 5, 7, 9, 11], [2, 6, 10, 14, 18, 22, 26, 30]])
def test_multi_penalty_fit_is_equivariant_under_column_order_and_response_scale(dev500, S):
    """|S| >= 4 (m >= 4 penalties), where no reference fixture reaches: the
    additive fit does not depend on the order of the conditioning columns
    (reversing them permutes the penalty blocks and the selected log sp the
    same way, residuals unchanged) and scales with the response (y -> 3 y:
    same log sp, residuals x 3, score x 9).  A block-offset or penalty-index
    mistake breaks the first, a response-dependent initial.sp the second."""
    data = np.array(dev500["data"])
    base, info = orc.regrxons(data, 0, S)
    # reverse the conditioning columns in place: sorted S now names them in
    # the opposite order
    d2 = data.copy()
    d2[:, S] = data[:, S[::-1]]
    r2, i2 = orc.regrxons(d2, 0, S)
    np.testing.assert_allclose(r2, base, rtol=0, atol=1e-7 * np.abs(base).max())
    assert abs(i2["score"] - info["score"]) <= 1e-9 * info["score"]
    flat = np.abs(info["log_sp"]) > 12  # effectively unpenalised or fully penalised blocks
    np.testing.assert_allclose(i2["log_sp"][::-1][~flat], info["log_sp"][~flat], atol=1e-4)
    d3 = data.copy()
    d3[:, 0] *= 3.0
    r3, i3 = orc.regrxons(d3, 0, S)
    np.testing.assert_allclose(r3, 3.0 * base, rtol=0, atol=1e-8 * np.abs(3.0 * base).max())
    assert abs(i3["score"] - 9.0 * info["score"]) <= 1e-9 * i3["score"]
    np.testing.assert_allclose(i3["log_sp"], info["log_sp"], atol=1e-6)
