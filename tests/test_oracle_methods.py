"""Pins the oracle's hsic.gamma, hsic.perm and dcc.perm restatements and its
counter-based permutation generator.

No reference fixture covers these methods (kernlab/energy are absent and no
golden file holds HSIC or permutation outputs), so each function is pinned
by an independent route: published known-answer vectors (Philox), a textbook
Schur-complement pivoted Cholesky on the explicit kernel matrix, literal
numpy transcriptions of kpcalg/R/hsicgamma.R:57-83 and hsicperm.R:61-88 (SVD
route), and brute-force energy::dcov.test sums over explicit centred
distance matrices with the same permutation indices.
"""
import itertools

import numpy as np
import pytest
from scipy import special, stats

from oracle import orc


# ------------------------------------------------------------------ Philox
# Random123 known-answer vectors for philox4x32 with 10 rounds.
@pytest.mark.parametrize("ctr,key,want", [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
])
def test_philox_known_answers(ctr, key, want):
    assert orc.philox4x32_10(ctr, key) == want


def test_permutation_is_a_deterministic_bijection():
    for n in (2, 3, 17, 351):
        for r in range(5):
            p = orc.permutation(987654321, r, n)
            assert sorted(p.tolist()) == list(range(n))
            assert np.array_equal(p, orc.permutation(987654321, r, n))
    assert not np.array_equal(orc.permutation(1, 0, 50), orc.permutation(1, 1, 50))
    assert not np.array_equal(orc.permutation(1, 0, 50), orc.permutation(2, 0, 50))


def test_permutation_follows_durstenfeld_on_philox_draws():
    # restate the shuffle from raw Philox words: draw i of replicate r uses
    # counter {i >> 1, r, 0x4b504331, 0}, words (1,0) for even i, (3,2) for odd
    seed, rep, n = 0x0123456789ABCDEF, 7, 33
    key = [seed & 0xFFFFFFFF, seed >> 32]
    perm = list(range(n))
    for i in range(n - 1, 0, -1):
        w = orc.philox4x32_10([i >> 1, rep, 0x4b504331, 0], key)
        u = (w[3] << 32 | w[2]) if i & 1 else (w[1] << 32 | w[0])
        j = (u * (i + 1)) >> 64
        perm[i], perm[j] = perm[j], perm[i]
    assert perm == orc.permutation(seed, rep, n).tolist()


def test_permutation_is_uniform_on_small_n():
    n, R = 4, 24000
    counts = {}
    for r in range(R):
        key = tuple(orc.permutation(42, r, n).tolist())
        counts[key] = counts.get(key, 0) + 1
    assert len(counts) == 24
    chi2 = sum((c - R / 24) ** 2 / (R / 24) for c in counts.values())
    assert stats.chi2.sf(chi2, 23) > 1e-4


def test_task_seed_is_an_orientation_free_identity():
    s = orc.task_seed(5, "hsic.perm", 3, 9, [1, 4])
    assert s == orc.task_seed(5, "hsic.perm", 9, 3, [1, 4])
    others = {orc.task_seed(5, "dcc.perm", 3, 9, [1, 4]), orc.task_seed(6, "hsic.perm", 3, 9, [1, 4]),
              orc.task_seed(5, "hsic.perm", 3, 9, [1, 5]), orc.task_seed(5, "hsic.perm", 3, 8, [1, 4]),
              orc.task_seed(5, "hsic.perm", 3, 9, [1])}
    assert s not in others and len(others) == 5


# ------------------------------------------------------------------ inchol
def inchol_schur(x, sig, tol, maxr):
    """Pivoted incomplete Cholesky in Schur-complement form on the explicit
    kernel matrix: column = residual kernel column / sqrt(its pivot)."""
    K = np.exp(-(1.0 / sig) * (x[:, None] - x[None, :]) ** 2)
    n = len(x)
    G = np.zeros((n, 0))
    d = np.ones(n)
    piv, p, res = [], 0, 1.0
    while res > tol and G.shape[1] < maxr:
        s = K[:, p] - G @ G[p, :]
        if not s[p] > 0:
            break
        g = s / np.sqrt(s[p])
        G = np.column_stack([G, g])
        d = d - g * g
        piv.append(p)
        p = int(np.argmax(d))
        res = d[p]
    return G, piv, K


@pytest.mark.parametrize("n,sig,maxr,seed", [(60, 1.0, 35, 0), (150, 1.0, 35, 1), (150, 0.2, 35, 2),
                                             (200, 0.05, 12, 3), (90, 4.0, 35, 4)])
def test_inchol_matches_schur_complement_form(n, sig, maxr, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_t(4, size=n)
    G, piv, margin = orc.inchol(x, sig=sig, max_rank=maxr)
    G2, piv2, K = inchol_schur(x, sig, 1e-3, maxr)
    if margin > 1e-10:
        assert list(piv) == piv2
        assert np.abs(G - G2).max() <= 1e-11
    resid = np.diag(K) - np.sum(G * G, axis=1)
    assert G.shape[1] == maxr or resid.max() <= 1e-3
    assert G.shape[1] == maxr or G.shape[1] < maxr and resid.max() <= 1e-3
    # exact interpolation on the pivot columns: (K - G G^T)[:, piv] = 0
    assert np.abs((K - G @ G.T)[:, piv]).max() <= 1e-12
    assert piv[0] == 0


def test_inchol_full_rank_is_the_pivoted_cholesky():
    rng = np.random.default_rng(7)
    x = rng.normal(size=12) * 3.0
    G, piv, _ = orc.inchol(x, sig=1.0, tol=-1.0, max_rank=12)
    K = np.exp(-(x[:, None] - x[None, :]) ** 2)
    assert G.shape[1] >= 11
    assert np.abs(K - G @ G.T).max() <= 1e-12


# -------------------------------------------------------------- hsic.gamma
def kpcalg_hsic_gamma_numpy(Gx, Gy):
    """kpcalg/R/hsicgamma.R:61-83 transcribed (explicit H, SVD route)."""
    n = Gx.shape[0]
    H = np.eye(n) - np.full((n, n), 1.0 / n)
    Ux, Sx, _ = np.linalg.svd(H @ Gx, full_matrices=False)
    Uy, Sy, _ = np.linalg.svd(H @ Gy, full_matrices=False)
    MSx, MSy = np.diag(Sx ** 2), np.diag(Sy ** 2)
    hsic = np.trace(Ux @ ((MSx @ Ux.T) @ (Uy @ MSy)) @ Uy.T) / n ** 2
    K, L = Gx @ Gx.T, Gy @ Gy.T
    mux = 1 / (n * (n - 1)) * np.sum(K - np.diag(np.diag(K)))
    muy = 1 / (n * (n - 1)) * np.sum(L - np.diag(np.diag(L)))
    mean = 1 / n * (1 + mux * muy - mux - muy)
    var = ((2 * (n - 4) * (n - 5) / (n * (n - 1) * (n - 2) * (n - 3)))
           * np.trace(Ux @ ((MSx @ Ux.T) @ (Ux @ MSx)) @ Ux.T)
           * np.trace(Uy @ ((MSy @ Uy.T) @ (Uy @ MSy)) @ Uy.T) / n ** 4)
    alpha, beta = mean ** 2 / var, var / mean
    return 1 - special.gammainc(alpha, hsic / beta), hsic, mean, var


@pytest.mark.parametrize("seed,dep,sig", [(0, 1.0, 1.0), (1, 0.0, 1.0), (2, 0.3, 0.5), (3, 0.1, 2.0)])
def test_hsic_gamma_matches_kpcalg_transcription(seed, dep, sig):
    rng = np.random.default_rng(seed)
    n = 150
    x = rng.normal(size=n)
    y = dep * np.sin(2 * x) + rng.normal(size=n)
    Gx, _, _ = orc.inchol(x, sig=sig)
    Gy, _, _ = orc.inchol(y, sig=sig)
    p, hsic, mean, var = kpcalg_hsic_gamma_numpy(Gx, Gy)
    r = orc.hsic_gamma(x, y, sig=sig)
    for got, want in ((r["statistic"], hsic), (r["mean"], mean), (r["variance"], var)):
        assert abs(got - want) <= 1e-12 * abs(want)
    assert abs(r["p.value"] - p) <= 1e-10 * max(p, 1e-300) + 1e-14


def test_hsic_gamma_is_symmetric_and_shift_invariant():
    rng = np.random.default_rng(11)
    x = rng.normal(size=100)
    y = x ** 2 + rng.normal(size=100)
    a = orc.hsic_gamma(x, y)
    b = orc.hsic_gamma(y, x)
    assert abs(a["statistic"] - b["statistic"]) <= 1e-13 * a["statistic"]
    assert abs(a["p.value"] - b["p.value"]) <= 1e-12


# --------------------------------------------------------------- hsic.perm
def test_hsic_perm_matches_kpcalg_transcription():
    rng = np.random.default_rng(5)
    n, R, seed = 80, 60, 99
    x = rng.normal(size=n)
    y = 0.2 * np.cos(x) + rng.normal(size=n)
    Gx, _, _ = orc.inchol(x)
    Gy, _, _ = orc.inchol(y)
    H = np.eye(n) - 1.0 / n
    Ux, Sx, _ = np.linalg.svd(H @ Gx, full_matrices=False)
    Uy, Sy, _ = np.linalg.svd(H @ Gy, full_matrices=False)
    MSx, MSy = np.diag(Sx ** 2), np.diag(Sy ** 2)
    hsic = np.trace(Ux @ ((MSx @ Ux.T) @ (Uy @ MSy)) @ Uy.T) / n ** 2
    reps = []
    for r in range(R):
        perm = orc.permutation(seed, r, n)
        Up = Uy[perm, :]
        reps.append(np.trace(Ux @ ((MSx @ Ux.T) @ (Up @ MSy)) @ Up.T) / n ** 2)
    reps = np.array(reps)
    got = orc.hsic_perm(x, y, replicates=R, seed=seed)
    assert abs(got["statistic"] - hsic) <= 1e-12 * hsic
    gaps = np.abs(reps - hsic) / hsic
    assert gaps.min() > 1e-9  # no replicate sits on the observed value
    assert got["p.value"] == np.mean(np.append(reps, hsic) >= hsic)
    assert abs(got["mean"] - reps.mean()) <= 1e-12 * reps.mean()
    assert abs(got["variance"] - reps.var(ddof=1)) <= 1e-9 * reps.var(ddof=1)
    off = orc.hsic_perm(x, y, replicates=R, seed=seed, include_observed=False)
    assert off["p.value"] == np.mean(reps >= hsic)


# ---------------------------------------------------------------- dcc.perm
def energy_dcov_perm_numpy(x, y, R, seed, index=1.0):
    """energy::dcov.test statistic n V^2 and permutation p with explicit
    double-centred distance matrices."""
    n = len(x)

    def centred(v):
        D = np.abs(v[:, None] - v[None, :]) ** index
        return D - D.mean(axis=1, keepdims=True) - D.mean(axis=0, keepdims=True) + D.mean()

    A, B = centred(x), centred(y)
    obs = np.sum(A * B)
    reps = np.array([np.sum(A * B[np.ix_(p, p)])
                     for p in (orc.permutation(seed, r, n) for r in range(R))])
    return obs / n, reps, obs


@pytest.mark.parametrize("index", [1.0, 0.5, 1.7])
def test_dcc_perm_matches_brute_force(index):
    rng = np.random.default_rng(int(index * 10))
    n, R, seed = 70, 80, 2024
    x = rng.normal(size=n)
    y = 0.3 * np.abs(x) + rng.normal(size=n)
    stat, reps, obs = energy_dcov_perm_numpy(x, y, R, seed, index)
    got = orc.dcc_perm(x, y, index=index, replicates=R, seed=seed)
    assert abs(got["statistic"] - stat) <= 1e-11 * abs(stat)
    assert np.abs(reps - obs).min() / abs(obs) > 1e-9
    assert got["p.value"] == (1 + np.sum(reps >= obs)) / (R + 1)
    assert abs(got["mean"] - reps.mean() / n) <= 1e-10 * abs(reps.mean() / n)


def test_dcc_perm_statistic_is_n_times_dcov_squared():
    # n V_n^2 with V_n^2 = mean(A * B): the textbook distance covariance of
    # Szekely, Rizzo & Bakirov (2007), computed from its pairwise definition
    rng = np.random.default_rng(3)
    n = 40
    x, y = rng.normal(size=n), rng.normal(size=n)
    a = np.abs(x[:, None] - x[None, :])
    b = np.abs(y[:, None] - y[None, :])
    v2 = (np.mean(a * b) + a.mean() * b.mean()
          - 2 * np.mean(a.mean(axis=1) * b.mean(axis=1)))
    got = orc.dcc_perm(x, y, replicates=0)
    assert abs(got["statistic"] - n * v2) <= 1e-12 * n * v2
    assert got["p.value"] == 1.0  # only the observed value: (0 + 1) / (0 + 1)


def test_permutation_p_value_is_exact_under_a_planted_dependence():
    rng = np.random.default_rng(8)
    x = rng.normal(size=60)
    r = orc.dcc_perm(x, x ** 2, replicates=199, seed=1)
    assert r["p.value"] == 1.0 / 200.0
    h = orc.hsic_perm(x, x ** 2, replicates=199, seed=1)
    assert h["p.value"] == 1.0 / 200.0


# ------------------------------------------------------------ ci_test glue
@pytest.mark.parametrize("method", ["hsic.gamma", "hsic.perm", "dcc.perm"])
def test_ci_test_at_level_zero_uses_raw_columns(method):
    rng = np.random.default_rng(17)
    data = rng.normal(size=(90, 4))
    data[:, 1] += 0.4 * np.tanh(data[:, 0])
    r = orc.ci_test(data, 0, 1, [], method=method, replicates=50, seed=3)
    lo, hi = data[:, 0], data[:, 1]
    seed = orc.task_seed(3, method, 0, 1, [])
    if method == "hsic.gamma":
        want = orc.hsic_gamma(lo, hi)
    elif method == "hsic.perm":
        want = orc.hsic_perm(lo, hi, replicates=50, seed=seed)
    else:
        want = orc.dcc_perm(lo, hi, replicates=50, seed=seed)
    assert r["p.value"] == want["p.value"]
    assert r["statistic"] == want["statistic"]
    # orientation-free: (y, x | S) is the same test
    r2 = orc.ci_test(data, 1, 0, [], method=method, replicates=50, seed=3)
    assert r2["p.value"] == r["p.value"]
