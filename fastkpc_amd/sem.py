"""Synthetic nonlinear SEM generators for the benchmark and parity workloads.

A random DAG in topological order 1..p with independent edge probability;
each node x_j = sum over parents k of (w_jk * tanh(x_k) + 0.5 * sin(x_k)) plus
noise, with w ~ U[0.5, 1.5] times a random sign; root nodes are N(0, 1).
numpy's PCG64 stream makes every draw reproducible from the seed.
"""
import numpy as np

CONFIGS = {
    # n=200, p=8, hsic.gamma workload (CPU-runnable case)
    "sem200x8": dict(n=200, p=8, edge_prob=0.3, noise="fixed", noise_sd=0.5, seed=1),
    # n=351, p=48, dcc.gamma headline workload
    "sem351x48": dict(n=351, p=48, edge_prob=0.08, noise="uniform_var", var_lo=0.25, var_hi=1.0,
                      seed=2),
    # n=500, p=50, batched-form workload with 10% Student-t(3) noise nodes
    "sem500x50": dict(n=500, p=50, edge_prob=0.06, noise="t3_mix", t_frac=0.1, var_lo=0.25,
                      var_hi=1.0, seed=3),
    # n=2000, p=64, permutation / large-additive stress workload, mixed links
    "sem2000x64": dict(n=2000, p=64, edge_prob=0.05, noise="uniform_var", var_lo=0.25,
                       var_hi=1.0, seed=5, link="mixed"),
}


def random_dag(p, edge_prob, rng):
    """Upper-triangular parent mask: parents of j are k < j with mask[k, j]."""
    u = rng.random((p, p))
    return np.triu(u < edge_prob, k=1)


def generate(n, p, edge_prob, seed, noise="fixed", noise_sd=0.5, var_lo=0.25, var_hi=1.0,
             t_frac=0.0, link="tanh_sin"):
    rng = np.random.Generator(np.random.PCG64(seed))
    mask = random_dag(p, edge_prob, rng)
    w = rng.uniform(0.5, 1.5, size=(p, p)) * np.where(rng.random((p, p)) < 0.5, -1.0, 1.0)
    if noise == "fixed":
        sd = np.full(p, noise_sd)
    else:
        sd = np.sqrt(rng.uniform(var_lo, var_hi, size=p))
    heavy = np.zeros(p, dtype=bool)
    if noise == "t3_mix":
        k = max(1, int(round(t_frac * p)))
        heavy[rng.choice(p, size=k, replace=False)] = True
    links = rng.integers(0, 3, size=(p, p))
    X = np.zeros((n, p))
    for j in range(p):
        parents = np.nonzero(mask[:, j])[0]
        if len(parents) == 0:
            X[:, j] = rng.standard_normal(n)
            continue
        acc = np.zeros(n)
        for k in parents:
            xk = X[:, k]
            if link == "tanh_sin":
                acc += w[k, j] * np.tanh(xk) + 0.5 * np.sin(xk)
            else:  # mixed tanh / quadratic / sin link per edge
                kind = links[k, j]
                if kind == 0:
                    acc += w[k, j] * np.tanh(xk)
                elif kind == 1:
                    acc += 0.3 * w[k, j] * xk * xk
                else:
                    acc += w[k, j] * np.sin(xk)
        if heavy[j]:
            eps = rng.standard_t(3, size=n) * sd[j]
        else:
            eps = rng.standard_normal(n) * sd[j]
        X[:, j] = acc + eps
    return np.asfortranarray(X), mask


def config_data(name):
    c = dict(CONFIGS[name])
    data, mask = generate(**c)
    return data, mask, c


def regression_batch(fits=1024, n=500, s_lo=1, s_hi=6, seed=4):
    """Standalone regrXonS batch: fit f draws |S_f| ~ U{s_lo..s_hi} predictors
    x_k ~ U[-2, 2] and the target y = sum_k sin(2 x_k) + 0.3 x_k^2 + N(0, 0.3^2).
    Returns the n x P data matrix (per fit: its predictors, then its target),
    the target columns and the conditioning sets (0-based, ascending)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    cols, targets, sets = [], [], []
    for _ in range(fits):
        s = int(rng.integers(s_lo, s_hi + 1))
        X = rng.uniform(-2.0, 2.0, size=(n, s))
        y = np.sum(np.sin(2.0 * X) + 0.3 * X * X, axis=1) + rng.normal(0.0, 0.3, size=n)
        base = len(cols)
        cols.extend(X[:, k] for k in range(s))
        cols.append(y)
        sets.append(list(range(base, base + s)))
        targets.append(base + s)
    return np.asfortranarray(np.column_stack(cols)), np.asarray(targets, dtype=np.int32), sets
