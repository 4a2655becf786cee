"""ctypes binding of the CPU oracle liborc.so.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py to check or time against the reference
algorithm.  The product (fastkpc_amd) never imports this module.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborc.so")
_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)


class OrcSkeletonOptions(ctypes.Structure):
    _fields_ = [("ci_method", ctypes.c_int), ("alpha", ctypes.c_double), ("m_max", ctypes.c_int),
                ("num_col", ctypes.c_int), ("hsic_sig", ctypes.c_double),
                ("replicates", ctypes.c_int), ("perm_seed", ctypes.c_ulonglong),
                ("threads", ctypes.c_int), ("index", ctypes.c_double),
                ("include_observed", ctypes.c_int), ("r_stream", ctypes.c_int)]


METHODS = {"dcc.gamma": 0, "hsic.gamma": 1, "dcc.perm": 2, "hsic.perm": 3}


def options(method="dcc.gamma", alpha=0.1, m_max=62, num_col=35, sig=1.0, replicates=100,
            seed=0, threads=0, index=1.0, include_observed=True, r_stream=False):
    return OrcSkeletonOptions(METHODS[method], alpha, m_max, num_col, sig, replicates, seed,
                              threads, index, 1 if include_observed else 0, 1 if r_stream else 0)


class OrcRng(ctypes.Structure):
    _fields_ = [("mt", ctypes.c_uint * 624), ("mti", ctypes.c_int)]


class OrcSkeletonResult(ctypes.Structure):
    _fields_ = [("p", ctypes.c_int), ("levels", ctypes.c_int),
                ("adjacency", ctypes.POINTER(ctypes.c_ubyte)), ("pmax", _D),
                ("sepset_len", _I), ("sepset", _I), ("n_edgetests", ctypes.c_int * 64),
                ("trace_len", ctypes.c_int), ("trace_cap", ctypes.c_int),
                ("trace_level", _I), ("trace_x", _I), ("trace_y", _I), ("trace_slen", _I),
                ("trace_S", _I), ("trace_p", _D), ("trace_del", ctypes.POINTER(ctypes.c_ubyte)),
                ("trace_margin", _D), ("trace_stat", _D), ("trace_mean", _D),
                ("trace_var", _D), ("rng_out", OrcRng)]


class OrcSetup(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("q", ctypes.c_int), ("m", ctypes.c_int), ("X", _D),
                ("S", ctypes.POINTER(_D)), ("off", _I), ("dim", _I), ("rank", _I)]


class OrcMagicControl(ctypes.Structure):
    _fields_ = [("tol", ctypes.c_double), ("max_step_halving_single", ctypes.c_int),
                ("max_step_halving_multi", ctypes.c_int), ("max_iterations", ctypes.c_int),
                ("max_newton_step", ctypes.c_double), ("boundary_probe_step", ctypes.c_double),
                ("max_boundary_probes", ctypes.c_int), ("rank_tol", ctypes.c_double)]


class OrcMagicResult(ctypes.Structure):
    _fields_ = [("score", ctypes.c_double), ("rss", ctypes.c_double), ("edf", ctypes.c_double),
                ("log_sp", ctypes.c_double * 64), ("iterations", ctypes.c_int),
                ("evaluations", ctypes.c_int), ("step_failed", ctypes.c_int),
                ("margin", ctypes.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.run(["make", "-s", "-C", HERE, "liborc.so"], check=True)
        L = ctypes.CDLL(LIB_PATH)
        L.orc_dcov_gamma.argtypes = [_D, _D, ctypes.c_int, ctypes.c_int, _D]
        L.orc_dcov_component.argtypes = [_D, ctypes.c_int, ctypes.c_int, _D, _D, _D, _D]
        L.orc_pgamma_lower.argtypes = [ctypes.c_double, ctypes.c_double]
        L.orc_pgamma_lower.restype = ctypes.c_double
        L.orc_gamma_pvalue.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.orc_gamma_pvalue.restype = ctypes.c_double
        L.orc_skeleton.argtypes = [_D, ctypes.c_int, ctypes.c_int,
                                   ctypes.POINTER(OrcSkeletonOptions),
                                   ctypes.POINTER(OrcSkeletonResult)]
        L.orc_skeleton_result_free.argtypes = [ctypes.POINTER(OrcSkeletonResult)]
        L.orc_magic_default_control.argtypes = [ctypes.POINTER(OrcMagicControl)]
        L.orc_regr_setup.argtypes = [_D, ctypes.c_int, ctypes.c_int, _I, ctypes.c_int,
                                     ctypes.POINTER(OrcSetup)]
        L.orc_setup_free.argtypes = [ctypes.POINTER(OrcSetup)]
        L.orc_magic_fit.argtypes = [ctypes.POINTER(OrcSetup), _D, ctypes.POINTER(OrcMagicControl),
                                    _D, ctypes.POINTER(OrcMagicResult)]
        L.orc_magic_eval.argtypes = [ctypes.POINTER(OrcSetup), _D, _D, ctypes.c_double, _D, _D, _D]
        L.orc_regrxons.argtypes = [_D, ctypes.c_int, ctypes.c_int, ctypes.c_int, _I, ctypes.c_int,
                                   ctypes.POINTER(OrcMagicControl), _D,
                                   ctypes.POINTER(OrcMagicResult)]
        L.orc_regrxons_at_sp.argtypes = [_D, ctypes.c_int, ctypes.c_int, ctypes.c_int, _I,
                                         ctypes.c_int, _D, ctypes.c_double, _D, _D]
        L.orc_ci_test.argtypes = [_D, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _I,
                                  ctypes.c_int, ctypes.POINTER(OrcSkeletonOptions), _D]
        _U = ctypes.POINTER(ctypes.c_uint)
        L.orc_philox4x32_10.argtypes = [_U, _U, _U]
        L.orc_task_seed.argtypes = [ctypes.c_ulonglong, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    _I, ctypes.c_int]
        L.orc_task_seed.restype = ctypes.c_ulonglong
        L.orc_permutation.argtypes = [ctypes.c_ulonglong, ctypes.c_int, ctypes.c_int, _I]
        L.orc_inchol.argtypes = [_D, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                 _D, _I, _D]
        L.orc_hsic_component.argtypes = [_D, ctypes.c_int, ctypes.c_double, ctypes.c_int, _D, _I,
                                         _D, _D, _D]
        L.orc_hsic_gamma_from_factors.argtypes = [_D, ctypes.c_int, ctypes.c_double,
                                                  ctypes.c_double, _D, ctypes.c_int,
                                                  ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                                  _D]
        L.orc_hsic_perm_from_factors.argtypes = [_D, ctypes.c_int, _D, ctypes.c_int, ctypes.c_int,
                                                 ctypes.c_int, ctypes.c_int, ctypes.c_ulonglong,
                                                 _D]
        L.orc_dcc_perm.argtypes = [_D, _D, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_ulonglong, _D]
        L.orc_rng_set_seed.argtypes = [ctypes.POINTER(OrcRng), ctypes.c_uint]
        L.orc_rng_word.argtypes = [ctypes.POINTER(OrcRng)]
        L.orc_rng_word.restype = ctypes.c_uint
        L.orc_unif_rand.argtypes = [ctypes.POINTER(OrcRng)]
        L.orc_unif_rand.restype = ctypes.c_double
        L.orc_unif_index.argtypes = [ctypes.POINTER(OrcRng), ctypes.c_double, ctypes.c_int]
        L.orc_unif_index.restype = ctypes.c_double
        L.orc_r_sample.argtypes = [ctypes.POINTER(OrcRng), ctypes.c_int, ctypes.c_int, _I, _I]
        L.orc_energy_permute.argtypes = [ctypes.POINTER(OrcRng), _I, ctypes.c_int]
        L.orc_r_perm_table.argtypes = [ctypes.POINTER(OrcRng), ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int, _I, _I]
        L.orc_dcc_perm_table.argtypes = [_D, _D, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_ulonglong, _I, _D]
        L.orc_hsic_perm_table.argtypes = [_D, ctypes.c_int, _D, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_int, ctypes.c_ulonglong, _I, _D]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_D)


def dcov_gamma(x, y, num_col=35):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    out = np.zeros(5)
    rc = lib().orc_dcov_gamma(_d(x), _d(y), len(x), num_col, _d(out))
    if rc != 0:
        raise RuntimeError("orc_dcov_gamma failed %d" % rc)
    return {"p.value": out[0], "nV2": out[1], "mean": out[2], "variance": out[3]}


def dcov_component(x, k=35):
    x = np.ascontiguousarray(x, dtype=np.float64)
    n = len(x)
    vals = np.zeros(k)
    vecs = np.zeros(n * k)
    dsum = np.zeros(1)
    mom = np.zeros(1)
    rc = lib().orc_dcov_component(_d(x), n, k, _d(vals), _d(vecs), _d(dsum), _d(mom))
    if rc != 0:
        raise RuntimeError("orc_dcov_component failed %d" % rc)
    return vals, vecs.reshape(k, n).T, dsum[0], mom[0]


def pgamma_lower(x, a):
    return lib().orc_pgamma_lower(float(x), float(a))


def gamma_pvalue(stat, mean, var):
    return lib().orc_gamma_pvalue(float(stat), float(mean), float(var))


def default_control():
    c = OrcMagicControl()
    lib().orc_magic_default_control(ctypes.byref(c))
    return c


def regrxons(data, target, S, control=None, ordered=False):
    """regrXonS of column `target` on the columns S (0-based); S is sorted
    unless ordered=True, which keeps the given column order (the order of the
    smooth terms in the GAM formula)."""
    d = np.asfortranarray(data, dtype=np.float64)
    n, p = d.shape
    Sa = np.ascontiguousarray(list(S) if ordered else sorted(S), dtype=np.int32)
    resid = np.zeros(n)
    res = OrcMagicResult()
    ctl = control or default_control()
    rc = lib().orc_regrxons(_d(d), n, n, int(target), Sa.ctypes.data_as(_I), len(Sa),
                            ctypes.byref(ctl), _d(resid), ctypes.byref(res))
    if rc != 0:
        raise RuntimeError("orc_regrxons failed %d" % rc)
    m = len(Sa) if len(Sa) > 2 else 1
    return resid, {"score": res.score, "rss": res.rss, "edf": res.edf,
                   "log_sp": np.array(res.log_sp[:m]), "iterations": res.iterations,
                   "evaluations": res.evaluations, "step_failed": res.step_failed,
                   "margin": res.margin}


def regrxons_at_sp(data, target, S, log_sp, rank_tol=1.4901161193847656e-8):
    """The oracle's final regrXonS fit at a given log(sp): (residual, {score, rss, edf})."""
    d = np.asfortranarray(data, dtype=np.float64)
    n, p = d.shape
    Sa = np.ascontiguousarray(sorted(S), dtype=np.int32)
    m = len(Sa) if len(Sa) > 2 else 1
    lsp = np.ascontiguousarray(log_sp, dtype=np.float64).reshape(m)
    resid = np.zeros(n)
    out = np.zeros(3)
    rc = lib().orc_regrxons_at_sp(_d(d), n, n, int(target), Sa.ctypes.data_as(_I), len(Sa),
                                  _d(lsp), rank_tol, _d(resid), _d(out))
    if rc != 0:
        raise RuntimeError("orc_regrxons_at_sp failed %d" % rc)
    return resid, {"score": out[0], "rss": out[1], "edf": out[2]}


def gcv_eval(data, target, S, log_sp, rank_tol=1.4901161193847656e-8):
    d = np.asfortranarray(data, dtype=np.float64)
    n, p = d.shape
    Sa = np.ascontiguousarray(sorted(S), dtype=np.int32)
    st = OrcSetup()
    rc = lib().orc_regr_setup(_d(d), n, n, Sa.ctypes.data_as(_I), len(Sa), ctypes.byref(st))
    if rc != 0:
        raise RuntimeError("orc_regr_setup failed %d" % rc)
    try:
        m = st.m
        lsp = np.ascontiguousarray(log_sp, dtype=np.float64).reshape(m)
        y = np.ascontiguousarray(d[:, target])
        score = np.zeros(1)
        grad = np.zeros(m)
        hess = np.zeros(m * m)
        rc = lib().orc_magic_eval(ctypes.byref(st), _d(y), _d(lsp), rank_tol, _d(score), _d(grad),
                                  _d(hess))
        if rc != 0:
            raise RuntimeError("orc_magic_eval failed %d" % rc)
        return score[0], grad, hess.reshape(m, m)
    finally:
        lib().orc_setup_free(ctypes.byref(st))


def ci_test(data, x, y, S, num_col=35, method="dcc.gamma", sig=1.0, replicates=100, seed=0,
            index=1.0, include_observed=True):
    """kernelCItest(x, y, S) on the oracle; 0-based x, y, S."""
    d = np.asfortranarray(data, dtype=np.float64)
    n, p = d.shape
    Sa = np.ascontiguousarray(S, dtype=np.int32)
    opt = options(method, 0.1, p - 2, num_col, sig, replicates, seed, 0, index, include_observed)
    out = np.zeros(5)
    rc = lib().orc_ci_test(_d(d), n, p, int(x), int(y), Sa.ctypes.data_as(_I), len(Sa),
                           ctypes.byref(opt), _d(out))
    if rc != 0:
        raise RuntimeError("orc_ci_test failed %d" % rc)
    return {"p.value": out[0], "statistic": out[1], "nV2": out[1], "mean": out[2],
            "variance": out[3], "margin": out[4]}


def philox4x32_10(ctr, key):
    c = (ctypes.c_uint * 4)(*[int(v) & 0xFFFFFFFF for v in ctr])
    k = (ctypes.c_uint * 2)(*[int(v) & 0xFFFFFFFF for v in key])
    o = (ctypes.c_uint * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return [int(v) for v in o]


def task_seed(user_seed, method, x, y, S):
    Sa = np.ascontiguousarray(S, dtype=np.int32)
    return int(lib().orc_task_seed(int(user_seed), METHODS[method], int(x), int(y),
                                   Sa.ctypes.data_as(_I), len(Sa)))


def permutation(seed, replicate, n):
    out = np.zeros(n, dtype=np.int32)
    lib().orc_permutation(int(seed), int(replicate), int(n), out.ctypes.data_as(_I))
    return out


def inchol(x, sig=1.0, tol=0.001, max_rank=35):
    """Incomplete Cholesky factor (uncentred, rank columns), pivots, margin."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    n = len(x)
    g = np.zeros(n * max_rank)
    piv = np.zeros(max_rank, dtype=np.int32)
    mg = np.zeros(1)
    r = lib().orc_inchol(_d(x), n, 1.0 / sig, tol, max_rank, _d(g), piv.ctypes.data_as(_I),
                         _d(mg))
    return g.reshape(max_rank, n).T[:, :r].copy(), piv[:r].copy(), mg[0]


def hsic_component(x, sig=1.0, max_rank=35):
    x = np.ascontiguousarray(x, dtype=np.float64)
    n = len(x)
    g = np.zeros(n * max_rank)
    rank = np.zeros(1, dtype=np.int32)
    off = np.zeros(1)
    mom = np.zeros(1)
    mg = np.zeros(1)
    rc = lib().orc_hsic_component(_d(x), n, float(sig), max_rank, _d(g),
                                  rank.ctypes.data_as(_I), _d(off), _d(mom), _d(mg))
    if rc != 0:
        raise RuntimeError("orc_hsic_component failed %d" % rc)
    r = int(rank[0])
    return g.reshape(max_rank, n).T[:, :r].copy(), off[0], mom[0], mg[0]


def hsic_gamma(x, y, sig=1.0, num_col=35):
    gx, ox, mx, _ = hsic_component(x, sig, num_col)
    gy, oy, my, _ = hsic_component(y, sig, num_col)
    n = len(x)
    out = np.zeros(5)
    fx = np.ascontiguousarray(gx.T).ravel()
    fy = np.ascontiguousarray(gy.T).ravel()
    lib().orc_hsic_gamma_from_factors(_d(fx), gx.shape[1], ox, mx, _d(fy), gy.shape[1], oy, my,
                                      n, _d(out))
    return {"p.value": out[0], "statistic": out[1], "mean": out[2], "variance": out[3]}


def hsic_perm(x, y, sig=1.0, num_col=35, replicates=100, seed=0, include_observed=True):
    gx, _, _, _ = hsic_component(x, sig, num_col)
    gy, _, _, _ = hsic_component(y, sig, num_col)
    n = len(x)
    out = np.zeros(5)
    fx = np.ascontiguousarray(gx.T).ravel()
    fy = np.ascontiguousarray(gy.T).ravel()
    lib().orc_hsic_perm_from_factors(_d(fx), gx.shape[1], _d(fy), gy.shape[1], n, replicates,
                                     1 if include_observed else 0, int(seed), _d(out))
    return {"p.value": out[0], "statistic": out[1], "mean": out[2], "variance": out[3],
            "margin": out[4]}


def dcc_perm(x, y, index=1.0, replicates=100, seed=0, include_observed=True):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    out = np.zeros(5)
    rc = lib().orc_dcc_perm(_d(x), _d(y), len(x), float(index), int(replicates),
                            1 if include_observed else 0, int(seed), _d(out))
    if rc != 0:
        raise RuntimeError("orc_dcc_perm failed %d" % rc)
    return {"p.value": out[0], "statistic": out[1], "mean": out[2], "variance": out[3],
            "margin": out[4]}


def skeleton(data, alpha=0.1, m_max=None, num_col=35, threads=8, method="dcc.gamma", sig=1.0,
             replicates=100, seed=0, index=1.0, include_observed=True, r_stream=False):
    """pcalg stable skeleton; r_stream: permutation indices from R's global
    stream after set.seed(seed), tests in pcalg's execution order."""
    d = np.asfortranarray(data, dtype=np.float64)
    n, p = d.shape
    opt = options(method, alpha, p - 2 if m_max is None else m_max, num_col, sig, replicates,
                  seed, threads, index, include_observed, r_stream)
    res = OrcSkeletonResult()
    rc = lib().orc_skeleton(_d(d), n, p, ctypes.byref(opt), ctypes.byref(res))
    if rc != 0:
        lib().orc_skeleton_result_free(ctypes.byref(res))
        raise RuntimeError("orc_skeleton failed %d" % rc)
    try:
        L = res.trace_len
        adj = np.ctypeslib.as_array(res.adjacency, (p * p,)).reshape(p, p).astype(bool).copy()
        pmax = np.ctypeslib.as_array(res.pmax, (p * p,)).reshape(p, p).copy()
        slen = np.ctypeslib.as_array(res.sepset_len, (p * p,)).reshape(p, p).copy()
        sep = np.ctypeslib.as_array(res.sepset, (p * p * 64,)).reshape(p, p, 64)
        sepsets = [[None if slen[i, j] < 0 else [int(v) + 1 for v in sep[i, j, :slen[i, j]]]
                    for j in range(p)] for i in range(p)]
        tS = np.ctypeslib.as_array(res.trace_S, (L * 64,)).reshape(L, 64)
        ts = np.ctypeslib.as_array(res.trace_slen, (L,))
        tasks = {
            "level": np.ctypeslib.as_array(res.trace_level, (L,)).copy(),
            "x": np.ctypeslib.as_array(res.trace_x, (L,)).copy() + 1,
            "y": np.ctypeslib.as_array(res.trace_y, (L,)).copy() + 1,
            "S_key": ["|".join(str(v + 1) for v in tS[i, :ts[i]]) for i in range(L)],
            "p_used": np.ctypeslib.as_array(res.trace_p, (L,)).copy(),
            "native_edge_deleted": np.ctypeslib.as_array(res.trace_del, (L,)).astype(bool).copy(),
            "margin": np.ctypeslib.as_array(res.trace_margin, (L,)).copy(),
            "statistic": np.ctypeslib.as_array(res.trace_stat, (L,)).copy(),
            "mean": np.ctypeslib.as_array(res.trace_mean, (L,)).copy(),
            "variance": np.ctypeslib.as_array(res.trace_var, (L,)).copy(),
        }
        out = {"adjacency": adj, "pMax": pmax, "sepsets": sepsets,
               "n.edgetests": list(res.n_edgetests)[:res.levels], "tasks": tasks,
               "levels": res.levels}
        if r_stream:
            out["r_state"] = [int(res.rng_out.mti)] + [int(v) for v in res.rng_out.mt]
        return out
    finally:
        lib().orc_skeleton_result_free(ctypes.byref(res))


class RStream:
    """R's global stream after set.seed(seed) (Mersenne-Twister, Inversion,
    Rejection), on the oracle's restatement."""

    def __init__(self, seed):
        self.r = OrcRng()
        lib().orc_rng_set_seed(ctypes.byref(self.r), int(seed) & 0xFFFFFFFF)

    def word(self):
        return int(lib().orc_rng_word(ctypes.byref(self.r)))

    def runif(self, k=1):
        return [lib().orc_unif_rand(ctypes.byref(self.r)) for _ in range(k)]

    def unif_index(self, dn, rounding=False):
        return lib().orc_unif_index(ctypes.byref(self.r), float(dn), 1 if rounding else 0)

    def sample(self, n, rounding=False):
        pool = np.zeros(n, dtype=np.int32)
        out = np.zeros(n, dtype=np.int32)
        lib().orc_r_sample(ctypes.byref(self.r), n, 1 if rounding else 0,
                           pool.ctypes.data_as(_I), out.ctypes.data_as(_I))
        return out

    def energy_permute(self, J):
        J = np.ascontiguousarray(J, dtype=np.int32)
        lib().orc_energy_permute(ctypes.byref(self.r), J.ctypes.data_as(_I), len(J))
        return J

    def perm_table(self, method, n, R):
        pool = np.zeros(n, dtype=np.int32)
        out = np.zeros(R * n, dtype=np.int32)
        lib().orc_r_perm_table(ctypes.byref(self.r), METHODS[method], n, R,
                               pool.ctypes.data_as(_I), out.ctypes.data_as(_I))
        return out.reshape(R, n)

    def state(self):
        """.Random.seed[2:626]: position, then the 624 words (as unsigned)."""
        return [int(self.r.mti)] + [int(v) for v in self.r.mt]


def dcc_perm_table(x, y, table, index=1.0, include_observed=True):
    """dcc.perm with an explicit R x n index table (y permuted as y[table[r]])."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    tab = np.ascontiguousarray(table, dtype=np.int32)
    out = np.zeros(5)
    rc = lib().orc_dcc_perm_table(_d(x), _d(y), len(x), float(index), tab.shape[0],
                                  1 if include_observed else 0, 0, tab.ctypes.data_as(_I), _d(out))
    if rc != 0:
        raise RuntimeError("orc_dcc_perm_table failed %d" % rc)
    return {"p.value": out[0], "statistic": out[1], "mean": out[2], "variance": out[3],
            "margin": out[4]}


def hsic_perm_table(x, y, table, sig=1.0, num_col=35, include_observed=True):
    """hsic.perm with an explicit R x n index table (rows of y's factor permuted)."""
    gx, _, _, _ = hsic_component(x, sig, num_col)
    gy, _, _, _ = hsic_component(y, sig, num_col)
    n = len(x)
    tab = np.ascontiguousarray(table, dtype=np.int32)
    out = np.zeros(5)
    fx = np.ascontiguousarray(gx.T).ravel()
    fy = np.ascontiguousarray(gy.T).ravel()
    lib().orc_hsic_perm_table(_d(fx), gx.shape[1], _d(fy), gy.shape[1], n, tab.shape[0],
                              1 if include_observed else 0, 0, tab.ctypes.data_as(_I), _d(out))
    return {"p.value": out[0], "statistic": out[1], "mean": out[2], "variance": out[3],
            "margin": out[4]}


class OrcOrientOptions(ctypes.Structure):
    _fields_ = [("ci_method", ctypes.c_int), ("alpha", ctypes.c_double), ("num_col", ctypes.c_int),
                ("hsic_sig", ctypes.c_double), ("replicates", ctypes.c_int),
                ("include_observed", ctypes.c_int), ("index", ctypes.c_double),
                ("perm_seed", ctypes.c_ulonglong), ("r_stream", ctypes.c_int),
                ("r_state_in", ctypes.POINTER(OrcRng)), ("orient_collider", ctypes.c_int),
                ("solve_confl", ctypes.c_int), ("rules", ctypes.c_int * 3)]


class OrcOrientResult(ctypes.Structure):
    _fields_ = [("pdag", _I), ("n_calls", ctypes.c_int), ("call_cap", ctypes.c_int),
                ("call_v", _I), ("call_slen", _I), ("call_S", _I), ("call_plen", _I),
                ("call_parents", _I), ("call_reject", _I), ("n_tests", ctypes.c_int),
                ("test_cap", ctypes.c_int), ("test_call", _I), ("test_other", _I),
                ("test_out", _D), ("rng_out", OrcRng)]


def sepset_arrays(sepsets, p):
    """1-based sepset lists (None = no sepset) -> (len p x p, 0-based p x p x 64)."""
    slen = np.full(p * p, -1, dtype=np.int32)
    sep = np.zeros(p * p * 64, dtype=np.int32)
    for i in range(p):
        for j in range(p):
            s = sepsets[i][j]
            if s is not None:
                slen[i * p + j] = len(s)
                sep[(i * p + j) * 64:(i * p + j) * 64 + len(s)] = [int(v) - 1 for v in s]
    return slen, sep


def wanpdag(data, adjacency, sepsets, method="dcc.gamma", alpha=0.1, num_col=35, sig=1.0,
            replicates=100, seed=0, index=1.0, include_observed=True, r_stream=False,
            r_state=None, orient_collider=True, solve_confl=False, rules=(True, True, True)):
    """kpcalg udag2wanpdag on the oracle: pdag (p x p ints), the regrVonPS
    calls and their residual tests (0-based variables)."""
    d = np.asfortranarray(data, dtype=np.float64)
    n, p = d.shape
    adj = np.ascontiguousarray(np.asarray(adjacency, dtype=np.uint8).reshape(p, p))
    slen, sep = sepset_arrays(sepsets, p)
    L = lib()
    if not hasattr(L, "_orient_ready"):
        L.orc_wanpdag.argtypes = [_D, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_ubyte),
                                  _I, _I, ctypes.POINTER(OrcOrientOptions),
                                  ctypes.POINTER(OrcOrientResult)]
        L.orc_orient_result_free.argtypes = [ctypes.POINTER(OrcOrientResult)]
        L._orient_ready = True
    st = None
    if r_state is not None:
        st = OrcRng()
        st.mti = int(r_state[0])
        for i in range(624):
            st.mt[i] = int(r_state[1 + i]) & 0xFFFFFFFF
    o = OrcOrientOptions(METHODS[method], alpha, num_col, sig, replicates,
                         1 if include_observed else 0, index, seed, 1 if r_stream else 0,
                         ctypes.pointer(st) if st is not None else None,
                         1 if orient_collider else 0, 1 if solve_confl else 0,
                         (ctypes.c_int * 3)(*[1 if r else 0 for r in rules]))
    res = OrcOrientResult()
    rc = L.orc_wanpdag(_d(d), n, p, adj.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)),
                       slen.ctypes.data_as(_I), sep.ctypes.data_as(_I), ctypes.byref(o),
                       ctypes.byref(res))
    try:
        if rc != 0:
            raise RuntimeError("orc_wanpdag failed %d" % rc)
        pdag = np.ctypeslib.as_array(res.pdag, (p * p,)).reshape(p, p).copy()
        calls = []
        for k in range(res.n_calls):
            calls.append({"V": res.call_v[k],
                          "S": [res.call_S[k * 64 + i] for i in range(res.call_slen[k])],
                          "parents": [res.call_parents[k * 64 + i] for i in range(res.call_plen[k])],
                          "reject": res.call_reject[k]})
        tests = []
        for k in range(res.n_tests):
            tests.append({"call": res.test_call[k], "other": res.test_other[k],
                          "out": [res.test_out[k * 4 + i] for i in range(4)]})
        out = {"pdag": pdag, "calls": calls, "tests": tests}
        if r_stream:
            out["r_state"] = [int(res.rng_out.mti)] + [int(v) for v in res.rng_out.mt]
        return out
    finally:
        L.orc_orient_result_free(ctypes.byref(res))
