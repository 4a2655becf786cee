"""Device arenas reused across calls and threads on MI355X.

The library keeps the device blocks of finished calls and hands them to later
calls of the same thread; kpc_empty_cache frees them.  Results must not depend
on which blocks a call receives:

* a skeleton run after kpc_empty_cache equals the runs before it, bit for bit;
* regrXonS batches and dcov.gamma component batches issued from two threads at
  once (ctypes drops the GIL, so the calls overlap on the device) equal the
  same calls made one after another.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def fk():
    import fastkpc_amd
    from fastkpc_amd import _lib
    assert _lib.lib().kpc_device_count() >= 1, "no GPU visible"
    return fastkpc_amd


def test_empty_cache_between_skeleton_runs():
    f = fk()
    from fastkpc_amd.sem import config_data
    data, _, _ = config_data("sem200x8")
    opts = {"numCol": 35, "ci_method": "hsic.gamma"}
    a = f.fastkpc_compatible_cuda_skeleton(data, 0.1, options=opts)
    b = f.fastkpc_compatible_cuda_skeleton(data, 0.1, options=opts)
    f.empty_cache()
    c = f.fastkpc_compatible_cuda_skeleton(data, 0.1, options=opts)
    for r in (b, c):
        assert r["n.edgetests"] == a["n.edgetests"]
        assert np.array_equal(r["adjacency"], a["adjacency"])
        assert np.array_equal(r["tasks"]["p_used"], a["tasks"]["p_used"])


def test_concurrent_threads_equal_sequential_calls():
    f = fk()
    from fastkpc_amd.sem import regression_batch
    data, targets, sets = regression_batch(fits=96, n=300, seed=11)
    halves = [list(range(0, 48)), list(range(48, 96))]
    rng = np.random.default_rng(5)
    xs = [rng.standard_normal((400, 40)), rng.standard_normal((650, 25))]

    def fit(h):
        idx = halves[h]
        return f.regrxons_mixed(data, targets[idx], [sets[i] for i in idx])

    def comps(h):
        return f.dcov_components(xs[h], 20)

    want_fit = [fit(h) for h in (0, 1)]
    want_comp = [comps(h) for h in (0, 1)]
    f.empty_cache()
    with ThreadPoolExecutor(max_workers=2) as pool:
        for _ in range(3):
            got_fit = list(pool.map(fit, (0, 1)))
            got_comp = list(pool.map(comps, (1, 0)))[::-1]
            for g, w in zip(got_fit, want_fit):
                assert np.array_equal(g[0], w[0])
                assert np.array_equal(g[1][:, :8], w[1][:, :8])
            for g, w in zip(got_comp, want_comp):
                for a, b in zip(g, w):
                    assert np.array_equal(a, b)
