"""The process-wide result and target-state caches (the reference's
CompactResultCache / TargetStateCache and cache control,
fastkpc/src/full_cuda_ci_one_call.cpp:1443-1755, 6812-6900) on MI355X.

* A skeleton replayed with the result cache runs no test and returns the cold
  run's adjacency, sepsets and trace bit for bit.
* Fits cached by one run serve another whose tests differ (numCol 35 then
  30): each reused fit replays its final step at the cached log sp, so
  residuals, and hence every p-value, statistic and margin, equal a cold
  run's bit for bit.
* Control: info, reset, configure with the reference's bounds; per-dataset
  entry counts; a different dataset misses.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def fk():
    import fastkpc_amd
    from fastkpc_amd import _lib
    assert _lib.lib().kpc_device_count() >= 1, "no GPU visible"
    return fastkpc_amd


def same(a, b):
    assert a["n.edgetests"] == b["n.edgetests"]
    assert np.array_equal(a["adjacency"], b["adjacency"])
    assert a["sepsets"] == b["sepsets"]
    for k in ("p_used", "method_statistic", "method_mean", "method_variance", "decision_margin",
              "task_index"):
        assert np.array_equal(a["tasks"][k], b["tasks"][k]), k
    assert a["tasks"]["S_key"] == b["tasks"]["S_key"]


@pytest.mark.parametrize("method", ["dcc.gamma", "hsic.gamma", "dcc.perm"])
def test_result_cache_replays_a_skeleton_without_tests(method):
    f = fk()
    from fastkpc_amd.sem import config_data
    data, _, _ = config_data("sem200x8")
    f.cache_control("reset")
    opts = {"numCol": 35, "ci_method": method, "cache": True,
            "permutation_params": {"replicates": 50, "seed": 3}}
    cold = f.fastkpc_compatible_cuda_skeleton(data, 0.1, options=opts)
    assert cold["summary"]["result_cache_hit_count"] == 0
    warm = f.fastkpc_compatible_cuda_skeleton(data, 0.1, options=opts)
    same(cold, warm)
    assert warm["summary"]["physical_tests_evaluated"] == 0
    # every test the replay asked for (speculative ones included) was a hit
    assert warm["summary"]["result_cache_hit_count"] >= sum(cold["n.edgetests"])
    st = f.cache_state(data)
    assert st["dataset_entries"] == st["entries"] >= sum(cold["n.edgetests"])
    other = f.cache_state(data[::-1].copy())
    assert other["dataset_entries"] == 0 and other["dataset_key"] != st["dataset_key"]
    # a different method is a different test: no hits
    plain = f.fastkpc_compatible_cuda_skeleton(
        data, 0.1, options={"numCol": 35, "ci_method": "hsic.perm" if method != "hsic.perm" else
                            "dcc.gamma", "cache": True,
                            "permutation_params": {"replicates": 50, "seed": 3}})
    assert plain["summary"]["result_cache_hit_count"] == 0


def test_target_state_cache_replays_fits_bit_for_bit():
    """Fits cached by a numCol = 35 run serve a numCol = 30 run (different
    tests, so no result-cache hits; same designs and targets): its skeleton
    equals a cold numCol = 30 run bit for bit, and every reused fit skipped
    the optimiser."""
    f = fk()
    from fastkpc_amd.sem import config_data
    data, _, _ = config_data("sem351x48")
    cold = f.fastkpc_compatible_cuda_skeleton(data, 0.1, options={"numCol": 30})
    f.cache_control("reset")
    f.fastkpc_compatible_cuda_skeleton(data, 0.1, options={"numCol": 35, "cache": True})
    warm = f.fastkpc_compatible_cuda_skeleton(data, 0.1, options={"numCol": 30, "cache": True})
    same(cold, warm)
    sw, sc = warm["summary"], cold["summary"]
    assert sw["result_cache_hit_count"] == 0
    assert sw["residual_fits"] == sc["residual_fits"]
    hits = sw["target_cache_hit_count"]
    assert hits >= 0.5 * sc["residual_fits"], (hits, sc["residual_fits"])
    # only the fits that missed ran the optimiser
    assert sw["magic_evaluations"] < sc["magic_evaluations"]
    f.cache_control("reset")


def test_cache_control_actions_and_bounds():
    f = fk()
    from fastkpc_amd import _lib
    r = f.cache_control("reset")
    assert r["entries"] == 0 and r["target_entries"] == 0 and r["requests"] == 0
    e0 = r["cache_epoch"]
    assert f.cache_control("reset")["cache_epoch"] == e0 + 1
    c = f.cache_control("configure", 1048576)
    assert c["capacity"] == 1048576
    for bad in (0, 1048577):
        with pytest.raises(_lib.KpcLibraryError):
            f.cache_control("configure", bad)
    with pytest.raises(_lib.KpcLibraryError):
        f.cache_control("configure_target", 524289)
    assert f.cache_control("configure_target", 524288)["target_capacity"] == 524288
    with pytest.raises(_lib.KpcLibraryError):  # info takes no capacity
        _lib.check(_lib.lib().kpc_cache_control(0, 5, _lib.ctypes.byref(_lib.KpcCacheInfo())))
    f.cache_control("configure", 262144)
    f.cache_control("configure_target", 131072)
    f.cache_control("reset")
