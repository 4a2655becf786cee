"""GCV trial slots on MI355X: the additive |S| = 3, 4 fits whose step-halving
trials and infinity probes are scored several at a time (one per wave) must
equal the one-at-a-time optimiser bit for bit.

KPC_MAGIC_SLOTS (read at every GCV launch) switches between them: 1 runs
the serial kernel, 4 the slot kernel with one point per wave, 5 (the
default) adds a fifth point on wave 0 at |S| = 3 (two points interleaved on
one wave), whose work goes to the workgroup's buffers for the try accepted
at the previous iteration.  Compared on
sem351x48 fits that include the slowest fits of its skeleton (123 Newton
iterations with five tries in most of them), on cached-fit replays, and on
a whole skeleton: residuals, every optimiser field of the fit record
(status, score, rss, edf, iterations, evaluations, step failure, rank,
log(sp), decision margin) and every trace column.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu



def record_fields(info, S_list):
    """The defined result fields of each fit record: status, score, rss, edf,
    iterations, evaluations, step failure, rank, the |S| log(sp) and the
    decision margin (other slots are device timings or unused)."""
    return [np.concatenate([info[i, :8], info[i, 8:8 + len(S)], info[i, 30:31]])
            for i, S in enumerate(S_list)]


def fk():
    import fastkpc_amd
    from fastkpc_amd import _lib
    assert _lib.lib().kpc_device_count() >= 1, "no GPU visible"
    return fastkpc_amd


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.int64)


class slots:
    def __init__(self, v):
        self.v = str(v)

    def __enter__(self):
        self.old = os.environ.get("KPC_MAGIC_SLOTS")
        os.environ["KPC_MAGIC_SLOTS"] = self.v

    def __exit__(self, *a):
        if self.old is None:
            os.environ.pop("KPC_MAGIC_SLOTS", None)
        else:
            os.environ["KPC_MAGIC_SLOTS"] = self.old


@pytest.fixture(scope="module")
def sem351():
    from fastkpc_amd.sem import config_data
    return config_data("sem351x48")[0]


def fit_sets():
    rng = np.random.default_rng(11)
    T, SL = [], []
    for k in range(150):
        s = 3 + k % 3
        cols = rng.choice(48, size=s + 1, replace=False)
        T.append(int(cols[0]))
        SL.append(sorted(int(c) for c in cols[1:]))
    # the slowest fits of the sem351x48 skeleton (five tries in most iterations)
    for t, S in [(2, [10, 21, 37]), (47, [2, 11, 17]), (42, [1, 5, 6]), (4, [32, 34, 36, 40])]:
        T.append(t)
        SL.append(S)
    return T, SL


def test_slot_fits_equal_serial(sem351):
    api = fk()
    T, SL = fit_sets()
    with slots(1):
        r1, i1 = api.regrxons_mixed(sem351, T, SL)
    with slots(4):
        r4, i4 = api.regrxons_mixed(sem351, T, SL)
    with slots(5):
        r5, i5 = api.regrxons_mixed(sem351, T, SL)
    print("fits", len(T), "evaluations", int(i1[:, 5].sum()), "max", int(i1[:, 5].max()))
    assert np.all(i1[:, 0] == 0)
    assert i1[:, 5].max() >= 500  # the 583-evaluation fit is in the set
    for r, i in ((r4, i4), (r5, i5)):
        assert np.array_equal(bits(r1), bits(r))
        for a, b in zip(record_fields(i1, SL), record_fields(i, SL)):
            assert np.array_equal(bits(a), bits(b))


def test_slot_skeleton_equals_serial(sem351):
    api = fk()
    out = {}
    for v in (1, 5):
        with slots(v):
            out[v] = api.fastkpc_compatible_cuda_skeleton(sem351, 0.1)
    a, b = out[1], out[5]
    assert np.array_equal(np.asarray(a["adjacency"]), np.asarray(b["adjacency"]))
    assert np.array_equal(bits(a["pMax"]), bits(b["pMax"]))
    assert a["summary"]["magic_evaluations"] == b["summary"]["magic_evaluations"]
    for k, v in a["tasks"].items():
        x, y = np.asarray(v), np.asarray(b["tasks"][k])
        if x.dtype.kind == "f":
            assert np.array_equal(bits(x), bits(y)), k
        else:
            assert np.array_equal(x, y), k


def test_slot_cached_replay_equals_serial(sem351):
    # target-state cache hits replay the final fit at the cached log(sp) with
    # no optimiser: fits cached by a numCol = 35 run serve a numCol = 30 run,
    # on the slot kernel's replay path and on the serial kernel's
    api = fk()
    res = {}
    for v in (1, 5):
        with slots(v):
            api.cache_control("reset")
            api.fastkpc_compatible_cuda_skeleton(sem351, 0.1, options={"numCol": 35, "cache": True})
            res[v] = api.fastkpc_compatible_cuda_skeleton(sem351, 0.1,
                                                          options={"numCol": 30, "cache": True})
    api.cache_control("reset")
    a, b = res[1], res[5]
    assert a["summary"]["target_cache_hit_count"] > 0
    assert a["summary"]["target_cache_hit_count"] == b["summary"]["target_cache_hit_count"]
    assert np.array_equal(bits(a["pMax"]), bits(b["pMax"]))
    for k, v in a["tasks"].items():
        x, y = np.asarray(v), np.asarray(b["tasks"][k])
        if x.dtype.kind == "f":
            assert np.array_equal(bits(x), bits(y)), k
