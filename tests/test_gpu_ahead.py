"""The next level's designs and fits built ahead on a third stream.

During level L (L >= 2) the skeleton may build level L + 1's additive designs
and residual fits for every (target, S) the level-start graph allows (a
superset of what level L + 1 can use, since adjacency only shrinks within a
level); level L + 1 claims the slots its tests use.  The fits are the same
computations on the same inputs, so every output must equal the skeleton run
with the chain off (KPC_AHEAD_LIMIT=0), bit for bit.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(data, method, limit):
    import fastkpc_amd as fk

    old = os.environ.get("KPC_AHEAD_LIMIT")
    os.environ["KPC_AHEAD_LIMIT"] = str(limit)
    try:
        return fk.fastkpc_compatible_cuda_skeleton(
            data, 0.1, options={"numCol": 35, "ci_method": method, "index": 1})
    finally:
        if old is None:
            del os.environ["KPC_AHEAD_LIMIT"]
        else:
            os.environ["KPC_AHEAD_LIMIT"] = old


def _same(a, b):
    np.testing.assert_array_equal(np.asarray(a["adjacency"]), np.asarray(b["adjacency"]))
    np.testing.assert_array_equal(np.asarray(a["n.edgetests"]), np.asarray(b["n.edgetests"]))
    ta, tb = a["tasks"], b["tasks"]
    assert sorted(ta) == sorted(tb)
    for k in ta:
        x, y = np.asarray(ta[k]), np.asarray(tb[k])
        if x.dtype.kind == "f":
            np.testing.assert_array_equal(x, y, err_msg=k)
        elif x.dtype.kind in "iub":
            np.testing.assert_array_equal(x, y, err_msg=k)
    sa, sb = a["summary"], b["summary"]
    for k in ("level_fits", "level_components", "level_evaluations", "level_physical"):
        assert list(sa[k]) == list(sb[k]), k


@pytest.mark.parametrize("cfg", ["sem351x48", "sem500x50"])
def test_level_built_ahead_equals_on_demand(cfg):
    from fastkpc_amd.sem import config_data

    data = config_data(cfg)[0]
    on = _run(data, "dcc.gamma", 512)
    off = _run(data, "dcc.gamma", 0)
    _same(on, off)
    # the skeleton reaches levels whose superset is small enough to build ahead
    assert len(on["n.edgetests"]) >= 5 and on["n.edgetests"][4] > 0


def test_every_level_built_ahead_equals_on_demand():
    # a limit that admits level 3's superset as well (chains at levels 2 and 3)
    from fastkpc_amd.sem import config_data

    data = config_data("sem351x48")[0]
    _same(_run(data, "dcc.gamma", 1 << 20), _run(data, "dcc.gamma", 0))
