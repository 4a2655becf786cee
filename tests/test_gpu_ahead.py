"""The next level's designs and fits built ahead on a third stream.

During level L (L >= 2) the skeleton may build level L + 1's additive designs
and residual fits for every (target, S) the level-start graph allows (a
superset of what level L + 1 can use, since adjacency only shrinks within a
level); level L + 1 claims the slots its tests use.  Each chain (this one and
level 1's, built during level 0) also computes its fits' components (Lanczos,
incomplete Cholesky or dcc.perm row means), claimed with the fits.  The work
is the same computations on the same inputs, so every output must equal the
skeleton run with the chains off (KPC_AHEAD_LIMIT=0, KPC_AHEAD_COMPONENTS=0),
bit for bit.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(data, method, limit, components=True, num_col=35):
    import fastkpc_amd as fk

    env = {"KPC_AHEAD_LIMIT": str(limit), "KPC_AHEAD_COMPONENTS": "1" if components else "0"}
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        return fk.fastkpc_compatible_cuda_skeleton(
            data, 0.1, options={"numCol": num_col, "ci_method": method, "index": 1,
                                "permutation_params": {"replicates": 60, "seed": 7}})
    finally:
        for k, v in old.items():
            if v is None:
                del os.environ[k]
            else:
                os.environ[k] = v


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


@pytest.mark.parametrize("cfg,method", [("sem351x48", "dcc.gamma"), ("sem200x8", "hsic.gamma"),
                                        ("sem200x8", "hsic.perm"), ("sem200x8", "dcc.perm")])
def test_components_built_ahead_equal_on_demand(cfg, method):
    from fastkpc_amd.sem import config_data

    data = config_data(cfg)[0]
    nc = 35 if cfg == "sem351x48" else 20
    _same(_run(data, method, 512, True, nc), _run(data, method, 0, False, nc))
