"""hipGraph-captured waves and the launch-shape autotuner on MI355X.

* KPC_GRAPH=1 captures each wave's device pipeline into one hipGraph: the
  skeleton (adjacency, n.edgetests, every trace p-value) must be bit-identical
  to the eager launches, over several graph runs interleaved with eager runs.
* The default two launch lanes (tests split by conditioning set over two
  streams) must give the skeleton of one lane (KPC_LANES=1), bit for bit.
* Every autotuner candidate (dcc.perm replicates per workgroup, hsic.perm
  replicates per unit, permutation generation width) must give bit-identical
  p-values, statistics, means and variances; the tuner's table must hold a
  measured choice after a workload large enough to time.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def fk():
    import fastkpc_amd
    from fastkpc_amd import _lib
    assert _lib.lib().kpc_device_count() >= 1, "no GPU visible"
    return fastkpc_amd


class env:
    def __init__(self, **kv):
        self.kv = kv

    def __enter__(self):
        self.old = {k: os.environ.get(k) for k in self.kv}
        os.environ.update({k: str(v) for k, v in self.kv.items()})

    def __exit__(self, *a):
        for k, v in self.old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def same_skeleton(a, b):
    assert a["n.edgetests"] == b["n.edgetests"]
    assert np.array_equal(a["adjacency"], b["adjacency"])
    assert np.array_equal(a["tasks"]["p_used"], b["tasks"]["p_used"])
    assert np.array_equal(a["tasks"]["native_edge_deleted"], b["tasks"]["native_edge_deleted"])


@pytest.mark.parametrize("cfg,method", [("sem351x48", "dcc.gamma"), ("sem200x8", "hsic.gamma")])
def test_graph_waves_equal_eager(cfg, method):
    f = fk()
    from fastkpc_amd.sem import config_data
    data, _, _ = config_data(cfg)
    opts = {"numCol": 35, "ci_method": method}
    eager = f.fastkpc_compatible_cuda_skeleton(data, 0.1, options=opts)
    assert eager["summary"]["graph_launches"] == 0
    for _ in range(3):
        with env(KPC_GRAPH=1):
            g = f.fastkpc_compatible_cuda_skeleton(data, 0.1, options=opts)
        assert g["summary"]["graph_launches"] > 0
        same_skeleton(eager, g)
        again = f.fastkpc_compatible_cuda_skeleton(data, 0.1, options=opts)
        same_skeleton(eager, again)


@pytest.mark.parametrize("cfg,method", [("sem351x48", "dcc.gamma"), ("sem200x8", "hsic.gamma"),
                                        ("sem500x50", "dcc.gamma")])
def test_two_lanes_equal_one_lane(cfg, method):
    """A wave's tests split by conditioning set over two streams (the default)
    give the same skeleton, bit for bit, as one stream (KPC_LANES=1); levels 1-2
    split per bivariate smooth, levels 3+ share the univariate smooths; waves
    of fewer than 64 tests stay on one lane."""
    f = fk()
    from fastkpc_amd.sem import config_data
    data, _, _ = config_data(cfg)
    opts = {"numCol": 35, "ci_method": method}
    with env(KPC_LANES=1):
        one = f.fastkpc_compatible_cuda_skeleton(data, 0.1, options=opts)
    for _ in range(2):
        two = f.fastkpc_compatible_cuda_skeleton(data, 0.1, options=opts)
        same_skeleton(one, two)
        assert two["summary"]["level_fits"] == one["summary"]["level_fits"]
    if cfg == "sem351x48":
        assert one["tasks"]["conditioning_size"].max() >= 3  # the shared-smooth levels ran


def perm_batch(method, n=600, p=10, seed=21, R=200):
    from fastkpc_amd.sem import generate
    data = generate(n=n, p=p, edge_prob=0.3, seed=seed, link="mixed")[0]
    rng = np.random.default_rng(5)
    xs, ys, Ss = [], [], []
    for _ in range(12):
        v = rng.permutation(p)
        xs.append(int(v[0]))
        ys.append(int(v[1]))
        Ss.append(sorted(int(a) for a in v[2:4]))
    return lambda: fk().ci_batch(data, xs, ys, np.array(Ss, np.int32), num_col=35,
                                 ci_method=method, replicates=R, seed=9)


def test_dcc_perm_tuned_shapes_are_bit_identical():
    run = perm_batch("dcc.perm")
    outs = []
    for shape in ("16x8", "16x12"):
        for w in ("8", "3"):
            with env(KPC_DCC_SHAPE=shape, KPC_PERM_GEN_WIDTH=w):
                outs.append(run())
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])


def test_hsic_perm_tuned_blocks_are_bit_identical():
    run = perm_batch("hsic.perm")
    outs = []
    for b in ("16", "8", "32", "5"):
        with env(KPC_HSIC_PERM_BLOCK=b):
            outs.append(run())
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])


def test_tuner_measures_and_chooses():
    f = fk()
    run = perm_batch("dcc.perm", n=1200, R=400)
    first = run()
    for _ in range(3):
        assert np.array_equal(run(), first)
    table = f.tuning_report()
    keys = [k for k in table if k.startswith("dcc_perm_rb/n1280")]
    assert keys, table
    chosen, cands = table[keys[0]]
    assert chosen in (8, 12) and all(v is not None and v > 0 for v in cands.values()), table
