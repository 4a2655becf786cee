"""The stored oracle traces used by the GPU parity tests are the oracle's
current output (re-run here on the CPU)."""
import os

import numpy as np
import pytest

from conftest import GOLDEN, trace_map
from oracle import orc


@pytest.mark.slow
def test_stored_sem351x48_trace_is_current_oracle_output():
    from fastkpc_amd.sem import config_data
    data, _, _ = config_data("sem351x48")
    gold = np.load(os.path.join(GOLDEN, "oracle_sem351x48.npz"))
    res = orc.skeleton(data, alpha=0.1, m_max=46, num_col=35, threads=8)
    assert res["n.edgetests"] == [int(v) for v in gold["n_edgetests"]]
    assert np.array_equal(res["adjacency"], gold["adjacency"])
    t = res["tasks"]
    mine = trace_map(t["level"], t["x"], t["y"], t["S_key"], t["p_used"])
    ref = trace_map(gold["level"], gold["x"], gold["y"], gold["S"], gold["p"])
    assert mine == ref


def test_reference_trace_edge_columns_are_min_max_of_the_tested_pair():
    """The reference's own decision trace (dev500x50 fixture, 1-based) names
    the tested edge as edge_x = min(x, y), edge_y = max(x, y) of the tested
    pair, whichever orientation the test conditioned from; the library's
    trace derives its edge_x / edge_y columns the same way."""
    z = np.load(os.path.join(GOLDEN, "dev500x50_v2.npz"))
    tx, ty = z["dt_tested_x"], z["dt_tested_y"]
    assert len(tx) > 1000 and np.any(tx > ty) and np.any(tx < ty)
    assert np.array_equal(z["dt_edge_x"], np.minimum(tx, ty))
    assert np.array_equal(z["dt_edge_y"], np.maximum(tx, ty))
    assert tx.min() >= 1 and ty.max() <= z["data"].shape[1]
