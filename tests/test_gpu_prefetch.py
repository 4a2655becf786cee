"""Level-1 work built ahead while level 0 runs.

Every column's univariate smooth (k = 10), level 1's designs S = {s} and its
residual fits (t | {s}) are computed on the second launch lane during level
0; level 1's tests (and later levels' additive terms) claim them.  A set-up
fails when its column has fewer unique values than the basis size (mgcv's
"fewer unique covariate combinations" error); such a failure must be
reported exactly where a design uses the column, as when the smooth was built
on demand: a column cut off at level 0 never raises, a column still adjacent
at level 1 does.  The plain form builds everything on demand and must agree.

The data make the level-0 decisions certain: alpha = 1e-12 deletes every
edge of an independent column (its dcov p-values cannot come that close to
zero at n = 80), while columns 0-3 form a tight chain whose edges survive.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

OPTS = {"numCol": 20, "ci_method": "dcc.gamma", "index": 1}


def _data(seed, tie_last_to_first):
    rng = np.random.default_rng(seed)
    n = 80
    x0 = rng.normal(size=n)
    x1 = x0 + 0.05 * rng.normal(size=n)
    x2 = x1 + 0.05 * rng.normal(size=n)
    x3 = np.sin(2.0 * x0) + 0.05 * rng.normal(size=n)
    src = x0 if tie_last_to_first else rng.normal(size=n)
    x4 = np.digitize(src, [-0.5, 0.5]).astype(np.float64)  # three unique values
    return np.column_stack([x0, x1, x2, x3, x4])


def test_failed_smooth_of_a_column_cut_off_at_level_0_is_not_an_error():
    import fastkpc_amd as fk

    res = fk.fastkpc_compatible_cuda_skeleton(_data(11, False), 1e-12, options=OPTS)
    adj = np.asarray(res["adjacency"])
    assert not adj[4].any() and not adj[:, 4].any()
    # the chain reached level 1, where the other columns' smooths were used
    assert len(res["n.edgetests"]) >= 2 and res["n.edgetests"][1] > 0


def test_failed_smooth_of_a_column_used_at_level_1_is_an_error():
    import fastkpc_amd as fk
    from fastkpc_amd._lib import KpcLibraryError

    with pytest.raises(KpcLibraryError, match="thin-plate setup failed"):
        fk.fastkpc_compatible_cuda_skeleton(_data(11, True), 1e-12, options=OPTS)


def test_plain_form_builds_its_own_smooths_and_agrees():
    import fastkpc_amd as fk

    data = _data(11, False)
    a = fk.fastkpc_compatible_cuda_skeleton(data, 1e-12, options=OPTS)
    b = fk.fastkpc_compatible_cuda_skeleton(data, 1e-12, options=dict(OPTS, mode="plain"))
    np.testing.assert_array_equal(np.asarray(a["adjacency"]), np.asarray(b["adjacency"]))
    np.testing.assert_array_equal(np.asarray(a["tasks"]["p_used"]), np.asarray(b["tasks"]["p_used"]))
