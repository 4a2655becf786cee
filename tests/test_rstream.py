"""R's global stream on the oracle (CPU): the restatement of R's
Mersenne-Twister seeding, unif_rand, R_unif_index and sample(), energy's
permute(), and the order the R-stream skeleton draws tables in.

Pins, none of which retypes the oracle's own formulas:
* outputs R prints for set.seed(s); runif(k) and set.seed(s); sample(10)
  under R >= 3.6 ("Rejection" sampling) and R < 3.6 ("Rounding");
* the Mersenne-Twister words against CPython's own MT19937 (random.setstate
  with the state set.seed leaves);
* energy's documented identity permute(J) == rev(sample(J)) under Rounding
  sampling (energy's utilities.c comment), and dcov.test's table rows as
  successive permute() calls on one vector;
* the R-stream oracle skeleton against an independent Python replay of pcalg's
  execution order (which tests run, in which order, drawing which tables).
"""
import os
import random
import sys

import numpy as np
import pytest

from conftest import GOLDEN
from oracle import orc


@pytest.mark.parametrize("seed,want", [
    (1, [0.2655087, 0.3721239, 0.5728534]),
    (42, [0.9148060]),
    (123, [0.2875775, 0.7883051, 0.4089769]),
])
def test_runif_matches_r(seed, want):
    got = orc.RStream(seed).runif(len(want))
    assert [round(v, 7) for v in got] == want


@pytest.mark.parametrize("seed,rounding,want", [
    (1, False, [9, 4, 7, 1, 2, 5, 3, 10, 6, 8]),
    (42, False, [1, 5, 10, 8, 2, 4, 6, 9, 7, 3]),
    (123, False, [3, 10, 2, 8, 6, 9, 1, 7, 5, 4]),
    (1, True, [3, 4, 5, 7, 2, 8, 9, 6, 10, 1]),
    (123, True, [3, 8, 4, 7, 6, 1, 10, 9, 2, 5]),
])
def test_sample_matches_r(seed, rounding, want):
    assert (orc.RStream(seed).sample(10, rounding) + 1).tolist() == want


@pytest.mark.parametrize("seed", [0, 1, 2024, 2 ** 31 - 1])
def test_mersenne_twister_words_match_cpython(seed):
    rs = orc.RStream(seed)
    st = rs.state()
    py = random.Random()
    py.setstate((3, tuple(st[1:]) + (st[0],), None))
    for _ in range(2000):  # crosses three block regenerations
        assert rs.word() == py.getrandbits(32)


def test_energy_permute_is_reversed_rounding_sample():
    n = 37
    for seed in (3, 17, 99):
        J = orc.RStream(seed).energy_permute(np.arange(n))
        s = orc.RStream(seed).sample(n, rounding=True)
        assert np.array_equal(J, s[::-1])


def test_dcc_perm_table_mutates_one_vector_and_hsic_table_resamples():
    n, R = 23, 6
    tab = orc.RStream(7).perm_table("dcc.perm", n, R)
    rs = orc.RStream(7)
    J = np.arange(n)
    for r in range(R):
        J = rs.energy_permute(J)
        assert np.array_equal(tab[r], J)
    tab = orc.RStream(7).perm_table("hsic.perm", n, R)
    rs = orc.RStream(7)
    for r in range(R):
        assert np.array_equal(tab[r], rs.sample(n))
    for row in tab:
        assert sorted(row.tolist()) == list(range(n))


@pytest.mark.slow
@pytest.mark.parametrize("method", ["dcc.perm", "hsic.perm"])
def test_r_stream_skeleton_draws_in_pcalg_order(method):
    """Replays the oracle's R-stream skeleton from its own trace: the trace's
    tests sorted into pcalg's order (level, then x, then y, then S) walk one
    stream from set.seed(2024); each test's p-value and statistic must be the
    oracle's CI test on the table drawn at that point, and the stream must
    end where the skeleton left it."""
    sys.path.insert(0, GOLDEN)
    from make_oracle_goldens import PERM_SEED, workload_data
    data = workload_data("sem200x8")
    res = orc.skeleton(data, method=method, replicates=40, seed=PERM_SEED, r_stream=True)
    t = res["tasks"]
    rows = sorted(range(len(t["x"])), key=lambda i: (
        t["level"][i], t["x"][i], t["y"][i],
        [int(v) for v in t["S_key"][i].split("|")] if t["S_key"][i] else []))
    rs = orc.RStream(PERM_SEED)
    n = data.shape[0]
    for i in rows:
        x, y = int(t["x"][i]) - 1, int(t["y"][i]) - 1
        S = [int(v) - 1 for v in t["S_key"][i].split("|")] if t["S_key"][i] else []
        rx = orc.regrxons(data, x, S)[0] if S else data[:, x]
        ry = orc.regrxons(data, y, S)[0] if S else data[:, y]
        tab = rs.perm_table(method, n, 40)
        ref = (orc.dcc_perm_table(rx, ry, tab) if method == "dcc.perm"
               else orc.hsic_perm_table(rx, ry, tab))
        assert t["p_used"][i] == ref["p.value"], (i, t["p_used"][i], ref)
        assert abs(t["statistic"][i] - ref["statistic"]) <= 1e-12 * abs(ref["statistic"])
    assert res["r_state"] == rs.state()
