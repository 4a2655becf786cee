"""Run the C oracle on the parity workloads and store its skeleton traces.

    python tests/golden/make_oracle_goldens.py [name ...]

Writes tests/golden/oracle_<name>.npz (adjacency, sepsets, n.edgetests and the
canonical trace with p-values, statistics, null means and variances, and
decision margins) for each workload in
WORKLOADS (data source, CI method, permutation replicates).  The GPU parity tests compare the HIP path
against these; tests/test_oracle_goldens.py re-runs the oracle on the small
workloads to show the stored traces are the oracle's current output.
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from fastkpc_amd.sem import config_data  # noqa: E402
from oracle import orc  # noqa: E402


def save(name, data, res, extra=None):
    p = data.shape[1]
    sep = np.full((p, p, 65), -1, dtype=np.int32)
    for i in range(p):
        for j in range(p):
            s = res["sepsets"][i][j]
            if s is not None:
                sep[i, j, 0] = len(s)
                sep[i, j, 1:1 + len(s)] = s
    t = res["tasks"]
    np.savez_compressed(os.path.join(HERE, "oracle_%s.npz" % name),
                        adjacency=res["adjacency"], pmax=res["pMax"], sepsets=sep,
                        n_edgetests=np.array(res["n.edgetests"], dtype=np.int32),
                        level=t["level"], x=t["x"], y=t["y"], S=np.array(t["S_key"]),
                        p=t["p_used"], deleted=t["native_edge_deleted"], margin=t["margin"],
                        stat=t["statistic"], mean=t["mean"], var=t["variance"], **(extra or {}))


# name -> (data source, ci_method, replicates); permutation workloads use seed 2024
WORKLOADS = {
    "dev500x50": ("dev500x50", "dcc.gamma", 0),
    "sem351x48": ("sem351x48", "dcc.gamma", 0),
    "sem351x48_hsic_gamma": ("sem351x48", "hsic.gamma", 0),
    "sem500x50": ("sem500x50", "dcc.gamma", 0),
    "sem200x8_hsic_gamma": ("sem200x8", "hsic.gamma", 0),
    "sem200x8_hsic_perm": ("sem200x8", "hsic.perm", 100),
    "sem200x8_dcc_perm": ("sem200x8", "dcc.perm", 100),
    "mixed200x14_hsic_perm": ("mixed200x14", "hsic.perm", 100),
    "mixed200x14_dcc_perm": ("mixed200x14", "dcc.perm", 100),
    # permutation indices from R's global stream after set.seed(PERM_SEED),
    # tests in pcalg's execution order (the reference's strict route)
    "sem200x8_dcc_perm_R": ("sem200x8", "dcc.perm", 100, "R"),
    "sem200x8_hsic_perm_R": ("sem200x8", "hsic.perm", 100, "R"),
    "mixed200x14_dcc_perm_R": ("mixed200x14", "dcc.perm", 100, "R"),
    "mixed200x14_hsic_perm_R": ("mixed200x14", "hsic.perm", 100, "R"),
}
PERM_SEED = 2024


def workload_data(src):
    if src == "dev500x50":
        return np.load(os.path.join(HERE, "dev500x50_v2.npz"))["data"]
    if src == "mixed200x14":
        from fastkpc_amd.sem import generate
        return generate(n=200, p=14, edge_prob=0.2, seed=14, link="mixed")[0]
    return config_data(src)[0]


def main(names):
    for name in names:
        t0 = time.time()
        src, method, reps = WORKLOADS[name][:3]
        r_stream = len(WORKLOADS[name]) > 3
        data = workload_data(src)
        res = orc.skeleton(data, alpha=0.1, m_max=data.shape[1] - 2, num_col=35, threads=8,
                           method=method, replicates=reps, seed=PERM_SEED, r_stream=r_stream)
        save(name, data, res,
             {"r_state": np.array(res["r_state"], dtype=np.uint32)} if r_stream else None)
        print(name, "tests", sum(res["n.edgetests"]), res["n.edgetests"], "%.1fs" % (time.time() - t0))


if __name__ == "__main__":
    main(sys.argv[1:] or list(WORKLOADS))
