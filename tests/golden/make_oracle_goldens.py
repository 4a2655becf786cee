"""Run the C oracle on the parity workloads and store its skeleton traces.

    python tests/golden/make_oracle_goldens.py [name ...]

Writes tests/golden/oracle_<name>.npz (adjacency, sepsets, n.edgetests and the
canonical trace with p-values).  The GPU parity tests compare the HIP path
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


def save(name, data, res):
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
                        p=t["p_used"], deleted=t["native_edge_deleted"], margin=t["margin"])


def main(names):
    for name in names:
        t0 = time.time()
        if name == "dev500x50":
            data = np.load(os.path.join(HERE, "dev500x50_v2.npz"))["data"]
        else:
            data, _, _ = config_data(name)
        res = orc.skeleton(data, alpha=0.1, m_max=data.shape[1] - 2, num_col=35, threads=8)
        save(name, data, res)
        print(name, "tests", sum(res["n.edgetests"]), res["n.edgetests"], "%.1fs" % (time.time() - t0))


if __name__ == "__main__":
    main(sys.argv[1:] or ["dev500x50", "sem351x48"])
