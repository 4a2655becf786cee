"""Oracle values for a stratified sample of the sem2000x64 permutation
skeletons (n=2000, p=64, 1000 replicates, seed 2024), where a full oracle
skeleton is out of reach on a CPU (tens of thousands of tests at seconds
each).

    python tests/golden/make_config5_sample.py TRACE_dcc_perm.npz TRACE_hsic_perm.npz

Each TRACE file is a GPU skeleton trace of one method (level, x, y, S, p) for
that workload.  Per method and level the script takes up to ROWS_PER_LEVEL
rows (the level's first and last rows, its deleting rows first, then evenly
spaced rows), runs the oracle's kernelCItest on each with the same seeded
permutation indices, and writes tests/golden/oracle_sem2000x64_sample.npz
(method, level, x, y, S_key and the oracle's p, statistic, mean, variance and
decision margin).  The GPU test then requires every sampled row in its own
trace with an identical p and statistic, mean and variance within 1e-9.
"""
import os
import sys
from multiprocessing import Pool

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

ROWS_PER_LEVEL = 6
SEED = 2024
REPLICATES = 1000


def pick(tr):
    rows = []
    for lev in sorted(set(tr["level"].tolist())):
        idx = np.nonzero(tr["level"] == lev)[0]
        dels = idx[tr["p"][idx] >= 0.1]
        keep = list(dels[:2]) + [idx[0], idx[-1]]
        for v in np.linspace(0, len(idx) - 1, ROWS_PER_LEVEL).astype(int):
            if len(set(keep)) >= ROWS_PER_LEVEL:
                break
            keep.append(idx[v])
        rows += sorted(set(int(k) for k in keep))[:ROWS_PER_LEVEL]
    return rows


def run(job):
    from fastkpc_amd.sem import config_data
    from oracle import orc
    method, x, y, skey = job
    data, _, _ = config_data("sem2000x64")
    S = [int(v) - 1 for v in skey.split("|")] if skey else []
    r = orc.ci_test(data, x - 1, y - 1, S, method=method, replicates=REPLICATES, seed=SEED)
    return r["p.value"], r["statistic"], r["mean"], r["variance"], r["margin"]


def main(paths):
    jobs, meta = [], []
    for path in paths:
        method = "dcc.perm" if "dcc" in os.path.basename(path) else "hsic.perm"
        tr = np.load(path)
        for i in pick(tr):
            jobs.append((method, int(tr["x"][i]), int(tr["y"][i]), str(tr["S"][i])))
            meta.append(int(tr["level"][i]))
    with Pool(8) as pool:
        out = pool.map(run, jobs, chunksize=1)
    out = np.array(out)
    np.savez_compressed(os.path.join(HERE, "oracle_sem2000x64_sample.npz"),
                        method=np.array([j[0] for j in jobs]), level=np.array(meta, np.int32),
                        x=np.array([j[1] for j in jobs], np.int32),
                        y=np.array([j[2] for j in jobs], np.int32),
                        S=np.array([j[3] for j in jobs]), p=out[:, 0], stat=out[:, 1],
                        mean=out[:, 2], var=out[:, 3], margin=out[:, 4],
                        replicates=np.int32(REPLICATES), seed=np.int64(SEED))
    print("wrote", len(jobs), "rows")


if __name__ == "__main__":
    main(sys.argv[1:])
