"""Turn the reference's committed RDS test fixtures into small golden files.

Run in a container where the reference tree is mounted (it is read as data
only; nothing in it is executed):

    python tests/golden/extract_fixtures.py /root/reference

Sources (relative to the reference root):
  fastkpc/tests/fixtures/legacy_dcov_gamma_oracle_v1.rds      -> dcov_gamma_oracle_v1.json
  fastkpc/tests/fixtures/full_cuda_ci_development_500x50_v2.rds -> dev500x50_v2.npz
  fastkpc/tests/fixtures/default_inf_production_cuda_v1.json    -> copied counts only
"""
import json
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from rds_reader import read_rds  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def _scalar(obj):
    v = obj.value[0]
    return v


def extract_dcov(root):
    src = os.path.join(root, "fastkpc/tests/fixtures/legacy_dcov_gamma_oracle_v1.rds")
    cases = []
    for case in read_rds(src).value:
        meta = case["meta"]
        res = case["residuals"]
        cases.append({
            "case_id": _scalar(meta["case_id"]),
            "n": _scalar(meta["n"]),
            "numCol": _scalar(meta["numCol"]),
            "index": _scalar(meta["index"]),
            "alpha": _scalar(meta["alpha"]),
            "p_value": _scalar(meta["p.value"]),
            "delete_edge": bool(_scalar(meta["delete_edge"])),
            "nV2": _scalar(meta["nV2"]),
            "nV2Mean": _scalar(meta["nV2Mean"]),
            "nV2Variance": _scalar(meta["nV2Variance"]),
            "rx": [float(v).hex() for v in res["rx"].value],
            "ry": [float(v).hex() for v in res["ry"].value],
        })
    out = os.path.join(HERE, "dcov_gamma_oracle_v1.json")
    with open(out, "w") as fh:
        json.dump({"source": "fastkpc/tests/fixtures/legacy_dcov_gamma_oracle_v1.rds",
                   "encoding": "residual vectors as float.hex strings",
                   "cases": cases}, fh, indent=1)
    print("wrote", out, len(cases), "cases")


def extract_dev500(root):
    src = os.path.join(root, "fastkpc/tests/fixtures/full_cuda_ci_development_500x50_v2.rds")
    fx = read_rds(src)
    cfg = fx["configuration"]
    data = np.array(fx["data"].value, dtype=np.float64).reshape(50, 500).T  # column-major
    ref = fx["oracle"]["reference"]
    p = 50
    adj = np.array([bool(v) for v in ref["adjacency"].value]).reshape(p, p).T
    pmax = np.array(ref["pMax"].value, dtype=np.float64).reshape(p, p).T
    # sepsets: list of p lists of p entries (NULL or 1-based int vector)
    sep_rows = []
    for i, row in enumerate(ref["sepsets"].value):
        for j, entry in enumerate(row.value):
            if entry is None:
                continue
            vals = entry.value if hasattr(entry, "value") else entry
            sep_rows.append([i, j, len(vals)] + [int(v) for v in vals] + [0] * (8 - len(vals)))
    sep = np.array(sep_rows, dtype=np.int32).reshape(-1, 11)
    lt = fx["oracle"]["logical_trace"]
    dt = fx["oracle"]["deletion_trace"]
    bs = fx["baseline_summary"]

    def skey(vals):
        return np.array([s if s is not None else "" for s in vals])

    np.savez_compressed(
        os.path.join(HERE, "dev500x50_v2.npz"),
        data=data,
        n_edgetests=np.array(ref["n.edgetests"].value, dtype=np.int32),
        adjacency=adj,
        pmax=pmax,
        sepsets=sep,
        lt_x=np.array(lt["x"].value, dtype=np.int32),
        lt_y=np.array(lt["y"].value, dtype=np.int32),
        lt_level=np.array(lt["level"].value, dtype=np.int32),
        lt_S=skey(lt["S_key"].value),
        lt_p=np.array(lt["p_value"].value, dtype=np.float64),
        lt_del=np.array([bool(v) for v in lt["deletes_edge"].value]),
        lt_seq=np.array(lt["logical_sequence_id"].value, dtype=np.int32),
        lt_task_index=np.array(lt["source_task_index"].value, dtype=np.int32),
        dt_level=np.array(dt["level"].value, dtype=np.int32),
        dt_edge_x=np.array(dt["edge_x"].value, dtype=np.int32),
        dt_edge_y=np.array(dt["edge_y"].value, dtype=np.int32),
        dt_tested_x=np.array(dt["tested_x"].value, dtype=np.int32),
        dt_tested_y=np.array(dt["tested_y"].value, dtype=np.int32),
        dt_S=skey(dt["S_key"].value),
        dt_p=np.array(dt["p_value"].value, dtype=np.float64),
        alpha=np.float64(float(_scalar(cfg["alpha"]))),
        num_col=np.int32(_scalar(cfg["num_col"])),
        resolved_max=np.int32(_scalar(cfg["resolved_max_conditioning_size"])),
        baseline_elapsed_sec=np.float64(_scalar(fx["baseline_elapsed_sec"])),
        # whole-run totals of the reference's levels data frame
        bs_levels=np.int32(_scalar(bs["levels"])),
        bs_tasks_planned=np.int64(_scalar(bs["tasks_planned"])),
        bs_tests_replayed=np.int64(_scalar(bs["tests_replayed"])),
        bs_tasks_ignored_after_delete=np.int64(_scalar(bs["tasks_ignored_after_delete"])),
        bs_deletions=np.int64(_scalar(bs["deletions"])),
    )
    print("wrote dev500x50_v2.npz")


def extract_counts(root):
    src = os.path.join(root, "fastkpc/tests/fixtures/default_inf_production_cuda_v1.json")
    with open(src) as fh:
        d = json.load(fh)
    keep = {k: d[k] for k in ("n", "p", "alpha", "resolved_max_conditioning_size",
                              "natural_stop_level", "n_edgetests", "logical_test_count",
                              "physical_test_count", "unique_target_key_count",
                              "native_setup_count")}
    keep["source"] = "fastkpc/tests/fixtures/default_inf_production_cuda_v1.json"
    with open(os.path.join(HERE, "default_inf_351x48_counts.json"), "w") as fh:
        json.dump(keep, fh, indent=1)
    print("wrote default_inf_351x48_counts.json")


if __name__ == "__main__":
    root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    extract_dcov(root)
    extract_dev500(root)
    extract_counts(root)
