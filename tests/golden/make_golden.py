"""Writes the committed fixtures of tests/golden/.

* ds_11948.npz  - the reference's only shipped input (SpaceOptimization/dataset/{colbeg,colend,csr}_11948.dat,
                  "BC"), converted to int32 arrays (beg, end, adj) so that tests and the GPU box can use it
                  without /root/reference. Its known answer (980,136 fill-ins) is recorded by the reference
                  in SpaceOptimization/with_supernodes/space_optimization_log_1_264_GPU_warp_centric.dat:2.
* golden.json   - known answers: the reference's fill count for the dataset, plus the values of the oracle
                  (after the oracle's own pins) for the dataset and the 64x64 grid workload.

Run from the repository root: python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
DATASET = "/root/reference/SpaceOptimization/dataset"
OUT = os.path.join(ROOT, "tests", "golden")


def read_ints(path):
    with open(path) as f:
        return np.array(f.read().split(), dtype=np.int64)


def main():
    import oracle
    import gsofatpds_amd as g

    files = {k: os.path.join(DATASET, f"{k}_11948.dat") for k in ("colbeg", "colend", "csr")}
    sha = {k: hashlib.sha256(open(p, "rb").read()).hexdigest() for k, p in files.items()}
    csr = read_ints(files["csr"])
    E, n = int(csr[0]), int(csr[1])
    adj = csr[2:2 + E].astype(np.int32)
    beg = read_ints(files["colbeg"])[:n].astype(np.int32)
    end = read_ints(files["colend"])[:n].astype(np.int32)
    np.savez_compressed(os.path.join(OUT, "ds_11948.npz"), beg=beg, end=end, adj=adj)

    golden = {"ds_11948": {"source": "SpaceOptimization/dataset/{colbeg,colend,csr}_11948.dat",
                           "sha256": sha, "n": n, "nnz": E,
                           "reference_fill_count": 980136,
                           "reference_log": "SpaceOptimization/with_supernodes/"
                                            "space_optimization_log_1_264_GPU_warp_centric.dat:2"}}
    ds = g.Pattern(n, beg, end, adj)
    for S in (128, 1024):
        o = oracle.symbfact(ds, chunk=S)
        golden["ds_11948"][f"oracle_S{S}"] = {k: int(o[k]) for k in ("fill", "nnz_F", "W", "R", "nsup")}
    p = g.generate_workload(g.WL_GRID2D_4096)
    o = oracle.symbfact(p, chunk=128, structure=True)
    h = hashlib.sha256()
    h.update(o["F_rowptr"].astype(np.int64).tobytes())
    h.update(o["F_cols"].astype(np.int32).tobytes())
    golden["grid2d_4096"] = {"n": p.n, "nnz": p.nnz,
                             "oracle_S128": {k: int(o[k]) for k in ("fill", "nnz_F", "W", "R", "nsup")},
                             "F_sha256": h.hexdigest(),
                             "nzu_sha256": hashlib.sha256(o["nzu"].astype(np.int32).tobytes()).hexdigest(),
                             "sup_flag_sha256": hashlib.sha256(o["sup_flag"].astype(np.uint8).tobytes()).hexdigest()}
    with open(os.path.join(OUT, "golden.json"), "w") as f:
        json.dump(golden, f, indent=1, sort_keys=True)
    print(json.dumps(golden, indent=1))


if __name__ == "__main__":
    main()
