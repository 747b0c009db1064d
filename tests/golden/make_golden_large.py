"""Oracle values for the 110,592-row and 659,033-row workloads (tests/golden/golden_large.json):
totals, digests of nzU and the supernode flags, and digests of the explicit L/U structure.

Too slow for a test (the label-correcting restatement takes minutes at these sizes), so they are
computed once here and committed; the GPU tests compare the engine against them at full size.
Run from the repository root: python tests/golden/make_golden_large.py [grid3d_110592] [grid2d_659033]
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
OUT = os.path.join(ROOT, "tests", "golden", "golden_large.json")


def main(names):
    import oracle
    import gsofatpds_amd as g
    wl = {"grid3d_110592": g.WL_GRID3D_110592, "grid2d_659033": g.WL_GRID2D_659033}
    data = json.load(open(OUT)) if os.path.exists(OUT) else {}
    for name in names:
        p = g.generate_workload(wl[name])
        t0 = time.time()
        entry = {"n": p.n, "nnz": p.nnz}
        for S in (128,):
            o = oracle.symbfact(p, chunk=S)
            entry[f"oracle_S{S}"] = {k: int(o[k]) for k in ("fill", "nnz_F", "W", "R", "nsup", "expansions",
                                                           "edge_checks")}
            entry[f"nzu_sha256_S{S}"] = hashlib.sha256(o["nzu"].astype(np.int32).tobytes()).hexdigest()
            entry[f"sup_flag_sha256_S{S}"] = hashlib.sha256(o["sup_flag"].astype(np.uint8).tobytes()).hexdigest()
        # the explicit structure: strict U by rows and strict L by columns, as digests
        o = oracle.symbfact(p, chunk=128, structure=True, stats=False)
        parts = oracle.split_structure_fast(p.n, o["F_rowptr"], o["F_cols"])
        del o
        entry["U_rows_sha256"], entry["L_cols_sha256"] = oracle.structure_digests(*parts)
        entry["nnz_U"], entry["nnz_L"] = int(parts[0][-1]), int(parts[2][-1])
        del parts
        entry["oracle_seconds"] = time.time() - t0
        data[name] = entry
        with open(OUT, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        print(name, json.dumps(entry), flush=True)


if __name__ == "__main__":
    main(sys.argv[1:] or ["grid3d_110592", "grid2d_659033"])
