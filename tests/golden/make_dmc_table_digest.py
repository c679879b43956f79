"""Writes tests/golden/dmc_tables_digest.json: SHA-256 digests and summary
counts of the reference's DMC tables (Atom3D
atom3d/mesh_extractor/tables.py: dmc_table, num_vd_table, check_table).
The tables are parsed as literal data (nothing is imported or executed).
Run in a container that has the reference at /root/reference."""
import ast
import hashlib
import json
import os

import numpy as np

SRC = "/root/reference/atom3d/mesh_extractor/tables.py"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "dmc_tables_digest.json")


def main():
    tables = {}
    for node in ast.parse(open(SRC).read()).body:
        if isinstance(node, ast.Assign) and node.targets[0].id in ("dmc_table", "num_vd_table", "check_table"):
            tables[node.targets[0].id] = np.array(ast.literal_eval(node.value), np.int64)
    out = {}
    for k, a in tables.items():
        out[k] = {"shape": list(a.shape), "sha256": hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest(),
                  "sum": int(a.sum())}
    out["num_vd_histogram"] = np.bincount(tables["num_vd_table"], minlength=5).tolist()
    out["check_cases"] = np.nonzero(tables["check_table"][:, 0] == 1)[0].tolist()
    json.dump(out, open(OUT, "w"), indent=1)
    print(OUT)


if __name__ == "__main__":
    main()
