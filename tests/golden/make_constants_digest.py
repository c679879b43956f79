"""Writes tests/golden/constants_digest.json: SHA-256 digests of the
reference's literal neighbour tables (Atom3D atom3d/constants.py:47-52
NEIGHBOR_OFFSETS_6 and :73-93 EDGE_NEIGHBOR_OFFSETS) as int32 arrays.  The
tables are parsed as literal data (nothing is imported or executed).  Run in
a container that has the reference at /root/reference."""
import ast
import hashlib
import json
import os

import numpy as np

SRC = "/root/reference/atom3d/constants.py"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "constants_digest.json")
NAMES = ("NEIGHBOR_OFFSETS_6", "EDGE_NEIGHBOR_OFFSETS")


def main():
    out = {}
    for node in ast.parse(open(SRC).read()).body:
        if isinstance(node, ast.Assign) and getattr(node.targets[0], "id", None) in NAMES:
            a = np.array(ast.literal_eval(node.value.args[0]), np.int32)
            out[node.targets[0].id] = {"shape": list(a.shape),
                                       "sha256": hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()}
    assert set(out) == set(NAMES), out.keys()
    json.dump(out, open(OUT, "w"), indent=1)
    print(OUT)


if __name__ == "__main__":
    main()
