"""The generated DMC tables equal the reference's (digest of its values) and
satisfy their definition."""
import hashlib
import json
import os

import numpy as np

from atom3d_amd.mesh_extractor import tables as T

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "dmc_tables_digest.json")))


def test_tables_match_reference_digest():
    for name in ("dmc_table", "num_vd_table", "check_table"):
        a = np.ascontiguousarray(getattr(T, name), np.int64)
        assert list(a.shape) == GOLD[name]["shape"]
        assert int(a.sum()) == GOLD[name]["sum"]
        assert hashlib.sha256(a.tobytes()).hexdigest() == GOLD[name]["sha256"], name
    assert np.bincount(T.num_vd_table, minlength=5).tolist() == GOLD["num_vd_histogram"]


def test_groups_partition_the_crossing_edges():
    for c in range(256):
        inside = [(c >> i) & 1 for i in range(8)]
        cross = {e for e, (a, b) in enumerate(T.EDGE_CORNERS) if inside[a] != inside[b]}
        gs = [set(int(x) for x in row if x >= 0) for row in T.dmc_table[c] if row[0] >= 0]
        assert len(gs) == T.num_vd_table[c]
        assert set().union(*gs) == cross if gs else not cross
        assert sum(len(g) for g in gs) == len(cross)
        # complement cases have the same crossing edges
        gc = [set(int(x) for x in row if x >= 0) for row in T.dmc_table[255 - c] if row[0] >= 0]
        assert set().union(*gc) == cross if gc else not cross
