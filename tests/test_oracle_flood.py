"""Pinning the flood-fill oracle (no GPU): reachability against scipy's
connected-component labelling; dam / collision / water / dry against a direct
neighbourhood definition."""
import numpy as np
import pytest
from scipy import ndimage

from oracle import flood

RANK = {6: 1, 18: 2, 26: 3}


def _dil(x, conn):
    st = ndimage.generate_binary_structure(3, RANK[conn])
    return ndimage.binary_dilation(x, structure=st)


@pytest.mark.parametrize("conn", [6, 18, 26])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_dense_labels_match_connected_components(conn, seed):
    rng = np.random.default_rng(seed)
    occ = rng.random((19, 23, 17)) < (0.25 if conn == 6 else 0.45)
    free = np.argwhere(~occ)
    start = free[rng.integers(len(free))]
    mask = flood.flood_fill_dense(occ, start, conn)
    lab, _ = ndimage.label(~occ, structure=ndimage.generate_binary_structure(3, RANK[conn]))
    reached = lab == lab[tuple(start)]
    np.testing.assert_array_equal(mask > 0, reached)
    dam = occ & _dil(reached, conn)
    col = reached & _dil(occ, conn)
    np.testing.assert_array_equal(mask == -1, dam)
    np.testing.assert_array_equal(mask == 1, col)
    np.testing.assert_array_equal(mask == 2, reached & ~col)
    np.testing.assert_array_equal(mask == -2, ~reached & ~dam)


def test_start_outside_grid_leaves_everything_dry():
    occ = np.zeros((4, 5, 6), bool)
    mask = flood.flood_fill_dense(occ, (9, 0, 0), 26)
    assert np.all(mask == -2)


def test_sparse_pipeline_rules():
    # a hollow box shell: water outside, dry inside, dam = the shell
    R = 40
    g = np.zeros((R, R, R), bool)
    g[10:30, 10:30, 10:30] = True
    g[11:29, 11:29, 11:29] = False
    dam_coords = np.argwhere(g)
    for conn in (6, 26):
        water, dry, col, tdam, dmask, ext = flood.flood_fill_sparse(dam_coords, R, (0, 0, 0), padding=2,
                                                                    connectivity=conn)
        inside = {tuple(x) for x in np.argwhere(np.pad(np.ones((18, 18, 18), bool), 11)[:R, :R, :R])}
        assert {tuple(x) for x in dry} == inside          # the cavity stays dry
        assert len(tdam) == len(dam_coords) and dmask.all()
        assert len(ext) == 0
        lex = water[:, 0] * R * R + water[:, 1] * R + water[:, 2]
        assert np.all(np.diff(lex) > 0)                    # lexicographic order
        nb, s = flood.neighborhood_signs(tdam, water, dry, R, 1)
        assert set(np.unique(s)) <= {-1, 0, 1}
        assert (s == 0).sum() == len(tdam)


def test_occupied_source_gives_empty():
    dam = np.array([[3, 3, 3], [3, 3, 4]])
    out = flood.flood_fill_sparse(dam, 16, source=(3, 3, 3), padding=2)
    assert all(len(o) == 0 for o in out)
