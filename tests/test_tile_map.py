"""The fused kernel's block -> tile map (evaluated on the host through a
test hook) covers every tile of every chunk exactly once, with and without
the XCD-aware ordering."""
import numpy as np
import pytest

from mesobasic_amd import lbm


@pytest.mark.parametrize("nty,ntz,ncx", [(16, 8, 8), (32, 16, 16), (2, 4, 4), (8, 1, 3), (3, 2, 2),
                                         (1, 5, 7), (24, 3, 8), (5, 8, 1), (7, 7, 7)])
def test_tile_map_is_a_bijection(nty, ntz, ncx):
    ty, tz, cx = lbm.probe_tile_map(nty, ntz, ncx)
    assert ty.min() >= 0 and ty.max() < nty
    assert tz.min() >= 0 and tz.max() < ntz
    assert cx.min() >= 0 and cx.max() < ncx
    ids = (cx.astype(np.int64) * ntz + tz) * nty + ty
    assert np.unique(ids).size == nty * ntz * ncx


def test_tile_map_groups_y_neighbours_on_one_xcd():
    nty, ntz, ncx = 16, 8, 8
    ty, tz, cx = lbm.probe_tile_map(nty, ntz, ncx)
    b = np.arange(nty * ntz * ncx)
    # blocks on one XCD (same b % 8) that share (tz, cx) hold all y-tiles
    for g in range(8):
        sel = b % 8 == g
        col = {(int(z), int(x)) for z, x in zip(tz[sel], cx[sel])}
        for z, x in col:
            assert set(ty[sel & (tz == z) & (cx == x)]) == set(range(nty))


def test_tile_map_exhaustive_small_grids():
    for nty in range(1, 13):
        for ntz in range(1, 9):
            for ncx in range(1, 9):
                ty, tz, cx = lbm.probe_tile_map(nty, ntz, ncx)
                ids = (cx.astype(np.int64) * ntz + tz) * nty + ty
                assert ids.min() >= 0 and ids.max() < nty * ntz * ncx
                assert np.unique(ids).size == nty * ntz * ncx
