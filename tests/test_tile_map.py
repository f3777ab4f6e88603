"""The fused kernel's block -> tile map (evaluated on the host through a
test hook) covers every tile of every chunk exactly once, with and without
the XCD-aware ordering."""
import numpy as np
import pytest

from mesobasic_amd import lbm


@pytest.mark.parametrize("order", [0, 1, 2, 3])
@pytest.mark.parametrize("nty,ntz,ncx", [(16, 8, 8), (32, 16, 16), (2, 4, 4), (8, 1, 3), (3, 2, 2),
                                         (1, 5, 7), (24, 3, 8), (5, 8, 1), (7, 7, 7)])
def test_tile_map_is_a_bijection(nty, ntz, ncx, order):
    ty, tz, cx = lbm.probe_tile_map(nty, ntz, ncx, order)
    assert ty.min() >= 0 and ty.max() < nty
    assert tz.min() >= 0 and tz.max() < ntz
    assert cx.min() >= 0 and cx.max() < ncx
    ids = (cx.astype(np.int64) * ntz + tz) * nty + ty
    assert np.unique(ids).size == nty * ntz * ncx


@pytest.mark.parametrize("nty,ntz,ncx", [(16, 8, 8), (256, 1, 4), (512, 1, 8), (3, 2, 4)])
def test_tile_map_gives_each_xcd_a_contiguous_run(nty, ntz, ncx):
    # blocks run on XCD b % 8; when the block count is a multiple of 8 each
    # XCD walks a contiguous run of linear tiles (consecutive y-tiles), in
    # dispatch order
    n = nty * ntz * ncx
    assert n % 8 == 0
    ty, tz, cx = lbm.probe_tile_map(nty, ntz, ncx, 1)
    lin = (cx.astype(np.int64) * ntz + tz) * nty + ty
    b = np.arange(n)
    for g in range(8):
        run = lin[b % 8 == g]
        assert np.array_equal(run, np.arange(g * n // 8, (g + 1) * n // 8))


def test_tile_map_y_interleave_keeps_columns_on_one_xcd():
    nty, ntz, ncx = 16, 8, 8
    ty, tz, cx = lbm.probe_tile_map(nty, ntz, ncx, 2)
    b = np.arange(nty * ntz * ncx)
    for g in range(8):
        sel = b % 8 == g
        for z, x in {(int(z), int(x)) for z, x in zip(tz[sel], cx[sel])}:
            assert set(ty[sel & (tz == z) & (cx == x)]) == set(range(nty))


def test_tile_map_band_order_keeps_neighbours_on_one_xcd():
    nty, ntz, ncx = 64, 8, 8
    ty, tz, cx = lbm.probe_tile_map(nty, ntz, ncx, 3)
    b = np.arange(nty * ntz * ncx)
    for g in range(8):
        sel = b % 8 == g
        assert set(ty[sel]) == set(range(8 * g, 8 * g + 8))
        assert set(tz[sel]) == set(range(ntz))


def test_tile_map_exhaustive_small_grids():
    for order in (0, 1, 2, 3):
        for nty in range(1, 13):
            for ntz in range(1, 9):
                for ncx in range(1, 9):
                    ty, tz, cx = lbm.probe_tile_map(nty, ntz, ncx, order)
                    ids = (cx.astype(np.int64) * ntz + tz) * nty + ty
                    assert ids.min() >= 0 and ids.max() < nty * ntz * ncx
                    assert np.unique(ids).size == nty * ntz * ncx
