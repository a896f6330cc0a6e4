"""The oracle's tile schedule: the locality order in which the step visits
tiles (which, with the workgroup walk, fixes the RNG stream assignment)."""
from __future__ import annotations

import numpy as np
import pytest

import cudadsmc_amd as dm
import oracle


def _sim(dims, tile_cells, n_streams=1 << 14):
    s = dm.settings_default()
    s.grid_dim[:] = dims
    s.bmin[:] = (0.0, 0.0, 0.0)
    s.bmax[:] = (0.4, 0.2, 0.2)
    z = np.zeros(n_streams, np.uint32)
    return oracle.OracleSim(s, {"tile_cells": tile_cells, "threads": 256, "n_streams": n_streams}, (z, z, z))


def _rows(order, dims, tc):
    X, Y, _ = dims
    c = order.astype(np.int64) * tc
    return (c // X) % Y, c // (X * Y), c % X


@pytest.mark.parametrize("dims,tc", [((256, 128, 128), 128), ((64, 32, 32), 256), ((64, 32, 32), 32),
                                     ((100, 30, 20), 64)])
def test_tile_order_is_a_permutation(dims, tc):
    o = _sim(dims, tc)
    order = o.tile_order()
    ntiles = -(-dims[0] * dims[1] * dims[2] // tc)
    assert len(order) == ntiles
    assert np.array_equal(np.sort(order), np.arange(ntiles))


def test_tile_order_walks_patches_of_rows():
    # 256 x 128 x 128 cells, 128-cell tiles: two tiles per x row; a patch of
    # 8 x 8 rows holds 128 tiles
    dims, tc = (256, 128, 128), 128
    order = _sim(dims, tc).tile_order()
    for p in range(4):
        y, z, x = _rows(order[128 * p:128 * (p + 1)], dims, tc)
        assert len(set(zip(y.tolist(), z.tolist()))) == 64
        assert y.max() - y.min() == 7 and z.max() - z.min() == 7
        assert sorted(set(x.tolist())) == [0, 128]
    # consecutive patches advance in y first, then in z
    y0, z0, _ = _rows(order[:128], dims, tc)
    y1, z1, _ = _rows(order[128:256], dims, tc)
    assert y1.min() == y0.max() + 1 and z1.min() == z0.min()
    # a tile's y and z neighbours mostly sit in the same patch
    pos = np.empty(len(order), np.int64)
    pos[order] = np.arange(len(order))
    t = np.arange(len(order))
    same = 0
    total = 0
    for nb in (t + 2, t - 2, t + 256, t - 256):
        ok = (nb >= 0) & (nb < len(order))
        same += int(np.sum(pos[t[ok]] // 128 == pos[nb[ok]] // 128))
        total += int(ok.sum())
    assert same / total > 0.85


def test_tile_order_of_multi_row_tiles():
    # 64-cell rows, 256-cell tiles: a tile covers four rows of one z plane
    dims, tc = (64, 32, 32), 256
    order = _sim(dims, tc).tile_order()
    y, z, x = _rows(order, dims, tc)
    assert np.all(x == 0) and np.all(y % 4 == 0)


@pytest.mark.parametrize("dims,tc,streams", [((256, 128, 128), 128, 1 << 20), ((64, 32, 32), 32, 1 << 14),
                                             ((64, 32, 32), 256, 1 << 12), ((100, 30, 20), 64, 768 * 256)])
def test_tile_walk_partitions_the_order(dims, tc, streams):
    o = _sim(dims, tc, streams)
    G = streams // 256
    ntiles = -(-dims[0] * dims[1] * dims[2] // tc)
    seen = np.zeros(ntiles, np.int64)
    for vw in range(G):
        js = o.tile_walk(vw)
        assert np.all(np.diff(js.astype(np.int64)) > 0)  # increasing: one pass over the order
        seen[js] += 1
        if G % 8 == 0 and len(js):
            # every position of a workgroup lies in a patch of its XCD
            assert np.all((js // 128) % 8 == vw % 8)
    assert np.all(seen == 1)


def test_tile_walk_deals_patches_to_all_xcds():
    # the 1M-stream layout of the 33.5M-particle case: each XCD gets patches
    # spread over the whole order, so the body's tiles are shared out
    o = _sim((256, 128, 128), 128, 1 << 20)
    first = [o.tile_walk(vw)[0] for vw in range(8)]
    assert first == [128 * x for x in range(8)]
    for x in range(8):
        js = np.concatenate([o.tile_walk(vw) for vw in range(x, 4096, 8)])
        assert len(js) == 32768 // 8
        # a share of every sixteenth of the domain
        assert np.all(np.bincount(js // 2048, minlength=16) == 256)
