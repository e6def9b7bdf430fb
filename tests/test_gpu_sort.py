"""ct_sort_intersections_by_edge (geometry_utils.py:564-574) on the GPU against the oracle's
np.lexsort restatement, bit for bit.

The kernel ranks rows a tile at a time with the t values of the tile and a halo of rows
either side in LDS; a segment whose rows reach past that window is ranked from global
memory.  The rows here mix segment sizes from 1 to 5000 (both paths, groups straddling tile
edges), repeated start points (equal t: the stable order decides), NaN start points (NaN
sorts last) and segments of zero length (every t is 0).
"""

import sys

import numpy as np
import pytest

from oracle import refcpu as R

pytestmark = pytest.mark.gpu


def _rows(seed: int):
    rng = np.random.default_rng(seed)
    sizes = np.concatenate([
        rng.integers(1, 40, 3000),        # short segments, ranked from LDS
        [255, 256, 257, 300, 700, 1023, 1024, 1025, 2048, 5000],  # around and past the halo
        rng.integers(1, 4, 2000),
    ])
    rng.shuffle(sizes)
    n_seg = len(sizes) + 50
    seg_of = np.sort(rng.choice(n_seg, len(sizes), replace=False))  # some segments have no rows
    a = rng.uniform(-10, 10, (n_seg, 2))
    b = a + rng.normal(size=(n_seg, 2))
    b[::97] = a[::97]  # zero-length segments
    edge_coords = np.stack([a, b], axis=1)
    ii = np.repeat(seg_of, sizes).astype(np.int64)
    n = len(ii)
    xy = rng.uniform(-10, 10, (n, 2, 2))
    # equal start points within a segment: copy an earlier row's
    dup = rng.random(n) < 0.1
    src = np.maximum(np.arange(n) - rng.integers(1, 5, n), 0)
    same = dup & (ii[src] == ii)
    xy[same, 0] = xy[src[same], 0]
    xy[rng.random(n) < 0.01, 0, 0] = np.nan
    jj = rng.integers(0, 1 << 40, n).astype(np.int64)
    return ii, jj, xy, edge_coords


@pytest.mark.parametrize("seed", [0, 1])
def test_sort_intersections_by_edge_matches_lexsort(seed):
    from numba_celltree_amd import algorithms as A

    ii, jj, xy, edge_coords = _rows(seed)
    print(f"[sort] seed {seed}: {len(ii)} rows, largest segment {np.bincount(ii).max()}", file=sys.stderr,
          flush=True)
    gi, gj, gxy = A.sort_intersections_by_edge(ii, jj, xy, edge_coords)
    oi, oj, oxy = R.sort_intersections_by_edge(ii, jj, xy, edge_coords)
    assert np.array_equal(gi, oi)
    assert np.array_equal(gj, oj)
    assert np.array_equal(gxy.view(np.int64), oxy.view(np.int64))
