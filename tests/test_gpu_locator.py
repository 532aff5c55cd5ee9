"""Cell search through the cell locator (a uniform grid of candidate lists)
against the oracle's search over every cell in order
(World::FindCellContaining, World.cpp:28-37).

The multi-zone box has 33 cells: a 6 x 6 x 6 grid whose voxel faces fall
on the cube faces.  The sliced slab has 402 cells bounded along x only: a
one-axis grid.  Points sit exactly on voxel and cell faces, one ulp to
either side, on the grid's outer faces, outside the grid and at random; the
cell index must be the oracle's exactly.
"""
import numpy as np
import pytest

from minimc_amd import abi, configs
from oracle.oracle import Oracle, SoA
from tests.gpu_util import require_gpu
from tests.test_gpu_multigroup import sliced_slab

pytestmark = pytest.mark.gpu


def _soa(p):
    n = len(p)
    return SoA(n, x=p[:, 0].copy(), y=p[:, 1].copy(), z=p[:, 2].copy(),
               u=np.ones(n), v=np.zeros(n), w=np.zeros(n), E=np.full(n, 1e-6),
               rng=np.arange(1, n + 1))


def _edge_values(edges):
    v = []
    for e in edges:
        v += [e, np.nextafter(e, -np.inf), np.nextafter(e, np.inf), e - 1e-9, e + 1e-9]
    return np.array(v)


def _points(rng, axis_values, n_random, span):
    n = len(axis_values)
    # every axis value on each axis in turn, the others random in the box
    out = []
    for a in range(3):
        p = rng.uniform(-span, span, size=(n, 3))
        p[:, a] = axis_values
        out.append(p)
    # corners and edges: all three coordinates from the edge values
    out.append(rng.choice(axis_values, size=(4 * n, 3)))
    out.append(rng.uniform(-span, span, size=(n_random, 3)))
    return np.concatenate(out)


def _compare(pr, pts):
    from minimc_amd.tracker import Tracker
    o = Oracle(pr)
    t = Tracker(pr, max_in_flight=1 << 12)
    a = o.eval_event(abi.MMC_EVAL_FIND_CELL, _soa(pts))
    b = t.eval_event(abi.MMC_EVAL_FIND_CELL, _soa(pts))
    t.close()
    assert np.array_equal(a.cell, b.cell), np.flatnonzero(a.cell != b.cell)[:10]
    return a.cell


def test_multizone_cell_search_matches_oracle():
    require_gpu()
    pr = configs.c5_multizone(histories=1000)
    rng = np.random.default_rng(11)
    voxel_faces = np.linspace(-30.0, 30.0, 7)        # the 6 x 6 x 6 grid
    vals = np.concatenate([_edge_values(voxel_faces), _edge_values([-35.0, 35.0])])
    cells = _compare(pr, _points(rng, vals, 20000, 36.0))
    # every cube and every void cell is found somewhere
    assert set(range(len(pr.cells))) <= set(cells.tolist())


def test_sliced_slab_cell_search_matches_oracle():
    require_gpu()
    pr = sliced_slab()
    rng = np.random.default_rng(12)
    planes = np.array([s["a"] for s in pr.surfaces if s["kind"] == abi.MMC_PLANEX])
    xs = np.concatenate([_edge_values(planes), rng.uniform(-2.0, 22.0, 4000)])
    p = rng.uniform(-1e6, 1e6, size=(len(xs), 3))
    p[:, 0] = xs
    cells = _compare(pr, p)
    assert set(range(len(pr.cells))) <= set(cells.tolist())
