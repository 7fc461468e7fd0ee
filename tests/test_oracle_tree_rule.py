"""Pins the oracle's QTree rebuild rule (src/cuda/geometry.cu:1267-1393) with an
independent cell histogram.

The reference counts points per cell of a 16384 x 16384 grid over the map
(cuda_build_qtree, geometry.cu:1270-1279), then merges bottom-up: a block with
more than grid_capacity points becomes a BRANCH whose children are the
sub-blocks that are branches themselves or halving LEAF boxes of the block
(cuda_merge_qtree, geometry.cu:1295-1393). Here the cells are computed in numpy
from the points and every node of the oracle's tree is checked against them:

* a branch is a block holding more than cap points, with the block's box
  (lo + x * (hi - lo) / xdim, geometry.cu:1306-1311);
* a leaf child of a branch above the finest level is a sub-block holding at
  most cap points, with the halving box of geometry.cu:1318-1334;
* the data hold sub-blocks with exactly cap and cap + 1 points, so a `>=`
  in place of `>` (or the reverse) changes the tree and fails.
"""
import numpy as np
import pytest

import oracle_lib as O
from stcdb_amd import HostTrace, workload

ONE_DIM = 16384


def cells(xy, box):
    lo, hi = box
    cx = ((xy[:, 0] - lo[0]) / (hi[0] - lo[0]) * ONE_DIM).astype(np.int64)
    cy = ((xy[:, 1] - lo[1]) / (hi[1] - lo[1]) * ONE_DIM).astype(np.int64)
    return cx, cy


def run(cap, n, steps=3, **kw):
    w = workload("c1", grid_capacity=cap, num_objects=n, **kw)
    orc = O.Oracle(w["cfg"], w["box"])
    tr = HostTrace(w["gen"])
    xy = None
    for t, xy in tr.steps(steps):
        orc.step(xy, t)
    return w, orc, xy


@pytest.mark.parametrize("cap,n", [(3, 5000), (12, 5000), (40, 8000)])
def test_tree_split_rule_from_cell_histogram(cap, n):
    w, orc, xy = run(cap, n)
    lo, hi = w["box"]
    W, H = hi[0] - lo[0], hi[1] - lo[1]
    cx, cy = cells(xy, w["box"])
    assert cx.min() >= 0 and cy.min() >= 0 and cx.max() < ONE_DIM and cy.max() < ONE_DIM
    boxes, mids, typ, ch = orc.nodes()

    def block_of(b):
        xdim = int(round(W / (b[2] - b[0])))
        assert xdim & (xdim - 1) == 0 and xdim <= ONE_DIM
        return xdim, int(round((b[0] - lo[0]) / W * xdim)), int(round((b[1] - lo[1]) / H * xdim))

    def count(xdim, x, y):
        g = ONE_DIM // xdim
        return int(np.count_nonzero((cx // g == x) & (cy // g == y)))

    seen = {"cap": 0, "cap+1": 0, "leaves_checked": 0}
    branches = np.nonzero(typ == 1)[0]
    assert typ[0] == 1 and block_of(boxes[0]) == (1, 0, 0)
    for i in branches:
        xdim, x, y = block_of(boxes[i])
        b = boxes[i]
        # the block's box, evaluated in the reference's operation order
        assert b[0] == lo[0] + x * W / xdim and b[2] == lo[0] + (x + 1) * W / xdim
        assert b[1] == lo[1] + y * H / xdim and b[3] == lo[1] + (y + 1) * H / xdim
        assert mids[i][0] == (b[0] + b[2]) / 2 and mids[i][1] == (b[1] + b[3]) / 2
        c = count(xdim, x, y)
        assert c > cap, (i, xdim, x, y, c)
        seen["cap+1"] += c == cap + 1
        xhalf, yhalf = mids[i][0] - b[0], mids[i][1] - b[1]
        for k in range(4):
            v = int(ch[i][k])
            sub = (2 * xdim, 2 * x + k % 2, 2 * y + k // 2)
            if typ[v] == 1:
                assert xdim < ONE_DIM and block_of(boxes[v]) == sub
                continue
            assert typ[v] == 2
            cb = boxes[v]
            assert cb[0] == b[0] + (k % 2) * xhalf and cb[2] == mids[i][0] + (k % 2) * xhalf
            assert cb[1] == b[1] + (k // 2) * yhalf and cb[3] == mids[i][1] + (k // 2) * yhalf
            if xdim < ONE_DIM:
                sc = count(*sub)
                assert sc <= cap, (v, sub, sc)
                seen["cap"] += sc == cap
                seen["leaves_checked"] += 1
    # both sides of the threshold occur in the data
    assert seen["cap"] > 0 and seen["cap+1"] > 0 and seen["leaves_checked"] > 100, seen


def test_leaf_member_counts_follow_the_cells():
    """a leaf's reported member count is the number of points whose descent
    ends in it; away from ulp slivers that is its sub-block's cell count"""
    cap = 12
    w, orc, xy = run(cap, 5000)
    lo, hi = w["box"]
    lb, lc = orc.leaves()
    assert lc.sum() == len(xy)
    leaf, _ = orc.point_leaf()
    assert np.array_equal(np.bincount(leaf, minlength=len(lc)), lc)
