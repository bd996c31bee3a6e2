"""The oracle's memory regions (OMP_NUM_THREADS > 1 in SCORE).

SCORE splits the domain into T regions (ompshar_ca2regions,
SCORE_GridPartitioning.cpp:69-166), each with its own Inside and Border lists.
A region's growth step touches only its own cells; an infection that crosses
into another region is parked in a halo (SCORE_Kernel.cpp:4494-4555) and, after
the barrier, the receiving region takes the parked infections up neighbour by
neighbour in its mynborregions order (ompshar_synchronize_haloregions_rx,
:4777-4872).  So for a cell that two fronts reach in the same step:

* a front of the cell's own region wins over any halo (its update precedes the
  barrier), whatever the pool order;
* between halos, the neighbour that comes first in the receiving region's
  order (myself, six faces -x +x -y +y -z +z, then edges, then corners,
  :2028-2080) wins;
* with one region the list order decides, which follows the pool order.

The expected winners below follow from those rules and the symmetry of the
hand-built domains, not from the oracle's code; the region grid follows from
smallestYFactor (SCORE_Math.cpp:97-115) worked by hand."""
import math

import numpy as np
import pytest

from oracle.oracle import Oracle, OracleError, smallest_y_factor
from score_amd import abi
from score_amd.synth import DomainArrays, ParamsSpec, make_voronoi_domain


def lin(x, y, z, nx, ny):
    return x + y * nx + z * nx * ny


def one_grain_domain(n, sites, same_ori=True):
    """One deformed grain filling an n^3 box and nuclei at the given sites, all
    with the same orientation (so every front advances at the same rate)."""
    N = n ** 3
    qdef = [1.0, 0.0, 0.0, 0.0]
    h = math.radians(40.0) / 2
    qrx = [math.cos(h), 0.0, 0.0, math.sin(h)]
    return DomainArrays(n, n, n, 0.5e-6, np.zeros(N, np.uint32), np.array([qdef, qrx]),
                        np.array([0], np.uint32), np.array([1e15]),
                        np.array([1] * len(sites), np.uint32),
                        np.array([lin(*s, n, n) for s in sites], np.uint32))


def claim_of(o, cell, nlists):
    """Step until `cell` is claimed; return (list index, rx grain) of the one
    ACTIVE record that holds it."""
    while o.grainmap()[cell] != abi.CURRENTLY_INFECTED:
        assert not o.step(1)
    hits = []
    for li in range(nlists):
        fr = o.front(li)
        for k in np.nonzero((fr["cell"] == cell) & (fr["activity"] == 1))[0]:
            hits.append((li, int(fr["rxg"][k])))
    assert len(hits) == 1, hits
    return hits[0]


# ------------------------------------------------------------ region grid

@pytest.mark.parametrize("T,npy", [(1, 1), (2, 1), (3, 1), (4, 2), (5, 1), (6, 2), (7, 1), (8, 2), (9, 3),
                                   (12, 3), (16, 2), (18, 3), (24, 3), (30, 5), (36, 3), (64, 2)])
def test_smallest_y_factor_by_hand(T, npy):
    """smallestYFactor keeps the last listed prime dividing T before the next
    prime reaches the cofactor: 16 -> 2 x 8 (4 x 4 is not a prime split),
    30 -> 5 x 6, 36 -> 3 x 12."""
    assert smallest_y_factor(T) == npy
    assert T % npy == 0 and npy <= T // npy


def test_region_boxes_and_capacities():
    """T = 4 on 20 x 30 x 25: npy = npz = 2, nry = 15, nrz = 12, the last z slab
    takes the remainder; region id = iz * npy + iy.  Each region's Border list
    holds 2(ab + ac + bc) of its own box and its Inside list floor(0.2 V)
    (SCORE_Kernel.cpp:3589, :3679)."""
    dom = make_voronoi_domain(20, 30, 25, n_grains=6, n_nuclei=10, seed=3)
    o = Oracle(dom, ParamsSpec(n_regions=4))
    (npy, npz), boxes = o.regions()
    assert (npy, npz) == (2, 2)
    want = [(0, 0, 0, 19, 14, 11), (0, 15, 0, 19, 29, 11), (0, 0, 12, 19, 14, 24), (0, 15, 12, 19, 29, 24)]
    assert [tuple(b) for b in boxes] == want
    for r, b in enumerate(want):
        ex, ey, ez = b[3] - b[0] + 1, b[4] - b[1] + 1, b[5] - b[2] + 1
        caps = o.capacities(r)
        assert caps[0] == int(0.2 * ex * ey * ez)
        assert caps[2] == 2 * (ex * ey + ex * ez + ey * ez)
    o.close()


def test_region_count_limits():
    dom = make_voronoi_domain(16, 24, 24, n_grains=4, n_nuclei=4, seed=1)
    for bad in (-1, 1025, 3):  # 3 regions along z: 8 cells thick, not more than 10
        with pytest.raises(OracleError):
            Oracle(dom, ParamsSpec(n_regions=bad))
    Oracle(dom, ParamsSpec(n_regions=2)).close()   # 1 x 2: 12 cells thick
    Oracle(dom, ParamsSpec(n_regions=4)).close()   # 2 x 2: 12 x 12


def test_one_region_is_the_default():
    dom = make_voronoi_domain(24, 22, 20, n_grains=8, n_nuclei=30, seed=5)
    runs = []
    for T in (0, 1):
        o = Oracle(dom, ParamsSpec(n_regions=T))
        o.solve_rxgrowth()
        runs.append((o.trajectory().tobytes(), o.grainmap().tobytes(), o.status()[0].tobytes()))
        o.close()
    assert runs[0] == runs[1]


# ------------------------------------------------------------ precedence

@pytest.mark.parametrize("T", [2, 3])
@pytest.mark.parametrize("first", ["below", "above"])
def test_own_region_front_beats_the_halo(first, T):
    """T = 2 on 24^3 (T = 3 on 36^3): regions z in [0,11], [12,23] (, [24,35]).  Nucleus L at (5,5,10) in
    region 0 and U at (5,5,14) in region 1 are two cells from c = (5,5,12), the
    bottom layer of region 1.  Their face children (5,5,11) and (5,5,13) complete
    in the same step (same orientation, same shape factor) and both aim at c:
    U's child directly in region 1's update, L's child through region 0's halo,
    taken up only after the barrier.  So U owns c with two regions, whatever the
    pool order; with one region the earlier pool entry's child comes first in
    the Inside list and owns c."""
    n = 12 * T
    L, U = (5, 5, 10), (5, 5, 14)
    sites = [L, U] if first == "below" else [U, L]
    iL, iU = sites.index(L), sites.index(U)
    c = lin(5, 5, 12, n, n)
    dom = one_grain_domain(n, sites)
    o1 = Oracle(dom, ParamsSpec(n_regions=1))
    o2 = Oracle(dom, ParamsSpec(n_regions=T))
    assert claim_of(o1, c, 2) == (abi.LIST_INSIDE, 0)             # pool order
    # region 1's own front, in its Border list (z = 12 is its lower face)
    assert claim_of(o2, c, 2 * T) == (2 * 1 + abi.LIST_BORDER, iU)
    for o in (o1, o2):
        o.solve_rxgrowth()
    assert o1.grainmap()[c] == 1 + 0 and o2.grainmap()[c] == 1 + iU
    o1.close()
    o2.close()


@pytest.mark.parametrize("T", [4, 9])
@pytest.mark.parametrize("first", ["y", "z"])
def test_face_halo_order_decides_between_neighbours(first, T):
    """T = 4 on 24^3 (T = 9 on 36^3): npy = npz = 2 (3), slabs 12 cells thick.
    c = (5,12,12) is the corner cell of region R = (iy 1, iz 1).  Nucleus Y at
    (5,10,12) lies in R's -y neighbour, Z at (5,12,10) in its -z neighbour, both
    two cells from c; their face children (5,11,12) and (5,12,11) complete in the
    same step and park c in their halos.  R takes up its -y neighbour (order 3)
    before its -z neighbour (order 5), so Y owns c whatever the pool order; with
    one region the pool order decides.  With three slabs a region's -y and +y
    neighbours differ, so the halo each neighbour keeps towards R (the opposite
    direction) is the only one holding c."""
    n = 12 * T // (2 if T == 4 else 3)
    Y, Z = (5, 10, 12), (5, 12, 10)
    sites = [Y, Z] if first == "y" else [Z, Y]
    iY = sites.index(Y)
    c = lin(5, 12, 12, n, n)
    dom = one_grain_domain(n, sites)
    o1 = Oracle(dom, ParamsSpec(n_regions=1))
    o4 = Oracle(dom, ParamsSpec(n_regions=T))
    npy = smallest_y_factor(T)
    assert claim_of(o1, c, 2) == (abi.LIST_INSIDE, 0)
    assert claim_of(o4, c, 2 * T) == (2 * (npy + 1) + abi.LIST_BORDER, iY)
    o1.close()
    o4.close()


def test_status_is_region_major():
    """ompcrit_clarify_status_infectedcells walks region 0's Inside and Border
    lists, then region 1's (SCORE_Io_Characterization.cpp:955-994): with two
    regions split at z = 12 every status entry of region 0 precedes region 1's."""
    n = 24
    dom = make_voronoi_domain(n, n, n, n_grains=6, n_nuclei=40, seed=11)
    o = Oracle(dom, ParamsSpec(n_regions=2))
    o.step(6)
    ix, gid = o.status()
    z = ix // (n * n)
    assert len(z) > 0 and (z < 12).any() and (z >= 12).any()
    k = int(np.argmax(z >= 12))
    assert (z[:k] < 12).all() and (z[k:] >= 12).all()
    o.close()


@pytest.mark.parametrize("T", [2, 4, 6, 8])
def test_regions_conserve_cells(T):
    """Whatever the region count, every cell is counted once: a cell leaves its
    deformed grain's count when it completes, so the deformed and rx counts add
    up to the box, the rx counts to Xcells, and the map agrees with both (a
    claimed cell still counts for its deformed grain)."""
    dom = make_voronoi_domain(26, 48, 50, n_grains=10, n_nuclei=60, seed=T)
    o = Oracle(dom, ParamsSpec(n_regions=T))
    r = o.solve_rxgrowth()
    dc, rc = o.counts()
    assert int(dc.sum()) + int(rc.sum()) == dom.N
    assert float(rc.sum()) == r["xcells"]
    g = o.grainmap()
    assert int((g < dom.n_defg).sum()) + int((g == abi.CURRENTLY_INFECTED).sum()) == int(dc.sum())
    rx = g[(g >= dom.n_defg) & (g != abi.CURRENTLY_INFECTED)] - dom.n_defg
    assert np.array_equal(np.bincount(rx, minlength=dom.n_rxg), rc)
    o.close()
