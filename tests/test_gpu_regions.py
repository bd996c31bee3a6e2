"""GPU growth machine with several memory regions (SCORE with OMP_NUM_THREADS =
T > 1) against the oracle: every region's Inside and Border lists slot by slot,
grain maps, counts, status vectors and trajectories bit for bit, at T = 2, 3, 4,
8 (and 9 for the hand-built halo-order case), on small domains and on the
100^3 domain."""
import numpy as np
import pytest

from score_amd import abi, synth
from score_amd.synth import ParamsSpec

from oracle.oracle import Oracle, smallest_y_factor
from test_gpu_parity import assert_front_equal, assert_traj_equal
from test_oracle_regions import one_grain_domain, lin

pytestmark = pytest.mark.gpu


def _gm(dom, prm):
    from score_amd.growth import GrowthMachine
    return GrowthMachine(dom, prm)


def assert_lists_equal(o, g, T, what):
    for lst in range(2 * T):
        assert_front_equal(o.front(lst), g.front(lst), f"{what} list{lst}")


def assert_final_equal_regions(o, g, T):
    ro, rg = o.result(), g.result()
    for k in ("t", "X", "xcells", "step", "completed", "sum_sv", "nuclei_placed", "pmax"):
        assert ro[k] == rg[k], (k, ro[k], rg[k])
    assert_traj_equal(o.trajectory(), g.trajectory())
    mo, mg = o.grainmap(), g.grainmap()
    assert np.array_equal(mo, mg), f"grain maps differ in {(mo != mg).sum()} cells"
    do, ro_ = o.counts()
    dg, rg_ = g.counts()
    assert np.array_equal(do, dg) and np.array_equal(ro_, rg_)
    assert_lists_equal(o, g, T, "final")
    for a, b in zip(o.status(0.5), g.status(0.5)):
        assert np.array_equal(a, b)


def region_domain(T, seed, n_grains=8, n_nuclei=30):
    """a small Voronoi domain whose slabs are 12 to 14 cells thick at T regions"""
    npy = smallest_y_factor(T)
    npz = T // npy
    rng = np.random.default_rng(seed)
    ny = 12 * npy + int(rng.integers(0, 3))
    nz = 12 * npz + int(rng.integers(0, 3))
    nx = int(rng.integers(14, 30))
    return synth.make_voronoi_domain(nx, ny, nz, n_grains=n_grains, n_nuclei=n_nuclei, seed=seed)


@pytest.mark.parametrize("T", [2, 3, 4, 8])
def test_regions_stepwise(require_gpu, T):
    dom = region_domain(T, seed=10 + T)
    prm = ParamsSpec(n_regions=T)
    print(f"regions stepwise T={T} {dom.nx}x{dom.ny}x{dom.nz}", flush=True)
    o, g = Oracle(dom, prm), _gm(dom, prm)
    assert_lists_equal(o, g, T, "nucleation")
    k = 0
    while True:
        do, dg = o.step(1), g.step(1)
        assert do == dg, k
        if do:
            break
        assert_lists_equal(o, g, T, f"step{k}")
        k += 1
    assert_final_equal_regions(o, g, T)
    o.close()
    g.close()


@pytest.mark.parametrize("T,chunks", [(2, (3, 8, 5)), (4, (8,)), (6, (2, 7)), (8, (5, 1, 8))])
def test_regions_in_launch_windows(require_gpu, T, chunks):
    dom = region_domain(T, seed=40 + T, n_grains=10, n_nuclei=45)
    prm = ParamsSpec(n_regions=T)
    print(f"regions windows T={T} {dom.nx}x{dom.ny}x{dom.nz} chunks={chunks}", flush=True)
    o, g = Oracle(dom, prm), _gm(dom, prm)
    k = 0
    while True:
        n = chunks[k % len(chunks)]
        do, dg = o.step(n), g.step(n)
        assert do == dg, k
        assert_traj_equal(o.trajectory(), g.trajectory())
        assert_lists_equal(o, g, T, f"chunk{k}")
        if do:
            break
        k += 1
    assert_final_equal_regions(o, g, T)
    o.close()
    g.close()


@pytest.mark.parametrize("T,chunks", [(2, (20, 13)), (4, (33,)), (8, (11, 27))])
def test_regions_chunks_of_several_windows(require_gpu, T, chunks):
    """chunks longer than a launch window: several windows per call, and chunks
    that end inside a window"""
    dom = region_domain(T, seed=60 + T, n_grains=10, n_nuclei=45)
    prm = ParamsSpec(n_regions=T)
    print(f"regions chunks of several windows T={T} {dom.nx}x{dom.ny}x{dom.nz} chunks={chunks}", flush=True)
    o, g = Oracle(dom, prm), _gm(dom, prm)
    k = 0
    while True:
        n = chunks[k % len(chunks)]
        do, dg = o.step(n), g.step(n)
        assert do == dg, k
        assert_traj_equal(o.trajectory(), g.trajectory())
        assert_lists_equal(o, g, T, f"chunk{k}")
        if do:
            break
        k += 1
    assert_final_equal_regions(o, g, T)
    o.close()
    g.close()


@pytest.mark.parametrize("T", [2, 3])
@pytest.mark.parametrize("first", ["below", "above"])
def test_own_region_front_beats_the_halo(require_gpu, first, T):
    """the hand-built conflict of the oracle's pin: the cell's own region wins
    over a halo in the same step, whatever the pool order"""
    n = 12 * T
    L, U = (5, 5, 10), (5, 5, 14)
    sites = [L, U] if first == "below" else [U, L]
    dom = one_grain_domain(n, sites)
    prm = ParamsSpec(n_regions=T)
    o, g = Oracle(dom, prm), _gm(dom, prm)
    k = 0
    while not o.step(1):
        assert not g.step(1)
        assert_lists_equal(o, g, T, f"step{k}")
        k += 1
    assert g.step(1)
    assert_final_equal_regions(o, g, T)
    assert g.grainmap()[lin(5, 5, 12, n, n)] == 1 + sites.index(U)


@pytest.mark.parametrize("T", [4, 9])
@pytest.mark.parametrize("first", ["y", "z"])
def test_face_halo_order_decides_between_neighbours(require_gpu, first, T):
    n = 12 * T // (2 if T == 4 else 3)
    Y, Z = (5, 10, 12), (5, 12, 10)
    sites = [Y, Z] if first == "y" else [Z, Y]
    dom = one_grain_domain(n, sites)
    prm = ParamsSpec(n_regions=T)
    o, g = Oracle(dom, prm), _gm(dom, prm)
    k = 0
    while not o.step(1):
        assert not g.step(1)
        assert_lists_equal(o, g, T, f"step{k}")
        k += 1
    assert g.step(1)
    assert_final_equal_regions(o, g, T)
    assert g.grainmap()[lin(5, 12, 12, n, n)] == 1 + sites.index(Y)


@pytest.mark.parametrize("T", [2, 4])
def test_regions_defragmentation(require_gpu, T):
    """every region compacts both its lists at the same Xcells thresholds"""
    dom = region_domain(T, seed=70 + T, n_grains=9, n_nuclei=40)
    prm = ParamsSpec(n_regions=T, defrag_at=(0.1 * dom.N, 0.3 * dom.N, 0.6 * dom.N))
    o, g = Oracle(dom, prm), _gm(dom, prm)
    k = 0
    while True:
        do, dg = o.step(3), g.step(3)
        assert do == dg, k
        assert_lists_equal(o, g, T, f"chunk{k}")
        if do:
            break
        k += 1
    assert_final_equal_regions(o, g, T)


@pytest.mark.parametrize("T", [2, 3, 4, 8])
def test_regions_voronoi100(require_gpu, T):
    """the 100^3 domain (50 grains, 500 nuclei) run to X = 0.99 with T regions"""
    dom = synth.make_voronoi_domain(100, 100, 100, n_grains=50, n_nuclei=500, seed=1)
    prm = ParamsSpec(n_regions=T)
    o, g = Oracle(dom, prm), _gm(dom, prm)
    o.solve_rxgrowth()
    g.solve_rxgrowth()
    assert_final_equal_regions(o, g, T)


@pytest.mark.parametrize("seed", range(12))
def test_regions_mixed_settings(require_gpu, seed):
    """seeded draws of region count, mobility model, recovery, drag and
    defragmentation, stepped in launch windows of drawn lengths"""
    rng = np.random.default_rng(500 + seed)
    T = int(rng.choice([2, 3, 4, 5, 6, 8, 9, 12]))
    dom = region_domain(T, seed=900 + seed, n_grains=int(rng.integers(2, 12)), n_nuclei=int(rng.integers(3, 40)))
    kw = {"n_regions": T, "nmax": 3000,
          "mobility_model": abi.MOB_SEBALDGOTTSTEIN if rng.random() < 0.3 else abi.MOB_ROLLETTHOLM,
          "rh_hagb_m0": float(10 ** rng.uniform(-0.5, 1.5))}
    rec = int(rng.integers(0, 5))
    if rec == 1:
        kw.update(recovery_mode=1, rec_vacancy_d0=float(10 ** rng.uniform(-7, -5)))
    elif rec == 3:
        kw.update(recovery_mode=3)
    elif rec == 4:
        kw.update(recovery_mode=4, rec_tau=float(rng.uniform(20.0, 2000.0)))
    if rng.random() < 0.4:
        kw.update(zener_mode=2, zener_alpha=1.5, zener_gamma=0.324, disp_time=(0.0, 500.0, 5000.0),
                  disp_fr=(0.0, 5e4, 2e4))
    if rng.random() < 0.4:
        kw["defrag_at"] = tuple(sorted(float(x) * dom.N for x in rng.uniform(0.05, 0.9, 2)))
    prm = ParamsSpec(**kw)
    chunks = tuple(int(c) for c in rng.integers(1, 9, 3))
    print(f"regions mixed {seed}: T={T} {dom.nx}x{dom.ny}x{dom.nz} model={prm.mobility_model} rec={prm.recovery_mode} "
          f"zener={prm.zener_mode} defrag={len(prm.defrag_at)} chunks={chunks}", flush=True)
    o, g = Oracle(dom, prm), _gm(dom, prm)
    k = 0
    while True:
        n = chunks[k % len(chunks)]
        do, dg = o.step(n), g.step(n)
        assert do == dg, k
        assert_traj_equal(o.trajectory(), g.trajectory())
        if do:
            break
        k += 1
    assert_final_equal_regions(o, g, T)
