"""Synthetic inputs: bucketed Voronoi labelling equals brute force, pools are
valid, everything is a pure function of the seed."""
import numpy as np

from score_amd import synth


def test_voronoi_matches_brute_force():
    for seed, shape, n, stretch in [(1, (23, 17, 11), 37, (1, 1, 1)), (2, (30, 30, 12), 20, (4, 1, 0.25)),
                                    (3, (16, 16, 16), 1, (1, 1, 1)), (4, (40, 9, 7), 60, (1, 2, 1))]:
        xyz = synth.points(seed, n, *shape)
        a = synth.voronoi(*shape, xyz, stretch=stretch)
        b = synth.voronoi(*shape, xyz, stretch=stretch, brute_force=True)
        assert np.array_equal(a, b)


def test_voronoi_periodic_xy_not_z():
    # one seed near x=0 and one near x=nx: the cells at the far x face belong to the first by wrap
    xyz = np.array([[0.2, 5.0, 5.0], [6.0, 5.0, 5.0]])
    g = synth.voronoi(20, 10, 10, xyz).reshape(10, 10, 20)
    assert g[5, 5, 19] == 0
    xyz = np.array([[5.0, 5.0, 0.2], [5.0, 5.0, 6.0]])
    g = synth.voronoi(10, 10, 20, xyz).reshape(20, 10, 10)
    assert g[19, 5, 5] == 1


def test_pools_valid_and_seeded():
    d1 = synth.make_voronoi_domain(32, 32, 32, n_grains=20, n_nuclei=300, seed=9)
    d2 = synth.make_voronoi_domain(32, 32, 32, n_grains=20, n_nuclei=300, seed=9)
    d3 = synth.make_voronoi_domain(32, 32, 32, n_grains=20, n_nuclei=300, seed=10)
    assert np.array_equal(d1.cellgrid, d2.cellgrid) and np.array_equal(d1.rxg_site, d2.rxg_site)
    assert not np.array_equal(d1.rxg_site, d3.rxg_site)
    assert len(np.unique(d1.rxg_site)) == 300 and d1.rxg_site.max() < d1.N
    assert np.allclose(np.linalg.norm(d1.ori_quat, axis=1), 1.0, atol=1e-12)
    assert d1.defg_rho.min() >= 1e14 and d1.defg_rho.max() < 1e15
    assert set(np.unique(d1.cellgrid)) <= set(range(20))


def test_quaternions_uniform():
    q = synth.quaternions(5, 20000)
    # uniform on the unit 3-sphere: E[qi^2] = 1/4 for every component, E|qi| = 4/(3 pi)
    assert np.allclose((q ** 2).mean(axis=0), 0.25, atol=0.01)
    assert abs(np.abs(q[:, 0]).mean() - 4 / (3 * np.pi)) < 0.01
