"""SCORE's input deck and domain setup (score_amd/domain_setup.py over
libscore_setup.so) against the reference's own code and rules.

Pinned bit for bit against the reference's SCORE_Random.cpp / SCORE_Math.cpp
(oracle/_ref, built where the reference tree is present): the Mersenne Twister
stream of every seed, the Shoemake nucleus orientations and the Bunge-to-
quaternion conversion.  The composition -- which draw becomes which cuboid
offset, grain pick, nucleus site and orientation (SCORE_StructureSynthesis.cpp:
355-529, SCORE_NucleationModels.cpp:207-282) -- is checked by replaying it from
the reference's own streams.  The reader is checked on crafted decks for the
rules of io::readDataBlock/geTReal/geTInt (SCORE_Io.cpp:265-432) and on the
decks the reference ships."""
import glob
import math
import os
import sys

import numpy as np
import pytest

from score_amd import abi
from score_amd import domain_setup as ds

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
import make_uds_deck as mk  # noqa: E402

from oracle import ref  # noqa: E402
from oracle.oracle import Oracle  # noqa: E402

needs_ref = pytest.mark.skipif(not ref.available(), reason="oracle/_ref not built (the reference tree is absent)")
DECK = os.path.join(ROOT, "tests", "golden", "cuboid_fe.uds")
SEEDS = [-11, -12, -1, 1, 0, 5489, -46356, 2147483647, -2147483648, 4357]


def f32(x: float) -> float:
    return float(np.float32(x))


@needs_ref
@pytest.mark.parametrize("seed", SEEDS)
def test_mersenne_twister_is_the_references(seed):
    got = ds.MersenneTwister(seed).draws(3000)  # crosses four regenerations of the state
    want = ref.mt_stream(seed, 3000)
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))
    assert got.min() >= 0.0 and got.max() <= 1.0 - np.finfo(float).eps


@needs_ref
def test_shoemake_is_the_references():
    g = ds.MersenneTwister(-7)
    got = np.array([g.shoemake() for _ in range(500)])
    want = ref.shoemake(-7, 500)
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))
    assert np.allclose(np.linalg.norm(got, axis=1), 1.0, atol=1e-15)


def write(tmp_path, text, name="deck.uds"):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_reader_rules(tmp_path):
    text = ('|| First\t|| *(ss)\t|| Parameter, Value\n'
            '"A"\t\t"1.5"\t\t// a comment after the value\n'
            '\n'
            '"Broken\t\t"7"\n'                       # a key with one quote, as in the reference's own decks
            '"Tail"\t"2x"\n'
            '"Sp"\t"3 "\n'
            '   || indented header line ends the block\n'
            '"Hidden"\t"9"\n'
            '|| First\t|| *(ss)\t|| a second block of the same name is never read\n'
            '"A"\t"2.5"\n'
            '|| Tab\t|| *(ff)\t|| time, value\n'
            '0.1\t\t20\n'
            '\n'
            '5.8\t\t600.7 extra tokens are ignored\n'
            '|| Bad\t|| *(ff)\t|| x\n'
            '1.0\n'
            '|| Typo\t|| (ff)\t|| no star\n'
            '1 2\n')
    u = ds.UdsFile(write(tmp_path, text))
    assert u.real("First", "A") == 1.5
    assert u.int("First", "Broken") == 7
    assert u.shape("First") == (4, 2)
    with pytest.raises(ds.SetupError) as e:
        u.real("First", "Tail")                     # geTReal: something's behind the double
    assert e.value.code == ds.ERR_FORMAT
    with pytest.raises(ds.SetupError):
        u.real("First", "Hidden")                   # after the indented "||" line
    with pytest.raises(ds.SetupError):
        u.real("First", "Missing")
    assert u.shape("Tab") == (2, 2)
    assert u.table_real("Tab", 0, 0) == f32(0.1) != 0.1   # 'f' columns are read as float
    assert u.table_real("Tab", 1, 1) == f32(600.7)
    for blk in ("Bad", "Typo", "Nowhere"):
        with pytest.raises(ds.SetupError) as e:
            u.shape(blk)
        assert e.value.code == ds.ERR_FORMAT
    with pytest.raises(ds.SetupError) as e:
        ds.UdsFile(str(tmp_path / "absent.uds"))
    assert e.value.code == ds.ERR_FILE


@pytest.mark.skipif(not os.path.isdir("/root/reference/msmse2019"), reason="the reference tree is absent")
def test_reference_decks():
    decks = sorted(glob.glob("/root/reference/msmse2019/*.uds"))
    assert decks
    for d in decks:
        u = ds.UdsFile(d)
        assert u.real("RuntimeControl", "XMAX") == 1.0 and u.int("RuntimeControl", "NMAX") == 100000
        assert u.real("EnsembleDefinition", "CellSize") == 0.25
        assert u.shape("AnnealingSchedule") == (13, 2) and u.table_real("AnnealingSchedule", 1, 1) == 600.0
        assert u.shape("DeformedGrainsPool")[1] == 4
        with pytest.raises(ds.SetupError) as e:
            ds.Ensemble(u)                          # these decks synthesize from EBSD maps
        assert e.value.code == ds.ERR_UNSUPPORTED
    for d in glob.glob("/root/reference/*/SCORE.Input.TemplateIron.uds"):
        with pytest.raises(ds.SetupError) as e:     # the legacy layout has no RuntimeControl block
            ds.Ensemble.from_uds(d)
        assert e.value.code == ds.ERR_FORMAT


def test_committed_deck_is_the_generators():
    assert open(DECK).read() == mk.deck_text()


@needs_ref
def test_world_pool_quaternions_and_merge():
    e = ds.Ensemble.from_uds(DECK)
    rows = mk.pool_rows()
    assert e.info.n_world_defg == len(rows) and e.info.n_world_ori == len(rows) - 2
    bunge = np.array([[f32(float(x)) / 180.0 * math.pi for x in r[:3]] for r in rows])
    q = ref.euler2quat(bunge)
    oris = [e.world_defg(g)[0] for g in range(len(rows))]
    assert oris[1] == oris[0] and oris[7] == oris[0]       # within 0.5 degree: merged into row 0
    assert len(set(oris)) == len(rows) - 2
    for g in range(len(rows)):
        o, rho = e.world_defg(g)
        assert rho == f32(float(rows[g][3]))
        if g not in (1, 7):
            assert np.array_equal(e.world_quat(o).view(np.uint64), q[g].view(np.uint64)), g


@needs_ref
@pytest.mark.parametrize("jobid", [0, 3])
def test_domain_replays_scores_draws(jobid):
    """Replay the domain's setup from the reference's own streams: three cuboid
    offsets, one grain pick per cuboid, then per nucleus a site draw (a site
    already taken costs one draw), a pool draw and three Shoemake draws from the
    math stream."""
    e = ds.Ensemble.from_uds(DECK)
    dom, prm, drawn = e.domain(jobid)
    seed = -(jobid + 1) - 10
    assert drawn.local_seed == seed
    info = e.info
    mt = ref.mt_stream(seed, 20000)
    k = 0
    dg = info.defg_mean
    off = [-1.0 * mt[k + a] * dg[a] for a in range(3)]
    k += 3
    assert tuple(off) == drawn.offset
    box = [info.cellsize * n for n in (info.nx, info.ny, info.nz)]
    ngr = [int((box[a] - off[a]) / dg[a]) + 1 for a in range(3)]
    assert tuple(ngr) == drawn.ngr
    npool = info.n_world_defg
    local, first = [], {}
    for s in range(ngr[0] * ngr[1] * ngr[2]):
        lucky = int(mt[k] * npool)
        k += 1
        first.setdefault(lucky, len(first))
        local.append(first[lucky])
    assert dom.n_defg == len(first) == drawn.n_defg
    cs = info.cellsize
    g = [np.array([int(((i * cs) + 0.5 * cs - off[a]) / dg[a]) for i in range(n)])
         for a, n in enumerate((info.nx, info.ny, info.nz))]
    gxyz = g[0][None, None, :] + g[1][None, :, None] * ngr[0] + g[2][:, None, None] * (ngr[0] * ngr[1])
    assert np.array_equal(dom.cellgrid, np.array(local, np.uint32)[gxyz.reshape(-1)])
    # the deformed grains' orientations and densities come from the world pool in pick order
    world_of = {v: w for w, v in first.items()}
    for i in range(dom.n_defg):
        o, rho = e.world_defg(world_of[i])
        assert np.array_equal(dom.ori_quat[dom.defg_ori[i]], e.world_quat(o)) and dom.defg_rho[i] == rho
    # nuclei
    N = dom.N
    sites, coll = [], 0
    while len(sites) < info.n_nuclei:
        site = int(mt[k] * N)
        k += 1
        if site in sites:
            coll += 1
            continue
        k += 1                                   # the pool index drawn for the nucleus, unused
        sites.append(site)
    assert list(dom.rxg_site) == sites and drawn.collisions == coll and drawn.nuclei_placed == len(sites)
    qn = ref.shoemake(seed, len(sites))
    assert np.array_equal(dom.ori_quat[dom.rxg_ori].view(np.uint64), qn.view(np.uint64))


def test_seeds_and_domains():
    e = ds.Ensemble.from_uds(DECK)
    a, _, da = e.domain(0)
    b, _, db = e.domain(1)
    assert (da.local_seed, db.local_seed) == (-11, -12)
    assert not np.array_equal(a.rxg_site, b.rxg_site)
    c, _, dc = e.domain(0, seed=-77)
    d, _, dd = e.domain(2, seed=-77)
    assert dc.local_seed == dd.local_seed == -77
    assert np.array_equal(c.cellgrid, d.cellgrid) and np.array_equal(c.rxg_site, d.rxg_site)
    a2, _, _ = e.domain(0)
    assert np.array_equal(a.cellgrid, a2.cellgrid) and np.array_equal(a.ori_quat, a2.ori_quat)


def test_params_from_deck(tmp_path):
    e = ds.Ensemble.from_uds(DECK)
    dom, prm, _ = e.domain(0)
    assert prm.mobility_model == abi.MOB_ROLLETTHOLM
    assert prm.rh_hagb_m0 == 0.1 and prm.rh_hagb_hact_ev == 2.8 and prm.G0 == 6.92e10 and prm.b0 == 2.48e-10
    assert prm.sched_time == (0.0, f32(360000.0)) and prm.sched_temp == (f32(726.85) + 273.15,) * 2
    assert prm.defrag_at == (f32(0.30) * dom.N, f32(0.60) * dom.N)
    assert (prm.xmax, prm.tmax, prm.nmax, prm.maxfill) == (0.99, 360000.0, 100000, 0.1)
    assert dom.cellsize == 0.5 * 1e-6
    # the C view and ParamsSpec.to_c() hold the same doubles (enthalpies times echarge)
    import ctypes as C
    lib = ds.load()
    h = C.c_void_p()
    assert lib.score_setup_domain_create(e._h, 0, ds.DEFAULT_SEED, C.byref(h)) == 0
    d, p = abi.Domain(), abi.Params()
    assert lib.score_setup_domain_view(h, C.byref(d), C.byref(p)) == 0
    q, keep = prm.to_c()
    for f, _ in abi.Params._fields_:
        if f in ("sched_time", "sched_temp", "defrag_at", "disp_time", "disp_fr", "grainlog_at"):
            continue
        assert getattr(p, f) == getattr(q, f), f
    lib.score_setup_domain_free(h)
    # values out of SCORE's recommended ranges are reset, defragmentation can be off
    e2 = ds.Ensemble.from_uds(write(tmp_path, mk.deck_text({("RuntimeControl", "MAXFILLPERSTEP"): "0.5",
                                                            ("RuntimeControl", "InitialCacheSizeRXCells"): "0.9",
                                                            ("RuntimeControl", "OnTheFlyDefragmentation"): "0"})))
    _, p2, _ = e2.domain(0)
    assert p2.maxfill == 0.1 and p2.rel_cell_caching == 0.10 and p2.defrag_at == ()
    # Zener drag from the deck: alpha, gamma and the f/r table (float columns)
    assert prm.zener_mode == 0 and prm.disp_time == ()
    e3 = ds.Ensemble.from_uds(write(tmp_path, mk.deck_text({("DragParameter", "ZenerConsider"): "2"})))
    _, p3, _ = e3.domain(0)
    assert (p3.zener_mode, p3.zener_alpha, p3.zener_gamma) == (2, 1.5, 0.324)
    assert p3.disp_time == (0.0, f32(360000.0)) and p3.disp_fr == (0.0, 0.0)
    # recovery from the deck: RecoveryConsider 1-3 selects a model, any other value
    # none; activation enthalpies in eV times echarge, as for the mobilities
    assert prm.recovery_mode == 0
    for mode, want in (("1", 1), ("2", 2), ("3", 3), ("7", 0)):
        e4 = ds.Ensemble.from_uds(write(tmp_path, mk.deck_text({("RecoveryParameter", "RecoveryConsider"): mode,
                                                                ("RecoveryParameter", "RecoveryVacancyDiffPreexp"): "2.5e-5",
                                                                ("RecoveryParameter", "RecoverySoluteConcentration"): "0.01"})))
        _, p4, _ = e4.domain(0)
        assert p4.recovery_mode == want
        assert (p4.rec_vacancy_d0, p4.rec_vacancy_hact_ev, p4.rec_solute_d0, p4.rec_solute_hact_ev) == (2.5e-5, 2.9, 1.0, 1.6)
        assert (p4.rec_solute_ls, p4.rec_solute_conc, p4.rec_alpha3, p4.rec_kappa2, p4.rec_c3, p4.rec_c5) == \
            (1.0, 0.01, 1.0, 0.1, 1.0, 0.65)
        h = C.c_void_p()
        assert lib.score_setup_domain_create(e4._h, 0, ds.DEFAULT_SEED, C.byref(h)) == 0
        dv, pv = abi.Domain(), abi.Params()
        assert lib.score_setup_domain_view(h, C.byref(dv), C.byref(pv)) == 0
        q, keep = p4.to_c()
        for f in ("recovery_mode", "rec_vacancy_d0", "rec_vacancy_hact", "rec_solute_d0", "rec_solute_hact",
                  "rec_solute_ls", "rec_solute_conc", "rec_alpha3", "rec_kappa2", "rec_c3", "rec_c5"):
            assert getattr(pv, f) == getattr(q, f), f
        lib.score_setup_domain_free(h)
        e4.close()


@pytest.mark.parametrize("override,code", [
    ((("NucleationModeling", "CSRNucleation"), "2"), ds.ERR_UNSUPPORTED),
    ((("NucleationModeling", "GBNucleation"), "1"), ds.ERR_UNSUPPORTED),
    ((("EnsembleDefinition", "DefStructureSynthesis"), "4"), ds.ERR_UNSUPPORTED),
    ((("RuntimeControl", "XMAX"), "1.5"), ds.ERR_RANGE),
    ((("RuntimeControl", "NMAX"), "0"), ds.ERR_RANGE),
    ((("EnsembleDefinition", "3DCAEdgeLengthInCellsND"), "40"), ds.ERR_RANGE),
    ((("EnsembleDefinition", "MeanDefGrainSizeinRD"), "4.0"), ds.ERR_RANGE),
    ((("NucleationModeling", "IncubationTimeModel"), "2"), ds.ERR_RANGE),
    ((("BoundaryMobilityModel", "RHModelHAGBm0"), "0.0"), ds.ERR_RANGE),
    ((("MaterialProperties", "ZeroKelvinShearModulusG0"), "abc"), ds.ERR_FORMAT),
])
def test_rejected_decks(tmp_path, override, code):
    with pytest.raises(ds.SetupError) as e:
        ds.Ensemble.from_uds(write(tmp_path, mk.deck_text(dict([override]))))
    assert e.value.code == code, str(e.value)


def test_bad_pools(tmp_path):
    with pytest.raises(ds.SetupError) as e:
        ds.Ensemble.from_uds(write(tmp_path, mk.deck_text(pool=[("10.0", "20.0", "30.0", "1.0E+10")])))
    assert e.value.code == ds.ERR_RANGE                 # rho at or below the recrystallized value
    with pytest.raises(ds.SetupError) as e:
        ds.Ensemble.from_uds(write(tmp_path, mk.deck_text(pool=[("400.0", "20.0", "30.0", "1.0E+15")])))
    assert e.value.code == ds.ERR_RANGE                 # Bunge angle beyond 360 degrees
    with pytest.raises(ds.SetupError) as e:
        ds.Ensemble.from_uds(write(tmp_path, mk.deck_text(schedule=((0.0, 700.0), (0.0, 700.0)))))
    assert e.value.code == ds.ERR_RANGE                 # time not increasing


def test_deck_domain_grows_on_the_oracle():
    dom, prm, _ = ds.domain_from_uds(DECK, jobid=1)
    o = Oracle(dom, prm)
    r = o.solve_rxgrowth()
    assert r["completed"] and r["X"] > 0.99
    dc, rc = o.counts()
    # a cell on the front still counts for its deformed grain until it completes
    assert int(dc.sum()) + int(rc.sum()) == dom.N
    assert int(rc.sum()) == r["xcells"] and (rc > 0).sum() == dom.n_rxg
    o.close()
