"""Ensemble post-processing (score_amd/postprocess.py) against the reference's own
example/jmak outputs: the rediscretized kinetics and the final grain-size
distribution SCORE computed from that run's grain-evolution log
(tests/golden/jmak_postprocess.json, made by make_postprocess_fixture.py).  The
reference printed its log times with six significant digits, and every
interpolated volume inherits that rounding (about 1e-5 relative), so the
rediscretized values agree to 2e-5; the grain sizes are integers and agree to
the printed precision.  Then the same analysis from a growth run's trajectory."""
import json
import math
import os

import numpy as np

from score_amd import postprocess as pp
from score_amd import synth
from oracle.oracle import Oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIX = json.load(open(os.path.join(ROOT, "tests", "golden", "jmak_postprocess.json")))


def jmak_log():
    lg = FIX["log"]
    nuc = np.asarray(FIX["nucleus_final_counts"], np.float64)
    counts = np.zeros((1 + len(nuc), len(lg["times"])))
    counts[0] = lg["deformed_counts"]
    return pp.GrainLog(N=FIX["N"], ndefg=1, times=np.asarray(lg["times"]), counts=counts,
                       final_counts=np.concatenate([[lg["deformed_counts"][-1]], nuc]), tsimend=FIX["tsimend"])


def test_rediscretized_kinetics_match_reference():
    times = pp.rediscr_times(0.0, FIX["tsimend"], FIX["nslots"])
    vol, vref = pp.kinetics([jmak_log()], times)
    rows = pp.kinetics_rows(times, vol, vref)
    assert vref == FIX["N"]
    for n, t, v, x, lnt, lnx in FIX["kinetics_every_4th"]:
        r = rows[n]
        assert math.isclose(r[1], t, rel_tol=1e-5), (n, r[1], t)
        assert math.isclose(r[2], v, rel_tol=2e-5), (n, r[2], v)
        assert math.isclose(r[3], x, rel_tol=2e-5), (n, r[3], x)
        assert (r[4] is None) == (lnt is None) and (lnt is None or abs(r[4] - lnt) < 2e-5)
        assert (r[5] is None) == (lnx is None)
        if lnx is not None and x < 0.99:
            assert abs(r[5] - lnx) < 2e-3, (n, r[5], lnx)


def test_final_grain_size_distribution_matches_reference():
    g = pp.final_gsd([jmak_log()], FIX["tsimend"])
    h = FIX["gsd_header"]
    assert g["ngrains"] == int(h["NGrains"])
    for k, key in (("vref", "Vref"), ("mean", "Mean"), ("median", "Median"), ("stddev", "StdDev")):
        assert math.isclose(g[k], h[key], rel_tol=5e-6), (k, g[k], h[key])
    ref = FIX["gsd_rows"]
    assert [r[1] for r in g["rows"]] == ref["volume"]
    assert np.allclose([r[4] for r in g["rows"]], ref["ln_v_over_mean"], atol=5e-6)
    assert np.allclose([r[5] for r in g["rows"]], ref["cumsum_percent"], rtol=5e-6)


def test_kinetics_from_a_growth_trajectory(tmp_path):
    dom = synth.make_voronoi_domain(32, 30, 28, n_grains=6, n_nuclei=15, seed=3)
    o = Oracle(dom, synth.ParamsSpec())
    r = o.solve_rxgrowth()
    tr = o.trajectory()
    dc, rc = o.counts()
    xlog = [0.0] + [0.01 * i * dom.N for i in range(1, 101)]
    lg = pp.grain_log_from_trajectory(tr, dom.N, dc, rc, xlog)
    # logged where Xcells first reached each threshold, at the step's start time
    k = 0
    for row in tr:
        if k < len(xlog) and row["xcells"] >= xlog[k]:
            assert lg.times[k] == row["t"] and lg.counts[0, k] == dom.N - row["xcells"]
            k += 1
    assert k == len(lg.times)
    times = pp.rediscr_times(0.0, lg.tsimend, 200)
    vol, vref = pp.kinetics([lg], times)
    x = vol / vref
    assert np.all(np.diff(x) >= -1e-12) and x[-1] == r["xcells"] / dom.N
    # at the log times themselves the interpolation returns the logged volume
    vlog, _ = pp.kinetics([lg], lg.times[1:])
    assert np.allclose(vlog, dom.N - lg.counts[0, 1:len(lg.times)])
    g = pp.final_gsd([lg], lg.tsimend)
    assert g["ngrains"] == dom.n_rxg and g["vref"] == float(rc.sum())
    pp.write_kinetics_csv(str(tmp_path / "k.csv"), pp.kinetics_rows(times, vol, vref), vref)
    pp.write_final_gsd_csv(str(tmp_path / "g.csv"), g)
    lines = open(tmp_path / "k.csv").read().splitlines()
    assert lines[0].startswith("Step;RediscrTime/s;SumRediscrVolume--") and len(lines) == 201
    o.close()


def _qmul(p, q):
    a0, a1, a2, a3 = p
    b0, b1, b2, b3 = q
    return np.array([a0 * b0 - a1 * b1 - a2 * b2 - a3 * b3, a0 * b1 + a1 * b0 + a2 * b3 - a3 * b2,
                     a0 * b2 - a1 * b3 + a2 * b0 + a3 * b1, a0 * b3 + a1 * b2 - a2 * b1 + a3 * b0])


def _qaxis(axis, deg):
    a = np.asarray(axis, float) / np.linalg.norm(axis)
    h = math.radians(deg) / 2
    return np.array([math.cos(h), *(math.sin(h) * a)])


def _bunge_quat(b):
    """Bunge (phi1, Phi, phi2) as the rotation Z(phi1) X(Phi) Z(phi2), from axis-angle products"""
    return _qmul(_qmul(_qaxis([0, 0, 1], math.degrees(b[0])), _qaxis([1, 0, 0], math.degrees(b[1]))),
                 _qaxis([0, 0, 1], math.degrees(b[2])))


def _jmak_ensemble(tmp_path, ideals=None):
    import importlib.util
    from score_amd import domain_setup as ds
    spec = importlib.util.spec_from_file_location("mk", os.path.join(ROOT, "tests", "golden", "make_uds_deck.py"))
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    ideals = ideals or [tuple(c) for c in FIX["ideal_components"]]
    p = tmp_path / "deck.uds"
    p.write_text(mk.deck_text(ideals=ideals))
    return ds.Ensemble.from_uds(str(p))


def test_texture_classes_of_the_reference_run(tmp_path):
    """ens_get_closest_standardlage on the orientations SCORE logged for the jmak run's
    grains (six-digit Bunge angles) gives the class it logged with them"""
    e = _jmak_ensemble(tmp_path)
    ideals = e.ideals()
    assert [round(math.degrees(x), 9) for x in ideals[0][0]] == [60.0, 55.0, 45.0]
    assert math.isclose(ideals[1][2], math.radians(2.0), rel_tol=1e-15)
    for b1, b2, b3, cls in [FIX["deformed_bunge_class"]] + FIX["nuclei_bunge_class"]:
        assert e.closest_ideal(_bunge_quat((b1, b2, b3))) == cls
    # the deck's quaternion for a component is the Bunge rotation (up to sign)
    for b, q, _ in ideals:
        r = _bunge_quat(b)
        assert min(np.abs(q - r).max(), np.abs(q + r).max()) < 1e-15
    e.close()


def test_closest_ideal_rule(tmp_path):
    """within the scatter of a component (cubic symmetry reduced), the closest one; the
    last of equally close ones; RANDOM (0) beyond every scatter"""
    e = _jmak_ensemble(tmp_path, ideals=[("A", 0.0, 0.0, 0.0, 10.0), ("B", 20.0, 0.0, 0.0, 15.0),
                                         ("C", 0.0, 0.0, 0.0, 10.0)])
    z = lambda d: _qaxis([0, 0, 1], d)
    assert e.closest_ideal(z(0.0)) == 3                 # A and C coincide: the later wins
    assert e.closest_ideal(z(4.0)) == 3                 # 4 deg from A/C, 16 from B
    assert e.closest_ideal(z(11.0)) == 2                # beyond A/C's scatter, 9 deg from B
    assert e.closest_ideal(z(90.0 + 4.0)) == 3          # a cube symmetry away
    assert e.closest_ideal(z(36.0)) == 0                # 16 deg from B (scatter 15), 36 from A: RANDOM
    assert e.closest_ideal(_qaxis([1, 1, 1], 25.0)) == 0
    assert e.closest_ideal(_qaxis([1, 0, 0], 9.9)) == 3
    assert e.closest_ideal(_qaxis([1, 0, 0], 10.1)) == 0
    e.close()


def test_rediscretized_macrotexture_matches_reference(tmp_path):
    """postprocess_rediscr_macrotexture on the jmak run's log: the deformed grain in
    class 1, every nucleus in class 2 (the classes SCORE logged), against SCORE's own
    Rediscretized.Macrotexture.csv: the nuclei's class to 2e-5 as the recrystallized
    volume of the kinetics, the deformed class, which is N minus it, to 2e-5 of N"""
    lg = FIX["log"]
    counts = np.array([lg["deformed_counts"], lg["nuclei_sum_counts"]], np.float64)
    log = pp.GrainLog(N=FIX["N"], ndefg=1, times=np.asarray(lg["times"]), counts=counts,
                      final_counts=np.array([lg["deformed_counts"][-1], sum(FIX["nucleus_final_counts"])], float),
                      tsimend=FIX["tsimend"], classes=np.array([1, 2]))
    times = pp.rediscr_times(0.0, FIX["tsimend"], FIX["nslots"])
    vol, vref = pp.macrotexture([log], times, 3)
    assert vref == FIX["N"]
    for row in FIX["macrotexture_every_4th"]:
        n, t, vals = row[0], row[1], row[2:]
        assert math.isclose(times[n], t, rel_tol=1e-5)
        assert vol[n, 0] == vals[0] == 0.0
        assert abs(vol[n, 1] - vals[1]) <= 2e-5 * FIX["N"], (n, vol[n, 1], vals[1])
        assert math.isclose(vol[n, 2], vals[2], rel_tol=2e-5), (n, vol[n, 2], vals[2])
    e = _jmak_ensemble(tmp_path)
    path = tmp_path / "tex.csv"
    pp.write_macrotexture_csv(str(path), times, vol, [b for b, _, _ in e.ideals()], vref)
    lines = path.read_text().splitlines()
    assert lines[0] == FIX["macrotexture_header"] and len(lines) == 1 + FIX["nslots"]
    assert lines[-1].endswith(";") and lines[-1].split(";")[0] == str(FIX["nslots"] - 1)
    e.close()


def test_macrotexture_from_a_per_grain_log():
    """the full per-grain log of a growth run: counts where the trajectory reaches each
    threshold, every column summing to N; the classes' volumes then sum to N at every
    time and the deformed rows give the same kinetics as the aggregated log"""
    dom = synth.make_voronoi_domain(32, 30, 28, n_grains=6, n_nuclei=15, seed=3)
    xlog = [0.0] + [0.05 * i * dom.N for i in range(1, 21)]
    o = Oracle(dom, synth.ParamsSpec(grainlog_at=tuple(xlog)))
    o.solve_rxgrowth()
    classes = np.arange(dom.n_defg + dom.n_rxg) % 3
    lg = pp.grain_log_from_machine(o, dom.N, classes)
    tr = o.trajectory()
    agg = pp.grain_log_from_trajectory(tr, dom.N, *o.counts(), xlog)
    assert np.array_equal(lg.times, agg.times)
    assert np.all(lg.counts.sum(axis=0) == dom.N)
    assert np.array_equal(lg.counts[:dom.n_defg].sum(axis=0), agg.counts[0, :len(agg.times)])
    assert lg.tsimend == o.result()["t"] and np.array_equal(lg.final_counts, np.concatenate(o.counts()))
    times = pp.rediscr_times(0.0, lg.tsimend, 50)
    vol, vref = pp.macrotexture([lg, lg], times, 3)
    assert vref == 2 * dom.N and np.allclose(vol.sum(axis=1), 2 * dom.N, rtol=1e-12)
    kin, _ = pp.kinetics([agg], times)
    defvol = sum(np.array([lg.interp_cell_count(g, t) for t in times]) for g in range(dom.n_defg))
    assert np.allclose(dom.N - defvol, kin, rtol=1e-12)
    o.close()
