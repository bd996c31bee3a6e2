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
