"""Write tests/golden/voronoi512_seed1.json: the oracle's complete run of the
512^3 workload (8,000 Voronoi grains, 100,000 nuclei, seed 1) reduced to
hashes and the trajectory, so the GPU run at full size can be compared bit for
bit without rerunning the serial oracle (about two minutes of CPU time).

    python tests/golden/make_voronoi512_fixture.py
"""
import hashlib
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle.oracle import Oracle  # noqa: E402
from score_amd import synth  # noqa: E402


def digest(a) -> str:
    return hashlib.sha256(a.tobytes()).hexdigest()


def main():
    t0 = time.time()
    dom = synth.workload("voronoi512", seed=1)
    o = Oracle(dom, synth.ParamsSpec())
    r = o.solve_rxgrowth()
    tr = o.trajectory()
    dc, rc = o.counts()
    tr2 = tr.copy()
    tr2["npending"] = 0
    out = {
        "workload": "voronoi512", "seed": 1,
        "domain_sha256": digest(dom.cellgrid), "sites_sha256": digest(dom.rxg_site),
        "grainmap_sha256": digest(o.grainmap()),
        "trajectory_sha256": digest(tr2),
        "defg_counts_sha256": digest(dc), "rxg_counts_sha256": digest(rc),
        "steps": int(r["step"]), "t_end": r["t"], "X": r["X"], "sum_sv": int(r["sum_sv"]),
        "nuclei_placed": int(r["nuclei_placed"]),
        "oracle_seconds": time.time() - t0,
    }
    with open(os.path.join(HERE, "voronoi512_seed1.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(out)


if __name__ == "__main__":
    main()
