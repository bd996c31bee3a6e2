"""Writes tests/golden/jmak_raw_maps.json: the file names and per-file statistics of
the four grain-ID maps the reference ships with its example/jmak run
(SCORE.1.3DRawData.tar: 200^3 uint32 stacks written by
ompcrit_write_voxeldata_coloring_grainids), with the X of their steps from the same
run's SingleGrainData.csv.  Run in a container that has the reference mounted:

    python tests/golden/make_raw_maps_fixture.py /root/reference
"""
import json
import os
import re
import sys
import tarfile

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
JMAK = os.path.join(REF, "example", "jmak")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "jmak_raw_maps.json")

rows = {}
for line in open(os.path.join(JMAK, "SCORE.1.MyRank.0.JobID.0.SingleGrainData.csv")):
    f = line.rstrip("\n").split(";")
    if f[0] in ("Step", "X"):
        rows[f[0]] = [v for v in f[9:] if v != ""]
x_of_step = {int(s): float(x) for s, x in zip(rows["Step"], rows["X"])}

maps = []
with tarfile.open(os.path.join(JMAK, "SCORE.1.3DRawData.tar")) as tf:
    for m in sorted(tf.getmembers(), key=lambda m: m.name):
        name = os.path.basename(m.name)
        g = re.fullmatch(r"SCORE\.(\d+)\.JobID\.(\d+)\.Iter\.(\d+)\.X\.(\d+)\.(\w+)\.raw", name)
        data = np.frombuffer(tf.extractfile(m).read(), dtype="<u4")
        step = int(g.group(3))
        rx = data[data >= 3]
        ids, counts = np.unique(rx, return_counts=True)
        maps.append({
            "name": name, "simid": int(g.group(1)), "jobid": int(g.group(2)), "step": step,
            "postfix": g.group(5), "X": x_of_step[step], "bytes": int(m.size),
            "n_infected": int((data == 1).sum()), "n_deformed": int((data == 2).sum()),
            "n_recrystallized": int(rx.size), "n_grains": int(ids.size),
            "min": int(data.min()), "max": int(data.max()),
            "grain_volumes_sorted": sorted(int(c) for c in counts) if step == max(x_of_step) else None,
        })
json.dump({"source": "example/jmak/SCORE.1.3DRawData.tar of the reference (200^3, 1 deformed grain, 947 nuclei)",
           "nx": 200, "ny": 200, "nz": 200, "maps": maps}, open(OUT, "w"), indent=1)
print("wrote", OUT)
