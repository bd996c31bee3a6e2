"""Writes tests/golden/jmak_postprocess.json from the reference's example/jmak run:
the grain-evolution log of its one domain (SingleGrainData.csv: log times and
cell counts of the deformed grain and of every nucleus at each log point), its
end time, and the two rediscretized outputs the reference computed from them
(Rediscretized.Kinetics.csv, Rediscretized.FinalGSD.csv, Rediscretized.Macrotexture.csv),
with the orientation and texture class SCORE logged for every grain and the
deck's ideal components.

    python tests/golden/make_postprocess_fixture.py /root/reference
"""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
J = os.path.join(REF, "example", "jmak")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "jmak_postprocess.json")

lines = [l.rstrip("\n").split(";") for l in open(os.path.join(J, "SCORE.1.MyRank.0.JobID.0.SingleGrainData.csv"))]
rows = {l[0]: l for l in lines[:5]}
first = 9
nlog = len([v for v in rows["Step"][first:] if v != ""])
times = [float(v) for v in rows["Time"][first:first + nlog]]
steps = [int(v) for v in rows["Step"][first:first + nlog]]
X = [float(v) for v in rows["X"][first:first + nlog]]
deformed = [int(v) for v in lines[5][first:first + nlog]]          # grain 0, the only deformed grain
# the nuclei's counts at the last log point (X = 1: their final volumes)
nuclei_final = [int(l[first + nlog - 1]) for l in lines[7:] if l and l[0] != ""]
# the nuclei's summed counts at every log point (all of them share one texture class)
nuc_rows = [l for l in lines[7:] if l and l[0] != ""]
nuclei_sum = [sum(int(l[first + j]) for l in nuc_rows) for j in range(nlog)]
# orientation (Bunge, rad, six digits) and texture class of the deformed grain and of the nuclei
deformed_ori = [float(v) for v in lines[5][5:8]] + [int(lines[5][8])]
nuclei_ori = sorted({tuple([float(v) for v in l[5:8]] + [int(l[8])]) for l in nuc_rows})
ideals = []
in_block = False
for l in open(os.path.join(J, "SCORE.Input.TemplateIron.uds")):
    if l.startswith("||"):
        in_block = "IdealComponents" in l
        continue
    f = l.split()
    if in_block and len(f) == 5:
        ideals.append([f[0].strip('"')] + [float(x) for x in f[1:]])
tex_lines = open(os.path.join(J, "SCORE.1.Rediscretized.Macrotexture.csv")).read().splitlines()
tex = [[int(r.split(";")[0])] + [float(x) for x in r.split(";")[1:] if x != ""] for r in tex_lines[1:]]

kin = []
for l in open(os.path.join(J, "SCORE.1.Rediscretized.Kinetics.csv")).read().splitlines()[1:]:
    f = l.split(";")
    kin.append([int(f[0])] + [float(x) if x != "" else None for x in f[1:6]])
gsd_lines = open(os.path.join(J, "SCORE.1.Rediscretized.FinalGSD.csv")).read().splitlines()
head = dict(kv.split("=") for kv in gsd_lines[0].split(","))
gsd = [[float(x) for x in l.split(";")] for l in gsd_lines[2:] if l]
tsimend = float(head["Time"].rstrip("s"))

json.dump({
    "source": "example/jmak of the reference: SingleGrainData.csv, Rediscretized.Kinetics.csv, Rediscretized.FinalGSD.csv",
    "N": 8000000, "ndefg": 1, "tsimend": tsimend, "nslots": len(kin),
    "log": {"steps": steps, "times": times, "X": X, "deformed_counts": deformed, "nuclei_sum_counts": nuclei_sum},
    "deformed_bunge_class": deformed_ori, "nuclei_bunge_class": [list(t) for t in nuclei_ori],
    "ideal_components": ideals,
    "macrotexture_header": tex_lines[0],
    "macrotexture_every_4th": [r for r in tex if r[0] % 4 == 0 or r[0] == len(tex) - 1],
    "nucleus_final_counts": nuclei_final,
    "kinetics_every_4th": [r for r in kin if r[0] % 4 == 0 or r[0] == len(kin) - 1],
    "gsd_header": {k: float(v.rstrip("s")) for k, v in head.items()},
    "gsd_rows": {"volume": [r[1] for r in gsd], "incubation": [r[2] for r in gsd], "ideal": [int(r[3]) for r in gsd],
                 "ln_v_over_mean": [r[4] for r in gsd], "cumsum_percent": [r[5] for r in gsd]},
}, open(OUT, "w"))
print("wrote", OUT, "log points", nlog, "nuclei", len(nuclei_final), "kinetics rows", len(kin), "gsd rows", len(gsd))
