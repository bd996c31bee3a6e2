"""Write tests/golden/jmak_stats.json from the outputs the reference ships in
example/jmak (a 200^3 run of SCORE: one deformed grain, 947 nuclei, 309 steps).

Only data is extracted: the per-step front statistics (Step;Time;X;Sv;...;
nextSlotNeverActiveRXFront, SCORE.1.MyRank.0.JobID.0.RXFrontStats.csv), the
final grain-size list (SCORE.1.Rediscretized.FinalGSD.csv) and the run summary
(SCORE.1.ProfilingLog.csv).  Run in the container that holds the reference:

    python tests/golden/make_jmak_fixture.py
"""
import json
import os

SRC = "/root/reference/example/jmak"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "jmak_stats.json")


def main():
    steps, times, X, Sv, nslots = [], [], [], [], []
    with open(os.path.join(SRC, "SCORE.1.MyRank.0.JobID.0.RXFrontStats.csv")) as f:
        next(f)
        for line in f:
            c = line.strip().split(";")
            steps.append(int(c[0]))
            times.append(float(c[1]))
            X.append(float(c[2]))
            Sv.append(int(c[3]))
            nslots.append(int(c[7]))
    vols = []
    with open(os.path.join(SRC, "SCORE.1.Rediscretized.FinalGSD.csv")) as f:
        head = next(f).strip()
        next(f)
        for line in f:
            c = line.strip().split(";")
            if len(c) > 1 and c[1]:
                vols.append(int(c[1]))
    summary = dict(kv.split("=") for kv in head.split(","))
    with open(os.path.join(SRC, "SCORE.1.ProfilingLog.csv")) as f:
        keys = next(f).strip().split(";")
        vals = next(f).strip().split(";")
    prof = dict(zip(keys, vals))
    out = {
        "source": "example/jmak of the reference (2015 MPI-only build, legacy input SCORE.Input.TemplateIron.uds)",
        "setup": {"nx": 200, "ny": 200, "nz": 200, "cellsize_um": 1.0, "n_defg": 1, "rho0": 1e15,
                  "n_nuclei": int(summary["NGrains"]), "T_C": 500.0, "RHModelHAGBm0": 300.0, "RHModelHAGBHact_eV": 2.45,
                  "G0": 85.0e9, "b0": 2.5e-10, "maxfill": 0.1, "InitialCacheSizeRXCells": 0.2},
        "steps": steps, "time": times, "X": X, "Sv": Sv, "nextSlotNeverActive": nslots,
        "final_grain_volumes": sorted(vols),
        "gsd_summary": {"mean": float(summary["Mean"]), "median": float(summary["Median"]),
                        "std": float(summary["StdDev"]), "n": int(summary["NGrains"])},
        "tend": float(prof["tend"]), "iterations": int(prof["Iterations"]),
        "mpi_wtime_growth_s": float(prof["MPIWTimeGrowth"]),
    }
    with open(OUT, "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print(f"wrote {OUT}: {len(steps)} steps, {len(vols)} grains")


if __name__ == "__main__":
    main()
