"""Ensemble post-processing of growth runs: rediscretized kinetics, macrotexture
and the final grain-size distribution, as SCORE's solitary-unit analysis
computes them (ensembleHdl::postprocess_*, SCORE_SolitaryUnitAveraging.cpp:51-93,
:546-1063) from each domain's grain-evolution log (caHdl::log_grainevo,
SCORE_OnTheFlyCharacterization.cpp:478-530).

A domain's log holds, at each log point, the time and the cell count of every
grain (deformed grains first, then nuclei).  SCORE logs in the master section of
the first step whose Xcells reaches each output threshold, before t advances
(SCORE_OnTheFlyCharacterization.cpp:646-651).  From the growth machine's
trajectory the same points give the time and Xcells; the deformed grains then
hold N - Xcells cells in all, which is all the kinetics needs, and the nuclei's
final counts give the final grain sizes.  The macrotexture needs every grain's
count at every log point: the growth machine records them itself where SCORE's
master section logs (grain_log_from_machine).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

SMALLEST_TIME_LOGARITMIZE = 1e-8  # SCORE_Defs_Physics.h


@dataclasses.dataclass
class GrainLog:
    """mygrainevolution of one domain: log times, cell counts per grain (rows:
    deformed grains, then nuclei; columns: log points), the counts after the run
    (mydefgpool/myrxgpool cellcount), and the end time."""
    N: int
    ndefg: int
    times: np.ndarray
    counts: np.ndarray
    final_counts: np.ndarray
    tsimend: float
    tincub: np.ndarray | None = None
    ideal: np.ndarray | None = None
    classes: np.ndarray | None = None   # texture class of every row (closestideal; 0 = RANDOM)

    def interp_cell_count(self, gid: int, when: float) -> float:
        """caHdl::get_interpCellCount (SCORE_SolitaryUnitAveraging.cpp:1304-1363)"""
        if when >= self.tsimend:
            return float(self.final_counts[gid])
        n = len(self.times)
        i = 0
        while i < n and self.times[i] < when:
            i += 1
        if 0 < i < n:
            ti1, c1 = self.times[i - 1], float(self.counts[gid, i - 1])
            ti, c = self.times[i], float(self.counts[gid, i])
            return c1 + ((when - ti1) * ((c - c1) / (ti - ti1)))
        if i == 0:
            return float(self.counts[gid, 0])
        return float(self.counts[gid, n - 1])


def grain_log_from_trajectory(traj: np.ndarray, N: int, defg_counts, rxg_counts, xlog) -> GrainLog:
    """A domain's log from the growth machine's trajectory: log points at the
    first step whose Xcells reaches each threshold of `xlog` (in cells, ascending;
    one per step), recorded with the step's start time; one aggregated deformed
    row of N - Xcells cells, nucleus rows unknown until the end (their final
    counts; enough for the kinetics and the final grain sizes)."""
    defg_counts = np.asarray(defg_counts, np.float64)
    rxg_counts = np.asarray(rxg_counts, np.float64)
    times, defrow = [], []
    k = 0
    for r in traj:
        if k < len(xlog) and r["xcells"] >= xlog[k]:
            times.append(float(r["t"]))
            defrow.append(float(N) - float(r["xcells"]))
            k += 1
    nlog = len(times)
    counts = np.zeros((1 + len(rxg_counts), max(nlog, 1)))
    counts[0, :nlog] = defrow
    final = np.concatenate([[defg_counts.sum()], rxg_counts])
    tend = float(traj["t"][-1] + traj["dt"][-1]) if len(traj) else 0.0
    return GrainLog(N=N, ndefg=1, times=np.asarray(times), counts=counts, final_counts=final, tsimend=tend)


def grain_log_from_machine(machine, N: int, classes=None) -> GrainLog:
    """A domain's full log (mygrainevolution) as the growth machine (or the
    oracle) recorded it during its run: at the first step whose Xcells reached
    each log point (score_gm_params.grainlog_at, one per step) the cell count of
    every grain, deformed grains first, with the step's start time, as SCORE's
    master section logs before t advances (log_OUTPUT_Diehl,
    SCORE_OnTheFlyCharacterization.cpp:646-651)."""
    rows, counts = machine.grainlog()
    dg, rg = machine.counts()
    final = np.concatenate([dg, rg]).astype(np.float64)
    cols = counts.T.astype(np.float64) if len(rows) else np.zeros((len(final), 1))
    return GrainLog(N=N, ndefg=len(dg), times=np.asarray(rows["t"], np.float64), counts=cols, final_counts=final,
                    tsimend=float(machine.result()["t"]),
                    classes=None if classes is None else np.asarray(classes, np.int64))


def rediscr_times(tmin: float, tmax: float, nslots: int) -> np.ndarray:
    """postprocess_initrediscrtimes: equidistant slots, each the end of its interval"""
    dt = (tmax - tmin) / float(nslots - 1)
    return np.array([tmin + (float(n + 1) * dt) for n in range(nslots)])


def kinetics(logs, times) -> tuple[np.ndarray, float]:
    """postprocess_rediscr_kinetics: recrystallized volume of the ensemble at each
    time (box volume minus the interpolated deformed volume, summed over domains)
    and the total box volume"""
    vol = np.zeros(len(times))
    vref = 0.0
    for lg in logs:
        vref += float(lg.N)
        dv = np.zeros(len(times))
        for dg in range(lg.ndefg):
            for n, t in enumerate(times):
                dv[n] = dv[n] + lg.interp_cell_count(dg, t)
        vol += float(lg.N) - dv
    return vol, vref


def kinetics_rows(times, vol, vref):
    """rows (n, t, volume, X, ln t or None, ln(-ln(1-X)) or None)"""
    rows = []
    for n, (t, v) in enumerate(zip(times, vol)):
        x = v * (1.0 / vref)
        lnt = math.log(t) if t > SMALLEST_TIME_LOGARITMIZE else None
        lnx = math.log(-1.0 * math.log(1.0 - x)) if (1.0 - x) > SMALLEST_TIME_LOGARITMIZE else None
        rows.append((n, float(t), float(v), x, lnt, lnx))
    return rows


def write_kinetics_csv(path: str, rows, vref: float) -> None:
    with open(path, "w") as f:
        f.write(f"Step;RediscrTime/s;SumRediscrVolume--{vref:g}--;X;ln(t);ln(-ln(1-X))\n")
        for n, t, v, x, lnt, lnx in rows:
            f.write(f"{n};{t:g};{v:g};{x:g};" + ("" if lnt is None else f"{lnt:g}") + ";" +
                    ("" if lnx is None else f"{lnx:g}") + "\n")


def macrotexture(logs, times, nclasses: int) -> tuple[np.ndarray, float]:
    """postprocess_rediscr_macrotexture (SCORE_SolitaryUnitAveraging.cpp:673-806):
    the interpolated volume of every grain (deformed, then nuclei) summed into its
    texture class at each time, per domain and then over the ensemble; rows are
    times, columns classes (0 = RANDOM).  Also the total box volume."""
    ens = np.zeros((len(times), nclasses))
    vref = 0.0
    for lg in logs:
        if lg.classes is None:
            raise ValueError("the log has no texture classes")
        vref = vref + float(lg.N)
        dom = np.zeros((len(times), nclasses))
        for g in range(len(lg.final_counts)):
            c = int(lg.classes[g])
            for n, t in enumerate(times):
                dom[n, c] += lg.interp_cell_count(g, t)
        ens += dom
    return ens, vref


def write_macrotexture_csv(path: str, times, vol, ideals_bunge, vref: float) -> None:
    """SCORE.<simid>.Rediscretized.Macrotexture.csv: `ideals_bunge` holds each ideal
    component's Bunge angles in rad, printed in degrees (RAD2DEG, SCORE_Math.h:58)"""
    deg = lambda r: (r) / (3.1415926535897932384626433832795) * (180.0)
    with open(path, "w") as f:
        f.write("Step;RediscrTime/s;RANDOM-phi1-PHI-phi2")
        for i, b in enumerate(ideals_bunge):
            f.write(f";{i}-{deg(b[0]):g}-{deg(b[1]):g}-{deg(b[2]):g}")
        f.write(f";SumRediscrVolume={vref:g}\n")
        for n, t in enumerate(times):
            f.write(f"{n};{float(t):g}" + "".join(f";{float(v):g}" for v in vol[n]) + ";\n")


def final_gsd(logs, now: float) -> dict:
    """postprocess_rediscr_finalgrainsizedistribution (site-saturated nuclei): the
    nuclei's volumes at `now`, sorted ascending, with the statistics SCORE prints"""
    g = []
    for rank, lg in enumerate(logs):
        nrxg = len(lg.final_counts) - lg.ndefg
        for r in range(nrxg):
            g.append((lg.interp_cell_count(lg.ndefg + r, now),
                      0.0 if lg.tincub is None else float(lg.tincub[r]),
                      0 if lg.ideal is None else int(lg.ideal[r]), 0))
    g.sort(key=lambda e: e[0])
    vols = np.array([e[0] for e in g], np.float64)
    total = float(sum(e[0] for e in g))
    n = len(g)
    mean = total / float(n)
    median = -1.0
    if n > 3:
        m = n // 2
        median = vols[m] if n % 2 == 1 else 0.5 * (vols[m] + vols[m - 1])
    ssum = 0.0
    for v in vols:
        ssum = ssum + (v - mean) * (v - mean)
    std = math.pow(ssum / float(n), 0.5)
    rows, cum = [], 0.0
    for vol, tinc, ideal, rank in g:
        cum = cum + vol
        norm = vol / mean
        prob = cum / total * 100.0
        rows.append((rank, vol, tinc, ideal, math.log(norm) if norm > SMALLEST_TIME_LOGARITMIZE else None, prob))
    return {"time": now, "ngrains": n, "vref": total, "mean": mean, "median": float(median), "stddev": std,
            "min": float(vols.min()), "max": float(vols.max()), "rows": rows}


def write_final_gsd_csv(path: str, gsd: dict) -> None:
    with open(path, "w") as f:
        f.write(f"Time={gsd['time']:g}s,NGrains={gsd['ngrains']},Vref={gsd['vref']:g},Mean={gsd['mean']:g},"
                f"Median={gsd['median']:g},StdDev={gsd['stddev']:g},Min={gsd['min']:g},Max={gsd['max']:g}\n")
        f.write("Rank;Volume/Cells;IncubationTime/s;Ideal(RANDOMis0);ln(Vi/Vav);CumSumi/Vref*100%\n")
        for rank, vol, tinc, ideal, lnv, prob in gsd["rows"]:
            if lnv is None:
                f.write(f"{rank};{vol:g};{tinc:g};{ideal};;\n")
            else:
                f.write(f"{rank};{vol:.12g};{tinc:.12g};{ideal};{lnv:.12g};{prob:.12g}\n")
