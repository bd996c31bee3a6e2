#!/usr/bin/env python3
"""Benchmark of the growth machine on one MI355X.

One bench step = one complete growth run (SCORE's solve_RXGROWTH) of the
512^3 Voronoi workload to X = 0.99, with the domain already resident in HBM
when the timed region starts (the handle is created, and the nuclei placed,
before it).  The metric is front cell-updates per second: the sum over all CA
steps of the ACTIVE front cells advanced (SCORE's Sv) divided by the wall time
of the run.  Kernel times come from HIP events recorded by the library on its
own stream; the CPU baseline is the serial oracle timed on a bounded prefix of
the same workload on the host.

    python bench.py [--steps K] [--warmup W] [--gpus 1]

Prints one JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="voronoi512")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="bound of the CPU-baseline sample (oracle steps on the same domain)")
    ap.add_argument("--no-cpu", action="store_true")
    a = ap.parse_args(argv)

    if a.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {a.gpus} is not supported (domains shard as "
              f"independent per-GPU queues, see DESIGN.md)", file=sys.stderr)
        return 2

    from score_amd import synth
    from score_amd.growth import GrowthMachine, device_count, same_trajectory

    if device_count() < 1:
        print("no gfx950 device visible", file=sys.stderr)
        return 3

    t0 = time.time()
    dom = synth.workload(a.workload, seed=a.seed)
    prm = synth.ParamsSpec()
    log(f"workload {a.workload} seed {a.seed}: {dom.nx}x{dom.ny}x{dom.nz}, {dom.n_defg} grains, "
        f"{dom.n_rxg} nuclei, generated in {time.time() - t0:.1f}s")

    runs = []
    prof_tot = None
    digest = None
    for rep in range(a.warmup + a.steps):
        timed = rep >= a.warmup
        gm = GrowthMachine(dom, prm)
        gm.set_profiling(True)
        t1 = time.perf_counter()
        res = gm.solve_rxgrowth()
        t2 = time.perf_counter()
        if not res["completed"]:
            raise RuntimeError("run did not complete")
        pf = gm.profile()
        gmap = gm.grainmap()
        h = hashlib.sha256(gmap.tobytes()).hexdigest()[:16]
        if digest is None:
            digest = h
        elif digest != h:
            raise RuntimeError("grain map differs between repetitions")
        gm.close()
        log(f"{'run' if timed else 'warmup'} {rep}: {res['step']} CA steps, X={res['X']:.6f}, "
            f"sum Sv={res['sum_sv']}, {t2 - t1:.3f}s, sweep {pf['ms_sweep']:.1f}ms infect {pf['ms_infect']:.1f}ms")
        if timed:
            runs.append((t2 - t1, res, pf))
            if prof_tot is None:
                prof_tot = dict(pf)
            else:
                for k, v in pf.items():
                    prof_tot[k] += v

    secs = sum(r[0] for r in runs)
    updates = sum(r[1]["sum_sv"] for r in runs)
    value = updates / secs
    ca_steps = runs[0][1]["step"]
    voxel_steps = dom.N * ca_steps * len(runs) / secs

    # dominant phase and its roofline (algorithmic bytes / event-timed duration)
    phases = {"sweep": (prof_tot["ms_sweep"], prof_tot["sweep_bytes"]),
              "infect": (prof_tot["ms_infect"], prof_tot["infect_bytes"])}
    dom_phase = max(phases, key=lambda k: phases[k][0])
    ms, nbytes = phases[dom_phase]
    launches = prof_tot["steps"]
    achieved = nbytes / (ms * 1e-3) / 1e9
    roofline = {
        "bound": "hbm", "kernel": "k_sweep" if dom_phase == "sweep" else "infection (k_propose..k_retire)",
        "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
        "avg_launch_us": round(ms * 1e3 / launches, 3), "alg_bytes_per_launch": int(nbytes / launches),
        "phase_ms_per_run": {k: round(v[0] / len(runs), 3) for k, v in phases.items()} | {
            "compact": round(prof_tot["ms_compact"] / len(runs), 3), "tail": round(prof_tot["ms_tail"] / len(runs), 3),
            "host_step_loop": round(prof_tot["ms_host"] / len(runs), 3)},
    }

    cpu = None
    if not a.no_cpu:
        from oracle.oracle import Oracle
        o = Oracle(dom, prm)
        k = 0
        sv = 0
        tc0 = time.perf_counter()
        while time.perf_counter() - tc0 < a.cpu_seconds:
            if o.step(1):
                break
            k += 1
        tc = time.perf_counter() - tc0
        tr = o.trajectory()
        sv = int(tr["sv"].sum())
        xr = float(tr["X"][-1]) if len(tr) else 0.0
        # the GPU trajectory must agree with the oracle's over the sampled prefix
        gm = GrowthMachine(dom, prm)
        gm.step(len(tr))
        gt = gm.trajectory()
        prefix_ok = same_trajectory(gt, tr)
        gm.close()
        o.close()
        cpu = {"value": sv / tc, "unit": "cell-updates/s", "cores": 1, "kind": "port",
               "sample": f"serial oracle, first {k} CA steps of the same {a.workload} domain "
                         f"(X={xr:.4f}, sum Sv={sv}) in {tc:.1f}s",
               "gpu_trajectory_matches_prefix": prefix_ok}
        log(f"cpu baseline: {cpu['value']:.3e} cell-updates/s over {k} steps; GPU prefix match {prefix_ok}")

    out = {
        "metric": "cell-updates/s", "value": value, "unit": "cell-updates/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": 1e3 * secs / len(runs), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{a.workload}: {dom.nx}x{dom.ny}x{dom.nz} voxels, dx=0.5um, {dom.n_defg} Voronoi "
                               f"grains, {dom.n_rxg} site-saturated nuclei, Fe/Humphreys, T=1000K, run to X=0.99",
                   "seed": a.seed, "ca_steps_per_run": ca_steps, "sum_sv_per_run": runs[0][1]["sum_sv"],
                   "grainmap_sha256_16": digest, "parallelism": "1 GPU"},
        "voxel_steps_per_s": voxel_steps,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if cpu:
        out["speedup_vs_cpu_baseline"] = value / cpu["value"]
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
