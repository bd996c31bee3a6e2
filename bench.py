#!/usr/bin/env python3
"""Benchmark of the growth machine on one MI355X.

One bench step = one complete growth run (SCORE's solve_RXGROWTH) of the
512^3 Voronoi workload to X = 0.99, with the domain already resident in HBM
when the timed region starts (the handle is created, and the nuclei placed,
before it).  The metric is front cell-updates per second: the sum over all CA
steps of the ACTIVE front cells advanced (SCORE's Sv) divided by the wall time
of the run.

The timed runs record no events.  One more, untimed run records HIP events
between the kernels of every step on the library's own stream and the work
counters of each step; the roofline of the dominant kernel comes from it.  The
CPU baseline is the serial oracle timed on one complete run of the same
workload at half the edge.  The same line carries the drop-in call's parts
(create, run, outputs), the batched ensemble (512 domains of edges 64-192,
batched against sequential plain runs) and the served stream (Poisson
arrivals, edges 64-320, p50/p99 time to X = 0.99).

    python bench.py [--steps K] [--warmup W] [--gpus 1]
    python bench.py --copy-calibration      # a plain 1 GiB device copy, for PMC scaling
    python bench.py --batch 512 --inflight 12 # ensemble of independent domains, domains/hour
    python bench.py --serve 64 --inflight 12  # Poisson stream of requests, p50/p99 time to X=0.99

Prints one JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import dataclasses
import hashlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_voronoi512.json")


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def kernel_bytes(name: str, pf: dict) -> tuple[int, str]:
    """Algorithmic HBM bytes of one kernel over the profiled run, and the per-unit
    figures used.  Infection kernels use the reference algorithm's per-unit bytes
    (a completed cell reads its 26 neighbours' cell-grid entries, 104 B, and its
    record, 30 B; a valid attempt is a 4 B atomic; a new record is 30 B plus its
    4 B cell-grid write; retirement writes 4 B of grid and 8 B of counts).  Sweep
    kernels use what this design's sweep must touch: a 2 B completion-step code
    per occupied slot, 16 B of masks per 64-slot tile, 4 B per list entry, and on
    the steps whose scalars changed 100 B per re-anchored cell (the 64 B record
    read, its 32 B first sector written, its 2 B code read and written).  A
    retired cell's grid entry is not written during the run: its (cell, grain)
    pair goes to the retirement log with the step's parent data (8 B)."""
    full, att, new = pf["n_full"], pf["n_attempts"], pf["n_new"]
    scanned, tiles = pf["slots_scanned"], (pf["slots_scanned"] + 63) // 64
    table = {
        "k_reanchor": (100 * pf["n_reanchored"],
                       "100 B per re-anchored cell (record 64 read, 32 written, code 2 read and 2 written)"),
        "k_sweep": (2 * scanned + 16 * tiles, "2 B per occupied slot + 16 B per tile"),
        "k_compact": (16 * tiles + 4 * full, "16 B per tile + 4 B per FullRX entry"),
        "k_recycle": (4 * pf["n_recycled"], "4 B per recycled slot handed out"),
        "k_propose": (228 * full + 4 * att, "228 B per completed cell (26 cell-grid entries 104, record 64 read, "
                                           "parent data 48, attempt mask 4 and retirement-log entry 8 written) "
                                           "+ 4 B per valid attempt"),
        "k_decide": (28 * full + 4 * att, "28 B per completed cell (position 16, attempt mask 4, winner masks 8) "
                                          "+ 4 B per valid attempt (claim)"),
        "k_commit": (34 * new + 38 * full, "34 B per new record + 38 B per completed cell (record 30, counts 8)"),
    }
    return table.get(name, (0, "scan of per-block counts"))


def copy_calibration() -> int:
    import torch
    n = 1 << 30
    a = torch.ones(n // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n, "copy_bytes_written": n, "dispatches": 5}), flush=True)
    return 0


def stream_reference(nbytes: int) -> dict:
    """What one launch that streams nbytes reaches on this GPU: torch's float32 sum
    over a buffer of that size (reads it once), timed with events on torch's
    stream, as the yardstick of a pass that reads every occupied slot's code.
    It runs in a child process: torch brings its own HIP runtime, which is not
    mixed with the growth library's in one process."""
    import subprocess
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--stream-reference-bytes", str(int(nbytes))],
                       capture_output=True, text=True, timeout=300)
    if r.returncode != 0:
        raise RuntimeError(f"streaming yardstick failed: {r.stderr[-2000:]}")
    return json.loads(r.stdout.strip().splitlines()[-1])


def stream_reference_run(nbytes: int, reps: int = 50) -> dict:
    import torch
    a = torch.ones(max(nbytes // 4, 1), dtype=torch.float32, device="cuda")
    for _ in range(5):
        torch.sum(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        torch.sum(a)
    e1.record()
    torch.cuda.synchronize()
    us = e0.elapsed_time(e1) * 1e3 / reps
    del a
    return {"what": "torch.sum over a float32 buffer of the same bytes, one call per launch",
            "bytes": int(nbytes), "us": round(us, 2), "GBs": round(nbytes / (us * 1e-6) / 1e9, 1)}


def ensemble_domains(seed: int, n: int):
    """The batched-ensemble workload: n domains, edges U{64..192}, edge^3/2000
    grains, nucleus density log-uniform in [1e-5, 1e-3] per voxel, T uniform in
    [900, 1100] K, each with its own seed, run to X = 0.99."""
    import numpy as np
    from score_amd import synth
    rng = np.random.default_rng(seed)
    doms, prms = [], []
    for i in range(n):
        e = int(rng.integers(64, 193))
        dens = 10.0 ** rng.uniform(-5.0, -3.0)
        T = float(rng.uniform(900.0, 1100.0))
        doms.append(synth.make_voronoi_domain(e, e, e, n_grains=max(1, e ** 3 // 2000),
                                              n_nuclei=max(1, int(round(dens * e ** 3))), seed=seed * 100003 + i))
        prms.append(synth.ParamsSpec(sched_temp=(T, T)))
    return doms, prms


def served_domains(seed: int, n: int):
    """The served-stream workload: n requests, edges log-uniform in [64, 320],
    grain and nucleus densities of the 512^3 workload, T uniform in [900, 1100] K."""
    import numpy as np
    from score_amd import synth
    rng = np.random.default_rng(seed + 7)
    doms, prms = [], []
    for i in range(n):
        e = int(round(np.exp(rng.uniform(np.log(64.0), np.log(320.0)))))
        vol = e ** 3
        T = float(rng.uniform(900.0, 1100.0))
        doms.append(synth.make_voronoi_domain(e, e, e, n_grains=max(1, round(vol * 8000 / 512 ** 3)),
                                              n_nuclei=max(1, round(vol * 100000 / 512 ** 3)), seed=seed * 7919 + i))
        prms.append(synth.ParamsSpec(sched_temp=(T, T)))
    return doms, prms


def ensemble_section(seed: int, n: int, inflight: int) -> dict:
    """Batched ensemble (the rank's loop over its domains in SCORE's
    SIMULATE_myCAs, SCORE_Kernel.cpp:1693-1757): all n domains through
    score_gm_batch_run with `inflight` in flight, against the same n domains as
    sequential plain drop-in runs (one at a time); both give the same results."""
    import numpy as np
    from score_amd.growth import batch_run
    t0 = time.time()
    doms, prms = ensemble_domains(seed, n)
    voxels = sum(d.N for d in doms)
    log(f"ensemble of {n} domains, {voxels} voxels, generated in {time.time() - t0:.1f}s")
    batch_run(doms[:2], prms[:2], max_inflight=2)  # warm-up
    timings, outs = {}, {}
    for k in (1, inflight):
        t1 = time.perf_counter()
        outs[k] = batch_run(doms, prms, max_inflight=k)
        timings[k] = time.perf_counter() - t1
        log(f"ensemble, {k} in flight: {timings[k]:.2f}s, {n / timings[k] * 3600:.0f} domains/hour")
    same = all(r1["result"]["xcells"] == r2["result"]["xcells"] and r1["result"]["t"] == r2["result"]["t"]
               and np.array_equal(r1["defg_counts"], r2["defg_counts"])
               and np.array_equal(r1["rxg_counts"], r2["rxg_counts"])
               for r1, r2 in zip(outs[1], outs[inflight]))
    if not same:
        raise RuntimeError("batched results differ from one-at-a-time results")
    secs = np.array([r["seconds"] for r in outs[inflight]])
    lat = np.array([r["queued_seconds"] + r["seconds"] for r in outs[inflight]])
    sv = sum(r["result"]["sum_sv"] for r in outs[inflight])
    return {
        "metric": "domains/hour", "value": n / timings[inflight] * 3600.0, "unit": "domains/hour",
        "sequential_domains_per_hour": n / timings[1] * 3600.0,
        "speedup_vs_sequential": timings[1] / timings[inflight],
        "wall_s": timings[inflight], "sequential_wall_s": timings[1],
        "cell_updates_per_s": sv / timings[inflight],
        "domain_seconds_p50": float(np.percentile(secs, 50)), "domain_seconds_p99": float(np.percentile(secs, 99)),
        "time_to_x099_p50_s": float(np.percentile(lat, 50)), "time_to_x099_p99_s": float(np.percentile(lat, 99)),
        "results_equal_to_sequential": True,
        "workload": f"{n} Voronoi domains, edges U{{64..192}}, edge^3/2000 grains, nuclei log-U[1e-5,1e-3]/voxel, "
                    f"T~U[900,1100] K, to X=0.99, seed {seed}",
        "voxels": voxels, "max_inflight": inflight,
    }


def served_section(seed: int, n: int, inflight: int, rate: float = 0.0) -> dict:
    """Served stream through the request loop (score_gm_server_*): requests
    arrive as a Poisson process at `rate` per second, or at 80% of the rate the
    same requests reach when all are queued at once; latency is arrival to the
    X = 0.99 result.  The same arrivals are served one at a time for comparison,
    and every result equals its one-at-a-time run."""
    import numpy as np
    from score_amd.growth import Server, batch_run
    rng = np.random.default_rng(seed + 11)
    t0 = time.time()
    doms, prms = served_domains(seed, n)
    voxels = sum(d.N for d in doms)
    log(f"served stream of {n} requests, {voxels} voxels, generated in {time.time() - t0:.1f}s")
    batch_run(doms[:2], prms[:2], max_inflight=2)  # warm-up
    if not rate:
        t1 = time.perf_counter()
        batch_run(doms, prms, max_inflight=inflight)
        cap = n / (time.perf_counter() - t1)
        rate = 0.8 * cap
        log(f"capacity with {inflight} in flight: {cap:.1f} requests/s; arrival rate {rate:.1f}/s")
    gaps = rng.exponential(1.0 / rate, n)
    arrivals = np.cumsum(gaps) - gaps[0]

    def serve(k):
        with Server(max_inflight=k) as srv:
            reqs = []
            t1 = time.perf_counter()
            for i in range(n):
                dt = arrivals[i] - (time.perf_counter() - t1)
                if dt > 0:
                    time.sleep(dt)
                reqs.append(srv.submit(doms[i], prms[i]))
            for r in reqs:
                srv.wait(r)
            span = time.perf_counter() - t1
        if any(r.item.status != 0 for r in reqs):
            raise RuntimeError("a served request failed")
        lat = np.array([r.item.queued_seconds + r.item.seconds for r in reqs])
        return span, lat, [r.outcome() for r in reqs]

    span, lat, outs = serve(inflight)
    log(f"served with {inflight} in flight: {span:.2f}s, p50 {np.percentile(lat, 50):.3f}s, "
        f"p99 {np.percentile(lat, 99):.3f}s")
    span1, lat1, outs1 = serve(1)
    log(f"served one at a time: {span1:.2f}s, p50 {np.percentile(lat1, 50):.3f}s, p99 {np.percentile(lat1, 99):.3f}s")
    same = all(x["result"]["xcells"] == y["result"]["xcells"] and x["result"]["t"] == y["result"]["t"] and
               np.array_equal(x["defg_counts"], y["defg_counts"]) for x, y in zip(outs, outs1))
    if not same:
        raise RuntimeError("served results differ between the mix and one at a time")
    return {
        "metric": "domains/hour", "value": n / span * 3600.0, "unit": "domains/hour",
        "time_to_x099_p50_s": float(np.percentile(lat, 50)), "time_to_x099_p99_s": float(np.percentile(lat, 99)),
        "arrival_rate_per_s": rate, "wall_s": span,
        "one_at_a_time": {"domains_per_hour": n / span1 * 3600.0,
                          "time_to_x099_p50_s": float(np.percentile(lat1, 50)),
                          "time_to_x099_p99_s": float(np.percentile(lat1, 99))},
        "results_equal_alone_and_in_the_mix": True,
        "workload": f"{n} Voronoi domains arriving as a Poisson stream, edges log-U[64,320], grain and nucleus "
                    f"densities of the 512^3 workload, T~U[900,1100] K, to X=0.99, seed {seed}",
        "voxels": voxels, "max_inflight": inflight,
    }


ROLLED_SHEET = dict(nx=1024, ny=1024, nz=256, n_grains=200000, n_nuclei=1000000, arrival_steps=70, gb_bias=3.0)
ROLLED_SHEET_PARAMS = dict(recovery_mode=4, rec_tau=100.0, rh_hagb_m0=100.0)


def rolled_sheet_section(seed: int, reps: int = 1) -> dict:
    """The large rolled-sheet domain on one GPU: 1024x1024x256 cells, 200,000
    pancake grains (Voronoi metric stretched 4:1:0.25), rho log-normal (median
    5e14, sigma 0.5), 10^6 nuclei at sites biased 3:1 toward grain boundaries,
    half site-saturated and half arriving at a constant rate over steps 1-70
    (about the first 20% of the run's steps), rho recovering as exp(-t/100 s).
    The Rollett-Holm prefactor is 100 m^4/(J s) at 1000 K, so that the growth
    outruns the recovery (with the 512^3 workload's 0.1 the first time step is
    17 s and rho is gone within a few steps).  One untimed run, then `reps`
    timed ones; the grain map must repeat."""
    from score_amd import synth
    from score_amd.growth import GrowthMachine, footprint
    t0 = time.time()
    dom = synth.make_rolled_sheet_domain(seed=seed, **ROLLED_SHEET)
    prm = synth.ParamsSpec(**ROLLED_SHEET_PARAMS)
    log(f"rolled sheet {dom.nx}x{dom.ny}x{dom.nz}: {dom.n_defg} grains, {dom.n_rxg} nuclei, "
        f"generated in {time.time() - t0:.1f}s")
    out, digest = [], None
    for rep in range(reps + 1):
        tc0 = time.perf_counter()
        gm = GrowthMachine(dom, prm)
        tc1 = time.perf_counter()
        res = gm.solve_rxgrowth()
        t2 = time.perf_counter()
        h = hashlib.sha256(gm.grainmap().tobytes()).hexdigest()[:16]
        tr = gm.trajectory()
        gm.close()
        if digest is None:
            digest = h
        elif digest != h:
            raise RuntimeError("rolled-sheet grain map differs between repetitions")
        log(f"rolled sheet {'run' if rep else 'warmup'}: {res['step']} CA steps, X={res['X']:.4f}, "
            f"sum Sv={res['sum_sv']}, create {tc1 - tc0:.2f}s, run {t2 - tc1:.3f}s")
        if rep:
            out.append((t2 - tc1, tc1 - tc0, res, tr))
    secs = sum(o[0] for o in out) / len(out)
    res, tr = out[0][2], out[0][3]
    return {
        "metric": "cell-updates/s", "value": res["sum_sv"] / secs, "unit": "cell-updates/s",
        "run_s": secs, "create_s": sum(o[1] for o in out) / len(out),
        "ca_steps": res["step"], "X": res["X"], "sum_sv": res["sum_sv"], "t_end_s": res["t"],
        "nuclei_placed": res["nuclei_placed"], "nuclei_arrived_after_step0": int(tr["nplaced"].sum()),
        "windows": res["windows"], "device_bytes_upper": footprint(dom, prm), "grainmap_sha256_16": digest,
        "workload": "rolled sheet 1024x1024x256, 200000 pancake grains (4:1:0.25), rho log-normal(ln 5e14, 0.5), "
                    "1e6 nuclei (half site-saturated, half over steps 1-70, sites 3x biased to grain boundaries), "
                    "exponential recovery tau=100 s, Fe/Rollett-Holm m0=100, T=1000K, to X=0.99",
    }


def regions_section(dom, prm, T: int, reps: int = 1) -> dict:
    """The headline workload with SCORE's OMP_NUM_THREADS = T semantics (T memory
    regions, halo precedence): one untimed run, then `reps` timed ones; the grain
    map must repeat.  It differs from the one-region map, as SCORE's does."""
    from score_amd.growth import GrowthMachine
    prmT = dataclasses.replace(prm, n_regions=T)
    out, digest = [], None
    for rep in range(reps + 1):
        gm = GrowthMachine(dom, prmT)
        t1 = time.perf_counter()
        res = gm.solve_rxgrowth()
        t2 = time.perf_counter()
        h = hashlib.sha256(gm.grainmap().tobytes()).hexdigest()[:16]
        gm.close()
        if digest is None:
            digest = h
        elif digest != h:
            raise RuntimeError(f"grain map at {T} regions differs between repetitions")
        log(f"{T} regions {'run' if rep else 'warmup'}: {res['step']} CA steps, X={res['X']:.6f}, "
            f"sum Sv={res['sum_sv']}, {t2 - t1:.3f}s")
        if rep:
            out.append((t2 - t1, res))
    secs = sum(o[0] for o in out) / len(out)
    res = out[0][1]
    return {"metric": "cell-updates/s", "value": res["sum_sv"] / secs, "unit": "cell-updates/s", "n_regions": T,
            "run_s": secs, "ca_steps": res["step"], "X": res["X"], "sum_sv": res["sum_sv"],
            "windows": res["windows"], "grainmap_sha256_16": digest}


def batch_bench(a) -> int:
    from score_amd.growth import device_count
    if device_count() < 1:
        print("no gfx950 device visible", file=sys.stderr)
        return 3
    out = ensemble_section(a.seed, a.batch, a.inflight)
    out.update({"n_gpus": 1, "higher_is_better": True, "dtype": "f64", "data": "synthetic"})
    print(json.dumps(out), flush=True)
    return 0


def serve_bench(a) -> int:
    from score_amd.growth import device_count
    if device_count() < 1:
        print("no gfx950 device visible", file=sys.stderr)
        return 3
    out = served_section(a.seed, a.serve, a.inflight, a.rate)
    out.update({"n_gpus": 1, "higher_is_better": True, "dtype": "f64", "data": "synthetic"})
    print(json.dumps(out), flush=True)
    return 0


def host_cpu() -> dict:
    """The host's core count and CPU model, for the CPU baseline."""
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)), "cpu_model": model}


def cpu_full_run(name: str, seed: int, scale: float | None = None) -> dict:
    """A complete serial oracle run (to X = 0.99) of a named workload."""
    from oracle.oracle import Oracle
    from score_amd import synth
    dom = synth.workload(name, seed=seed, scale=scale)
    o = Oracle(dom, synth.ParamsSpec())
    t0 = time.perf_counter()
    r = o.solve_rxgrowth()
    secs = time.perf_counter() - t0
    o.close()
    return {"workload": f"{name}" + (f" scaled to {dom.nx}^3 ({dom.n_defg} grains, {dom.n_rxg} nuclei)" if scale else ""),
            "seconds": secs, "steps": r["step"], "sum_sv": r["sum_sv"], "cell_updates_per_s": r["sum_sv"] / secs}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="voronoi512")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=5.0,
                    help="bound of the CPU-baseline sample (oracle steps on the same domain)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-profile", action="store_true", help="skip the profiled run (no roofline)")
    ap.add_argument("--profile-run-only", action="store_true",
                    help="only the profiled run (one step per launch window, events between kernels); for rocprofv3")
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--stream-reference-bytes", type=int, default=0,
                    help="only the streaming yardstick (torch sum over this many bytes), as one JSON line")
    ap.add_argument("--profile-regions", type=int, default=0,
                    help="with --profile-run-only: profile the run with this many memory regions")
    ap.add_argument("--batch", type=int, default=0,
                    help="ensemble mode: this many independent domains, domains/hour batched vs one at a time")
    ap.add_argument("--inflight", type=int, default=12, help="domains in flight in ensemble and served modes")
    ap.add_argument("--serve", type=int, default=0,
                    help="served mode: this many requests arriving as a Poisson stream; p50/p99 time to X=0.99")
    ap.add_argument("--rate", type=float, default=0.0, help="arrival rate of the served mode, requests/s")
    ap.add_argument("--ensemble-domains", type=int, default=512,
                    help="domains of the batched-ensemble part of the default line (0: leave it out)")
    ap.add_argument("--served-requests", type=int, default=64,
                    help="requests of the served-stream part of the default line (0: leave it out)")
    ap.add_argument("--no-rolled-sheet", action="store_true",
                    help="leave the 1024x1024x256 rolled-sheet part out of the default line")
    ap.add_argument("--rolled-sheet", action="store_true", help="only the rolled-sheet domain")
    ap.add_argument("--regions", type=int, default=8,
                    help="also run the workload with SCORE's OMP_NUM_THREADS semantics at this many regions (0: skip)")
    a = ap.parse_args(argv)

    if a.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {a.gpus} is not supported: a domain runs on one GPU, "
              f"and an ensemble of domains shards as one independent queue per GPU", file=sys.stderr)
        return 2
    if a.copy_calibration:
        return copy_calibration()
    if a.stream_reference_bytes:
        print(json.dumps(stream_reference_run(a.stream_reference_bytes)), flush=True)
        return 0
    if a.batch:
        return batch_bench(a)
    if a.serve:
        return serve_bench(a)
    if a.rolled_sheet:
        from score_amd.growth import device_count as _dc
        if _dc() < 1:
            print("no gfx950 device visible", file=sys.stderr)
            return 3
        out = rolled_sheet_section(a.seed, reps=max(1, a.steps))
        out.update({"n_gpus": 1, "higher_is_better": True, "dtype": "f64", "data": "synthetic"})
        print(json.dumps(out), flush=True)
        return 0

    from score_amd import abi, synth
    from score_amd.growth import GrowthMachine, device_count, same_trajectory

    if device_count() < 1:
        print("no gfx950 device visible", file=sys.stderr)
        return 3

    t0 = time.time()
    dom = synth.workload(a.workload, seed=a.seed)
    prm = synth.ParamsSpec()
    log(f"workload {a.workload} seed {a.seed}: {dom.nx}x{dom.ny}x{dom.nz}, {dom.n_defg} grains, "
        f"{dom.n_rxg} nuclei, generated in {time.time() - t0:.1f}s")

    drop = []  # (create, run, outputs) seconds of the timed repetitions

    def one_run(profile: bool):
        """create (upload, nucleation), run, then read back what SCORE keeps
        (counts and grain map), then destroy: the call SCORE's adapter makes"""
        t0 = time.perf_counter()
        gm = GrowthMachine(dom, prm)
        tc = time.perf_counter()
        if profile:
            gm.set_profiling(True)
        t1 = time.perf_counter()
        res = gm.solve_rxgrowth()
        t2 = time.perf_counter()
        if not res["completed"]:
            raise RuntimeError("run did not complete")
        pf = gm.profile() if profile else None
        gm.counts()
        gmap = gm.grainmap(abi.MAP_RAW2GID)
        gm.close()
        t3 = time.perf_counter()
        h = hashlib.sha256(gmap.tobytes()).hexdigest()[:16]
        if not profile:
            drop.append((tc - t0, t2 - t1, t3 - t2))
        return t2 - t1, res, pf, h

    if a.profile_run_only:
        if a.profile_regions > 1:
            prm = dataclasses.replace(prm, n_regions=a.profile_regions)
        psecs, pres, pf, h = one_run(True)
        log(f"profiled run: {psecs:.3f}s, {pf['steps']} steps, grain map {h}; per-kernel ms "
            f"{json.dumps({k: round(v, 3) for k, v in pf['ms_kernel'].items() if v > 0.0})}")
        ref = stream_reference(int(kernel_bytes("k_sweep", pf)[0] / pf["steps"]))
        log(f"streaming yardstick over the sweep's bytes: {json.dumps(ref)}")
        return 0

    runs = []
    digest = None
    for rep in range(a.warmup + a.steps):
        timed = rep >= a.warmup
        secs, res, _, h = one_run(False)
        if digest is None:
            digest = h
        elif digest != h:
            raise RuntimeError("grain map differs between repetitions")
        log(f"{'run' if timed else 'warmup'} {rep}: {res['step']} CA steps, X={res['X']:.6f}, "
            f"sum Sv={res['sum_sv']}, {secs:.3f}s")
        if timed:
            runs.append((secs, res))
    drop_t = drop[a.warmup:]

    secs = sum(r[0] for r in runs)
    updates = sum(r[1]["sum_sv"] for r in runs)
    value = updates / secs
    ca_steps = runs[0][1]["step"]
    voxel_steps = dom.N * ca_steps * len(runs) / secs

    roofline = None
    if not a.no_profile:
        psecs, pres, pf, h = one_run(True)
        if h != digest:
            raise RuntimeError("profiled run differs")
        steps = pf["steps"]
        kms = pf["ms_kernel"]
        dom_k = max(kms, key=lambda k: kms[k])
        nbytes, units = kernel_bytes(dom_k, pf)
        launches = pf["n_reanchor_steps"] if dom_k == "k_reanchor" else steps
        avg_us = kms[dom_k] * 1e3 / launches
        achieved = nbytes / launches / (avg_us * 1e-6) / 1e9
        traffic = None
        if os.path.exists(TRAFFIC_FILE):
            tr = json.load(open(TRAFFIC_FILE))
            if tr.get("kernel") == dom_k:
                traffic = tr.get("bytes_per_launch")
        roofline = {
            "bound": "hbm", "kernel": dom_k,
            "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "avg_launch_us": round(avg_us, 3), "alg_bytes_per_launch": int(nbytes / launches), "alg_bytes_units": units,
            "kernel_ms_per_run": {k: round(v, 3) for k, v in kms.items() if v > 0.0},
            "device_ms_per_run": round(pf["ms_step"], 3), "host_loop_ms_per_run": round(pf["ms_host"], 3),
            "counters_per_run": {k: pf[k] for k in ("steps", "sum_sv", "slots_scanned", "n_full", "n_recycled",
                                                    "n_attempts", "n_new", "n_predicted", "n_reanchor_steps",
                                                    "n_reanchored", "pending_resolved")},
        }
        # the sweep beside a plain streaming pass over the same bytes
        sw_bytes = kernel_bytes("k_sweep", pf)[0] / steps
        sw_us = kms["k_sweep"] * 1e3 / steps
        ref = stream_reference(int(sw_bytes))
        roofline["sweep_vs_stream"] = {"sweep_bytes_per_launch": int(sw_bytes), "sweep_us": round(sw_us, 2),
                                       "sweep_GBs": round(sw_bytes / (sw_us * 1e-6) / 1e9, 1), "stream": ref}
        log(f"profiled run: {psecs:.3f}s; dominant kernel {dom_k} {avg_us:.1f} us/launch, "
            f"{achieved:.0f} GB/s algorithmic; per-kernel ms {json.dumps(roofline['kernel_ms_per_run'])}")

    cpu = None
    if not a.no_cpu:
        # the sample: a complete run (to X = 0.99) of the same workload at half the
        # edge (256^3, grain and nucleus counts scaled by volume), so that early and
        # late steps weigh as in the full run; beside it the complete 100^3 run of
        # the CPU workload and a prefix of the 512^3 domain itself, whose GPU
        # trajectory must agree with the oracle's
        from oracle.oracle import Oracle
        half = cpu_full_run(a.workload, a.seed, scale=0.5)
        log(f"cpu baseline: complete {half['workload']} in {half['seconds']:.1f}s, "
            f"{half['cell_updates_per_s']:.3e} cell-updates/s")
        c1 = cpu_full_run("voronoi100", a.seed)
        log(f"cpu: complete voronoi100 in {c1['seconds']:.1f}s, {c1['cell_updates_per_s']:.3e} cell-updates/s")
        o = Oracle(dom, prm)
        k = 0
        tc0 = time.perf_counter()
        while time.perf_counter() - tc0 < a.cpu_seconds:
            if o.step(1):
                break
            k += 1
        tc = time.perf_counter() - tc0
        tr = o.trajectory()
        sv = int(tr["sv"].sum())
        xr = float(tr["X"][-1]) if len(tr) else 0.0
        gm = GrowthMachine(dom, prm)
        gm.step(len(tr))
        prefix_ok = same_trajectory(gm.trajectory(), tr)
        gm.close()
        o.close()
        cpu = {"value": half["cell_updates_per_s"], "unit": "cell-updates/s", "cores": 1, "kind": "port",
               "sample": f"serial oracle, one complete run to X=0.99 of the same workload at half the edge: "
                         f"{half['workload']}, {half['steps']} CA steps, sum Sv={half['sum_sv']}, "
                         f"{half['seconds']:.1f}s",
               "host": host_cpu(),
               "voronoi100_complete_run": c1,
               "voronoi512_prefix": {"steps": k, "X": xr, "sum_sv": sv, "seconds": tc,
                                     "cell_updates_per_s": sv / tc, "gpu_trajectory_matches": prefix_ok}}
        log(f"cpu: {a.workload} prefix {sv / tc:.3e} cell-updates/s over {k} steps; GPU prefix match {prefix_ok}")

    out = {
        "metric": "cell-updates/s", "value": value, "unit": "cell-updates/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": 1e3 * secs / len(runs), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{a.workload}: {dom.nx}x{dom.ny}x{dom.nz} voxels, dx=0.5um, {dom.n_defg} Voronoi "
                               f"grains, {dom.n_rxg} site-saturated nuclei, Fe/Rollett-Holm, T=1000K, run to X=0.99",
                   "seed": a.seed, "ca_steps_per_run": ca_steps, "sum_sv_per_run": runs[0][1]["sum_sv"],
                   "grainmap_sha256_16": digest, "parallelism": "1 GPU"},
        "voxel_steps_per_s": voxel_steps,
        "drop_in": {
            "create_ms": 1e3 * sum(d[0] for d in drop_t) / len(drop_t),
            "run_ms": 1e3 * sum(d[1] for d in drop_t) / len(drop_t),
            "outputs_ms": 1e3 * sum(d[2] for d in drop_t) / len(drop_t),
            "drop_in_ms": 1e3 * sum(sum(d) for d in drop_t) / len(drop_t),
            "what": "create = host arrays to HBM, list set-up and nucleation; run = the timed growth loop; outputs = "
                    "grain counts and the 512 MiB grain map back to the host, then destroy (PCIe included)"},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if cpu:
        out["speedup_vs_cpu_baseline"] = value / cpu["value"]
    if a.ensemble_domains:
        out["batch"] = ensemble_section(a.seed, a.ensemble_domains, a.inflight)
    if a.served_requests:
        out["serve"] = served_section(a.seed, a.served_requests, a.inflight, a.rate)
    if a.regions > 1:
        out["regions"] = regions_section(dom, prm, a.regions)
    if not a.no_rolled_sheet:
        out["rolled_sheet"] = rolled_sheet_section(a.seed)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
