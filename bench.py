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
CPU baseline is the serial oracle timed on a bounded prefix of the same
workload on the host.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
    python bench.py --copy-calibration      # a plain 1 GiB device copy, for PMC scaling
    python bench.py --batch 512 --inflight 8 # ensemble of independent domains, domains/hour
    python bench.py --serve 64 --inflight 8  # Poisson stream of requests, p50/p99 time to X=0.99

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
    read, its 32 B first sector written, its 2 B code read and written)."""
    full, att, new = pf["n_full"], pf["n_attempts"], pf["n_new"]
    scanned, tiles = pf["slots_scanned"], (pf["slots_scanned"] + 63) // 64
    table = {
        "k_reanchor": (100 * pf["n_reanchored"],
                       "100 B per re-anchored cell (record 64 read, 32 written, code 2 read and 2 written)"),
        "k_sweep": (2 * scanned + 16 * tiles, "2 B per occupied slot + 16 B per tile"),
        "k_compact": (16 * tiles + 4 * full, "16 B per tile + 4 B per FullRX entry"),
        "k_recycle": (4 * pf["n_recycled"], "4 B per recycled slot handed out"),
        "k_propose": (220 * full + 4 * att, "220 B per completed cell (26 cell-grid entries 104, record 64 read, "
                                           "parent data 48 and attempt mask 4 written) + 4 B per valid attempt"),
        "k_decide": (28 * full + 4 * att, "28 B per completed cell (position 16, attempt mask 4, winner masks 8) "
                                          "+ 4 B per valid attempt (claim)"),
        "k_commit": (34 * new + 42 * full, "34 B per new record + 42 B per completed cell (record 30, retire 12)"),
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


def batch_bench(a) -> int:
    """Ensemble of independent domains (the rank's loop over its domains in
    SCORE's SIMULATE_myCAs): edges U{64..192}, edge^3/2000 grains, nucleus density
    log-uniform in [1e-5, 1e-3] per voxel, T uniform in [900, 1100] K, each run to
    X = 0.99.  Timed once with one domain at a time and once with `--inflight`
    domains in flight; both must give the same results."""
    import numpy as np
    from score_amd import synth
    from score_amd.growth import batch_run, device_count

    if device_count() < 1:
        print("no gfx950 device visible", file=sys.stderr)
        return 3
    rng = np.random.default_rng(a.seed)
    t0 = time.time()
    doms, prms = [], []
    for i in range(a.batch):
        e = int(rng.integers(64, 193))
        dens = 10.0 ** rng.uniform(-5.0, -3.0)
        T = float(rng.uniform(900.0, 1100.0))
        doms.append(synth.make_voronoi_domain(e, e, e, n_grains=max(1, e ** 3 // 2000),
                                              n_nuclei=max(1, int(round(dens * e ** 3))), seed=a.seed * 100003 + i))
        prms.append(synth.ParamsSpec(sched_temp=(T, T)))
    voxels = sum(d.N for d in doms)
    log(f"batch of {a.batch} domains, {voxels} voxels, generated in {time.time() - t0:.1f}s")
    batch_run(doms[:2], prms[:2], max_inflight=2)  # warm-up
    timings = {}
    outs = {}
    for k in (1, a.inflight):
        t1 = time.perf_counter()
        res = batch_run(doms, prms, max_inflight=k)
        timings[k] = time.perf_counter() - t1
        outs[k] = res
        log(f"max_inflight {k}: {timings[k]:.2f}s, {a.batch / timings[k] * 3600:.0f} domains/hour")
    same = all(r1["result"]["xcells"] == r2["result"]["xcells"] and r1["result"]["t"] == r2["result"]["t"]
               and np.array_equal(r1["defg_counts"], r2["defg_counts"])
               for r1, r2 in zip(outs[1], outs[a.inflight]))
    if not same:
        raise RuntimeError("batched results differ from one-at-a-time results")
    secs = np.array([r["seconds"] for r in outs[a.inflight]])
    sv = sum(r["result"]["sum_sv"] for r in outs[a.inflight])
    out = {
        "metric": "domains/hour", "value": a.batch / timings[a.inflight] * 3600.0, "unit": "domains/hour",
        "n_gpus": 1, "higher_is_better": True, "dtype": "f64", "data": "synthetic",
        "one_at_a_time_domains_per_hour": a.batch / timings[1] * 3600.0,
        "speedup_vs_one_at_a_time": timings[1] / timings[a.inflight],
        "cell_updates_per_s": sv / timings[a.inflight],
        "domain_seconds_p50": float(np.percentile(secs, 50)), "domain_seconds_p99": float(np.percentile(secs, 99)),
        "config": {"workload": f"ensemble of {a.batch} Voronoi domains, edges U{{64..192}}, edge^3/2000 grains, "
                               f"nuclei log-U[1e-5,1e-3]/voxel, T~U[900,1100] K, to X=0.99",
                   "voxels": voxels, "max_inflight": a.inflight, "seed": a.seed},
    }
    print(json.dumps(out), flush=True)
    return 0


def serve_bench(a) -> int:
    """Served stream: requests arrive as a Poisson process (edges log-uniform in
    [64, 320], grain and nucleus densities of the 512^3 workload, T uniform in
    [900, 1100] K) and run through the request loop (score_gm_server_*) with
    `--inflight` workers and the default memory budget.  The arrival rate is
    `--rate` per second, or 80% of the throughput the same requests reach when
    all are queued at once.  Latency is arrival to X = 0.99 result.  The same
    arrivals are then served one request at a time for comparison."""
    import numpy as np
    from score_amd import synth
    from score_amd.growth import Server, batch_run, device_count

    if device_count() < 1:
        print("no gfx950 device visible", file=sys.stderr)
        return 3
    rng = np.random.default_rng(a.seed + 7)
    t0 = time.time()
    doms, prms = [], []
    for i in range(a.serve):
        e = int(round(np.exp(rng.uniform(np.log(64.0), np.log(320.0)))))
        n = e ** 3
        T = float(rng.uniform(900.0, 1100.0))
        doms.append(synth.make_voronoi_domain(e, e, e, n_grains=max(1, round(n * 8000 / 512 ** 3)),
                                              n_nuclei=max(1, round(n * 100000 / 512 ** 3)), seed=a.seed * 7919 + i))
        prms.append(synth.ParamsSpec(sched_temp=(T, T)))
    voxels = sum(d.N for d in doms)
    log(f"served stream of {a.serve} requests, {voxels} voxels, generated in {time.time() - t0:.1f}s")
    batch_run(doms[:2], prms[:2], max_inflight=2)  # warm-up
    rate = a.rate
    if not rate:
        t1 = time.perf_counter()
        batch_run(doms, prms, max_inflight=a.inflight)
        cap = a.serve / (time.perf_counter() - t1)
        rate = 0.8 * cap
        log(f"capacity with {a.inflight} in flight: {cap:.1f} requests/s; arrival rate {rate:.1f}/s")
    gaps = rng.exponential(1.0 / rate, a.serve)
    arrivals = np.cumsum(gaps) - gaps[0]

    def serve(inflight):
        with Server(max_inflight=inflight) as srv:
            reqs = []
            t1 = time.perf_counter()
            for i in range(a.serve):
                dt = arrivals[i] - (time.perf_counter() - t1)
                if dt > 0:
                    time.sleep(dt)
                reqs.append(srv.submit(doms[i], prms[i]))
            for r in reqs:
                srv.wait(r)
            span = time.perf_counter() - t1
        lat = np.array([r.item.queued_seconds + r.item.seconds for r in reqs])
        if any(r.item.status != 0 for r in reqs):
            raise RuntimeError("a served request failed")
        return span, lat, [r.outcome() for r in reqs]

    span, lat, outs = serve(a.inflight)
    log(f"served with {a.inflight} in flight: {span:.2f}s, p50 {np.percentile(lat, 50):.3f}s, "
        f"p99 {np.percentile(lat, 99):.3f}s")
    span1, lat1, outs1 = serve(1)
    log(f"served one at a time: {span1:.2f}s, p50 {np.percentile(lat1, 50):.3f}s, p99 {np.percentile(lat1, 99):.3f}s")
    same = all(x["result"]["xcells"] == y["result"]["xcells"] and x["result"]["t"] == y["result"]["t"] and
               np.array_equal(x["defg_counts"], y["defg_counts"]) for x, y in zip(outs, outs1))
    if not same:
        raise RuntimeError("served results differ between the mix and one at a time")
    out = {
        "metric": "domains/hour", "value": a.serve / span * 3600.0, "unit": "domains/hour", "n_gpus": 1,
        "higher_is_better": True, "dtype": "f64", "data": "synthetic",
        "p50_time_to_x099_s": float(np.percentile(lat, 50)), "p99_time_to_x099_s": float(np.percentile(lat, 99)),
        "arrival_rate_per_s": rate,
        "one_at_a_time": {"domains_per_hour": a.serve / span1 * 3600.0,
                          "p50_time_to_x099_s": float(np.percentile(lat1, 50)),
                          "p99_time_to_x099_s": float(np.percentile(lat1, 99))},
        "config": {"workload": f"served stream of {a.serve} Voronoi domains, Poisson arrivals, edges log-U[64,320], "
                               f"grain and nucleus densities of the 512^3 workload, T~U[900,1100] K, to X=0.99",
                   "voxels": voxels, "max_inflight": a.inflight, "seed": a.seed},
    }
    print(json.dumps(out), flush=True)
    return 0


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
    ap.add_argument("--no-profile", action="store_true", help="skip the profiled run (no roofline)")
    ap.add_argument("--profile-run-only", action="store_true",
                    help="only the profiled run (one step per launch window, events between kernels); for rocprofv3")
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--batch", type=int, default=0,
                    help="ensemble mode: this many independent domains, domains/hour batched vs one at a time")
    ap.add_argument("--inflight", type=int, default=8, help="domains in flight in ensemble and served modes")
    ap.add_argument("--serve", type=int, default=0,
                    help="served mode: this many requests arriving as a Poisson stream; p50/p99 time to X=0.99")
    ap.add_argument("--rate", type=float, default=0.0, help="arrival rate of the served mode, requests/s")
    a = ap.parse_args(argv)

    if a.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {a.gpus} is not supported: a domain runs on one GPU, "
              f"and an ensemble of domains shards as one independent queue per GPU", file=sys.stderr)
        return 2
    if a.copy_calibration:
        return copy_calibration()
    if a.batch:
        return batch_bench(a)
    if a.serve:
        return serve_bench(a)

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

    def one_run(profile: bool):
        gm = GrowthMachine(dom, prm)
        if profile:
            gm.set_profiling(True)
        t1 = time.perf_counter()
        res = gm.solve_rxgrowth()
        t2 = time.perf_counter()
        if not res["completed"]:
            raise RuntimeError("run did not complete")
        pf = gm.profile() if profile else None
        h = hashlib.sha256(gm.grainmap().tobytes()).hexdigest()[:16]
        gm.close()
        return t2 - t1, res, pf, h

    if a.profile_run_only:
        psecs, pres, pf, h = one_run(True)
        log(f"profiled run: {psecs:.3f}s, {pf['steps']} steps, grain map {h}; per-kernel ms "
            f"{json.dumps({k: round(v, 3) for k, v in pf['ms_kernel'].items()})}")
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
        # effective bandwidth of the reference algorithm's bytes per phase
        sweep_ref = 30 * pf["sum_sv"] + (pf["slots_scanned"] - pf["sum_sv"]) + 4 * (pf["n_full"] + pf["n_recycled"])
        infect_ref = (pf["n_full"] * (26 * 4 + 30) + 4 * pf["n_attempts"] + pf["n_new"] * 34 + pf["n_full"] * 12)
        sweep_ms = sum(kms[k] for k in ("k_reanchor", "k_sweep", "k_sweep_scan", "k_compact"))
        infect_ms = sum(kms[k] for k in ("k_propose", "k_decide", "k_decide_scan", "k_commit"))
        roofline = {
            "bound": "hbm", "kernel": dom_k,
            "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "avg_launch_us": round(avg_us, 3), "alg_bytes_per_launch": int(nbytes / launches), "alg_bytes_units": units,
            "kernel_ms_per_run": {k: round(v, 3) for k, v in kms.items()},
            "device_ms_per_run": round(pf["ms_step"], 3), "host_loop_ms_per_run": round(pf["ms_host"], 3),
            "reference_bytes_effective_GBs": {
                "sweep": round(sweep_ref / (sweep_ms * 1e-3) / 1e9, 1),
                "infection": round(infect_ref / (infect_ms * 1e-3) / 1e9, 1),
                "step": round((sweep_ref + infect_ref) / (pf["ms_step"] * 1e-3) / 1e9, 1)},
            "counters_per_run": {k: pf[k] for k in ("steps", "sum_sv", "slots_scanned", "n_full", "n_recycled",
                                                    "n_attempts", "n_new", "n_predicted", "n_reanchor_steps",
                                                    "n_reanchored", "pending_resolved")},
        }
        log(f"profiled run: {psecs:.3f}s; dominant kernel {dom_k} {avg_us:.1f} us/launch, "
            f"{achieved:.0f} GB/s algorithmic; per-kernel ms {json.dumps(roofline['kernel_ms_per_run'])}")

    cpu = None
    if not a.no_cpu:
        from oracle.oracle import Oracle
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
        # the GPU trajectory must agree with the oracle's over the sampled prefix
        gm = GrowthMachine(dom, prm)
        gm.step(len(tr))
        prefix_ok = same_trajectory(gm.trajectory(), tr)
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
                               f"grains, {dom.n_rxg} site-saturated nuclei, Fe/Rollett-Holm, T=1000K, run to X=0.99",
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
