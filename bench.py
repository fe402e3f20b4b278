#!/usr/bin/env python3
"""Benchmark: zone-updates/s of the GRMHD fluid step on one MI355X, FP64 and
mixed precision, on the 3-level AMR Fishbone-Moncrief torus; serving p50/p99.

Workload (the AMR torus the metric is quoted on): a = 0.9375, Gamma = 13/9,
r_in = 6, r_max = 12, A_phi = max(rho/rho_max - 0.2, 0) scaled to beta = 100,
+-4 % PCG64(42) pressure perturbation, floors rho = 1e-7 r^-1.5,
eps = 1e-5 r^-1.5, CFL 0.4, on a log-r (1.2 - 40) x theta x phi grid of
16^3-zone blocks with two ghost layers over three levels (level 0 of 128 x 64
x 64 zones, refined where rho > 0.1 rho_max and 2:1 balanced: 1976 blocks,
56 / 384 / 1536 per level, 8.09 M zones).  The physics is the reference's
step (src/fluid/solver/integrator.cpp:262-275) in its self-consistent form
(csrc/sph_physics.hpp).  A step is one RK2 step of the whole tree.

Passes, each from freshly computed initial data (the state is resident in
HBM when timing starts):
  1. autotuning of the launch configuration (bhfluid.tune), per precision;
  2. timed: W warm-up steps, then K steps replayed from the captured HIP
     graph with an event between consecutive steps; value = zones / median
     step time (FP64; the mixed-precision figures beside it);
  3. accuracy: the same steps_acc steps in FP64, in mixed precision and in
     mixed precision with hardware FP32 reciprocal and square root ("fast")
     from the same initial data; L1(rho), mass and energy drift, max |div B|;
  4. per-kernel HIP-event breakdown (eager launches) and the dominant
     kernel's roofline from its launches timed alone;
  5. the uniform 3D torus (256 x 128 x 128), 200 FP64 steps, the last 20
     timed: zone-updates/s, HBM GB/s by the 296 B measure, finiteness and
     max |div B|;
  6. CPU baseline: the oracle (C++/OpenMP restatement) stepping the same tree
     from the same initial data on this host's cores;
  7. serving: requests with Poisson arrivals, grids 64^3 .. 192 x 96 x 96
     (uniform or two levels), spin ~ U(0, 0.99), beta in {30, 100, 300},
     100 steps each, continuous batching under an HBM block budget, mixed
     precision; p50 / p99 latency and aggregate zone-updates/s.
Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "black-hole-simulator_amd"))
sys.path.insert(0, str(ROOT / "tests"))

METRIC = "zone-updates/s FP64 vs mixed, AMR torus @L1(ρ)≤1e-5; p99 run latency"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFS = 78.6
# compulsory state traffic of one zone-update (FP64 state in both precisions):
# stage 1 reads U^n (8 words) and writes U* (8); stage 2 reads U* (8) and the
# stored hydro words of U^n (5) and writes U^{n+1} (8): 37 words
BYTES_PER_ZONE_UPDATE = 296
STAGE_BYTES = (16 * 8, 21 * 8)  # per zone-update, stage 1 and stage 2
PROFILES = ROOT / "profiles"


def c3_setup(sph):
    g = sph.c3_grid()
    blocks = sph.amr_blocks(g, 3)
    return g, blocks


def make_engine(sph, g, blocks, mixed):
    eng = sph.Engine(16, len(blocks))
    eng.set_precision(mixed)
    run = eng.add_run(sph.PHYSICS, g, blocks)
    run.init_torus(sph.TORUS, 42)
    return eng, run


def measured_traffic(kernel, precision):
    """Copy-calibrated HBM bytes per dispatch of ``kernel`` from the committed
    counter summary (profiles/traffic_c3_amr_torus.md), or None."""
    f = PROFILES / "traffic_c3_amr_torus.md"
    if not f.exists():
        return None
    for line in f.read_text().splitlines():
        cols = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cols) >= 6 and cols[0] == kernel and cols[1] == precision:
            try:
                return float(cols[5]) * 1e9
            except ValueError:
                return None
    return None


def measured_step_traffic_c2():
    """HBM GB per FP64 step of the uniform 3D torus from the committed
    counter summary (profiles/kernels_traffic_c2_uniform_torus.md), or None."""
    f = PROFILES / "kernels_traffic_c2_uniform_torus.md"
    if not f.exists():
        return None
    for line in f.read_text().splitlines():
        cols = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cols) >= 5 and cols[0] == "step" and cols[1] == "fp64":
            try:
                return float(cols[4])
            except ValueError:
                return None
    return None


def measured_flops(precision):
    f = PROFILES / "flops_c3_amr_torus.md"
    if not f.exists():
        return None
    for line in f.read_text().splitlines():
        cols = [c.strip() for c in line.strip().strip("|").split("|")]
        # | step | precision | ms / step | FP64 flops / step | ...
        if len(cols) >= 4 and cols[0] == "step" and cols[1] == precision:
            try:
                return float(cols[3])
            except ValueError:
                return None
    return None


def timed(sph, tune, g, blocks, mixed, steps, warmup, tune_log):
    eng, run = make_engine(sph, g, blocks, mixed)
    t = tune.tune(eng, [run], mixed, steps=8, use_cache=False, log=tune_log)
    # fresh initial data for the timed pass
    run.init_torus(sph.TORUS, 42)
    eng.step([run], warmup, want_dt=False)
    ms = eng.step_timed([run], steps)
    return eng, run, ms, t["config"]


def accuracy(sph, g, blocks, nsteps, cfg64, cfgmx, cfgfast):
    """FP64, mixed and mixed-fast runs of nsteps from the same initial data;
    the mixed runs' L1(rho), drifts and div B against the FP64 run (the
    mixed-fast figures under "fast")."""
    from bhfluid import tune
    out = {}
    states = {}
    for name, mixed, cfg in (("fp64", False, cfg64), ("mixed", True, cfgmx), ("fast", 2, cfgfast)):
        eng, run = make_engine(sph, g, blocks, mixed)
        tune.apply(eng, cfg)
        d0 = run.diagnostics()
        dts = eng.step([run], nsteps)[:, 0]
        run.c2p()
        d1 = run.diagnostics()
        U, W, Bf = run.state()
        rho = run.active(W)[:, 0].copy()
        states[name] = dict(rho=rho, d0=d0, d1=d1, t=float(dts.sum()), finite=bool(np.isfinite(run.active(W)).all()))
        del eng, run, U, W, Bf
    # L1 of rho weighted by each zone's coordinate volume, relative to the
    # FP64 run's
    w = sph.zone_volumes(g, blocks)
    a, b = states["mixed"]["rho"], states["fp64"]["rho"]
    l1 = float((np.abs(a - b) * w).sum() / (np.abs(b) * w).sum())
    f, m = states["fp64"], states["mixed"]
    out.update(steps=nsteps, L1_rho=l1,
               mass_drift=(m["d1"]["mass"] - f["d1"]["mass"]) / f["d1"]["mass"],
               energy_drift=(m["d1"]["energy"] - f["d1"]["energy"]) / f["d1"]["energy"],
               fp64_mass_change=(f["d1"]["mass"] - f["d0"]["mass"]) / f["d0"]["mass"],
               fp64_energy_change=(f["d1"]["energy"] - f["d0"]["energy"]) / f["d0"]["energy"],
               divb_rel_fp64=f["d1"]["divb_max"] / f["d1"]["divb_scale"],
               divb_rel_mixed=m["d1"]["divb_max"] / m["d1"]["divb_scale"],
               t_fp64=f["t"], t_mixed=m["t"], finite=f["finite"] and m["finite"],
               c2p_failures_fp64=f["d1"]["c2p_fail"], c2p_failures_mixed=m["d1"]["c2p_fail"])
    q = states["fast"]
    out["fast"] = dict(L1_rho=float((np.abs(q["rho"] - b) * w).sum() / (np.abs(b) * w).sum()),
                       mass_drift=(q["d1"]["mass"] - f["d1"]["mass"]) / f["d1"]["mass"],
                       energy_drift=(q["d1"]["energy"] - f["d1"]["energy"]) / f["d1"]["energy"],
                       divb_rel=q["d1"]["divb_max"] / q["d1"]["divb_scale"], t=q["t"], finite=q["finite"],
                       c2p_failures=q["d1"]["c2p_fail"])
    return out


def kernel_breakdown(sph, g, blocks, mixed, cfg, steps):
    """Per-kernel HIP-event times (eager launches), then the dominant kernel
    timed alone (events around its launches only)."""
    from bhfluid import tune
    eng, run = make_engine(sph, g, blocks, mixed)
    tune.apply(eng, cfg)
    eng.profile(True)
    eng.step([run], steps, want_dt=False)
    rep = eng.profile_report()
    dom = max((k for k in rep if k != "k_begin"), key=lambda k: rep[k]["total_ms"])
    eng.profile(True, only=dom)
    eng.step([run], steps, want_dt=False)
    alone = eng.profile_report()[dom]
    eng.profile(False)
    return rep, dom, alone


def cpu_baseline(sph, g, blocks, nsteps=2):
    """The oracle stepping the C3 tree from the device's initial data (copied
    to the host), FP64, all host threads; median over steps."""
    from sph_oracle import OracleRun
    eng, run = make_engine(sph, g, blocks, False)
    U, W, Bf = run.state()
    del eng, run
    p = sph.PHYSICS
    cfg = dict(bz=16, a=p["a"], Q=p["Q"], gamma=p["gamma"], cfl=p["cfl"], rho_fl=p["rho_floor"],
               eps_fl=p["eps_floor"], sigma_max=p["sigma_max"], w_max=p["gamma_max"], nr=g["nr"], nth=g["nth"],
               nph=g["nph"], r_in=g["r_in"], r_out=g["r_out"], nlev=g["nlevels"], riemann=sph.riemann_code(p))
    orc = OracleRun(cfg, blocks)
    orc.set_state(U, W, Bf)
    del U, W, Bf
    times = []
    for _ in range(nsteps):
        t0 = time.perf_counter()
        orc.step(False)
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    threads = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp:
        threads = min(threads, int(omp))
    zones = len(blocks) * 4096
    return {"value": zones / med, "unit": "zone-updates/s", "cores": threads, "kind": "port",
            "sample": f"{nsteps} FP64 steps of the oracle on the whole C3 tree ({len(blocks)} blocks, {zones} zones) "
                      f"from the same initial data; median {med:.2f} s per step"}


def c2_run(sph, tune, nsteps, timed_steps):
    """The uniform 3D torus (256 x 128 x 128, 1024 blocks, 4.19 M zones),
    FP64: the tuned launch configuration, then nsteps steps of which the last
    timed_steps are timed per step (graph replays, HIP events); finiteness,
    max |div B| and the mass change after the run."""
    g = sph.c2_grid()
    blocks = sph.uniform_blocks(g, 16)
    eng = sph.Engine(16, len(blocks))
    run = eng.add_run(sph.PHYSICS, g, blocks)
    run.init_torus(sph.TORUS, 42)
    cfg = tune.tune(eng, [run], False, steps=8, use_cache=False)["config"]
    run.init_torus(sph.TORUS, 42)
    d0 = run.diagnostics()
    eng.step([run], nsteps - timed_steps, want_dt=False)
    ms = eng.step_timed([run], timed_steps)
    run.c2p()
    d1 = run.diagnostics()
    U, W, Bf = run.state()
    finite = bool(np.isfinite(run.active(U)).all() and np.isfinite(run.active(W)).all())
    del U, W, Bf
    zones = len(blocks) * 4096
    med = float(np.median(ms))
    value = zones / (med * 1e-3)
    out = {"workload": "uniform 3D FM torus 256x128x128 (1024 blocks of 16^3), FP64, HLLD", "zones": zones,
           "steps": nsteps, "timed_steps": timed_steps, "ms_per_step": med, "value": value,
           "unit": "zone-updates/s", "hbm_GBps_296B": value * BYTES_PER_ZONE_UPDATE / 1e9,
           "hbm_frac_296B": value * BYTES_PER_ZONE_UPDATE / 1e9 / HBM_PEAK_GBS, "launch": cfg, "finite": finite,
           "divb_rel": d1["divb_max"] / d1["divb_scale"], "mass_change": (d1["mass"] - d0["mass"]) / d0["mass"]}
    gb = measured_step_traffic_c2()
    if gb:
        out.update(hbm_GB_per_step_rocprof=gb, hbm_GBps_rocprof=gb / (med * 1e-3),
                   hbm_frac_rocprof=gb / (med * 1e-3) / HBM_PEAK_GBS,
                   hbm_source="rocprofv3 FETCH_SIZE + WRITE_SIZE of every step kernel, copy-calibrated "
                              "(profiles/kernels_traffic_c2_uniform_torus.md)")
    del eng, run
    return out


def serving(sph, nreq, rate, budget):
    """The request stream at ``rate`` per second (Poisson arrivals, real
    time), then the same requests all queued at t = 0 (a closed batch: the
    engine's throughput when admission, not arrivals, limits the batch)."""
    from bhfluid.serve import Server, make_requests
    reqs = make_requests(nreq, rate)
    srv = Server(budget, mixed=2)
    s = srv.run(reqs)
    s.update(rate_per_s=rate, budget_blocks=budget, steps_per_request=reqs[0].nsteps,
             finite=all(np.isfinite(r.result["mass"]) for r in reqs),
             precision="mixed-fast (FP32 fluxes/EMFs with hardware reciprocal and sqrt, FP64 state)")
    del srv
    closed = make_requests(nreq, rate)
    for r in closed:
        r.arrival = 0.0
    srv = Server(budget, mixed=2)
    c = srv.run(closed, realtime=False)
    c.update(finite=all(np.isfinite(r.result["mass"]) for r in closed),
             latency_from="t = 0 (every request queued at the start)")
    s["closed_batch"] = c
    del srv
    return s


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--accuracy-steps", type=int, default=500)
    ap.add_argument("--serve-requests", type=int, default=256)
    ap.add_argument("--serve-rate", type=float, default=12.0)
    ap.add_argument("--serve-budget", type=int, default=4000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--c2-steps", type=int, default=200,
                    help="steps of the uniform 3D torus (256 x 128 x 128) run; 0 skips it")
    ap.add_argument("--kernels-only", choices=("fp64", "mixed", "fast", "ref"),
                    help="for counter collection: C3 initial data, then --steps eager steps of the step's kernels "
                         "in the given precision with the default launch configuration; or (ref) the reference-path "
                         "engine on the reference's own production tree (main.cpp:61-168: 8^3 root patches on "
                         "[-50, 50]^3, FM torus a = 0.9, Q = 0.1, Gamma = 5/3, one regrid), initial data and "
                         "regrid on the device, then --steps steps; prints one JSON line")
    ap.add_argument("--workload", choices=("c3", "c2"), default="c3",
                    help="with --kernels-only fp64/mixed/fast: the 3-level AMR torus (c3) or the uniform 3D torus "
                         "256 x 128 x 128 (c2)")
    ap.add_argument("--launch", default=None,
                    help="with --kernels-only: the launch configuration as JSON (as config.launch_fp64 or "
                         "config.launch_mixed of a bench line); default: the untuned one")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="for counter collection: three plain 1 GiB device copies (reads and writes of known size)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the step runs on a single MI355X (every box here has one); "
              "--gpus must be 1", file=sys.stderr)
        sys.exit(2)

    from bhfluid import sph, tune, device_info

    if args.copy_calibration:
        from bhfluid import copy_probe
        ms = copy_probe(1 << 30, 3)
        print(json.dumps({"mode": "copy-calibration", "bytes": 1 << 30, "reps": 3, "ms_per_copy": ms}))
        return
    if args.kernels_only == "ref":
        import bhfluid
        t = bhfluid.AmrTree((-50.0,) * 3, (50.0,) * 3, 8, 1.0, 0.9, 0.1, 5.0 / 3.0)
        bhfluid.torus_setup(t, regrid=True, device=True)
        rk = bhfluid.RK2Integrator(0.4)
        faults = 0
        for _ in range(args.steps):
            try:
                rk.step(t)
            except bhfluid.DomainError:
                faults += 1   # the reference would stop here; the step itself completed
        bhfluid.device_sync()
        print(json.dumps({"mode": "kernels-only", "workload": "reference-path main.cpp torus", "steps": args.steps,
                          "patches": int(t.nblocks), "steps_with_eos_faults": faults}))
        return
    if args.kernels_only:
        if args.workload == "c2":
            g = sph.c2_grid()
            blocks = sph.uniform_blocks(g, 16)
        else:
            g, blocks = c3_setup(sph)
        eng, run = make_engine(sph, g, blocks, {"fp64": 0, "mixed": 1, "fast": 2}[args.kernels_only])
        launch = dict(tune.DEFAULT, **json.loads(args.launch)) if args.launch else dict(tune.DEFAULT)
        tune.apply(eng, launch)
        eng.use_graph(False)
        eng.step([run], args.steps, want_dt=False)
        eng.sync()
        print(json.dumps({"mode": "kernels-only", "workload": args.workload, "precision": args.kernels_only,
                          "steps": args.steps,
                          "blocks": int(len(blocks)), "zones": int(len(blocks)) * 4096,
                          "launch": launch}))
        return

    g, blocks = c3_setup(sph)
    zones = len(blocks) * 4096
    per_level = [int((blocks[:, 0] == l).sum()) for l in range(3)]
    tune_log = {"fp64": [], "mixed": []}

    eng, run, ms64, cfg64 = timed(sph, tune, g, blocks, False, args.steps, args.warmup, tune_log["fp64"])
    del eng, run
    eng, run, msmx, cfgmx = timed(sph, tune, g, blocks, True, args.steps, args.warmup, tune_log["mixed"])
    del eng, run
    eng, run, msfa, cfgfa = timed(sph, tune, g, blocks, 2, args.steps, args.warmup, [])
    del eng, run
    med64, medmx = float(np.median(ms64)), float(np.median(msmx))
    value = zones / (med64 * 1e-3)
    value_mx = zones / (medmx * 1e-3)

    medfa = float(np.median(msfa))
    value_fa = zones / (medfa * 1e-3)
    acc = accuracy(sph, g, blocks, args.accuracy_steps, cfg64, cfgmx, cfgfa)
    acc_fast = acc.pop("fast")

    rep64, dom64, alone64 = kernel_breakdown(sph, g, blocks, False, cfg64, 4)
    repmx, dommx, alonemx = kernel_breakdown(sph, g, blocks, True, cfgmx, 4)
    # roofline of the dominant FP64 kernel: the per-stage share of the
    # compulsory 296 B per zone-update (148 B: the mean of the two stages)
    # over one launch, which covers every zone once
    algo = (STAGE_BYTES[0] + STAGE_BYTES[1]) / 2 * zones
    launches_per_step = max(1, round(rep64[dom64]["launches"] / 4))
    ach = algo / (alone64["avg_ms"] * 1e-3) / 1e9
    traffic = measured_traffic(dom64, "fp64")
    roofline = {"bound": "hbm", "kernel": dom64, "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": ach / HBM_PEAK_GBS, "traffic": traffic,
                "algo_bytes_per_launch": algo, "avg_launch_ms": alone64["avg_ms"],
                "launches_per_step": launches_per_step,
                "traffic_source": "rocprofv3 FETCH_SIZE + WRITE_SIZE per dispatch, copy-calibrated "
                                  "(profiles/traffic_c3_amr_torus.md)" if traffic else None}
    flops = measured_flops("fp64")
    step_roof = {"bytes_per_zone_update": BYTES_PER_ZONE_UPDATE,
                 "achieved_GBps_fp64": value * BYTES_PER_ZONE_UPDATE / 1e9,
                 "hbm_frac_fp64": value * BYTES_PER_ZONE_UPDATE / 1e9 / HBM_PEAK_GBS,
                 "achieved_GBps_mixed": value_mx * BYTES_PER_ZONE_UPDATE / 1e9,
                 "hbm_frac_mixed": value_mx * BYTES_PER_ZONE_UPDATE / 1e9 / HBM_PEAK_GBS}
    if flops:
        tfs = flops / (med64 * 1e-3) / 1e12
        step_roof.update(fp64_flops_per_step=flops, fp64_TFLOPs=tfs, fp64_peak=FP64_PEAK_TFS,
                         fp64_frac=tfs / FP64_PEAK_TFS,
                         fp64_source="rocprofv3 SQ_INSTS_VALU_FLOPS_FP64 (profiles/flops_c3_amr_torus.md)")

    c2 = c2_run(sph, tune, args.c2_steps, min(20, args.c2_steps)) if args.c2_steps > 0 else None
    cpu = None if args.no_cpu_baseline else cpu_baseline(sph, g, blocks)
    srv = None if args.no_serve else serving(sph, args.serve_requests, args.serve_rate, args.serve_budget)

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": med64,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: Fishbone-Moncrief torus initial data computed on the device (PCG64 seed 42 perturbation)",
        "config": {
            "workload": "3-level AMR FM torus (C3): a=0.9375, Gamma=13/9, r 1.2-40 log x theta x phi, level 0 "
                        "128x64x64, 16^3 blocks + 2 ghosts refined where rho>0.1 rho_max, CFL 0.4, PLM minmod + HLLD "
                        "+ CT + RK2, consistent preset",
            "blocks": int(len(blocks)), "blocks_per_level": per_level, "zones": zones,
            "ms_per_step_mean": float(np.mean(ms64)), "ms_per_step_p90": float(np.percentile(ms64, 90)),
            "timing": "median over K graph-replayed steps, HIP events between steps on the engine stream",
            "launch_fp64": cfg64, "launch_mixed": cfgmx, "device": device_info(),
            "kernels_fp64": rep64, "kernels_mixed": repmx,
            "kernels_source": "a separate pass of 4 steps with eager launches bracketed by HIP events",
        },
        "mixed": {"value": value_mx, "ms_per_step": medmx, "speedup_vs_fp64": value_mx / value,
                  "dtype": "f32 reconstruction/HLLD/EMF, f64 con2prim/sources/update/reductions",
                  "dominant_kernel": dommx, "dominant_avg_launch_ms": alonemx["avg_ms"], **acc},
        "mixed_fast": {"value": value_fa, "ms_per_step": medfa, "speedup_vs_fp64": value_fa / value,
                       "dtype": "as mixed, with f32 division and sqrt from the hardware reciprocal and "
                                "square-root instructions (a few ulp; not bit-identical to a CPU)",
                       "launch": cfgfa, "steps": args.accuracy_steps, **acc_fast},
        "roofline": roofline,
        "step_roofline": step_roof,
        "cpu_baseline": cpu,
        "c2": c2,
        "serving": srv,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
