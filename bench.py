#!/usr/bin/env python3
"""Throughput of the per-timestep PIC loop on one MI355X.

Workload (default): periodic 256^3 grid, 64 particles per cell per species,
electrons (q/m -1, q -1/64, v_th 1.0) and ions (q/m 0.01, q +1/64, v_th 0.1),
uniform B0z 0.5, dt 0.1, FP64 positions/velocities, seeded uniform Maxwellian
(PCG32 seed 42/52) generated on the device.  One step = charge deposit (fused
into the previous step), spectral Poisson solve, E, gather + Boris + wrap +
re-bin of all 2^31 particles.

Prints ONE JSON line on stdout (progress goes to stderr).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published


def log(*a):
    print(*a, file=sys.stderr, flush=True)


WORKLOADS = {
    # name: grid, mean ppc per species, precision, density hump amplitude
    "uniform256_ppc64_fp64": dict(n=(256, 256, 256), ppc=64, precision=0, hump=0.0),
    "uniform256_ppc64_mixed": dict(n=(256, 256, 256), ppc=64, precision=1, hump=0.0),
    "hump256_ppc64_mixed": dict(n=(256, 256, 256), ppc=64, precision=1, hump=9.0),
    "uniform128_ppc64_fp64": dict(n=(128, 128, 128), ppc=64, precision=0, hump=0.0),
    "uniform128_ppc64_mixed": dict(n=(128, 128, 128), ppc=64, precision=1, hump=0.0),
    "uniform32_ppc16_fp64": dict(n=(32, 32, 32), ppc=16, precision=0, hump=0.0),
}


def make_sim(ohp, wl, args, track_ids=False):
    n = wl["n"]
    ppc = wl["ppc"]
    npcl = ppc * n[0] * n[1] * n[2]
    cfg = ohp.SimConfig(n=n, qm=[-1.0, 0.01], q_macro=[-1.0 / ppc, 1.0 / ppc],
                        capacity=[npcl, npcl], dt=0.1, b0=(0.0, 0.0, 0.5),
                        precision=wl["precision"], slack=args.slack, track_ids=track_ids,
                        tile=tuple(args.tile), chunks=args.chunks)
    sim = ohp.Simulation(cfg)
    if wl["hump"] > 0:
        # n(r) = n0 (1 + 9 exp(-|r-c|^2 / (2 16^2))), mean ppc over the box
        sim.seed_hump_maxwell([npcl, npcl], [1.0, 0.1], wl["hump"], 16.0, seed=(42, 52),
                              budget=2052)
    else:
        sim.seed_uniform_maxwell([npcl, npcl], [1.0, 0.1], seed=(42, 52))
    return sim, cfg, npcl


def run_gpu(args, workload=None, timing=True):
    import ohhelppic3d_amd as ohp

    workload = workload or args.workload
    wl = WORKLOADS[workload]
    n = wl["n"]
    ppc = wl["ppc"]
    ng = n[0] * n[1] * n[2]
    t0 = time.time()
    sim, cfg, npcl = make_sim(ohp, wl, args)
    log(f"[bench] workload {workload}: grid {n}, {2 * npcl} particles, "
        f"est. footprint {ohp.estimate_footprint(cfg) / 1e9:.1f} GB")
    log(f"[bench] created + loaded in {time.time() - t0:.1f} s, tile {sim.tile}, "
        f"device bytes {sim.memory_footprint() / 1e9:.1f} GB")
    for i in range(args.warmup):
        sim.step(1)
        log(f"[bench] warmup step {i + 1}/{args.warmup} done")
    sim.synchronize()
    # energy drift is measured from the first diagnostics available: after
    # the warm-up, or after the first timed step without warm-up
    e_start = sim.diagnostics()["total_energy"] if args.warmup > 0 else None
    # per-kernel HIP events for the roofline (they add gaps between launches,
    # which only small grids notice)
    sim.enable_timing(timing)
    per_step = []
    t_start = time.perf_counter()
    for i in range(args.steps):
        ts = time.perf_counter()
        sim.step(1)
        per_step.append(time.perf_counter() - ts)
        if e_start is None:
            e_start = sim.diagnostics()["total_energy"]
        if (i + 1) % max(1, args.steps // 5) == 0:
            log(f"[bench] timed step {i + 1}/{args.steps}: {per_step[-1] * 1e3:.2f} ms")
    sim.synchronize()
    t_total = time.perf_counter() - t_start
    kt = sim.kernel_times()
    diag = sim.diagnostics()
    tile = list(sim.tile)
    sim.close()

    total_p = 2 * npcl
    value = total_p * args.steps / t_total
    push_ms, push_n = kt.get("push_fused", (0.0, 0))
    solve_ms, solve_n = kt.get("field_solve", (0.0, 0))
    push_avg = push_ms / max(push_n, 1)
    # particle state read and written per step: FP64 x, v (48 B); mixed
    # position words + FP32 v (24 B, the cell is implicit in the bin)
    sp_bytes = 48 if wl["precision"] == 0 else 24
    alg_bytes = 2 * sp_bytes * total_p + 24 * ng + 8 * ng
    achieved = alg_bytes / (push_avg * 1e-3) / 1e9 if push_avg > 0 else 0.0
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"traffic_push_fused_{workload}.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    res = {
        "metric": "particle-steps/s at 256^3x64ppc, FP64 vs mixed; p50/p99 ensemble latency",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_total / args.steps * 1e3,
        "ms_per_step_p50": float(np.median(per_step) * 1e3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if wl["precision"] == 0 else "f32+f64",
        "data": "synthetic (device-seeded PCG32 uniform Maxwellian)",
        "config": {"workload": workload, "grid": list(n), "ppc_per_species": ppc,
                   "species": 2, "particles": total_p, "dt": 0.1, "b0z": 0.5,
                   "tile": tile, "slack": args.slack},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "push_fused", "kernel_avg_ms": push_avg,
                     "alg_bytes_per_launch": alg_bytes},
        "field_solve_avg_ms": solve_ms / max(solve_n, 1),
        "energy": {"total": diag["total_energy"], "field": diag["field_energy"],
                   "repacks": diag["repacks"],
                   "drift_over_timed_steps": (abs(diag["total_energy"] - e_start) / abs(e_start)
                                              if e_start else None)},
    }
    return res


def mixed_error_vs_fp64(args):
    """FP64 and mixed runs from the same seeded state at 128^3 x 64 ppc:
    relative L2 of rho, phi and E after the first solve (the loaded state)
    and after the second (one push later), for both CIC weight orders."""
    import ohhelppic3d_amd as ohp

    out = {"setup": "128^3 x 64 ppc x 2 species, seed 42/52; step0 = fields of the "
                    "loaded state, step1 = fields after one push"}
    for wname, wo in (("reference_weights", ohp.WEIGHTS_REFERENCE),
                      ("standard_weights", ohp.WEIGHTS_STANDARD)):
        grids = {}
        for prec in ("fp64", "mixed"):
            wl = WORKLOADS[f"uniform128_ppc64_{prec}"]
            n, ppc = wl["n"], wl["ppc"]
            npcl = ppc * n[0] * n[1] * n[2]
            cfg = ohp.SimConfig(n=n, qm=[-1.0, 0.01], q_macro=[-1.0 / ppc, 1.0 / ppc],
                                capacity=[npcl, npcl], precision=wl["precision"],
                                slack=args.slack, weight_order=wo, tile=tuple(args.tile))
            sim = ohp.Simulation(cfg)
            sim.seed_uniform_maxwell([npcl, npcl], [1.0, 0.1], seed=(42, 52))
            res = []
            for _ in range(2):
                sim.step(1)
                res.append([sim.grid(w) for w in (ohp.FIELD_RHO, ohp.FIELD_PHI, ohp.FIELD_EX,
                                                  ohp.FIELD_EY, ohp.FIELD_EZ)])
            sim.close()
            grids[prec] = res
        out[wname] = {}
        for st in range(2):
            e = {}
            for k, a, b in zip(("rho", "phi", "Ex", "Ey", "Ez"), grids["mixed"][st],
                               grids["fp64"][st]):
                e[k] = float(np.linalg.norm(a - b) / np.linalg.norm(b))
            out[wname][f"step{st}"] = e
    return out


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo", "r") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def run_cpu_baseline(seconds: float):
    """The FP64 oracle (a CPU restatement of the reference step, gcc -O2
    -fopenmp) on the host: 32^3 grid, 16 ppc per species (2^20 particles), one
    MPI-free rank, timed on one thread and on the OpenMP default thread count
    (per-thread charge arrays for the deposit).  ``value`` is the threaded
    rate; the one-thread rate is reported beside it."""
    import oracle as orc

    n = (32, 32, 32)
    ppc = 16
    rng = orc.Pcg32(42, 52)
    sp = []
    for qm, vt in ((-1.0, 1.0), (0.01, 0.1)):
        x, y, z, vx, vy, vz = orc.load_uniform_maxwell(rng, ppc * 32 ** 3, n, vt)
        sp.append(dict(qm=qm, q=(1 if qm > 0 else -1) / ppc, x=x, y=y, z=z, vx=vx, vy=vy,
                       vz=vz))
    sim = orc.OracleSim(n, sp, 0.1, (0, 0, 0.5), 0)
    npart = 2 * ppc * 32 ** 3

    def timed(step, budget):
        step()  # warm-up
        times = []
        t_end = time.perf_counter() + budget
        while time.perf_counter() < t_end or len(times) < 3:
            t = time.perf_counter()
            step()
            times.append(time.perf_counter() - t)
        return float(np.median(times)), len(times)

    med1, k1 = timed(lambda: sim.step(1), 0.6 * seconds)
    used = [1]

    def threaded():
        used[0] = sim.step_threads(1, 0)

    medt, kt = timed(threaded, 0.4 * seconds)
    return {"value": npart / medt, "unit": "particle-steps/s", "cores": used[0],
            "kind": "port", "single_thread_value": npart / med1,
            "cpu_model": _cpu_model(), "ranks": 1, "compiler": "gcc -O2 -fopenmp",
            "sample": f"FP64 oracle step, 32^3 grid x 16 ppc x 2 species ({npart} particles): "
                      f"{used[0]} threads, {kt} steps, median {medt * 1e3:.1f} ms; "
                      f"1 thread, {k1} steps, median {med1 * 1e3:.1f} ms"}


def run_ensemble_bench(args):
    """128 independent mixed-precision instances (each axis from
    {16,24,32,48,64}, 8..64 ppc per species, v_th,e in [0.5,2], omega_ce in
    [0,1], dt 0.1) advanced as one group: autotuned tile/chunks, one launch
    per kernel and step replayed from a hipGraph.  Also times the same
    instances stepped one after another for the batching gain."""
    from ohhelppic3d_amd import ensemble

    specs = ensemble.sample_ensemble(args.ensemble_instances, seed=2024)
    log(f"[ensemble] {len(specs)} instances, {sum(s.particles for s in specs)} particles, "
        f"{sum(s.cells for s in specs)} cells")
    res = ensemble.run_ensemble(specs, steps=args.ensemble_steps, progress=log)
    tune = {"tile": tuple(res["tile"]), "chunks": res["chunks"]}
    log("[ensemble] same instances stepped one after another")
    seq = ensemble.run_sequential(specs, steps=args.ensemble_seq_steps, tune=tune)
    res["sequential"] = seq
    res["batching_gain"] = res["value"] / seq["value"]
    res["config"] = {"workload": f"ensemble{len(specs)}_mixed", "seed": 2024,
                     "axes": list(ensemble.AXES), "ppc_range": [8, 64], "dt": 0.1}
    return res


def run_serving_bench(args):
    """Serving loop: seeded requests (grids 32^3/64^3/96^3/128^3 weighted
    4:3:2:1, 16..64 ppc per species, 200..2000 steps, mixed precision)
    admitted against a 240 GB budget.  Capacity is measured with every
    request arriving at once; then Poisson arrivals at 50/80/100% of it."""
    import ohhelppic3d_amd as ohp
    from ohhelppic3d_amd import serve

    n = args.serve_requests
    probe = serve.sample_requests(n, rate=float("inf"), seed=7)
    cfgs = [r.spec.config() for r in probe[:16]]
    t = time.perf_counter()
    tune = ohp.autotune(cfgs, sample=8, steps=2)
    log(f"[serve] autotune: tile {tune['tile']}, chunks {tune['chunks']} "
        f"({time.perf_counter() - t:.1f} s)")
    loop = serve.ServingLoop(tile=tune["tile"], chunks=tune["chunks"], creators=args.serve_creators)
    log(f"[serve] capacity run: {n} requests at t=0")
    rep = loop.run(probe, progress=log)
    cap = rep.summary()
    cap_rps = cap["done"] / cap["makespan"]
    log(f"[serve] capacity {cap_rps:.3f} requests/s, "
        f"{cap['particle_steps_per_s']:.3e} particle-steps/s")
    levels = {"saturated": cap}
    for k, load in enumerate((0.5, 0.8, 1.0)):
        reqs = serve.sample_requests(n, rate=load * cap_rps, seed=8 + k)
        log(f"[serve] offered load {int(load * 100)}%: {n} requests at "
            f"{load * cap_rps:.3f}/s")
        levels[f"load{int(load * 100)}"] = loop.run(reqs, progress=log).summary()
    top = levels["load100"]
    return {
        "metric": "serving loop: p50/p99 time-to-completion and aggregate particle-steps/s",
        "value": top["particle_steps_per_s"], "unit": "particle-steps/s", "n_gpus": 1,
        "higher_is_better": True, "dtype": "f32+f64", "data": "synthetic (seeded requests)",
        "config": {"workload": f"serving{n}_mixed", "grids": [list(g) for g in serve.GRIDS],
                   "grid_weights": list(serve.GRID_WEIGHTS), "ppc_range": [16, 64],
                   "steps_range": [200, 2000], "budget_bytes": serve.HBM_BUDGET,
                   "tile": list(tune["tile"]), "chunks": tune["chunks"],
                   "creators": args.serve_creators, "batch": loop.batch},
        "capacity_requests_per_s": cap_rps, "levels": levels,
    }


def run_copy_calibration(gib: float = 2.0, reps: int = 5):
    """A plain device copy of known size: its FETCH_SIZE / WRITE_SIZE counters
    calibrate the per-dispatch HBM byte counts of the PIC kernels."""
    import torch

    n = int(gib * (1 << 30) // 8)
    a = torch.ones(n, dtype=torch.float64, device="cuda:0")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t) / reps
    return {"copy_bytes_read": 8 * n, "copy_bytes_written": 8 * n, "reps": reps,
            "copy_GBps": 16 * n / dt / 1e9}


def run_field_solve(n=(256, 256, 256), reps: int = 20):
    """The field solve alone (five FFT passes with the fused Poisson z pass,
    E, the energy sums) on a random charge grid: host-timed mean over `reps`
    solves; algorithmic bytes per FFT pass are one read and one write of the
    nx*ny*nz FP64 array."""
    import ohhelppic3d_amd as ohp

    n = tuple(int(v) for v in n)
    cfg = ohp.SimConfig(n=n, qm=(-1.0, 0.01), q_macro=(-1.0, 1.0), capacity=(64, 64), dt=0.1)
    with ohp.Simulation(cfg) as sim:
        sim.seed_uniform_maxwell([64, 64], [1.0, 0.1])
        rng = np.random.default_rng(1)
        sim.put_grid(ohp.FIELD_RHO, rng.standard_normal(n[0] * n[1] * n[2]))
        sim.field_solve()
        t = time.perf_counter()
        for _ in range(reps):
            sim.field_solve()
        sim.synchronize()
        ms = (time.perf_counter() - t) / reps * 1e3
    ng = n[0] * n[1] * n[2]
    return {"workload": f"field_solve_{n[0]}x{n[1]}x{n[2]}", "reps": reps, "ms_per_solve": ms,
            "fft_pass_bytes": 16 * ng, "efield_bytes": 40 * ng}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="uniform256_ppc64_fp64", choices=list(WORKLOADS))
    ap.add_argument("--slack", type=float, default=0.15)
    ap.add_argument("--tile", type=int, nargs=3, default=(0, 0, 0),
                    help="cells per tile (0 0 0: the library's default, 8^3 at 256^3)")
    ap.add_argument("--chunks", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--field-solve", action="store_true",
                    help="time the field solve alone on a random charge grid")
    ap.add_argument("--field-grid", type=int, nargs=3, default=(256, 256, 256))
    ap.add_argument("--field-reps", type=int, default=20)
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-small", action="store_true",
                    help="skip the GPU run of the CPU baseline's workload (32^3 x 16 ppc), "
                         "which follows the CPU baseline")
    ap.add_argument("--ensemble-instances", type=int, default=128)
    ap.add_argument("--ensemble-steps", type=int, default=500)
    ap.add_argument("--ensemble-seq-steps", type=int, default=10)
    ap.add_argument("--serve", action="store_true",
                    help="run the serving-loop workload instead (its own JSON line)")
    ap.add_argument("--serve-requests", type=int, default=48)
    ap.add_argument("--serve-creators", type=int, default=2,
                    help="worker threads creating admitted runs (0: in the loop)")
    args = ap.parse_args()
    if args.copy_calibration:
        print(json.dumps(run_copy_calibration()), flush=True)
        return
    if args.field_solve:
        print(json.dumps(run_field_solve(args.field_grid, args.field_reps)), flush=True)
        return
    if args.serve:
        print(json.dumps(run_serving_bench(args)), flush=True)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the library has no multi-GPU path "
              "(one MI355X holds the whole 256^3 domain)", file=sys.stderr)
        sys.exit(2)
    res = run_gpu(args)
    if not args.no_mixed and args.workload == "uniform256_ppc64_fp64":
        mixed = {}
        for name in ("uniform256_ppc64_mixed", "hump256_ppc64_mixed"):
            log(f"[bench] mixed-precision workload {name}")
            r = run_gpu(args, name)
            mixed[name] = {"value": r["value"], "ms_per_step": r["ms_per_step"],
                           "vs_fp64": r["value"] / res["value"], "roofline": r["roofline"],
                           "energy": r["energy"]}
        log("[bench] mixed vs FP64 one-step error at 128^3")
        mixed["error_vs_fp64"] = mixed_error_vs_fp64(args)
        res["mixed"] = mixed
    if not args.no_ensemble and args.workload == "uniform256_ppc64_fp64":
        res["ensemble"] = run_ensemble_bench(args)
    if not args.no_cpu:
        log("[bench] timing the CPU baseline (oracle)")
        res["cpu_baseline"] = run_cpu_baseline(args.cpu_seconds)
    else:
        res["cpu_baseline"] = None
    if not args.no_small and res["cpu_baseline"] and args.workload == "uniform256_ppc64_fp64":
        # the CPU baseline's own workload (32^3 x 16 ppc x 2 species, FP64) on
        # the GPU, stepped one call per step like the headline
        log("[bench] the CPU baseline's workload on the GPU")
        small = argparse.Namespace(**vars(args))
        small.steps, small.warmup = 100, 5
        r = run_gpu(small, "uniform32_ppc16_fp64", timing=False)
        cpu = res["cpu_baseline"]
        res["cpu_workload_on_gpu"] = {
            "workload": "uniform32_ppc16_fp64", "value": r["value"], "unit": "particle-steps/s",
            "ms_per_step": r["ms_per_step"], "steps": small.steps,
            "vs_cpu_baseline": r["value"] / cpu["value"]}
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
