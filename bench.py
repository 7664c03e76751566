"""Benchmark: 2D 2SHOC RK4 in double precision on a 1024 x 1024 grid.

Workload (the reference's 2D headline case): 1024 x 1024 points on
[-51.2, 51.2)^2 (h = 0.1), a = 1, s = -1, V = 0, MSD boundaries, chunks of
50 RK4 steps, initial state tanh(r) e^{i theta} + 1e-3 complex uniform noise
(seed 0).  The reference states k = 0.3 h^2 for this case, which is beyond
RK4's stability limit for 2SHOC (0.265 h^2/a; the solution overflows within
~100 steps); the timed run uses the reference driver's own
rule k = 0.8 * 0.75 * h^2 / (2 sqrt 2) = 0.2121 h^2 (NLSEmagic2D.m:131-136).

One bench "step" = one chunk of 50 RK4 steps on the device-resident state
(U, V already in HBM when the timed region starts).  `value` is grid-point
RK4 steps per second.  Also reported:
  roofline      per-kernel HIP-event timing vs the algorithmic HBM bytes of
                one fused kernel per RK4 stage (56/88/88/72 = 304 B per
                point-step); the default kernel fuses all four stages in LDS
                and moves 40 B per point-step, so `frac` can exceed 1 and
                `design_frac` gives the fraction on its own 40 B
  cpu_baseline  the reference's own CPU integrator (oracle/_ref, built from
                src/NLSEmagic2D/NLSEmagic2D_CPU.c), single thread, on a
                bounded sample of the same workload
Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
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

N = M = 1024
H = 0.1
CHUNK = 50
K_RULE = 0.8 * 0.75 / (2 * np.sqrt(2))  # k / h^2, reference auto rule for 2SHOC 2D
STAGE_BYTES = (56, 88, 88, 72)  # algorithmic HBM bytes per point, stages 1..4 (FP64)
HBM_PEAK_GBS = 8000.0
PROFILE_TRAFFIC = os.path.join(ROOT, "profiles", "fused2d_2shoc_f64_1024_traffic.json")
# HBM bytes per point and RK4 step that each fusion depth must move:
#  1: four stage launches (56+88+88+72); 2: stages 1-2 read U,V write Y_2,ktot
#  (56) then stages 3-4 read Y_2,U,V,ktot write U (72); 4: read U,V write U
DESIGN_BYTES = {1: 304, 2: 128, 4: 40}
KERNEL_NAME = {1: "fused2d_kernel<double,2SHOC,16,s,s> (4 launches per RK4 step)",
               2: "fused2d_kernel<double,2SHOC,16,1,2> + <..,3,4> (2 launches per RK4 step)",
               4: "fused2d_kernel<double,2SHOC,16,1,4> (1 launch per RK4 step)"}


def initial_state():
    x = -51.2 + H * np.arange(N)
    X, Y = np.meshgrid(x, x, indexing="ij")  # (N, M): second index contiguous, like meshgrid's y
    r = np.hypot(X, Y)
    rng = np.random.default_rng(0)
    noise = rng.uniform(-1, 1, (N, M)) + 1j * rng.uniform(-1, 1, (N, M))
    U = np.tanh(r) * np.exp(1j * np.arctan2(Y, X)) + 1e-3 * noise
    return U, np.zeros((N, M))


def cpu_baseline(U, V, k, budget_s=12.0):
    """Reference CPU integrator, 1 thread, chunks of 10 steps until budget_s."""
    from oracle import oracle as O
    use_ref = O.ref(2) is not None
    prev = os.sched_getaffinity(0)
    try:
        os.sched_setaffinity(0, {sorted(prev)[0]})
        steps = 0
        t0 = time.perf_counter()
        W = U
        while time.perf_counter() - t0 < budget_s:
            W = O.take_steps(W, V, -1.0, 1.0, H * H, 2, 10, k, 2, use_ref=use_ref)
            steps += 10
        dt = time.perf_counter() - t0
    finally:
        os.sched_setaffinity(0, prev)
    return {"value": N * M * steps / dt, "unit": "grid-point-steps/s", "cores": 1,
            "kind": "reference" if use_ref else "port",
            "sample": f"{steps} RK4 steps of the same 1024x1024 2SHOC FP64 problem ({dt:.1f} s, 1 thread)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40, help="timed chunks of 50 RK4 steps")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--graph", type=int, default=CHUNK, help="RK4 steps per hipGraph replay (0 = off)")
    ap.add_argument("--fusion", type=int, default=2, choices=(1, 2, 4),
                    help="RK4 stages per kernel launch (results are identical)")
    ap.add_argument("--tile", type=int, default=8, choices=(8, 16, 32), help="tile height in rows")
    ap.add_argument("--autotune", action="store_true", help="pick fusion and tile with the built-in tuner")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run a plain device copy of BYTES (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_probe:
        import nlsemagic_amd as nm
        ms = nm.copy_probe(args.copy_probe, reps=5)
        print(json.dumps({"copy_probe_bytes": args.copy_probe, "ms": ms,
                          "GBps_read_plus_write": 2 * args.copy_probe / ms / 1e6}))
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (got --gpus {args.gpus}); "
              "the integrator does not shard a grid across GPUs", file=sys.stderr)
        sys.exit(2)

    import nlsemagic_amd as nm

    U, V = initial_state()
    k = K_RULE * H * H
    sol = nm.Solver((N, M), 2, 2, 2, -1.0, 1.0, H * H, k)
    sol.set_graph(args.graph)
    sol.set_fusion(args.fusion)
    sol.set_tile(args.tile)
    sol.set_state(U, V)
    if args.autotune:
        args.fusion, args.tile, _ = sol.autotune(5)
    mass0 = float(np.sum(np.abs(U) ** 2))

    for _ in range(args.warmup):
        sol.step(CHUNK)
    sol.sync()
    t0 = time.perf_counter()
    dev_ms = sol.step_timed(CHUNK * args.steps)
    wall = time.perf_counter() - t0
    total_steps = CHUNK * args.steps
    value = N * M * total_steps / wall
    Uend = sol.state()
    mass_drift = float(np.sum(np.abs(Uend) ** 2)) / mass0 - 1.0
    finite = bool(np.all(np.isfinite(Uend)))

    # per-kernel timing (a HIP event around every launch, on the solver stream)
    ms_launch = sol.profile_launches(20)
    t_step = sum(ms_launch) * 1e-3
    bytes_step = sum(STAGE_BYTES) * N * M
    achieved = bytes_step / t_step / 1e9
    design_bytes = DESIGN_BYTES[args.fusion]
    traffic = None
    if os.path.exists(PROFILE_TRAFFIC):
        with open(PROFILE_TRAFFIC) as f:
            traffic = json.load(f).get(f"fusion{args.fusion}", {}).get("hbm_bytes_per_launch")
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": KERNEL_NAME[args.fusion],
                "algorithmic_bytes_per_point_step": sum(STAGE_BYTES),
                "launch_ms": [round(x, 5) for x in ms_launch],
                "design_bytes_per_point_step": design_bytes,
                "design_GBps": round(design_bytes * N * M / t_step / 1e9, 1),
                "design_frac": round(design_bytes * N * M / t_step / 1e9 / HBM_PEAK_GBS, 4)}
    sol.close()

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(U, V, k)

    line = {
        "metric": "RK4 grid-point updates/s & % HBM peak; batched-solve p50/p99 latency",
        "value": value,
        "unit": "grid-point-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "2D 2SHOC RK4 FP64, 1024x1024, MSD, dark vortex + 1e-3 noise (seed 0)",
                   "grid": [N, M], "h": H, "k_over_h2": round(K_RULE, 6), "chunk_size": CHUNK,
                   "rk4_steps_timed": total_steps, "device_ms": dev_ms, "hip_graph_steps": args.graph,
                   "stages_per_launch": args.fusion, "tile_rows": args.tile,
                   "mass_drift": mass_drift, "finite": finite},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if cpu:
        line["config"]["speedup_vs_cpu_reference"] = value / cpu["value"]
    print(json.dumps(line))


if __name__ == "__main__":
    main()
