"""Benchmark: 2D 2SHOC RK4 in double precision on a 1024 x 1024 grid.

Workload (the reference's 2D headline case): 1024 x 1024 points on
[-51.2, 51.2)^2 (h = 0.1), a = 1, s = -1, V = 0, MSD boundaries, chunks of
50 RK4 steps, initial state tanh(r) e^{i theta} + 1e-3 complex uniform noise
(seed 0).  The reference states k = 0.3 h^2 for this case, which is beyond
RK4's stability limit for 2SHOC (0.265 h^2/a; the solution overflows within
~100 steps, which --workload 2d-matrix reports); the timed run uses the reference driver's own
rule k = 0.8 * 0.75 * h^2 / (2 sqrt 2) = 0.2121 h^2 (NLSEmagic2D.m:131-136).

One bench "step" = one chunk of 50 RK4 steps on the device-resident state
(U, V already in HBM when the timed region starts).  `value` is grid-point
RK4 steps per second.  Also reported:
  roofline      mean launch time of the stage kernel (HIP events around the
                timed hipGraph replays) vs the algorithmic HBM bytes of one
                kernel per RK4 stage (56/88/88/72 = 304 B per point-step);
                the default kernel (row-sweep LDS tiles of 8 rows) fuses two
                stages per launch and moves 128 B per point-step, so `frac`
                can exceed 1 and `design_frac` gives the fraction on its own
                128 B
  cpu_baseline  the reference's own CPU integrator (oracle/_ref, built from
                src/NLSEmagic2D/NLSEmagic2D_CPU.c), single thread, on a
                bounded sample of the same workload
Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
from __future__ import annotations

import argparse
import dataclasses
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


def stage_bytes(precision):
    """Algorithmic HBM bytes per point of stages 1..4 for complex elements of
    e bytes and a real V of v bytes: 3e+v, 5e+v, 5e+v, 4e+v, plus the low
    part of a compensated state read and written by stage 4."""
    e, v = (16, 8) if precision == 2 else (8, 4)
    lo = 2 * e if precision == 3 else 0
    return (3 * e + v, 5 * e + v, 5 * e + v, 4 * e + v + lo)


def design_bytes(precision, fusion):
    """HBM bytes per point and RK4 step of 1, 2 or 4 stages per launch."""
    e, v = (16, 8) if precision == 2 else (8, 4)
    lo = 2 * e if precision == 3 else 0
    return {1: sum(stage_bytes(precision)), 2: 7 * e + 2 * v + lo, 4: 2 * e + v + lo}[fusion]
HBM_PEAK_GBS = 8000.0
PROFILE_TRAFFIC = os.path.join(ROOT, "profiles", "rows2d_2shoc_f64_1024_traffic.json")
KERNEL_FAMILY = {0: "fused2d_kernel", 1: "rows2d_kernel", 2: "strip2d_kernel"}


def kernel_name(fusion, kind, rows):
    fam = KERNEL_FAMILY[kind]
    if fusion == 1:
        return f"{fam}<double,2SHOC,s,s,rows={rows}> (4 launches per RK4 step)"
    if fusion == 2:
        return f"{fam}<double,2SHOC,1,2,rows={rows}> + <..,3,4,..> (2 launches per RK4 step)"
    return f"{fam}<double,2SHOC,1,4,rows={rows}> (1 launch per RK4 step)"


def initial_state():
    x = -51.2 + H * np.arange(N)
    X, Y = np.meshgrid(x, x, indexing="ij")  # (N, M): second index contiguous, like meshgrid's y
    r = np.hypot(X, Y)
    rng = np.random.default_rng(0)
    noise = rng.uniform(-1, 1, (N, M)) + 1j * rng.uniform(-1, 1, (N, M))
    U = np.tanh(r) * np.exp(1j * np.arctan2(Y, X)) + 1e-3 * noise
    return U, np.zeros((N, M))


def host_cpu():
    """CPU model and thread counts of the host the CPU baseline runs on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0))}


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
    return {"value": N * M * steps / dt, "unit": "grid-point-steps/s", "cores": 1, **host_cpu(),
            "kind": "reference" if use_ref else "port",
            "sample": f"{steps} RK4 steps of the same 1024x1024 2SHOC FP64 problem ({dt:.1f} s, 1 thread)"}, W, steps


def parity_vs_cpu(U, V, k, W, steps, precision=2):
    """The same `steps` steps from the same state on the GPU (drop-in
    solver path), compared with the CPU reference's result W."""
    import nlsemagic_amd as nm
    sol = nm.Solver((N, M), 2, precision, 2, -1.0, 1.0, H * H, k)
    sol.set_state(U, V)
    sol.step(steps)
    G = sol.state()
    sol.close()
    m0, mc, mg = (float(np.sum(np.abs(A) ** 2)) for A in (U, W, G))
    return {"steps": steps, "bit_identical": bool(np.array_equal(G.view(np.uint64), W.view(np.uint64))),
            "rel_l2_vs_cpu_reference": float(np.linalg.norm(G - W) / np.linalg.norm(W)),
            "mass_drift_gpu": mg / m0 - 1, "mass_drift_cpu_reference": mc / m0 - 1}


def batched_problems(nprob=512, seed=1):
    """The batched workload: per problem, in this order from one PCG64 stream
    seeded with 1: side n = round(exp(U[ln 64, ln 512])), s in {-1, +1},
    a ~ U[0.5, 1], omega ~ U[0.05, 0.2].  Square n x n grid, h = 0.1, centred;
    V = -omega^2 (x^2 + y^2); U0 = exp(-r^2/2) (s = +1) or tanh(r) e^{i theta}
    (s = -1); k = 0.8 * (3/4) h^2 / (2 sqrt 2 a) (NLSEmagic2D.m:129-136)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(nprob):
        n = int(round(np.exp(rng.uniform(np.log(64), np.log(512)))))
        s = float(rng.choice([-1.0, 1.0]))
        a = float(rng.uniform(0.5, 1.0))
        w = float(rng.uniform(0.05, 0.2))
        out.append(dict(n=n, s=s, a=a, w=w, k=0.8 * 0.75 * H * H / (2 * np.sqrt(2) * a)))
    return out


def batched_state(p):
    n = p["n"]
    x = H * (np.arange(n) - n / 2)
    X, Y = np.meshgrid(x, x, indexing="ij")
    r2 = X * X + Y * Y
    V = -p["w"] ** 2 * r2
    if p["s"] > 0:
        U = np.exp(-r2 / 2) + 0j
    else:
        U = np.tanh(np.sqrt(r2)) * np.exp(1j * np.arctan2(Y, X))
    return U, V


def run_batched(args):
    """512 uneven 2D 2SHOC problems, chunks of 100 steps, FP64 and FP32c."""
    import nlsemagic_amd as nm
    probs = batched_problems()
    shapes = [(p["n"], p["n"]) for p in probs]
    npts = sum(p["n"] ** 2 for p in probs)
    chunk = 100
    nchunks = args.steps
    res = {}
    finals = {}
    for prec in (2, 3):
        B = nm.Batch(shapes, [p["s"] for p in probs], [p["a"] for p in probs], [H * H] * len(probs),
                     [p["k"] for p in probs], method=2, precision=prec, BC=1)
        states = [batched_state(p) for p in probs]
        for i, (U, V) in enumerate(states):
            B.set_state(i, U, V)
        if args.batch_config:
            f, kind, t, g = (int(x) for x in args.batch_config.split(","))
            B.set_kernel(kind, t)
        else:
            B.autotune(3)
            f, kind, t, g = B.config()
        B.configure(f, t, g, chunk)
        B.step(chunk)  # warm-up: captures the chunk graph
        B.sync()
        for i, (U, V) in enumerate(states):
            B.set_state(i, U)
        lat = []
        for _ in range(nchunks):
            lat.append(B.step_timed(chunk))
        diag = B.diagnostics()
        finals[prec] = [B.state(i) for i in range(len(probs))] if prec == 3 or True else None
        m0 = np.array([np.sum(np.abs(U) ** 2) for U, _ in states])
        res[prec] = dict(rate=npts * chunk * nchunks / (sum(lat) * 1e-3), chunk_ms=lat, fusion=f, kind=kind, tile=t, grid=g,
                         mass=diag[:, 0], mass_drift=diag[:, 0] / m0 - 1, nonfinite=int(diag[:, 2].sum()))
        B.close()
    # FP32c vs FP64 per problem (the FP64 batch is the FP64 reference at full size)
    rel = np.array([np.linalg.norm(a - b) / np.linalg.norm(b) for a, b in zip(finals[3], finals[2])])
    mass_rel = res[3]["mass"] / res[2]["mass"] - 1
    # pin the full-size FP64 reference: the smallest problems against the CPU oracle, bit for bit
    from oracle import oracle as O
    order = np.argsort([p["n"] for p in probs])[:3]
    pinned = 0
    for i in order:
        p = probs[i]
        U, V = batched_state(p)
        exp = O.take_steps(U, V, p["s"], p["a"], H * H, 1, chunk * nchunks, p["k"], 2)
        pinned += int(np.array_equal(exp.view(np.uint64), finals[2][i].view(np.uint64)))
    pct = lambda x, q: float(np.percentile(x, q))
    line = {
        "metric": "RK4 grid-point updates/s & % HBM peak; batched-solve p50/p99 latency",
        "value": res[3]["rate"], "unit": "grid-point-steps/s", "n_gpus": 1, "steps": nchunks, "warmup": 1,
        "ms_per_step": float(np.mean(res[3]["chunk_ms"])), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 (compensated hi+lo state)", "data": "synthetic",
        "config": {"workload": "batched: 512 uneven 2D 2SHOC problems (sides 64-512, seed 1), chunks of 100 steps",
                   "points": npts, "fp64_rate": res[2]["rate"], "fp32c_over_fp64": res[3]["rate"] / res[2]["rate"],
                   "fp32c_rel_l2_max": float(rel.max()), "fp32c_rel_l2_p50": pct(rel, 50),
                   "fp32c_rel_l2_p99": pct(rel, 99),
                   "fp32c_problems_rel_l2_le_1e-6": f"{int(np.sum(rel <= 1e-6))}/{len(rel)}",
                   "fp32c_worst_problem": {"index": int(np.argmax(rel)), **{k: probs[int(np.argmax(rel))][k]
                                                                           for k in ("n", "s", "a", "w")},
                                           "max_abs_V": float(probs[int(np.argmax(rel))]["w"] ** 2 * 2 *
                                                              (H * probs[int(np.argmax(rel))]["n"] / 2) ** 2)},
                   "fp32c_mass_vs_fp64_max": float(np.abs(mass_rel).max()),
                   "fp64_mass_drift_max": float(np.abs(res[2]["mass_drift"]).max()),
                   "fp32c_mass_drift_max": float(np.abs(res[3]["mass_drift"]).max()),
                   "nonfinite": res[2]["nonfinite"] + res[3]["nonfinite"],
                   "chunk_ms_p50": {k: pct(v["chunk_ms"], 50) for k, v in (("fp64", res[2]), ("fp32c", res[3]))},
                   "chunk_ms_p99": {k: pct(v["chunk_ms"], 99) for k, v in (("fp64", res[2]), ("fp32c", res[3]))},
                   "tuned": {k: {"stages_per_launch": v["fusion"], "kernel_kind": v["kind"], "rows": v["tile"],
                                 "grid": v["grid"]} for k, v in (("fp64", res[2]), ("fp32c", res[3]))},
                   "fp64_pinned_to_oracle": f"{pinned}/3 smallest problems bit-identical after {chunk * nchunks} steps"},
    }
    print(json.dumps(line))


def dropin_host_rate(nm, U, V, k, calls=5):
    """The drop-in boundary as the reference's callers use it: one
    NLSE2D_TAKE_STEPS_HIP call per chunk of 50 steps on host buffers (upload
    over PCIe into the device buffers the library keeps from the previous
    call of this shape, the steps, download), timed on the host over `calls`
    calls after one warm-up call.  Not the headline value, which is measured
    with the state resident in HBM."""
    import ctypes as C
    L = nm.lib()
    dp = C.POINTER(C.c_double)
    ur = np.ascontiguousarray(U.real)
    ui = np.ascontiguousarray(U.imag)
    v = np.ascontiguousarray(V)
    outr = np.empty(U.shape)
    outi = np.empty(U.shape)
    args = (ur.ctypes.data_as(dp), ui.ctypes.data_as(dp), v.ctypes.data_as(dp), outr.ctypes.data_as(dp),
            outi.ctypes.data_as(dp), -1.0, 1.0, H * H, 2, CHUNK, k, N, M, 2)

    def call():
        rc = L.NLSE2D_TAKE_STEPS_HIP(*args)
        if rc != 0:
            raise RuntimeError(f"NLSE2D_TAKE_STEPS_HIP failed: {L.nlse_last_error().decode()}")

    call()
    t0 = time.perf_counter()
    for _ in range(calls):
        call()
    ms = (time.perf_counter() - t0) * 1e3 / calls
    return {"what": f"one NLSE2D_TAKE_STEPS_HIP call per chunk of {CHUNK} steps on host buffers "
                    f"(upload, steps, download; device buffers kept between calls), mean of {calls} calls",
            "ms_per_call": round(ms, 3), "rate": N * M * CHUNK / (ms * 1e-3)}


def run_2d_matrix(args):
    """The 2D configuration matrix on the headline grid (1024^2, MSD, the
    headline's initial state): {CD, 2SHOC} x {FP64, FP32, compensated FP32}
    at k = 0.3 h^2 (CD, the reference's stated step) and 0.2121 h^2 (2SHOC,
    the driver's rule), 2 warm-up and 20 timed chunks of 50 steps each (device
    time, HIP events); per variant the mean launch time, the fraction of HBM
    peak on the bytes its fused kernel moves, and the diagnostics.  Also 2SHOC
    at k = 0.3 h^2, beyond RK4's stability bound: the first step at which the
    state stops being finite.  CPU: the reference integrators (1 thread) on a
    bounded sample of each method."""
    import nlsemagic_amd as nm
    from oracle import oracle as O
    U, V = initial_state()
    res = {}
    final = {}
    for m, kh2 in ((1, 0.3), (2, K_RULE)):
        for prec in (2, 1, 3):
            sol = nm.Solver((N, M), m, prec, 2, -1.0, 1.0, H * H, kh2 * H * H)
            sol.set_graph(CHUNK)
            if args.fusion is not None:
                sol.set_fusion(args.fusion)
            if args.tile is not None or args.kernel != 1:
                sol.set_kernel(args.kernel, args.tile or 8)
            sol.set_state(U, V)
            sol.step(2 * CHUNK)
            sol.sync()
            ms = sol.step_timed(20 * CHUNK)
            fz, kind, rows = sol.config()
            launches = sol.profile_launches(10)
            t_launch = float(np.mean(launches)) * 1e-3
            mass, maxre, bad = sol.diagnostics()
            final[(m, prec)] = sol.state()
            dbytes = design_bytes(prec, fz)
            name = f"{'CD' if m == 1 else '2SHOC'}_{ {2: 'fp64', 1: 'fp32', 3: 'fp32c'}[prec] }"
            res[name] = {"rate": N * M * 20 * CHUNK / (ms * 1e-3), "k_over_h2": round(kh2, 6),
                         "launch_us": round(t_launch * 1e6, 2), "stages_per_launch": fz, "kernel_kind": kind,
                         "rows": rows, "design_bytes_per_point_step": dbytes,
                         "design_frac": round(dbytes * N * M / (t_launch * (4 // fz)) / 1e9 / HBM_PEAK_GBS, 4),
                         "mass_over_initial": mass / float(np.sum(np.abs(U) ** 2)), "nonfinite": int(bad)}
            sol.close()
    # single precision against the FP64 state after the same 1100 steps (the
    # FP64 state is bit-identical to the CPU reference's, as the tests show)
    for m in (1, 2):
        ref = final[(m, 2)]
        for prec in (1, 3):
            name = f"{'CD' if m == 1 else '2SHOC'}_{ {1: 'fp32', 3: 'fp32c'}[prec] }"
            res[name]["rel_l2_vs_fp64_after_1100_steps"] = float(np.linalg.norm(final[(m, prec)] - ref) /
                                                                 np.linalg.norm(ref))
    # 2SHOC at the stated k = 0.3 h^2: RK4 is unstable there for this stencil
    sol = nm.Solver((N, M), 2, 2, 2, -1.0, 1.0, H * H, 0.3 * H * H)
    sol.set_state(U, V)
    blow = None
    for c in range(200):
        sol.step(10)
        if sol.diagnostics()[2] > 0:
            blow = 10 * (c + 1)
            break
    sol.close()
    cpu = {}
    use_ref = O.ref(2) is not None
    for m, kh2 in ((1, 0.3), (2, K_RULE)):
        t0 = time.perf_counter()
        steps = 0
        while time.perf_counter() - t0 < 5.0:
            O.take_steps(U, V, -1.0, 1.0, H * H, 2, 10, kh2 * H * H, m, use_ref=use_ref)
            steps += 10
        cpu["CD" if m == 1 else "2SHOC"] = N * M * steps / (time.perf_counter() - t0)
    head = res["2SHOC_fp64"]
    line = {
        "metric": "RK4 grid-point updates/s & % HBM peak; batched-solve p50/p99 latency",
        "value": head["rate"], "unit": "grid-point-steps/s", "n_gpus": 1, "steps": 20, "warmup": 2,
        "ms_per_step": N * M * CHUNK / head["rate"] * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64 (value); f64/f32/f32c (matrix)", "data": "synthetic",
        "config": {"workload": "2D RK4 1024x1024, MSD, dark vortex + 1e-3 noise (seed 0): CD and 2SHOC x FP64/FP32/"
                               "FP32c, chunks of 50 steps",
                   "matrix": res,
                   "twoshoc_k0.3h2_first_nonfinite_step": blow,
                   "cpu_rate_1thread": cpu, "cpu_kind": "reference" if use_ref else "port",
                   "speedup_vs_cpu": {k.split("_")[0] + "_fp64": res[k]["rate"] / cpu[k.split("_")[0]]
                                      for k in ("CD_fp64", "2SHOC_fp64")}},
    }
    print(json.dumps(line))


def run_parity_2d(args):
    """The 2D configuration's whole run against the reference CPU integrator
    (oracle/_ref, src/NLSEmagic2D/NLSEmagic2D_CPU.c, one thread): 1024^2, MSD,
    the dark vortex + 1e-3 noise (seed 0), 5000 steps in chunks of 50, CD at
    the stated k = 0.3 h^2 and 2SHOC at the drivers' k = 0.2121 h^2 (0.3 h^2
    overflows 2SHOC).  FP64 on the GPU against the CPU's FP64 result (bits,
    relative L2), FP32 and compensated FP32 against it (relative L2), and the
    mass sum |U|^2 h^2 drift of each.  About 4 minutes of CPU time."""
    import nlsemagic_amd as nm
    from oracle import oracle as O
    U, V = initial_state()
    steps, chunk = args.steps if args.steps != 40 else 5000, CHUNK
    use_ref = O.ref(2) is not None
    mass = lambda A: float(np.sum(np.abs(A) ** 2)) * H * H
    m0 = mass(U)
    out = {}
    for method, kf in ((1, 0.3), (2, K_RULE)):
        k = kf * H * H
        name = "CD" if method == 1 else "2SHOC"
        t0 = time.perf_counter()
        W = U
        for done in range(0, steps, chunk):
            W = O.take_steps(W, V, -1.0, 1.0, H * H, 2, chunk, k, method, use_ref=use_ref)
            if (done + chunk) % 500 == 0:
                print(f"parity-2d: {name} CPU {done + chunk}/{steps} steps", file=sys.stderr, flush=True)
        cpu_s = time.perf_counter() - t0
        res = {"k_over_h2": kf, "cpu_seconds_1thread": cpu_s, "cpu_kind": "reference" if use_ref else "port",
               "mass_initial": m0, "mass_drift_cpu": mass(W) / m0 - 1}
        for prec, tag in ((2, "fp64"), (1, "fp32"), (3, "fp32c")):
            sol = nm.Solver((N, M), method, prec, 2, -1.0, 1.0, H * H, k)
            sol.set_graph(chunk)
            sol.set_state(U, V)
            sol.step(steps)
            G = sol.state()
            sol.close()
            r = {"rel_l2_vs_cpu_fp64": float(np.linalg.norm(G - W) / np.linalg.norm(W)),
                 "max_abs_diff": float(np.max(np.abs(G - W))), "mass_drift": mass(G) / m0 - 1,
                 "finite": bool(np.all(np.isfinite(G)))}
            if prec == 2:
                r["bit_identical"] = bool(np.array_equal(G.view(np.uint64), W.view(np.uint64)))
            res[tag] = r
            print(f"parity-2d: {name} {tag} done", file=sys.stderr, flush=True)
        out[name] = res
    print(json.dumps({"workload": "parity-2d: 1024x1024 MSD dark vortex + 1e-3 noise (seed 0), "
                                  f"{steps} steps in chunks of {chunk}, GPU vs the reference CPU integrator",
                      "steps": steps, "cpu_host": host_cpu(), "results": out}))


def vortex_rings_3d(n=256, L=25.6, nrings=4, radius=8.0, seed=0):
    """Product of dark vortex rings tanh(d/sqrt 2) e^{+-i phi} on [-L, L)^3:
    d is the distance to the ring core, phi the poloidal angle around it.
    Per ring, from one PCG64 stream seeded with `seed`: the unit normal
    (three normal draws, normalised), the centre (three uniform draws in
    [-6, 6]) and the winding (+1 or -1).  Returns U with shape (n, n, n)."""
    rng = np.random.default_rng(seed)
    h = 2 * L / n
    x = -L + h * np.arange(n)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    U = np.ones((n, n, n), complex)
    for _ in range(nrings):
        nv = rng.normal(size=3)
        nv /= np.linalg.norm(nv)
        c = rng.uniform(-6, 6, 3)
        w = float(rng.choice([-1.0, 1.0]))
        vx, vy, vz = X - c[0], Y - c[1], Z - c[2]
        z = vx * nv[0] + vy * nv[1] + vz * nv[2]
        rho = np.sqrt(np.maximum(vx * vx + vy * vy + vz * vz - z * z, 0.0))
        d = np.sqrt((rho - radius) ** 2 + z * z)
        U *= np.tanh(d / np.sqrt(2)) * np.exp(1j * w * np.arctan2(z, rho - radius))
    return U, h


VORT_BYTES = 96  # per point: velocity pass 16 B in + 24 B out, curl pass 24 B in + 32 B out


def time_vorticity(sol, h, veps=1e-3, reps=20):
    """Device vorticity of the resident 3D state (rmc_vorticity3D): wall time
    per call over `reps` back-to-back calls into device arrays (the two
    kernels run in a fraction of a millisecond, launch costs are a few us),
    and the host result of one more call for the parity check."""
    import nlsemagic_amd as nm
    n = int(np.prod(sol.shape))
    outs = [nm.DeviceArray(n) for _ in range(4)]
    ptrs = [d.address for d in outs]
    sol.vorticity_device(1 / (2 * h), veps, *ptrs)
    sol.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        sol.vorticity_device(1 / (2 * h), veps, *ptrs)
    sol.sync()
    dt = (time.perf_counter() - t0) / reps
    for d in outs:
        d.free()
    return dict(ms=dt * 1e3, rate=n / dt, gbs=VORT_BYTES * n / dt / 1e9, host=sol.vorticity(1 / (2 * h), veps))


PROFILE_VORT_TRAFFIC = os.path.join(ROOT, "profiles", "vorticity3d_256_traffic.json")


def run_vorticity(args):
    """3D vorticity (rmc_vorticity3D) of a 256^3 FP64 state: four dark
    vortex rings of radius 8 made on the device by the ring generator
    (NLSE_add_vortex_ring, opt_2D = 0) on the unit background, orientations,
    centres and charges from seed 0 (per ring: thetax and thetay uniform in
    [0, pi), the centre uniform in [-6, 6]^3, the charge +-1), then the
    vorticity (lh2 = 1/(2h), veps = 1e-3) into device arrays: 5 untimed and
    `steps` timed calls between HIP events on the solver's stream.  Checked
    bit for bit against the oracle; CPU: the oracle's C restatement, 1
    thread (the reference's own file needs MATLAB's mex.h)."""
    import nlsemagic_amd as nm
    from oracle import oracle as O
    n, h = 256, 0.2
    shape = (n, n, n)
    origin = (-25.6, -25.6, -25.6)
    sol = nm.Solver(shape, 2, 2, 1, -1.0, 1.0, h * h, 0.1 * h * h)
    rng = np.random.default_rng(0)
    t0 = time.perf_counter()
    sol.fill(1.0)
    for _ in range(4):
        tx, ty = rng.uniform(0, np.pi, 2)
        c = rng.uniform(-6, 6, 3)
        m = int(rng.choice([-1, 1]))
        sol.add_vortex_ring(m, -1.0, *c, tx, ty, 0.0, 0.0, 0.0, 0.0, 0.0, h, 1.0, -1.0, origin, 8.0)
    ic_s = time.perf_counter() - t0
    npts = n ** 3
    outs = [nm.DeviceArray(npts) for _ in range(4)]
    ptrs = [d.address for d in outs]
    lh2 = 1 / (2 * h)
    sol.vorticity_timed(lh2, 1e-3, 5, *ptrs)
    ms = sol.vorticity_timed(lh2, 1e-3, args.steps, *ptrs)
    got = [d.numpy().reshape(shape) for d in outs]
    for d in outs:
        d.free()
    U = sol.state()
    sol.close()
    t0 = time.perf_counter()
    exp = O.vorticity3d(U, lh2, 1e-3)
    cpu_s = time.perf_counter() - t0
    same = all(np.array_equal(a.view(np.uint64), b.view(np.uint64)) for a, b in zip(got, exp))
    alg, need = 96, 48
    achieved = alg * npts / (ms * 1e-3) / 1e9
    traffic = None
    if os.path.exists(PROFILE_VORT_TRAFFIC):
        with open(PROFILE_VORT_TRAFFIC) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    line = {
        "metric": "RK4 grid-point updates/s & % HBM peak; batched-solve p50/p99 latency",
        "value": npts / (ms * 1e-3), "unit": "vorticity points/s", "n_gpus": 1, "steps": args.steps, "warmup": 5,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "3D vorticity (rmc_vorticity3D), 256^3 FP64, four dark vortex rings made on the "
                               "device (seed 0)",
                   "ic_four_rings_s": ic_s, "bit_identical_to_oracle": bool(same)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "vort_march_kernel<double,16,16>", "launch_us": round(ms * 1e3, 2),
                     "algorithmic_bytes_per_point": alg, "design_bytes_per_point": need,
                     "design_frac": round(need * npts / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
        "cpu_baseline": {"value": npts / cpu_s, "unit": "vorticity points/s", "cores": 1, **host_cpu(), "kind": "port",
                         "sample": f"the whole 256^3 state once ({cpu_s:.2f} s, oracle C restatement)"},
    }
    print(json.dumps(line))


def run_grid(args):
    """1D (N = 2048, CD, MSD, dark soliton) or 3D (256^3, Dirichlet, vortex
    rings) single-problem workloads: device rate for FP64 / FP32 / FP32c and
    the CPU integrator on a bounded sample."""
    import nlsemagic_amd as nm
    from oracle import oracle as O
    out = {}
    if args.workload == "1d":
        n = 2048
        h = 100.0 / n
        x = -50 + h * np.arange(n)
        rng = np.random.default_rng(0)
        U = np.tanh(x / np.sqrt(2)) + 1e-3 * (rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n))
        V = np.zeros(n)
        shape, BC, methods, chunk, nchunks = (n,), 2, (1,), 100, 100
        ks = {1: 0.4 * h * h}
        desc = "1D CD RK4, N=2048 on [-50,50), MSD, dark soliton + 1e-3 noise (seed 0), 100 chunks of 100 steps"
    else:
        U, h = vortex_rings_3d()
        V = np.zeros(U.shape)
        shape, BC, methods, chunk, nchunks = U.shape, 1, (1, 2), 50, 20
        ks = {1: 0.8 * h * h / (3 * np.sqrt(2)), 2: 0.8 * 0.75 * h * h / (3 * np.sqrt(2))}
        desc = "3D RK4, 256^3 on [-25.6,25.6)^3, Dirichlet, four vortex rings (seed 0), 20 chunks of 50 steps"
    npts = int(np.prod(shape))
    for m in methods:
        for prec in (2, 1, 3):
            sol = nm.Solver(shape, m, prec, BC, -1.0, 1.0, h * h, ks[m])
            if args.kernel != 1 or args.tile is not None:  # 3d: plane-march rows; 1d: steps per launch
                sol.set_kernel(args.kernel, args.tile or 8)
            sol.set_state(U, V)
            sol.step(chunk)
            sol.sync()
            sol.set_state(U)
            lat = [sol.step_timed(chunk) for _ in range(nchunks)]
            mass, _, bad = sol.diagnostics()
            W = sol.state() if prec == 2 or args.workload == "1d" else None
            out[(m, prec)] = dict(rate=npts * chunk * nchunks / (sum(lat) * 1e-3), chunk_ms=float(np.median(lat)),
                                  mass=mass, bad=bad, W=W)
            if args.workload == "3d" and prec == 2 and m == methods[-1]:
                vort = time_vorticity(sol, h)
                vort["W"] = W
            sol.close()
    # CPU integrator on a bounded sample (1 thread): the reference for 1D, the restatement for 3D
    use_ref = args.workload == "1d" and O.ref(1) is not None
    cpu = {}
    full_size = {}
    for m in methods:
        steps = 2000 if args.workload == "1d" else 2
        t0 = time.perf_counter()
        Wc = O.take_steps(U, V, -1.0, 1.0, h * h, BC, steps, ks[m], m, use_ref=use_ref)
        cpu[m] = npts * steps / (time.perf_counter() - t0)
        if args.workload == "3d":
            # the same steps at full size on the GPU, FP64, against the CPU result
            sol = nm.Solver(shape, m, 2, BC, -1.0, 1.0, h * h, ks[m])
            sol.set_state(U, V)
            sol.step(steps)
            G = sol.state()
            sol.close()
            full_size["CD" if m == 1 else "2SHOC"] = {
                "steps": steps, "bit_identical": bool(np.array_equal(G.view(np.uint64), Wc.view(np.uint64))),
                "rel_l2": float(np.linalg.norm(G - Wc) / np.linalg.norm(Wc))}
    m0 = float(np.sum(np.abs(U) ** 2))
    fp64 = {m: out[(m, 2)] for m in methods}
    line = {
        "metric": "RK4 grid-point updates/s & % HBM peak; batched-solve p50/p99 latency",
        "value": out[(methods[-1], 2)]["rate"], "unit": "grid-point-steps/s", "n_gpus": 1, "steps": nchunks,
        "warmup": 1, "ms_per_step": out[(methods[-1], 2)]["chunk_ms"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": desc,
                   "rates": {f"{'CD' if m == 1 else '2SHOC'}_{ {2: 'fp64', 1: 'fp32', 3: 'fp32c'}[p] }": v["rate"]
                             for (m, p), v in out.items()},
                   "mass_drift_fp64": {('CD' if m == 1 else '2SHOC'): fp64[m]["mass"] / m0 - 1 for m in methods},
                   "nonfinite": int(sum(v["bad"] for v in out.values())),
                   "cpu_rate_1thread": {('CD' if m == 1 else '2SHOC'): cpu[m] for m in methods},
                   "cpu_host": host_cpu(),
                   "cpu_kind": "reference" if use_ref else "port"},
    }
    if args.workload == "3d":
        line["config"]["fp64_vs_cpu_full_size"] = full_size
        t0 = time.perf_counter()
        exp = O.vorticity3d(vort["W"], 1 / (2 * h), 1e-3)
        cpu_v = npts / (time.perf_counter() - t0)
        same = all(np.array_equal(a.view(np.uint64), b.view(np.uint64)) for a, b in zip(vort["host"], exp))
        line["config"]["vorticity"] = {
            "what": "rmc_vorticity3D of the 2SHOC FP64 state after the timed chunks (lh2 = 1/(2h), veps = 1e-3)",
            "ms_per_call": vort["ms"], "points_per_s": vort["rate"],
            "algorithmic_GBps": vort["gbs"], "hbm_frac": vort["gbs"] / HBM_PEAK_GBS,
            "bytes_per_point": VORT_BYTES, "bit_identical_to_oracle": bool(same),
            "cpu_points_per_s_1thread": cpu_v}
    if args.workload == "1d":
        exp = O.take_steps(U, V, -1.0, 1.0, h * h, BC, chunk * (nchunks + 0), ks[1], 1)
        line["config"]["fp64_bit_identical_to_oracle"] = bool(
            np.array_equal(exp.view(np.uint64), out[(1, 2)]["W"].view(np.uint64)))
        line["config"]["fp32c_rel_l2"] = float(np.linalg.norm(out[(1, 3)]["W"] - exp) / np.linalg.norm(exp))
    print(json.dumps(line))


def run_3d_counters(args):
    """A short run for hardware-counter passes (rocprofv3 --pmc): --steps RK4
    steps of one 3D variant (--method, --precision, --fusion, --tile) on the
    256^3 workload's state, with no timing and no CPU work."""
    import nlsemagic_amd as nm
    U, h = vortex_rings_3d()
    k = (0.8 if args.method == 1 else 0.8 * 0.75) * h * h / (3 * np.sqrt(2))
    sol = nm.Solver(U.shape, args.method, args.precision, 1, -1.0, 1.0, h * h, k)
    if args.fusion is not None:
        sol.set_fusion(args.fusion)
    sol.set_kernel(args.kernel, args.tile or 4)
    sol.set_state(U, np.zeros(U.shape))
    sol.step(args.steps)
    sol.sync()
    print(json.dumps({"workload": "3d-counters", "method": args.method, "precision": args.precision,
                      "config": sol.config(), "steps": args.steps}))
    sol.close()


def run_serving(args):
    """Poisson request streams (seed 2) at each --qps rate through the
    continuous-batching server, each compared with the same requests one at
    a time through the plain FP64 path.  The mix draws Dirichlet or MSD
    boundaries and, for half the requests, a harmonic trap (request_mix
    variety) unless --plain-mix is given."""
    import nlsemagic_amd as nm
    from nlsemagic_amd import serving as S
    warm = S.request_mix(12, qps=1000.0, seed=99)  # loads every kernel variant before timing
    for r in warm:
        r.steps = 20
        r.shape = tuple(min(n, 64) for n in r.shape)
        r.init = S.default_ic(r.shape, 0.1)
    S.Server(max_active=8, chunk_steps=10).run(warm)
    variety = not args.plain_mix
    points = {}
    for qps in [float(q) for q in str(args.qps).split(",")]:
        reqs = S.request_mix(args.requests, qps=qps, seed=2, variety=variety)
        srv = S.Server(budget_bytes=200e9, max_active=16, chunk_steps=200)
        stats = srv.run(reqs)
        served = {r.rid: r.result for r in reqs}
        lat_tagged = {p: [r.latency for r in reqs if r.precision == p] for p in (2, 3)}
        base = [S.Request(**{f.name: getattr(r, f.name) for f in dataclasses.fields(S.Request)
                             if f.name not in ("admitted", "finished", "result")}) for r in reqs]
        plain = S.run_one_at_a_time(base, precision=2)
        same = sum(int(np.array_equal(served[b.rid].view(np.uint64), b.result.view(np.uint64)))
                   for b in base if b.precision == 2)
        n64 = sum(1 for b in base if b.precision == 2)
        # each request's service time alone (FP64 drop-in: state built, steps,
        # download) against its latency in the server: the tail is set by the
        # largest requests' own compute when the ratio stays near 1
        alone = {b.rid: b.finished - b.admitted for b in base}
        worst = max(reqs, key=lambda r: r.latency)
        ratio64 = [r.latency / alone[r.rid] for r in reqs if r.precision == 2]
        points[qps] = {"server": stats, "one_at_a_time_fp64": plain,
                       "latency_p50_s": stats["latency_p50_s"], "latency_p99_s": stats["latency_p99_s"],
                       "latency_p50_by_precision": {("fp64" if p == 2 else "fp32c"): float(np.percentile(v, 50))
                                                    for p, v in lat_tagged.items() if v},
                       "msd_requests": sum(1 for r in reqs if r.BC == 2),
                       "trap_requests": sum(1 for r in reqs if r.make_V is not None),
                       "fp64_requests_identical_to_one_at_a_time": f"{same}/{n64}",
                       "slowest_request": {"rid": worst.rid, "shape": list(worst.shape), "method": worst.method,
                                           "precision": worst.precision, "steps": worst.steps,
                                           "latency_s": worst.latency, "alone_fp64_s": alone[worst.rid]},
                       "fp64_latency_over_alone_p50_p99": [float(np.percentile(ratio64, 50)),
                                                           float(np.percentile(ratio64, 99))]}
        print(f"serving at {qps} QPS: p50 {stats['latency_p50_s']:.3f} s, p99 {stats['latency_p99_s']:.3f} s",
              file=sys.stderr, flush=True)
    q0 = next(iter(points))
    first = points[q0]["server"]
    rates = ", ".join(f"{q:g}" for q in points)
    line = {
        "metric": "RK4 grid-point updates/s & % HBM peak; batched-solve p50/p99 latency",
        "value": first["throughput"], "unit": "grid-point-steps/s", "n_gpus": 1, "steps": args.requests,
        "warmup": 1, "ms_per_step": first["makespan_s"] * 1e3 / args.requests, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64 + f32c (per request)", "data": "synthetic",
        "config": {"workload": f"serving: {args.requests} Poisson requests at {rates} QPS (seed 2), mixed 1D/2D/3D "
                               "CD/2SHOC FP64/FP32c, 100-20000 steps, "
                               + ("Dirichlet or MSD, half with a harmonic trap, " if variety else "")
                               + "200 GB budget, initial states built on the device from parameters "
                               f"(value: the {q0:g} QPS stream)",
                   "qps_points": {f"{q:g}": v for q, v in points.items()},
                   "latency_p50_s": points[q0]["latency_p50_s"], "latency_p99_s": points[q0]["latency_p99_s"]},
    }
    print(json.dumps(line))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40, help="timed chunks of 50 RK4 steps")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--graph", type=int, default=CHUNK, help="RK4 steps per hipGraph replay (0 = off)")
    ap.add_argument("--fusion", type=int, default=None, choices=(1, 2, 4),
                    help="RK4 stages per kernel launch (results are identical)")
    ap.add_argument("--precision", type=int, default=2, choices=(1, 2, 3),
                    help="2d workload: 2 FP64 (the headline), 1 FP32, 3 compensated FP32")
    ap.add_argument("--method", type=int, default=2, choices=(1, 2), help="3d-counters: 1 CD, 2 2SHOC")
    ap.add_argument("--kernel", type=int, default=1, choices=(0, 1, 2),
                    help="2D stage kernel: 0 LDS tiles, 1 row-sweep tiles, 2 register strips")
    ap.add_argument("--tile", type=int, default=None,
                    help="rows per tile (strip segment) of that kernel (2d default 8); "
                         "3d: 4 (default) or 8 rows of the plane-marching tiles")
    ap.add_argument("--autotune", action="store_true", help="pick fusion and tile with the built-in tuner")
    ap.add_argument("--batch-config", default="", metavar="F,KIND,ROWS,GRID",
                    help="batched: stages per launch, kernel kind, rows, grid (skips the tuner)")
    ap.add_argument("--qps", default="4,16", help="serving: Poisson arrival rates, comma-separated")
    ap.add_argument("--plain-mix", action="store_true",
                    help="serving: Dirichlet, V = 0 requests only (no MSD or trap requests)")
    ap.add_argument("--requests", type=int, default=48, help="serving: number of requests")
    ap.add_argument("--workload", default="2d",
                    choices=("2d", "2d-matrix", "batched", "1d", "3d", "3d-counters", "serving", "vorticity",
                             "parity-2d"),
                    help="2d: the 1024^2 FP64 2SHOC headline; batched: 512 uneven problems, FP32c vs FP64")
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
    if args.workload == "serving":
        run_serving(args)
        return
    if args.workload == "parity-2d":
        run_parity_2d(args)
        return
    if args.workload in ("1d", "3d"):
        run_grid(args)
        return
    if args.workload == "3d-counters":
        run_3d_counters(args)
        return
    if args.workload == "2d-matrix":
        run_2d_matrix(args)
        return
    if args.workload == "vorticity":
        if args.steps == ap.get_default("steps"):
            args.steps = 20
        run_vorticity(args)
        return
    if args.workload == "batched":
        if args.steps == ap.get_default("steps"):
            args.steps = 10
        run_batched(args)
        return

    import nlsemagic_amd as nm

    U, V = initial_state()
    k = K_RULE * H * H
    sol = nm.Solver((N, M), 2, args.precision, 2, -1.0, 1.0, H * H, k)
    sol.set_graph(args.graph)
    if args.fusion is not None:
        sol.set_fusion(args.fusion)
    if args.tile is not None or args.kernel != 1:
        sol.set_kernel(args.kernel, args.tile or 8)
    sol.set_state(U, V)
    if args.autotune:
        sol.autotune(5)
    args.fusion, args.kernel, args.tile = sol.config()
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

    # dominant kernel: the timed region above is hipGraph replays of nothing but
    # the stage kernels, back to back on the solver's stream, so its HIP-event
    # time over the launch count is the mean launch duration
    nlaunch = total_steps * (4 // args.fusion)
    t_launch = dev_ms * 1e-3 / nlaunch
    alg_bytes = sum(stage_bytes(args.precision))
    bytes_launch = alg_bytes * N * M / (4 // args.fusion)
    achieved = bytes_launch / t_launch / 1e9
    dbytes = design_bytes(args.precision, args.fusion)
    traffic = valu = None
    if os.path.exists(PROFILE_TRAFFIC) and args.precision == 2:
        with open(PROFILE_TRAFFIC) as f:
            prof = json.load(f).get(f"fusion{args.fusion}", {})
        if prof.get("tile") == args.tile and prof.get("kernel", 0) == args.kernel:
            traffic = prof.get("hbm_bytes_per_launch")
            if prof.get("valu_insts_per_launch"):
                # VALU issue: wave64 instructions (FP64 or 32-bit) issue at one
                # per CU per clock; 256 CUs at 2.4 GHz
                cap = 256 * 2.4e9 * t_launch
                valu = {"insts_per_launch": prof["valu_insts_per_launch"], "issue_frac_at_2p4GHz":
                        round(prof["valu_insts_per_launch"] / cap, 4)}
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": kernel_name(args.fusion, args.kernel, args.tile),
                "launch_us": round(t_launch * 1e6, 2),
                "algorithmic_bytes_per_point_step": alg_bytes,
                "design_bytes_per_point_step": dbytes,
                "design_GBps": round(dbytes * N * M / (t_launch * (4 // args.fusion)) / 1e9, 1),
                "design_frac": round(dbytes * N * M / (t_launch * (4 // args.fusion)) / 1e9 / HBM_PEAK_GBS, 4),
                "launch_us_event_per_kernel": [round(x * 1e3, 2) for x in sol.profile_launches(20)],
                "valu": valu}
    sol.close()

    dropin = dropin_host_rate(nm, U, V, k) if args.precision == 2 else None

    cpu = parity = None
    if not args.no_cpu and args.precision == 2:
        cpu, W_cpu, cpu_steps = cpu_baseline(U, V, k)
        parity = parity_vs_cpu(U, V, k, W_cpu, cpu_steps)

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
        "dtype": {2: "f64", 1: "f32", 3: "f32 (compensated hi+lo state)"}[args.precision],
        "data": "synthetic",
        "config": {"workload": "2D 2SHOC RK4 FP64, 1024x1024, MSD, dark vortex + 1e-3 noise (seed 0)",
                   "grid": [N, M], "h": H, "k_over_h2": round(K_RULE, 6), "chunk_size": CHUNK,
                   "rk4_steps_timed": total_steps, "device_ms": dev_ms, "hip_graph_steps": args.graph,
                   "stages_per_launch": args.fusion, "kernel_kind": args.kernel, "tile_rows": args.tile,
                   "mass_drift": mass_drift, "finite": finite, "dropin_host_buffers": dropin},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if cpu:
        line["config"]["speedup_vs_cpu_reference"] = value / cpu["value"]
        line["config"]["parity_vs_cpu_reference"] = parity
    print(json.dumps(line))


if __name__ == "__main__":
    main()
