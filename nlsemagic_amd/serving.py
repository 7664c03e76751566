"""Continuous-batching request loop over device-resident NLSE solvers.

A request is one chunked solve (dimension, grid, CD/2SHOC, FP64 or
compensated FP32, step count, initial state and potential).  The server:

- admits arrived requests in arrival order, with backfill, while the sum of
  their device footprints (nlse_footprint_bytes) stays under the memory
  budget and fewer than ``max_active`` are running;
- keeps every admitted request's U, V and RK4 work arrays resident on the
  GPU until it finishes (one resident solver per request, each on its own
  HIP stream, so the work of different requests overlaps on the device);
- advances each active request by chunks of ``chunk_steps`` RK4 steps, one
  hipGraph replay per chunk, and enqueues the next chunk of a request as soon
  as its previous one completes (non-blocking stream queries), so requests
  that arrive while others run join the running mix without waiting for it
  to drain;
- batches the 2D requests: the active 2D requests of one method and
  precision take their chunks together as a group (nlse_group_*), one
  persistent launch per pair of RK4 stages over the tiles of all of them
  (each in its own buffers), the chunk being the smallest number of steps
  any of them has left (at most ``chunk_steps``); a 2D request that is ready
  while others of its kind are still working waits for them, so a request
  that arrives joins the next chunk of its kind's group;
- releases a finished request's memory at once and records its latency
  (finish - arrival).

Each request's result is bit-for-bit the result of running it alone: the
kernels are deterministic and requests share no state.

The reference has no serving layer; its drivers call one MEX integrator per
chunk in a loop (src/NLSEmagic2D/NLSEmagic2D.m:251-289).
"""
from __future__ import annotations

import dataclasses
import threading
import time
from concurrent.futures import ThreadPoolExecutor
from collections import deque
from typing import Callable, List, Optional

import numpy as np

from . import Group, Solver, footprint_bytes


@dataclasses.dataclass
class Request:
    rid: int
    arrival: float            # seconds after the start of the run
    shape: tuple              # (N,), (N, M) or (L, N, M)
    method: int               # 1 CD, 2 2SHOC
    precision: int            # 2 double, 3 compensated single
    BC: int
    s: float
    a: float
    h2: float
    k: float
    steps: int
    make_state: Optional[Callable[[], tuple]] = None  # -> (U0, V) host arrays, uploaded
    init: Optional[Callable[[Solver], None]] = None   # builds the state on the device instead
    make_V: Optional[Callable[[], np.ndarray]] = None  # with init: a potential uploaded after it (else V = 0)
    admitted: Optional[float] = None
    finished: Optional[float] = None
    result: Optional[np.ndarray] = None

    @property
    def points(self) -> int:
        return int(np.prod(self.shape))

    @property
    def work(self) -> int:
        """grid-point RK4 steps"""
        return self.points * self.steps

    @property
    def latency(self) -> float:
        return self.finished - self.arrival

    def initialize(self, sol: Solver):
        """Put the initial state into a fresh solver: on the device from the
        request's parameters (init) or from host arrays (make_state)."""
        if self.init is not None:
            self.init(sol)
            if self.make_V is not None:
                sol.set_state(sol.state(), self.make_V())
        else:
            U0, V = self.make_state()
            sol.set_state(U0, V)

    def initial_state(self):
        """(U0, V) on the host, for a path that takes arrays."""
        if self.init is None:
            return self.make_state()
        sol = Solver(self.shape, self.method, 2, self.BC, self.s, self.a, self.h2, self.k)
        try:
            self.init(sol)
            return sol.state(), (self.make_V() if self.make_V is not None else np.zeros(self.shape))
        finally:
            sol.close()


class Server:
    def __init__(self, budget_bytes: float = 200e9, max_active: int = 16, chunk_steps: int = 200,
                 keep_results: bool = True, group_2d: bool = True):
        self.budget = float(budget_bytes)
        self.max_active = int(max_active)
        self.chunk = int(chunk_steps)
        self.keep_results = keep_results
        self.group_2d = bool(group_2d)
        self.peak_bytes = 0
        self.peak_active = 0
        self.groups_formed = 0
        self.grouped_request_chunks = 0

    def _admit(self, r: Request, now: float):
        sol = Solver(r.shape, r.method, r.precision, r.BC, r.s, r.a, r.h2, r.k)
        sol.set_graph(min(self.chunk, r.steps))
        r.initialize(sol)
        r.admitted = now
        return sol

    def run(self, requests: List[Request]) -> dict:
        # Waiter threads sleep on each enqueued chunk's completion (a
        # blocking-sync event; the library call releases the GIL) and wake
        # the loop, which otherwise sleeps until the next arrival; every
        # waiter of a solver or group is joined before that solver or group
        # is closed.
        wake = threading.Event()
        pool = ThreadPoolExecutor(max_workers=self.max_active + 4)
        waiters = {}  # id(solver or group) -> futures of its waiters

        def watch(obj):
            f = pool.submit(obj.wait if isinstance(obj, Solver) else obj.sync)
            f.add_done_callback(lambda _: wake.set())
            waiters.setdefault(id(obj), []).append(f)

        def settle(obj):
            for f in waiters.pop(id(obj), []):
                f.result()

        try:
            return self._run(requests, wake, watch, settle)
        finally:
            pool.shutdown(wait=True)

    def _run(self, requests, wake, watch, settle) -> dict:
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        waiting: List[Request] = []
        active = {}  # rid -> [request, solver, steps_left, footprint]
        groups: List[Group] = []  # running groups of 2D requests
        grouped = {}              # id(group) -> its requests
        in_group = set()          # requests whose solver a running group holds
        used = 0
        t0 = time.perf_counter()
        clock = lambda: time.perf_counter() - t0
        while pending or waiting or active or groups:
            wake.clear()
            now = clock()
            while pending and pending[0].arrival <= now:
                waiting.append(pending.popleft())
            # admission: arrival order, later requests may backfill around one that does not fit
            still = []
            for r in waiting:
                fp = footprint_bytes(r.shape, r.method, r.precision)
                if len(active) < self.max_active and used + fp <= self.budget:
                    active[r.rid] = [r, self._admit(r, clock()), r.steps, fp]
                    used += fp
                else:
                    still.append(r)
            waiting = still
            self.peak_bytes = max(self.peak_bytes, used)
            self.peak_active = max(self.peak_active, len(active))
            progressed = False
            # running groups: a finished one hands its members back
            for g in [g for g in groups if g.done()]:
                settle(g)
                g.close()
                groups.remove(g)
                for rid in grouped.pop(id(g)):
                    in_group.discard(rid)
                progressed = True
            ready = {}  # (method, precision) -> 2D requests ready for a chunk
            busy = set()  # kinds with a 2D request still working on a chunk
            for rid in list(active):
                r, sol, left, fp = active[rid]
                if rid in in_group or not sol.done():
                    if self.group_2d and len(r.shape) == 2:
                        busy.add((r.method, r.precision))
                    continue
                progressed = True
                if left == 0:
                    if self.keep_results:
                        r.result = sol.state()
                    else:
                        sol.sync()
                    r.finished = clock()
                    settle(sol)
                    sol.close()
                    used -= fp
                    del active[rid]
                    continue
                if self.group_2d and len(r.shape) == 2:
                    ready.setdefault((r.method, r.precision), []).append(rid)
                    continue
                n = min(self.chunk, left)
                sol.step(n)
                watch(sol)
                active[rid][2] = left - n
            for kind, rids in ready.items():
                if kind in busy:
                    continue  # wait for the busy ones, then take the next chunk together
                if len(rids) == 1:
                    r, sol, left, fp = active[rids[0]]
                    n = min(self.chunk, left)
                    sol.step(n)
                    watch(sol)
                    active[rids[0]][2] = left - n
                    continue
                n = min([self.chunk] + [active[rid][2] for rid in rids])
                g = Group([active[rid][1] for rid in rids])
                g.step(n)
                watch(g)
                groups.append(g)
                grouped[id(g)] = rids
                for rid in rids:
                    in_group.add(rid)
                    active[rid][2] -= n
                self.groups_formed += 1
                self.grouped_request_chunks += len(rids)
            if not progressed:
                # until a chunk finishes or the next request arrives
                wait = pending[0].arrival - clock() if pending else 0.01
                wake.wait(max(0.0, min(wait, 0.01)))
        makespan = clock()
        return summarize(requests, makespan, peak_bytes=self.peak_bytes, peak_active=self.peak_active,
                         groups_formed=self.groups_formed, grouped_request_chunks=self.grouped_request_chunks)


def run_one_at_a_time(requests: List[Request], precision: Optional[int] = 2) -> dict:
    """The plain path: each request solved alone through the drop-in entry
    point (upload, all steps, download), in arrival order, one at a time.
    ``precision`` forces every request to that precision (None: keep tags)."""
    from . import take_steps
    t0 = time.perf_counter()
    for r in sorted(requests, key=lambda r: r.arrival):
        while time.perf_counter() - t0 < r.arrival:
            time.sleep(min(0.01, r.arrival - (time.perf_counter() - t0)))
        r.admitted = time.perf_counter() - t0
        U0, V = r.initial_state()
        prec = r.precision if precision is None else precision
        if prec == 3:
            prec = 1
        r.result = take_steps(U0, V, r.s, r.a, r.h2, r.BC, r.steps, r.k, precision=prec, method=r.method)
        r.finished = time.perf_counter() - t0
    return summarize(requests, time.perf_counter() - t0)


def summarize(requests: List[Request], makespan: float, **extra) -> dict:
    lat = np.array([r.latency for r in requests])
    work = float(sum(r.work for r in requests))
    out = {"requests": len(requests), "makespan_s": makespan, "throughput": work / makespan,
           "throughput_unit": "grid-point-steps/s", "latency_p50_s": float(np.percentile(lat, 50)),
           "latency_p99_s": float(np.percentile(lat, 99)), "latency_max_s": float(lat.max())}
    out.update(extra)
    return out


def klin(dim: int, method: int, h: float, a: float) -> float:
    """Linear RK4 stability bound used by the reference drivers: h^2/(d sqrt 2 a)
    for CD and 3/4 of it for 2SHOC (NLSEmagic1D.m:86-94, NLSEmagic2D.m:128-136,
    NLSEmagic3D.m:154-161)."""
    return (1.0 if method == 1 else 0.75) * h * h / (dim * np.sqrt(2) * a)


def auto_k(dim: int, method: int, h: float, a: float) -> float:
    """k = 0.8 klin, the reference drivers' default time step."""
    return 0.8 * klin(dim, method, h, a)


def plan_chunks(endtw: float, k: float, numframes: int):
    """The reference drivers' chunk planner (NLSEmagic1D.m:97-116): the end
    time is cut to a multiple of k, the steps are split into `numframes`
    chunks of ceil(steps/numframes) steps, the frame count is reduced when
    the rounding would add more than a chunk, and the few extra steps are
    taken.  Returns (chunk_size, numframes, steps, endt)."""
    endt = endtw - (endtw - np.floor(endtw / k) * k)  # MATLAB mod(endtw, k)
    steps = int(np.floor(endt / k))
    if numframes >= steps:
        numframes = steps
    chunk_size = int(np.ceil(steps / numframes))
    extra = chunk_size * numframes - steps
    if extra > chunk_size:
        numframes = int(np.ceil(steps / chunk_size))
        extra = chunk_size * numframes - steps
    if extra > 0:
        endt = chunk_size * numframes * k
        steps = steps + extra
    return chunk_size, numframes, steps, endt


def request_mix(n: int, qps: float, seed: int = 2, variety: bool = False) -> List[Request]:
    """The serving workload: Poisson arrivals at `qps`, uniform over
    {1D, 2D, 3D} x {CD, 2SHOC} x {compensated FP32, FP64}; sizes log-uniform
    in 1D [4096, 65536] points, 2D [128, 1024]^2, 3D [64, 192]^3; steps
    log-uniform in [100, 20000].  Per request, from one PCG64 stream seeded
    with `seed`, in this order: inter-arrival gap, dim, method, precision,
    size, steps.  s = -1, a = 1, V = 0, Dirichlet boundaries, k = 0.8 klin;
    h = 0.1 (1D, 2D) or 0.2 (3D); the initial state arrives as parameters
    and is built on the device (default_ic): a dark soliton (1D), a dark
    vortex (2D) or a dark vortex ring of radius a quarter of the box (3D).
    variety: per request, from a second stream seeded with (seed, 1), the
    boundary (Dirichlet or MSD, even odds) and, with even odds, a harmonic
    trap V = -omega^2 |x|^2 (omega ~ U[0.05, 0.1], make_V) uploaded after the
    state is built; the first stream's draws are unchanged."""
    rng = np.random.default_rng(seed)
    rng2 = np.random.default_rng([seed, 1])
    t = 0.0
    out = []
    lu = lambda lo, hi: float(np.exp(rng.uniform(np.log(lo), np.log(hi))))
    for i in range(n):
        t += float(rng.exponential(1.0 / qps))
        dim = int(rng.integers(1, 4))
        method = int(rng.integers(1, 3))
        precision = (2, 3)[int(rng.integers(0, 2))]
        if dim == 1:
            shape = (int(round(lu(4096, 65536))),)
        elif dim == 2:
            side = int(round(lu(128, 1024)))
            shape = (side, side)
        else:
            side = int(round(lu(64, 192)))
            shape = (side, side, side)
        steps = int(round(lu(100, 20000)))
        h = 0.1 if dim < 3 else 0.2
        BC, make_V = 1, None
        if variety:
            BC = (1, 2)[int(rng2.integers(0, 2))]
            if rng2.uniform() < 0.5:
                make_V = trap(shape, h, float(rng2.uniform(0.05, 0.1)))
        out.append(Request(i, t, shape, method, precision, BC, -1.0, 1.0, h * h, auto_k(dim, method, h, 1.0),
                           steps, init=default_ic(shape, h), make_V=make_V))
    return out


def trap(shape, h, omega):
    """The harmonic trap V = -omega^2 |x|^2 on the centred grid of spacing h."""
    def make_V():
        axes = [h * (np.arange(n) - (n - 1) / 2) for n in shape]
        return -omega * omega * sum(g * g for g in np.meshgrid(*axes, indexing="ij"))
    return make_V


def default_ic(shape, h):
    """The serving workload's initial state, built on the device by the
    reference's generators on the centred grid of spacing h (s = -1, a = 1,
    background amplitude 1):
      1D  the dark soliton tanh(x/sqrt 2) (NLSEmagic1D_FRS.m:207, c = 0);
      2D  a charge-1 dark vortex at the centre (NLS_add_vortex.m, opt = 0:
          tanh(r/sqrt 2) e^{i theta});
      3D  a charge-1 dark vortex ring of radius a quarter of the box in the
          x-y plane (NLSE_add_vortex_ring.m, opt_2D = 0) on the uniform
          background (NLSE3D.m:405-407)."""
    origin = tuple(-(n - 1) * h / 2 for n in shape)
    if len(shape) == 1:
        return lambda sol: sol.set_exact(2, h, origin, 0.0, a=1.0, s=-1.0, OM=-1.0, c=0.0, x0=0.0)
    if len(shape) == 2:
        o2 = (origin[0], origin[1])

        def init2(sol):
            sol.fill(1.0)
            sol.add_vortex_2d(1, -1.0, 0.0, 0.0, 0.0, 0.0, 0.0, h, 1.0, -1.0, 1.0, o2, 0)
        return init2
    L, N, M = shape
    o3 = (origin[1], origin[2], origin[0])  # (xmin, ymin, zmin): x along N, y along M, z along L

    def init3(sol):
        sol.fill(1.0)
        sol.add_vortex_ring(1, -1.0, 0.0, 0.0, 0.0, 0, 0, 0, 0, 0, 0, 0.0, h, 1.0, -1.0, o3, 0.25 * h * M)
    return init3
