"""Serving loop: host logic on the CPU; the GPU tests show every request's
result is the same alone and in the mix (bit for bit), and that admission
respects the memory budget."""
from __future__ import annotations

import numpy as np
import pytest

from nlsemagic_amd import serving as S


def test_auto_k_matches_the_reference_rule():
    # NLSEmagic2D.m:129-136: klin = h^2/(2 sqrt 2 a), 3/4 of it for 2SHOC, k = 0.8 klin
    assert np.isclose(S.auto_k(2, 1, 0.1, 1.0), 0.8 * 0.01 / (2 * np.sqrt(2)))
    assert np.isclose(S.auto_k(2, 2, 0.1, 1.0), 0.8 * 0.75 * 0.01 / (2 * np.sqrt(2)))
    assert np.isclose(S.auto_k(1, 1, 0.1, 0.5), 0.8 * 0.01 / (np.sqrt(2) * 0.5))
    assert np.isclose(S.auto_k(3, 2, 0.2, 1.0), 0.8 * 0.75 * 0.04 / (3 * np.sqrt(2)))


def test_auto_k_is_inside_the_rk4_stability_limit():
    """k * (largest |eigenvalue|) <= 2 sqrt 2 for every dim and method: the
    largest CD symbol is 4 d a / h^2 and 2SHOC's is 4/3 of it."""
    for d in (1, 2, 3):
        for m in (1, 2):
            lam = 4 * d / 0.01 * (4 / 3 if m == 2 else 1)
            assert S.auto_k(d, m, 0.1, 1.0) * lam < 2 * np.sqrt(2)


def test_plan_chunks_known_answers():
    # 10,000 steps in 100 frames: no adjustment
    assert S.plan_chunks(10000 * 0.001, 0.001, 100)[:3] == (100, 100, 10000)
    # more frames than steps: one step per frame
    c, f, n, _ = S.plan_chunks(0.05, 0.01, 50)
    assert (c, n) == (1, f)
    # 1000 steps in 300 frames: chunks of 4, frames cut to 250 so no chunk is wasted
    c, f, n, endt = S.plan_chunks(1.0, 0.001, 300)
    assert (c, f, n) == (4, 250, 1000) and np.isclose(endt, 1.0)
    # end time 10.3 with k = 1/4: cut to 10.25 (41 steps); 10 frames would need
    # chunks of 5 and 9 extra steps (more than a chunk), so 9 frames of 5 steps
    # are taken and the end time moves to 11.25
    assert S.plan_chunks(10.3, 0.25, 10) == (5, 9, 45, 11.25)


def test_request_mix_is_reproducible_and_in_range():
    a = S.request_mix(64, qps=4.0, seed=2)
    b = S.request_mix(64, qps=4.0, seed=2)
    assert [(r.arrival, r.shape, r.steps, r.method, r.precision) for r in a] == \
           [(r.arrival, r.shape, r.steps, r.method, r.precision) for r in b]
    for r in a:
        assert 100 <= r.steps <= 20000
        n = r.points
        assert {1: (4096, 65536), 2: (128 ** 2, 1024 ** 2), 3: (64 ** 3, 192 ** 3)}[len(r.shape)][0] <= n
        assert n <= {1: 65536, 2: 1024 ** 2, 3: 192 ** 3}[len(r.shape)]
    gaps = np.diff([0.0] + [r.arrival for r in a])
    assert 0.1 < gaps.mean() < 0.5  # Poisson at 4 QPS: mean gap 0.25 s
    assert {len(r.shape) for r in a} == {1, 2, 3} and {r.precision for r in a} == {2, 3}


def _small_mix(seed=5, n=10):
    reqs = S.request_mix(n, qps=200.0, seed=seed)
    for r in reqs:
        d = len(r.shape)
        r.shape = {1: (4100,), 2: (70, 90), 3: (20, 22, 24)}[d]
        r.steps = 30 + 17 * r.rid
        r.init = S.default_ic(r.shape, 0.1 if d < 3 else 0.2)
    return reqs


@pytest.mark.gpu
def test_each_result_is_the_same_alone_and_in_the_mix():
    import nlsemagic_amd as nm
    reqs = _small_mix()
    stats = S.Server(max_active=4, chunk_steps=16).run(reqs)
    assert stats["peak_active"] <= 4
    for r in reqs:
        sol = nm.Solver(r.shape, r.method, r.precision, r.BC, r.s, r.a, r.h2, r.k)
        r.initialize(sol)
        sol.step(r.steps)
        alone = sol.state()
        sol.close()
        print("CASE", r.rid, r.shape, r.method, r.precision, r.steps, flush=True)
        assert np.array_equal(alone.view(np.uint64), r.result.view(np.uint64)), r.rid


@pytest.mark.gpu
def test_admission_respects_the_memory_budget():
    import nlsemagic_amd as nm
    reqs = _small_mix(seed=6, n=8)
    biggest = max(nm.footprint_bytes(r.shape, r.method, r.precision) for r in reqs)
    srv = S.Server(budget_bytes=1.5 * biggest, max_active=16, chunk_steps=8)
    stats = srv.run(reqs)
    assert srv.peak_bytes <= 1.5 * biggest
    assert all(r.result is not None and np.all(np.isfinite(r.result)) for r in reqs)
    assert stats["requests"] == len(reqs)


@pytest.mark.gpu
def test_config_sizes_with_msd_and_traps_alone_and_in_the_mix():
    """Requests at the serving configuration's sizes (1D 4k-64k points, 2D
    128^2-1024^2, 3D 64^3-192^3, CD and 2SHOC, FP64 and compensated FP32),
    with MSD boundaries and harmonic traps in the mix (steps capped at 240 to
    bound the run): each result is bit for bit the result of the request
    run alone."""
    import nlsemagic_amd as nm
    reqs = S.request_mix(10, qps=100.0, seed=2, variety=True)
    for r in reqs:
        r.steps = min(r.steps, 240)
    assert {r.BC for r in reqs} == {1, 2} and {len(r.shape) for r in reqs} == {1, 2, 3}
    stats = S.Server(max_active=16, chunk_steps=64).run(reqs)
    assert stats["requests"] == len(reqs)
    for r in reqs:
        print("CASE", r.rid, r.shape, r.method, r.precision, r.BC, r.steps, flush=True)
        sol = nm.Solver(r.shape, r.method, r.precision, r.BC, r.s, r.a, r.h2, r.k)
        r.initialize(sol)
        sol.step(r.steps)
        alone = sol.state()
        sol.close()
        assert np.array_equal(alone.view(np.uint64), r.result.view(np.uint64)), r.rid
