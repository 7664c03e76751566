"""Groups of resident 2D solvers (nlse_group_*): members of different sizes,
parameters and boundaries advanced together by the batched kernels give,
bit for bit, their results alone, across odd and even group chunks mixed
with steps of their own; and the serving loop forms such groups."""
from __future__ import annotations

import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    import nlsemagic_amd as nm
    if nm.device_count() < 1:
        pytest.fail("no HIP device visible")
    return nm


PROBS = [((70, 90), 2, -1.0, 0.9, 0.01), ((33, 130), 1, 1.0, 0.7, 0.02), ((128, 128), 3, -1.0, 1.0, 0.01),
         ((3, 5), 2, -1.0, 0.8, 0.04)]


def _state(shape, seed):
    rng = np.random.default_rng(seed)
    U = (1 + 0.2 * rng.uniform(-1, 1, shape)) + 0.2j * rng.uniform(-1, 1, shape)
    V = 0.3 * rng.uniform(-1, 1, shape)
    return U, V


@pytest.mark.parametrize("prec", [2, 3])
@pytest.mark.parametrize("method", [1, 2])
def test_group_equals_alone(hip, method, prec):
    print("CASE group", method, prec, flush=True)
    sys.stdout.flush()
    mk = lambda shape, BC, s, a, h2: hip.Solver(shape, method, prec, BC, s, a, h2, 0.1 * h2 / a)
    grouped, alone = [], []
    for i, (shape, BC, s, a, h2) in enumerate(PROBS):
        U, V = _state(shape, 10 * i + method)
        for lst in (grouped, alone):
            sol = mk(shape, BC, s, a, h2)
            sol.set_state(U, V)
            lst.append(sol)
    plan = [("group", 5), ("own", 3), ("group", 4), ("group", 1), ("own", 2)]
    for what, n in plan:
        if what == "group":
            g = hip.Group(grouped)
            g.step(n)
            g.close()
        else:
            for sol in grouped:
                sol.step(n)
    total = sum(n for _, n in plan)
    for sol_g, sol_a in zip(grouped, alone):
        sol_a.step(total)
        a, b = sol_g.state(), sol_a.state()
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
        sol_g.close()
        sol_a.close()


def test_group_rejects_mixed_members(hip):
    a = hip.Solver((20, 20), 2, 2, 1, -1.0, 1.0, 0.01, 0.001)
    b = hip.Solver((20, 20), 1, 2, 1, -1.0, 1.0, 0.01, 0.001)
    c = hip.Solver((50,), 2, 2, 1, -1.0, 1.0, 0.01, 0.001)
    for members in ([a, b], [a, c], [a, a]):
        with pytest.raises(hip.HipError):
            hip.Group(members)
    for s in (a, b, c):
        s.close()


def test_server_groups_2d_requests(hip):
    """Eight 2D requests of one kind arriving together: the server advances
    them in groups, and each result equals the request alone."""
    from nlsemagic_amd import serving as S
    reqs = S.request_mix(8, qps=1000.0, seed=4)
    for i, r in enumerate(reqs):
        r.shape = (64 + 17 * i, 80 + 9 * i)
        r.method, r.precision, r.BC = 2, 2 if i % 2 == 0 else 3, (1, 2)[i % 2]
        r.h2 = 0.01
        r.k = S.auto_k(2, 2, 0.1, 1.0)
        r.steps = 50 + 13 * i
        r.init = S.default_ic(r.shape, 0.1)
    srv = S.Server(max_active=16, chunk_steps=32)
    stats = srv.run(reqs)
    assert stats["groups_formed"] > 0 and stats["grouped_request_chunks"] >= 2 * stats["groups_formed"]
    for r in reqs:
        sol = hip.Solver(r.shape, r.method, r.precision, r.BC, r.s, r.a, r.h2, r.k)
        r.initialize(sol)
        sol.step(r.steps)
        alone = sol.state()
        sol.close()
        assert np.array_equal(alone.view(np.uint64), r.result.view(np.uint64)), r.rid
