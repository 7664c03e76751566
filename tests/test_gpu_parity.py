"""HIP integrators vs the CPU oracle and the reference's golden vectors.

Every call goes through the C ABI (NLSE{1,2,3}D_TAKE_STEPS_HIP[_F]).  The bar
is bit-exact equality: the kernels keep the reference's operand order with FP
contraction off, so double results match the reference CPU integrators bit
for bit, and single-precision results match the reference's single-precision
integrators bit for bit.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

from conftest import bits_equal, golden_cases, rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    import nlsemagic_amd as nm
    if nm.device_count() < 1:
        pytest.fail("no HIP device visible")
    return nm


def _announce(*what):
    print("CASE", *what, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("case", golden_cases(), ids=lambda c: c[0]["name"])
def test_hip_matches_golden(hip, case):
    c, U, V, expected = case
    _announce(c["name"])
    got = hip.take_steps(U, V, c["s"], c["a"], c["h2"], c["BC"], c["chunk_size"], c["k"],
                         precision=c["precision"], method=c["method"])
    assert bits_equal(got, expected), c["name"]


def _cases():
    out = []
    seed = 1000
    shapes = {1: [(4,), (7,), (300,), (4096,), (4097,), (5000,), (513,)],
              2: [(3, 3), (17, 65), (33, 130), (50, 64), (18, 7)],
              3: [(3, 3, 3), (5, 7, 66), (9, 6, 65), (4, 17, 10)]}
    bcs = {1: [1, 2, 3, 4, 0], 2: [1, 2, 3, 0], 3: [1, 2, 3, 0]}
    for dim in (1, 2, 3):
        for shape in shapes[dim]:
            for method in (1, 2):
                for prec in (2, 1):
                    for BC in bcs[dim]:
                        seed += 1
                        out.append((dim, shape, method, prec, BC, seed))
    return out


@pytest.mark.parametrize("dim,shape,method,prec,BC,seed", _cases())
def test_hip_matches_oracle(hip, oracle, dim, shape, method, prec, BC, seed):
    _announce(dim, shape, method, prec, BC, seed)
    rng = np.random.default_rng(seed)
    U = (0.9 + 0.3 * rng.uniform(-1, 1, shape)) + 1j * (0.3 * rng.uniform(-1, 1, shape))
    V = 0.5 * rng.uniform(-1, 1, shape)
    h2 = 0.04
    k = 0.1 * h2
    s = -1.0 if BC != 3 else 0.8
    chunk = 4
    exp = oracle.take_steps(U, V, s, 0.9, h2, BC, chunk, k, method, precision=prec)
    got = hip.take_steps(U, V, s, 0.9, h2, BC, chunk, k, precision=prec, method=method)
    assert np.all(np.isfinite(exp))
    assert bits_equal(got, exp), f"max |diff| {np.max(np.abs(got - exp))}"


def test_hip_does_not_modify_inputs(hip):
    rng = np.random.default_rng(5)
    U = rng.uniform(-1, 1, (20, 30)) + 1j * rng.uniform(-1, 1, (20, 30))
    V = rng.uniform(-1, 1, (20, 30))
    U0, V0 = U.copy(), V.copy()
    hip.take_steps(U, V, -1.0, 1.0, 0.01, 2, 3, 0.001, precision=2, method=2)
    assert np.array_equal(U, U0) and np.array_equal(V, V0)


def test_real_input_gets_zero_imaginary_part(hip, oracle):
    """A real U is integrated as U + 0i (NLSEmagic2D_Take_Steps_GPU.c:20-28)."""
    rng = np.random.default_rng(6)
    U = 1 + 0.2 * rng.uniform(-1, 1, (16, 16))
    V = np.zeros((16, 16))
    got = hip.NLSEmagic2D_Take_Steps_GPU(U, V, -1.0, 1.0, 0.01, 1, 5, 0.002, 2, 1)
    exp = oracle.take_steps(U.astype(complex), V, -1.0, 1.0, 0.01, 1, 5, 0.002, 1)
    assert bits_equal(got, exp)


@pytest.mark.parametrize("fusion", [1, 2, 4])
@pytest.mark.parametrize("graph", [0, 3, 4])
def test_resident_solver_chunks_equal_one_call(hip, oracle, graph, fusion):
    """State kept on the GPU across chunks = one long call (2D 2SHOC, MSD),
    with and without hipGraph replay of the chunk loop, for every fusion
    depth (odd graph lengths exercise the U/Y1 buffer swap)."""
    rng = np.random.default_rng(8)
    shape = (40, 70)
    U = 1 + 0.1 * rng.uniform(-1, 1, shape) + 0.1j * rng.uniform(-1, 1, shape)
    V = 0.1 * rng.uniform(-1, 1, shape)
    s, a, h2, k = -1.0, 1.0, 0.01, 0.002
    sol = hip.Solver(shape, 2, 2, 2, s, a, h2, k)
    sol.set_graph(graph)
    sol.set_fusion(fusion)
    sol.set_state(U, V)
    for n in (7, 1, 9, 4):
        sol.step(n)
    got = sol.state()
    sol.close()
    exp = oracle.take_steps(U, V, s, a, h2, 2, 21, k, 2)
    assert bits_equal(got, exp)


def _fusion_cases():
    out = []
    seed = 5000
    for shape in [(3, 3), (4, 5), (17, 65), (33, 130), (50, 64), (18, 7), (31, 129), (65, 200), (9, 70), (33, 66)]:
        for method in (1, 2):
            for prec in (2, 1):
                for BC in (1, 2, 3, 0):
                    for fusion in (1, 2, 4):
                        seed += 1
                        out.append((shape, method, prec, BC, fusion, (8, 16, 32)[seed % 3], seed))
    return out


@pytest.mark.parametrize("shape,method,prec,BC,fusion,tile,seed", _fusion_cases())
def test_fused_2d_matches_oracle(hip, oracle, shape, method, prec, BC, fusion, tile, seed):
    """One, two and four RK4 stages per launch, on 8-, 16- and 32-row tiles,
    give the reference's bits."""
    _announce("fused", shape, method, prec, BC, fusion, tile, seed)
    rng = np.random.default_rng(seed)
    U = (0.9 + 0.3 * rng.uniform(-1, 1, shape)) + 1j * (0.3 * rng.uniform(-1, 1, shape))
    V = 0.5 * rng.uniform(-1, 1, shape)
    h2 = 0.04
    k = 0.1 * h2
    s = -1.0 if BC != 3 else 0.8
    exp = oracle.take_steps(U, V, s, 0.9, h2, BC, 3, k, method, precision=prec)
    sol = hip.Solver(shape, method, prec, BC, s, 0.9, h2, k)
    sol.set_fusion(fusion)
    sol.set_tile(tile)
    sol.set_state(U, V)
    sol.step(3)
    got = sol.state()
    sol.close()
    if prec == 1:  # the resident API widens single-precision state to double on download
        got = got.astype(np.complex64).astype(np.complex128)
    assert bits_equal(got, exp), f"max |diff| {np.max(np.abs(got - exp))}"


@pytest.mark.parametrize("fusion", [1, 2, 4])
@pytest.mark.parametrize("method", [1, 2])
def test_full_size_2d_bitwise_short_run(hip, oracle, method, fusion):
    """1024 x 1024 dark vortex, MSD, a few steps: bit-exact at full size."""
    n = 1024
    h = 0.1
    x = -51.2 + h * np.arange(n)
    X, Y = np.meshgrid(x, x, indexing="ij")
    r = np.hypot(X, Y)
    rng = np.random.default_rng(0)
    U = np.tanh(r) * np.exp(1j * np.arctan2(Y, X)) + 1e-3 * (rng.uniform(-1, 1, r.shape) + 1j * rng.uniform(-1, 1, r.shape))
    V = np.zeros(r.shape)
    k = (0.3 if method == 1 else 0.2121) * h * h
    _announce("full2d", method, fusion)
    sol = hip.Solver((n, n), method, 2, 2, -1.0, 1.0, h * h, k)
    sol.set_fusion(fusion)
    sol.set_state(U, V)
    sol.step(3)
    got = sol.state()
    sol.close()
    exp = oracle.take_steps(U, V, -1.0, 1.0, h * h, 2, 3, k, method)
    assert bits_equal(got, exp), rel_l2(got, exp)


def test_invalid_arguments_raise(hip):
    U = np.ones((8, 8), complex)
    V = np.zeros((8, 8))
    with pytest.raises(hip.HipError):
        hip.take_steps(U, V, -1, 1, 0.01, 2, 1, 0.001, precision=2, method=3)
    with pytest.raises(hip.HipError):
        hip.take_steps(np.ones((2, 8), complex), np.zeros((2, 8)), -1, 1, 0.01, 2, 1, 0.001, precision=2, method=1)


def _kernel_cases():
    """Row-sweep tiles (kind 1) and register strips (kind 2): grids whose sizes
    hit every decomposition edge (n mod rows == 1, M mod W == 1 for W = 48..62,
    a last strip or tile of two columns, one strip/segment/tile for the whole
    grid), both methods, both precisions, every BC, 1/2/4 stages per launch,
    every tile height and 3..32-row strip segments."""
    out = []
    seed = 9000
    shapes = [(3, 3), (4, 5), (3, 100), (100, 3), (9, 57), (17, 61), (33, 63), (5, 113), (7, 121),
              (40, 58), (25, 62), (12, 64), (65, 130), (34, 200), (10, 169), (13, 49), (21, 50)]
    strips = (3, 4, 8, 16, 32, 5)
    for shape in shapes:
        for method in (1, 2):
            for prec in (2, 1):
                for BC in (1, 2, 3, 0):
                    seed += 1
                    out.append((shape, method, prec, BC, (1, 2, 4)[seed % 3], 1, (8, 12, 16)[(seed // 3) % 3], seed))
                    seed += 1
                    out.append((shape, method, prec, BC, (1, 2, 4)[seed % 3], 2, strips[seed % len(strips)], seed))
    return out


@pytest.mark.parametrize("shape,method,prec,BC,fusion,kind,rows,seed", _kernel_cases())
def test_kernel_families_2d_match_oracle(hip, oracle, shape, method, prec, BC, fusion, kind, rows, seed):
    """The row-sweep tiles and the register-marching strips (DPP neighbours)
    give the reference's bits."""
    _announce("kernel", shape, method, prec, BC, fusion, kind, rows, seed)
    rng = np.random.default_rng(seed)
    U = (0.9 + 0.3 * rng.uniform(-1, 1, shape)) + 1j * (0.3 * rng.uniform(-1, 1, shape))
    V = 0.5 * rng.uniform(-1, 1, shape)
    h2 = 0.04
    k = 0.1 * h2
    s = -1.0 if BC != 3 else 0.8
    exp = oracle.take_steps(U, V, s, 0.9, h2, BC, 3, k, method, precision=prec)
    sol = hip.Solver(shape, method, prec, BC, s, 0.9, h2, k)
    sol.set_fusion(fusion)
    sol.set_kernel(kind, rows)
    assert sol.config() == (fusion, kind, rows)
    sol.set_state(U, V)
    sol.step(3)
    got = sol.state()
    sol.close()
    if prec == 1:
        got = got.astype(np.complex64).astype(np.complex128)
    assert bits_equal(got, exp), f"max |diff| {np.max(np.abs(got - exp))}"


@pytest.mark.parametrize("kind,fusion,rows", [(1, 1, 8), (1, 2, 8), (1, 2, 16), (1, 4, 12), (2, 1, 16), (2, 2, 10),
                                              (2, 4, 64)])
@pytest.mark.parametrize("method", [1, 2])
def test_kernel_families_full_size_2d_bitwise(hip, oracle, method, kind, fusion, rows):
    """1024 x 1024 dark vortex, MSD, a few steps through the row-sweep tiles
    and the register strips: bit-exact at full size."""
    n = 1024
    h = 0.1
    x = -51.2 + h * np.arange(n)
    X, Y = np.meshgrid(x, x, indexing="ij")
    r = np.hypot(X, Y)
    rng = np.random.default_rng(0)
    U = np.tanh(r) * np.exp(1j * np.arctan2(Y, X)) + 1e-3 * (rng.uniform(-1, 1, r.shape) + 1j * rng.uniform(-1, 1, r.shape))
    V = np.zeros(r.shape)
    k = (0.3 if method == 1 else 0.2121) * h * h
    _announce("kernel-full2d", method, kind, fusion, rows)
    sol = hip.Solver((n, n), method, 2, 2, -1.0, 1.0, h * h, k)
    sol.set_fusion(fusion)
    sol.set_kernel(kind, rows)
    sol.set_state(U, V)
    sol.step(3)
    got = sol.state()
    sol.close()
    exp = oracle.take_steps(U, V, -1.0, 1.0, h * h, 2, 3, k, method)
    assert bits_equal(got, exp), rel_l2(got, exp)


@pytest.mark.parametrize("fusion", [1, 2, 4])
@pytest.mark.parametrize("method", [1, 2])
def test_kernel_families_compensated_agree(hip, method, fusion):
    """Compensated single precision: LDS tiles, row-sweep tiles and register
    strips give the same bits (hi and lo parts), Dirichlet and MSD."""
    rng = np.random.default_rng(77 + method)
    shape = (70, 133)
    U = (0.9 + 0.3 * rng.uniform(-1, 1, shape)) + 1j * (0.3 * rng.uniform(-1, 1, shape))
    V = 0.5 * rng.uniform(-1, 1, shape)
    for BC in (1, 2):
        _announce("kernel-comp", method, fusion, BC)
        res = []
        for kind, rows in ((0, 8), (1, 12), (2, 8)):
            sol = hip.Solver(shape, method, 3, BC, -1.0, 0.9, 0.04, 0.004)
            sol.set_fusion(fusion)
            sol.set_kernel(kind, rows)
            sol.set_state(U, V)
            sol.step(5)
            res.append(sol.state())
            sol.close()
        assert bits_equal(res[0], res[1]) and bits_equal(res[0], res[2])


def _march3d_cases():
    """3D 2SHOC plane-marching stage kernel: several i chunks (32 planes),
    j tiles (4 or 8 rows) and k tiles (60 columns), a last k tile of one or
    two columns, one plane/row/column of interior only; Dirichlet, Lap0 and
    the zero boundary (MSD runs on the two-pass kernels)."""
    out = []
    seed = 20000
    for shape in [(70, 20, 130), (33, 9, 61), (40, 17, 121), (3, 30, 64), (65, 3, 3), (34, 11, 62)]:
        for prec in (2, 1):
            for BC in (1, 3, 0):
                for rows in (4, 8):
                    seed += 1
                    out.append((shape, prec, BC, rows, seed))
    return out


@pytest.mark.parametrize("shape,prec,BC,rows,seed", _march3d_cases())
def test_march3d_matches_oracle_and_two_pass(hip, oracle, shape, prec, BC, rows, seed):
    """The plane-marching 3D 2SHOC kernel (D on chip) gives the reference's
    bits, and the same bits as the two-pass kernels."""
    _announce("march3d", shape, prec, BC, rows, seed)
    rng = np.random.default_rng(seed)
    U = (0.9 + 0.3 * rng.uniform(-1, 1, shape)) + 1j * (0.3 * rng.uniform(-1, 1, shape))
    V = 0.5 * rng.uniform(-1, 1, shape)
    h2 = 0.04
    k = 0.1 * h2
    s = -1.0 if BC != 3 else 0.8
    exp = oracle.take_steps(U, V, s, 0.9, h2, BC, 3, k, 2, precision=prec)
    res = []
    for kind in (1, 0):
        sol = hip.Solver(shape, 2, prec, BC, s, 0.9, h2, k)
        sol.set_fusion(1)
        sol.set_kernel(kind, rows if kind == 1 else 8)
        sol.set_state(U, V)
        sol.step(3)
        got = sol.state()
        sol.close()
        if prec == 1:
            got = got.astype(np.complex64).astype(np.complex128)
        res.append(got)
    assert bits_equal(res[0], exp), f"max |diff| {np.max(np.abs(res[0] - exp))}"
    assert bits_equal(res[1], exp)


@pytest.mark.parametrize("shape,BC", [((70, 20, 130), 1), ((33, 9, 61), 3), ((40, 17, 121), 0)])
def test_march3d_compensated_matches_two_pass(hip, shape, BC):
    """Compensated single precision (hi + lo state) on the plane-marching
    kernel, whose update operands are loaded a plane ahead, equals the
    two-pass kernels bit for bit after several steps."""
    _announce("march3d comp", shape, BC)
    rng = np.random.default_rng(sum(shape) + BC)
    U = (0.9 + 0.3 * rng.uniform(-1, 1, shape)) + 1j * (0.3 * rng.uniform(-1, 1, shape))
    V = 0.5 * rng.uniform(-1, 1, shape)
    res = []
    for kind in (1, 0):
        sol = hip.Solver(shape, 2, 3, BC, -1.0, 0.9, 0.04, 0.004)
        sol.set_fusion(1)
        sol.set_kernel(kind, 4 if kind == 1 else 8)
        sol.set_state(U, V)
        sol.step(7)
        res.append(sol.state())
        sol.close()
    assert bits_equal(res[0], res[1])


def _block1d_cases():
    """1D time-blocked windows: grids smaller than, equal to and larger than
    one segment, steps per launch 1..max (CD 80, 2SHOC 32), chunk lengths that
    split into odd and even launch counts, every BC."""
    out = []
    seed = 30000
    for n in (5, 256, 257, 700, 2048, 5000):
        for method in (1, 2):
            for BC in (1, 2, 3, 4, 0):
                seed += 1
                S = (1, 3, 8, 17, 32 if method == 2 else 80)[seed % 5]
                steps = (1, 5, 37, 100)[seed % 4]
                out.append((n, method, BC, S, steps, (2, 1, 3)[seed % 3], seed))
    # the last segment would own one to three points: the MSD and
    # one-sided boundary formulas reach further inward than a stage's R, and
    # with one or two steps per launch that reach meets the window's stale edge
    for n, BC in ((257, 2), (258, 4), (257, 4), (259, 4), (513, 2)):
        for method in (1, 2):
            for S in (1, 2):
                seed += 1
                out.append((n, method, BC, S, 4 * S, 2, seed))
    return out


@pytest.mark.parametrize("n,method,BC,S,steps,prec,seed", _block1d_cases())
def test_block1d_matches_oracle(hip, oracle, n, method, BC, S, steps, prec, seed):
    """Time-blocked 1D windows give the reference's bits (compensated: the
    same bits as the one-workgroup kernel)."""
    _announce("block1d", n, method, BC, S, steps, prec, seed)
    rng = np.random.default_rng(seed)
    U = (0.9 + 0.3 * rng.uniform(-1, 1, n)) + 1j * (0.3 * rng.uniform(-1, 1, n))
    V = 0.5 * rng.uniform(-1, 1, n)
    h2 = 0.04
    k = 0.1 * h2
    s = -1.0 if BC != 3 else 0.8
    res = []
    for kind, rows, graph in ((1, S, 0), (1, S, 2 * S + 1), (0, 8, 0)):
        sol = hip.Solver((n,), method, prec, BC, s, 0.9, h2, k)
        sol.set_kernel(kind, rows)
        sol.set_graph(graph)
        sol.set_state(U, V)
        sol.step(steps)
        sol.step(steps)
        res.append(sol.state())
        sol.close()
    if prec != 3:
        exp = oracle.take_steps(U, V, s, 0.9, h2, BC, 2 * steps, k, method, precision=prec)
        if prec == 1:
            res = [r.astype(np.complex64).astype(np.complex128) for r in res]
        assert bits_equal(res[0], exp), f"max |diff| {np.max(np.abs(res[0] - exp))}"
    assert bits_equal(res[0], res[1]) and bits_equal(res[0], res[2])


def _pair3d_cases():
    """3D two-stage plane-marching launches (CD: 4-row x 60-column tiles,
    2SHOC: 4 x 56 with a halo of four): several i chunks (32 planes), a last
    j tile of one to three rows, a last k tile of one or two columns, one
    plane/row/column of interior only; Dirichlet, Lap0 and the zero
    boundary."""
    out = []
    seed = 40000
    for shape in [(70, 20, 130), (33, 9, 61), (40, 17, 121), (3, 30, 64), (65, 3, 3), (34, 11, 62), (5, 6, 7),
                  (36, 5, 57), (9, 13, 113), (4, 4, 58)]:
        for method in (1, 2):
            for prec in (2, 1):
                for BC in (1, 3, 0):
                    seed += 1
                    out.append((shape, method, prec, BC, seed))
    return out


@pytest.mark.parametrize("shape,method,prec,BC,seed", _pair3d_cases())
def test_pair3d_matches_oracle_and_one_stage(hip, oracle, shape, method, prec, BC, seed):
    """Two stages per launch in 3D (the first stage's output and, for 2SHOC,
    both stages' D kept on chip) give the reference's bits and the bits of
    one stage per launch."""
    _announce("pair3d", shape, method, prec, BC, seed)
    rng = np.random.default_rng(seed)
    U = (0.9 + 0.3 * rng.uniform(-1, 1, shape)) + 1j * (0.3 * rng.uniform(-1, 1, shape))
    V = 0.5 * rng.uniform(-1, 1, shape)
    h2 = 0.04
    k = 0.1 * h2
    s = -1.0 if BC != 3 else 0.8
    exp = oracle.take_steps(U, V, s, 0.9, h2, BC, 3, k, method, precision=prec)
    res = []
    for fusion in (2, 1):
        sol = hip.Solver(shape, method, prec, BC, s, 0.9, h2, k)
        sol.set_fusion(fusion)
        assert sol.config()[0] == fusion
        sol.set_state(U, V)
        sol.step(3)
        got = sol.state()
        sol.close()
        if prec == 1:
            got = got.astype(np.complex64).astype(np.complex128)
        res.append(got)
    assert bits_equal(res[0], exp), f"max |diff| {np.max(np.abs(res[0] - exp))}"
    assert bits_equal(res[1], exp)


@pytest.mark.parametrize("graph", [0, 3])
@pytest.mark.parametrize("BC", [1, 0])
@pytest.mark.parametrize("method", [1, 2])
def test_pair3d_resident_chunks_equal_one_call(hip, oracle, method, BC, graph):
    """State kept on the GPU across chunks of odd and even lengths, with and
    without hipGraph replay (an odd graph length is rounded up to even for the
    U/Y1 swap of a pair step), equals one long call."""
    _announce("pair3d chunks", method, BC, graph)
    rng = np.random.default_rng(41 + BC + graph + 7 * method)
    shape = (36, 10, 70)
    U = 1 + 0.1 * rng.uniform(-1, 1, shape) + 0.1j * rng.uniform(-1, 1, shape)
    V = 0.1 * rng.uniform(-1, 1, shape)
    s, a, h2, k = -1.0, 1.0, 0.04, 0.004
    sol = hip.Solver(shape, method, 2, BC, s, a, h2, k)
    sol.set_graph(graph)
    sol.set_fusion(2)
    sol.set_state(U, V)
    assert sol.config()[0] == 2
    for n in (7, 1, 9, 4):
        sol.step(n)
    got = sol.state()
    sol.close()
    exp = oracle.take_steps(U, V, s, a, h2, BC, 21, k, method)
    assert bits_equal(got, exp)


@pytest.mark.parametrize("method", [1, 2])
def test_pair3d_msd_runs_one_stage_per_launch(hip, oracle, method):
    """MSD in 3D pairs a boundary point with one in another plane, so a
    requested pair fusion runs one stage per launch (and stays bit-exact)."""
    shape = (9, 6, 65)
    rng = np.random.default_rng(3)
    U = (0.9 + 0.3 * rng.uniform(-1, 1, shape)) + 1j * (0.3 * rng.uniform(-1, 1, shape))
    V = 0.5 * rng.uniform(-1, 1, shape)
    sol = hip.Solver(shape, method, 2, 2, -1.0, 0.9, 0.04, 0.004)
    sol.set_fusion(2)
    assert sol.config()[0] == 1
    sol.set_state(U, V)
    sol.step(3)
    got = sol.state()
    sol.close()
    assert bits_equal(got, oracle.take_steps(U, V, -1.0, 0.9, 0.04, 2, 3, 0.004, method))


@pytest.mark.parametrize("method", [1, 2])
@pytest.mark.parametrize("shape,BC", [((70, 20, 130), 1), ((33, 9, 61), 3), ((40, 17, 121), 0)])
def test_pair3d_compensated_matches_one_stage(hip, shape, BC, method):
    """Compensated single precision (hi + lo state) through two stages per
    launch equals one stage per launch bit for bit after several steps."""
    _announce("pair3d comp", shape, BC, method)
    rng = np.random.default_rng(sum(shape) * 3 + BC)
    U = (0.9 + 0.3 * rng.uniform(-1, 1, shape)) + 1j * (0.3 * rng.uniform(-1, 1, shape))
    V = 0.5 * rng.uniform(-1, 1, shape)
    res = []
    for fusion in (2, 1):
        sol = hip.Solver(shape, method, 3, BC, -1.0, 0.9, 0.04, 0.004)
        sol.set_fusion(fusion)
        sol.set_state(U, V)
        sol.step(7)
        res.append(sol.state())
        sol.close()
    assert bits_equal(res[0], res[1])


@pytest.mark.parametrize("prec", [2, 1])
@pytest.mark.parametrize("method", [1, 2])
def test_pair3d_full_size_bitwise(hip, oracle, method, prec):
    """128^3 dark vortex ring, Lap0, two steps through two-stage launches
    (both precisions; single precision runs 2SHOC on 8-row tiles): bit-exact
    at a size with many i chunks and tiles in every direction."""
    n, h = 128, 0.2
    x = -12.8 + h * np.arange(n)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    rho = np.hypot(X, Y) - 4.0
    U = np.tanh(np.hypot(rho, Z)) * np.exp(1j * np.arctan2(Z, rho))
    V = np.zeros(U.shape)
    k = 0.05 * h * h
    _announce("pair3d full", n, method, prec)
    sol = hip.Solver((n, n, n), method, prec, 3, -1.0, 1.0, h * h, k)
    sol.set_fusion(2)
    assert sol.config()[0] == 2
    sol.set_state(U, V)
    sol.step(2)
    got = sol.state()
    sol.close()
    exp = oracle.take_steps(U, V, -1.0, 1.0, h * h, 3, 2, k, method, precision=prec)
    if prec == 1:
        got = got.astype(np.complex64).astype(np.complex128)
    assert bits_equal(got, exp), rel_l2(got, exp)


@pytest.mark.parametrize("method,prec,default", [(1, 2, 2), (2, 2, 2), (2, 1, 2), (2, 3, 2)])
def test_3d_default_fusion_and_autotune(hip, oracle, method, prec, default):
    """3D runs two stages per launch by default; the tuner picks between one
    and two on a saved copy of the state and the run continues from the
    untouched state."""
    shape = (40, 12, 70)
    rng = np.random.default_rng(17 + method + prec)
    U = 1 + 0.1 * rng.uniform(-1, 1, shape) + 0.1j * rng.uniform(-1, 1, shape)
    V = 0.1 * rng.uniform(-1, 1, shape)
    sol = hip.Solver(shape, method, prec, 1, -1.0, 1.0, 0.04, 0.004)
    assert sol.config()[0] == default
    sol.set_state(U, V)
    f, t, ms = sol.autotune(2)
    assert f in (1, 2) and ms > 0 and sol.config()[0] == f
    sol.step(3)
    got = sol.state()
    sol.close()
    if prec == 2:
        assert bits_equal(got, oracle.take_steps(U, V, -1.0, 1.0, 0.04, 1, 3, 0.004, method))
    else:
        assert rel_l2(got, oracle.take_steps(U, V, -1.0, 1.0, 0.04, 1, 3, 0.004, method)) < 1e-6


def test_config_reports_kernel_choice_in_every_dimension(hip):
    """The kernel kind and rows set by set_kernel are what config() reports,
    in 1D (steps per launch), 2D (rows per tile or strip segment) and 3D
    (plane-march rows)."""
    for shape, kind, rows in (((300,), 1, 17), ((300,), 0, 8), ((20, 30), 2, 12), ((20, 30), 1, 16),
                              ((10, 12, 70), 1, 8), ((10, 12, 70), 0, 8)):
        sol = hip.Solver(shape, 2, 2, 1, -1.0, 1.0, 0.04, 0.004)
        sol.set_kernel(kind, rows)
        assert sol.config()[1:] == (kind, rows), shape
        sol.close()
