"""Write the golden vectors of the reference CPU integrators.

Runs the reference's own NLSE1D_TAKE_STEPS[_F] / NLSE2D_TAKE_STEPS[_F]
(src/NLSEmagic1D/NLSEmagic1D_CPU.c, src/NLSEmagic2D/NLSEmagic2D_CPU.c),
compiled from /root/reference by ``make -C oracle ref`` into oracle/_ref/,
on small seeded inputs, and stores inputs, parameters and outputs in
``tests/golden/nlse_golden.npz`` plus a JSON manifest with SHA-256 digests.
The reference's 3D integrators include MATLAB's mex.h and are not built
here, so there are no 3D vectors (3D is pinned by the checks in
tests/test_oracle_math.py).

Usage:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402


def cases():
    """(name, shape, method, precision, BC, s, a, h2, k, chunk, seed, zero)

    zero: None, or (value, points) -- U set to `value` at the listed points
    before the run: MSD partners of zero modulus, where the reference divides
    by |U|^2 = 0 unguarded (NLSEmagic2D_CPU.c:64-68, :297-303) and produces
    NaN, and a modulus whose square underflows to 0 (x/0 = inf)."""
    out = []
    seed = 100
    for method in (1, 2):
        for prec in (2, 1):
            for BC in (1, 2, 3, 4, 0):
                seed += 1
                out.append((f"1d_m{method}_p{prec}_bc{BC}", (40,), method, prec, BC,
                            -1.0 if BC != 3 else 0.5, 0.9, 0.04, 0.006, 7, seed))
            for BC in (1, 2, 3, 0):
                for shape in ((13, 17), (5, 4)):
                    seed += 1
                    out.append((f"2d_m{method}_p{prec}_bc{BC}_{shape[0]}x{shape[1]}", shape, method, prec, BC,
                                -1.0 if BC != 1 else 0.7, 0.8, 0.01, 0.0015, 5, seed))
    # MSD with zero-modulus partners (chunks of one step, so the non-finite
    # values have not yet reached every point)
    for method in (1, 2):
        for prec in (2, 1):
            for tag, val in (("zero", 0.0), ("tiny", 1e-170)):
                if prec == 1 and tag == "tiny":
                    continue  # 1e-170 is already 0 in single precision
                seed += 1
                out.append((f"1d_m{method}_p{prec}_bc2_{tag}", (40,), method, prec, 2, -1.0, 0.9, 0.04, 0.006, 1,
                            seed, (val, [(1,), (38,)])))
                seed += 1
                out.append((f"2d_m{method}_p{prec}_bc2_{tag}_23x19", (23, 19), method, prec, 2, -1.0, 0.8, 0.01,
                            0.0015, 1, seed, (val, [(1, 7), (11, 1), (21, 12), (9, 17), (1, 1), (21, 17)])))
    return [c if len(c) == 12 else c + (None,) for c in out]


def inputs(shape, seed):
    rng = np.random.default_rng(seed)
    U = (1.0 + 0.3 * rng.uniform(-1, 1, shape)) + 1j * (0.2 * rng.uniform(-1, 1, shape))
    V = 0.5 * rng.uniform(-1, 1, shape)
    return U, V


def main():
    if O.ref(1) is None or O.ref(2) is None:
        raise SystemExit("oracle/_ref not built: run `make -C oracle ref` with /root/reference present")
    arrays = {}
    manifest = []
    for (name, shape, method, prec, BC, s, a, h2, k, chunk, seed, zero) in cases():
        U, V = inputs(shape, seed)
        if zero is not None:
            for pt in zero[1]:
                U[pt] = zero[0]
        out = O.take_steps(U, V, s, a, h2, BC, chunk, k, method, precision=prec, use_ref=True)
        arrays[f"{name}__U"] = U
        arrays[f"{name}__V"] = V
        arrays[f"{name}__out"] = out
        manifest.append(dict(name=name, shape=list(shape), method=method, precision=prec, BC=BC, s=s, a=a,
                             h2=h2, k=k, chunk_size=chunk, seed=seed,
                             sha256_out=hashlib.sha256(np.ascontiguousarray(out).tobytes()).hexdigest()))
    np.savez_compressed(os.path.join(HERE, "nlse_golden.npz"), **arrays)
    with open(os.path.join(HERE, "nlse_golden.json"), "w") as f:
        json.dump(dict(source="reference CPU integrators compiled by oracle/Makefile (g++ -O3 -ffp-contract=off)",
                       cases=manifest), f, indent=1)
    print(f"wrote {len(manifest)} cases")


if __name__ == "__main__":
    main()
