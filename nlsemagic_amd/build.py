"""Build nlsemagic_amd/libnlsemagic_hip.so with hipcc for gfx950.

Every HIP source is compiled with
  -ffp-contract=off                         no fused multiply-add: the FP64
                                            results must round exactly like
                                            the reference CPU integrators
  -fhip-fp32-correctly-rounded-divide-sqrt  IEEE single-precision division
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                            no packed FP32 instructions
Usage: python -m nlsemagic_amd.build  (or build() from Python)
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libnlsemagic_hip.so")
OBJDIR = os.path.join(HERE, "_build")
SOURCES = ["nlse1d.hip", "nlse2d.hip", "nlse2d_batch.hip", "nlse3d.hip", "nlse_capi.hip", "nlse_batch.hip", "nlse_diag.hip",
           "nlse_vort.hip", "nlse_ic.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("NLSE_OFFLOAD_ARCH", "gfx950")
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
]


def _compile(src: str) -> str:
    obj = os.path.join(OBJDIR, src.replace(".hip", ".o"))
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    warn = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if warn:
        print(warn, file=sys.stderr)
    return obj


MEXDIR = os.path.join(HERE, "mex")
# MATLAB-visible gateway names -> compile-time identity of nlse_mex_gateway.c
GATEWAYS = {f"NLSE{d}D_TAKE_STEPS_{m}_CUDA_{p}": [f"-DNLSE_DIM={d}", f"-DNLSE_METHOD={mi}", f"-DNLSE_PREC={pi}"]
            for d in (1, 2, 3) for m, mi in (("CD", 1), ("2SHOC", 2)) for p, pi in (("D", 2), ("F", 1))}
GATEWAYS.update({f"NLSEmagic{d}D_Take_Steps_GPU": [f"-DNLSE_DIM={d}", "-DNLSE_V2"] for d in (1, 2, 3)})
# gateways with a source of their own: name -> source under mex/
OTHER_GATEWAYS = {"rmc_vorticity3D": "nlse_mex_vorticity.c"}


def build_test_gateways(mex_include: str, extra_src=(), outdir=None) -> dict:
    """Compile every gateway name against a given mex.h directory (the
    test-only shim under tests/mex_shim when MATLAB/Octave are absent);
    returns {name: path of the shared object}."""
    outdir = outdir or os.path.join(MEXDIR, "_build")
    os.makedirs(outdir, exist_ok=True)
    inc = os.path.join(os.path.dirname(HERE), "include")
    out = {}
    jobs = [(name, defs, "nlse_mex_gateway.c") for name, defs in GATEWAYS.items()]
    jobs += [(name, [], src) for name, src in OTHER_GATEWAYS.items()]
    for name, defs, src in jobs:
        so = os.path.join(outdir, name + ".so")
        cmd = ["gcc", "-std=c99", "-O2", "-fPIC", "-shared", "-Wall", "-Werror", *defs, f"-I{mex_include}",
               f"-I{inc}", os.path.join(MEXDIR, src), *extra_src, f"-L{HERE}", "-lnlsemagic_hip",
               f"-Wl,-rpath,{HERE}", "-o", so]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"gateway {name} failed:\n{r.stderr}")
        out[name] = so
    return out


def build(verbose: bool = True) -> str:
    """Compile every HIP source and link the shared library; returns its path."""
    os.makedirs(OBJDIR, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(4, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build()
