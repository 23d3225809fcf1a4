"""Build libsixtrack_amd.so (host C library + HIP kernels for gfx950).

    python -m sixtracklib_amd.build          # or __graft_entry__.build()

Objects go to sixtracklib_amd/_build/, the library to
sixtracklib_amd/libsixtrack_amd.so (in-tree, so it travels with the
repository snapshot).  The tracking kernels are compiled twice:
``_exact`` without floating-point contraction (bit-identical to the
reference CPU tracker) and ``_fma`` with contraction (fused multiply-adds).
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsixtrack_amd.so")
ARCH = os.environ.get("SIXTRACKLIB_AMD_ARCH", "gfx950")

INC = ["-I", os.path.join(ROOT, "include"), "-I", CSRC]
# FP32 packed math stays off (no FP32 hot paths; keeps codegen predictable)
NO_PK_FP32 = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
HIP_COMMON = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH] + NO_PK_FP32 + INC
C_COMMON = ["-std=c99", "-O2", "-fPIC", "-fno-fast-math", "-ffp-contract=off",
            "-Wall", "-Wextra", "-Wno-unused-parameter"] + INC

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CC = os.environ.get("CC", "gcc")


def _jobs():
    j = []
    for src in ("buffer.c", "objects.c"):
        j.append(([CC] + C_COMMON + ["-c", os.path.join(CSRC, src), "-o",
                                      os.path.join(BUILD, src + ".o")],
                  os.path.join(BUILD, src + ".o")))
    for src in ("lattice_pack.cpp", "track_job.cpp", "track_job_new.cpp", "batch.cpp", "adjoint.cpp"):
        if not os.path.exists(os.path.join(CSRC, src)):
            continue
        j.append(([HIPCC] + HIP_COMMON + ["-ffp-contract=off", "-x", "hip", "-c",
                                          os.path.join(CSRC, src), "-o",
                                          os.path.join(BUILD, src + ".o")],
                  os.path.join(BUILD, src + ".o")))
    # "fast-honor-pragmas": contraction everywhere except in the blocks that
    # switch it off (the energy update and the drift's 1 + t/2), where the
    # reference's own rounding is kept
    for variant, contract in (("_exact", "off"), ("_fma", "fast-honor-pragmas")):
        for src in ("track_kernels.hip", "batch_kernels.hip", "adjoint_kernels.hip"):
            if not os.path.exists(os.path.join(CSRC, src)):
                continue
            out = os.path.join(BUILD, src + variant + ".o")
            defs = ["-DST_FMA_BUILD=1"] if variant == "_fma" else []
            # the element dispatch branches only on wave-uniform values:
            # leave those regions unstructurized (no flow blocks / copies)
            defs += ["-mllvm", "-structurizecfg-skip-uniform-regions=true"]
            # no machine-level hoisting of constants out of the element loop:
            # at the 64-VGPR budget of the lean kernel the hoisted
            # polynomial constants of sin/exp would be spilled to scratch
            # and reloaded at every Cavity (rematerialising them is free)
            defs += ["-mllvm", "-disable-machine-licm"]
            j.append(([HIPCC] + HIP_COMMON + defs + ["-ffp-contract=" + contract,
                                                     "-DST_KSUFFIX=" + variant, "-c",
                                              os.path.join(CSRC, src), "-o", out],
                      out))
    return j


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    lines = [ln for ln in p.stdout.splitlines() if "packed-fp32-ops" not in ln]
    return p.returncode, "\n".join(lines), cmd


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    jobs = _jobs()
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        results = list(ex.map(lambda jb: _run(jb[0]), jobs))
    for rc, out, cmd in results:
        if verbose and out.strip():
            print(out)
        if rc != 0:
            raise RuntimeError("compile failed: %s\n%s" % (" ".join(cmd), out))
    objs = [o for _, o in jobs]
    tmp = LIB + ".tmp"
    rc, out, cmd = _run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH,
                         "-o", tmp] + objs + ["-lm"])
    if rc != 0:
        raise RuntimeError("link failed: %s\n%s" % (" ".join(cmd), out))
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
