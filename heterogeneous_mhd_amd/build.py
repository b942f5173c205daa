"""Build recipe for libmhd_hip.so (gfx950) and the CPU checker libraries.

``python -m heterogeneous_mhd_amd.build`` compiles every HIP source under
``csrc/`` with hipcc into ``heterogeneous_mhd_amd/libmhd_hip.so`` (in-tree, so
the built library travels with the source tree), then runs ``make`` in
``oracle/`` for the CPU checkers.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "libmhd_hip.so"

SOURCES = ["mhd_fluid.hip", "mhd_advect.hip", "mhd_transpose.hip", "mhd_cfl.hip", "mhd_ic.hip",
           "mhd_capi.hip", "mhd_ensemble.hip", "mhd_corrector.hip",
           "mhd_fused.hip", "mhd_io.hip", "mhd_copy.hip",
           "mhd_advect_transpose.hip", "mhd_transpose_advect.hip"]

# -ffp-contract=off: no FMA contraction, so float/double rounding follows
# the reference's C promotion rules exactly (bitwise parity).
# -packed-fp32-ops off: packed FP32 instructions are not used.
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


def build_hip(verbose: bool = False, jobs: int | None = None) -> Path:
    """Compiles each source to an object in parallel (separate compilation:
    every kernel is launched from the file that defines it), then links."""
    objdir = PKG / "_build"
    objdir.mkdir(exist_ok=True)
    hc = hipcc()
    compile_flags = [f for f in HIPCC_FLAGS if f != "-shared"]
    cmds = []
    for s in SOURCES:
        obj = objdir / (Path(s).stem + ".o")
        cmds.append([hc, *compile_flags, "-I", str(ROOT / "include"), "-c", str(CSRC / s), "-o", str(obj)])
    jobs = jobs or min(len(cmds), os.cpu_count() or 1, 16)
    procs: list[tuple[list[str], subprocess.Popen]] = []
    failed = []
    pending = list(cmds)
    while pending or procs:
        while pending and len(procs) < jobs:
            cmd = pending.pop(0)
            if verbose:
                print(" ".join(cmd), flush=True)
            procs.append((cmd, subprocess.Popen(cmd)))
        cmd, p = procs.pop(0)
        if p.wait() != 0:
            failed.append(cmd[-3])
    if failed:
        raise RuntimeError(f"hipcc failed on: {', '.join(failed)}")
    tmp = LIB.with_suffix(".so.tmp")
    objs = [str(objdir / (Path(s).stem + ".o")) for s in SOURCES]
    cmd = [hc, *HIPCC_FLAGS, *objs, "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


def build_oracle(with_reference: bool | None = None, verbose: bool = False) -> None:
    targets = ["all"]
    refdir = Path("/root/reference/C")
    if with_reference is None:
        with_reference = refdir.is_dir()
    if with_reference:
        targets.append("ref")
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), *targets], check=True,
                   stdout=None if verbose else subprocess.DEVNULL)


def build_all(verbose: bool = False) -> None:
    build_hip(verbose)
    build_oracle(verbose=verbose)


if __name__ == "__main__":
    build_all(verbose="-v" in sys.argv)
