"""Build the gfx950 library libpfgpu.so (and the Fortran binding) in-tree.

    python -m plantfem_amd.build            # build everything
    python -m plantfem_amd.build --hip      # HIP library only

Objects go to plantfem_amd/_build/, the shared library to plantfem_amd/lib/.
Every call recompiles (the build doubles as the "does it build" check).
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
FSRC = PKG / "fortran"
BUILD = PKG / "_build"
LIB = PKG / "lib"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLANG = os.environ.get("FC", "/opt/rocm/llvm/bin/flang")
ARCH = os.environ.get("PF_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["scan.hip", "graph.hip", "assemble.hip", "bc.hip", "spmv.hip", "ilu.hip", "bicgstab.hip", "batch.hip", "multi.hip", "eig.hip", "noalloc.hip", "direct.hip", "capi.hip"]
# files whose arithmetic must follow the reference's operation order exactly
NO_CONTRACT = {"assemble.hip", "bc.hip", "ilu.hip"}

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-Wall",
    "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{ROOT / 'include'}",
]


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, capture_output=True, text=True)
    out = (res.stdout + res.stderr).strip()
    # hipcc's host pass does not know the packed-fp32 feature flag; that warning is noise
    lines = [l for l in out.splitlines() if "packed-fp32-ops" not in l]
    if res.returncode != 0:
        raise RuntimeError(f"command failed ({res.returncode}): {' '.join(cmd)}\n" + "\n".join(lines))
    if lines:
        print("\n".join(lines), file=sys.stderr)


def build_hip(jobs: int = 8) -> Path:
    BUILD.mkdir(exist_ok=True)
    LIB.mkdir(exist_ok=True)
    objs = []
    cmds = []
    for src in HIP_SOURCES:
        obj = BUILD / (Path(src).stem + ".o")
        flags = list(HIP_FLAGS)
        if src in NO_CONTRACT:
            flags.append("-ffp-contract=off")
        cmds.append([HIPCC, *flags, "-c", str(CSRC / src), "-o", str(obj)])
        objs.append(obj)
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        list(ex.map(_run, cmds))
    out = LIB / "libpfgpu.so"
    tmp = LIB / "libpfgpu.so.tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp),
          "-L/opt/rocm/lib", "-lrocsolver", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, out)
    return out


def build_fortran() -> Path | None:
    """Fortran FEMSolver_ binding (plantfem_gpu module) as libpfgpu_f.so + its .mod file."""
    src = FSRC / "plantfem_gpu.f90"
    if not src.exists():
        return None
    if not Path(FLANG).exists():
        raise RuntimeError(f"Fortran compiler not found at {FLANG}")
    mod_dir = LIB / "mod"
    mod_dir.mkdir(parents=True, exist_ok=True)
    out = LIB / "libpfgpu_f.so"
    _run([FLANG, "-O2", "-fPIC", "-shared", "-J", str(mod_dir), str(src), f"-L{LIB}", "-lpfgpu",
          "-Wl,-rpath,$ORIGIN", "-o", str(out)])
    ex_dir = FSRC / "examples"
    bin_dir = LIB / "bin"
    if ex_dir.exists():
        bin_dir.mkdir(exist_ok=True)
        for f in sorted(ex_dir.glob("*.f90")):
            _run([FLANG, "-O2", f"-I{mod_dir}", str(f), f"-L{LIB}", "-lpfgpu_f", "-lpfgpu",
                  "-Wl,-rpath,$ORIGIN/..", "-o", str(bin_dir / f.stem)])
    return out


def main(argv=None) -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--hip", action="store_true")
    ap.add_argument("--fortran", action="store_true")
    ap.add_argument("-j", type=int, default=8)
    a = ap.parse_args(argv)
    every = not (a.hip or a.fortran)
    if every or a.hip:
        print("built", build_hip(a.j))
    if every or a.fortran:
        print("built", build_fortran())


if __name__ == "__main__":
    main()
