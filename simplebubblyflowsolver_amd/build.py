"""Build libsbfs.so (HIP, gfx950) in-tree with hipcc.

    python -m simplebubblyflowsolver_amd.build [--jobs N] [--verbose]

Every .hip file under csrc/ is compiled to an object under _build/ and
linked into simplebubblyflowsolver_amd/libsbfs.so; the command-line driver
driver/main.c (C, gcc) is linked against it as simplebubblyflowsolver_amd/
sbfs_run.  Flags:
  -ffp-contract=off   no a*b+c contraction: the FP64 arithmetic keeps the
                      reference's rounding (bitwise parity with the oracle)
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      no packed FP32 instructions
"""
from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = ROOT / "_build"
LIB = PKG / "libsbfs.so"
DRIVER = PKG / "sbfs_run"
ARCH = os.environ.get("SBFS_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def cflags() -> list:
    return [f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
            "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wall", "-Wno-unused-command-line-argument", f"-I{ROOT / 'include'}", f"-I{CSRC}"]


def compile_one(src: Path, verbose: bool) -> Path:
    obj = BUILD / (src.stem + ".o")
    cmd = [hipcc()] + cflags() + ["-c", str(src), "-o", str(obj)]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"compiling {src.name} failed")
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if err and verbose:
        sys.stderr.write(err + "\n")
    return obj


def build(jobs: int = 8, verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip"))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: compile_one(s, verbose), srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-o", str(tmp)] + [str(o) for o in objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    build_driver(verbose)
    return LIB


def build_driver(verbose: bool = False) -> Path:
    """gcc the C driver against libsbfs.so; it finds the library next to
    itself ($ORIGIN), so the pair can move together."""
    cc = os.environ.get("CC", "gcc")
    tmp = DRIVER.with_name(DRIVER.name + ".tmp")
    cmd = [cc, "-std=c99", "-O2", "-Wall", "-Wextra", f"-I{ROOT / 'include'}", str(PKG / "driver" / "main.c"),
           f"-L{PKG}", "-lsbfs", "-Wl,-rpath,$ORIGIN", "-lm", "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, DRIVER)
    return DRIVER


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=min(8, os.cpu_count() or 1))
    ap.add_argument("--verbose", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, a.verbose))


if __name__ == "__main__":
    main()
