"""Build the gfx950 HIP library (and, for tests, the CPU oracle).

    python -m atom3d_amd.build            # library + oracle
    python -m atom3d_amd.build --lib      # library only

The library is compiled in-tree into ``atom3d_amd/_lib/libatom3d_amd.so`` with
hipcc for gfx950 only.  Flags that matter for parity:

* ``-ffp-contract=off``: no FMA contraction, so every FP32 operation of the
  geometry predicates is rounded once, in the reference's order;
* IEEE division and square root are hipcc's defaults for FP32 and are kept;
* ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed FP32 ops.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OUT_DIR = os.path.join(PKG, "_lib")
OBJ_DIR = os.path.join(PKG, "_lib", "obj")
LIB = os.path.join(OUT_DIR, "libatom3d_amd.so")
ORACLE_DIR = os.path.join(ROOT, "oracle")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


def sources() -> list[str]:
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def build_lib(verbose: bool = False) -> str:
    os.makedirs(OBJ_DIR, exist_ok=True)
    srcs = sources()
    objs = [os.path.join(OBJ_DIR, os.path.basename(s)[:-4] + ".o") for s in srcs]

    def comp(pair):
        src, obj = pair
        cmd = [HIPCC, *HIP_FLAGS, "-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        _run(cmd)

    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        list(ex.map(comp, zip(srcs, objs)))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


def build_oracle(verbose: bool = False) -> str:
    cmd = ["make", "-C", ORACLE_DIR, "-j8"]
    if verbose:
        print(" ".join(cmd), flush=True)
    _run(cmd)
    return os.path.join(ORACLE_DIR, "liba3d_oracle.so")


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib", action="store_true", help="build only the HIP library")
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args(argv)
    print(build_lib(a.verbose))
    if not a.lib:
        print(build_oracle(a.verbose))
    return 0


if __name__ == "__main__":
    sys.exit(main())
