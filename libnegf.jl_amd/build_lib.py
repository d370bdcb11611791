"""Build the gfx950 shared library libnegf_amd/libnegf_hip.so from csrc/*.hip.

Usage: python libnegf.jl_amd/build_lib.py [--jobs N]

Each translation unit is compiled with hipcc for gfx950 (in parallel), then
linked into one shared library next to the Python host package, where the
ctypes loader and the Julia shim look for it.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import pathlib
import subprocess
import sys

HERE = pathlib.Path(__file__).resolve().parent
CSRC = HERE / "csrc"
OUT = HERE / "libnegf_amd" / "libnegf_hip.so"
OBJ = HERE / "csrc" / "_obj"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # keep the compiler from forming packed-FP32 ops (see the design notes)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _compile(src: pathlib.Path) -> tuple:
    obj = OBJ / (src.stem + ".o")
    cmd = [HIPCC, *FLAGS, "-c", str(src), "-o", str(obj)]
    p = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in p.stderr.splitlines() if "packed-fp32-ops" not in l
                    and "not a recognized feature" not in l)
    return src.name, p.returncode, err, obj


def build(jobs: int = 8, verbose: bool = True) -> pathlib.Path:
    OBJ.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip"))
    objs = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for name, rc, err, obj in ex.map(_compile, srcs):
            if rc != 0:
                raise RuntimeError(f"hipcc failed on {name}:\n{err}")
            if verbose and err.strip():
                print(f"[{name}] {err.strip()}", file=sys.stderr)
            objs.append(str(obj))
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", str(tmp)]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"link failed:\n{p.stderr}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT}")
    return OUT


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    a = ap.parse_args()
    build(a.jobs)
