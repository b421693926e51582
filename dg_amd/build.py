"""Build recipe for libdgx.so (gfx950), in-tree.

    python -m dg_amd.build            # compile and link dg_amd/libdgx.so

Every HIP translation unit is compiled with hipcc for gfx950 with packed-FP32
instruction selection disabled; the limiter/sensor unit additionally without
FMA contraction (it must round like the CPU path).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libdgx.so")
ARCH = os.environ.get("DGX_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
CXX = shutil.which("g++") or "g++"

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]

SOURCES = [
    # (source, compiler, extra flags)
    ("mesh.cc", "cxx", []),
    ("capi.cc", "hip-host", []),
    ("rhs.hip", "hipcc", []),
    ("stage.hip", "hipcc", []),
    ("batch.hip", "hipcc", []),
    ("mlp.hip", "hipcc", []),
    ("limiter.hip", "hipcc", ["-ffp-contract=off"]),
    ("output.hip", "hipcc", ["-ffp-contract=off"]),
    ("compat.hip", "hipcc", []),
]


def _cmd(src: str, kind: str, extra: list[str], build_dir: str = BUILD) -> list[str]:
    obj = os.path.join(build_dir, os.path.basename(src) + ".o")
    path = os.path.join(CSRC, src)
    if kind == "cxx":
        return [CXX, "-O2", "-std=c++17", "-fPIC", "-c", path, "-o", obj] + extra
    if kind == "hip-host":
        return [CXX, "-O2", "-std=c++17", "-fPIC", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                "-c", path, "-o", obj] + extra
    return [HIPCC] + HIP_FLAGS + extra + ["-c", path, "-o", obj]


def build(verbose: bool = False, out: str = LIB, defines: tuple = ()) -> str:
    """Compile and link.  `defines` (NAME=VALUE strings) go to every HIP
    translation unit; a non-default `out` gets its own object folder."""
    build_dir = BUILD if out == LIB else os.path.join(BUILD, "variant_" + os.path.basename(out).replace(".", "_"))
    os.makedirs(build_dir, exist_ok=True)
    dflags = ["-D" + d for d in defines]
    cmds = [_cmd(s, k, e + (dflags if k == "hipcc" else []), build_dir) for s, k, e in SOURCES]

    def run(c):
        r = subprocess.run(c, capture_output=True, text=True)
        return c, r

    with ThreadPoolExecutor(max_workers=min(len(cmds), 8)) as ex:
        results = list(ex.map(run, cmds))
    for c, r in results:
        if verbose or r.returncode != 0:
            sys.stderr.write(" ".join(c) + "\n" + r.stdout + r.stderr)
        if r.returncode != 0:
            raise RuntimeError(f"compile failed: {' '.join(c)}")
    objs = [os.path.join(build_dir, os.path.basename(s) + ".o") for s, _, _ in SOURCES]
    tmp = out + ".tmp"
    link = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(link, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(link) + "\n" + r.stdout + r.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    import argparse

    ap = argparse.ArgumentParser()
    ap.add_argument("-v", action="store_true")
    ap.add_argument("--out", default=LIB)
    ap.add_argument("-D", dest="defines", action="append", default=[])
    a = ap.parse_args()
    print(build(verbose=a.v, out=os.path.abspath(a.out), defines=tuple(a.defines)))
