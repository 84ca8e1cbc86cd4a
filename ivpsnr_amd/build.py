"""Builds the gfx950 engine library ivpsnr_amd/libqmiv.so in-tree (hipcc cross-compiles; no GPU needed).

    python -m ivpsnr_amd.build        # compile everything that changed, then link

It also builds the command-line driver ivpsnr_amd/bin/qmiv (cli/qmiv_cli.cpp) against the library.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libqmiv.so")
CLI_SRC = os.path.join(HERE, "cli", "qmiv_cli.cpp")
CLI = os.path.join(HERE, "bin", "qmiv")
ARCH = os.environ.get("QMIV_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# Packed-FP32 VALU instructions stay disabled in every device build.
DEVICE_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-Wall",
                "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
# Host-only translation units (the HIP runtime API, no device code) go through g++.
HOST_HIP_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-D__HIP_PLATFORM_AMD__=1", "-I/opt/rocm/include"]
# The double-precision finalisation must round like the reference's ISO C++ build.
FINALIZE_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-ffp-contract=off", "-fno-fast-math"]

HEADERS = ["qmiv_internal.h", "finalize.h", "host_internal.h", "png_reader.h"]


def _units():
    units = []
    for f in sorted(os.listdir(CSRC)):
        src = os.path.join(CSRC, f)
        obj = os.path.join(OUT_DIR, f + ".o")
        if f.endswith(".hip"):
            units.append((src, obj, [HIPCC] + DEVICE_FLAGS + ["-c", src, "-o", obj]))
        elif f == "finalize.cpp":
            units.append((src, obj, ["g++"] + FINALIZE_FLAGS + ["-c", src, "-o", obj]))
        elif f.endswith(".cpp"):
            units.append((src, obj, ["g++"] + HOST_HIP_FLAGS + ["-c", src, "-o", obj]))
    return units


def _stale(src: str, obj: str) -> bool:
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    deps = [src] + [os.path.join(CSRC, h) for h in HEADERS] + [os.path.join(HERE, "..", "include", "qmiv.h")]
    return any(os.path.exists(d) and os.path.getmtime(d) > t for d in deps)


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r.stderr


def build(force: bool = False, verbose: bool = False) -> str:
    """Compile the engine and link libqmiv.so. Returns the library path."""
    os.makedirs(OUT_DIR, exist_ok=True)
    units = _units()
    todo = [u for u in units if force or _stale(u[0], u[1])]
    with cf.ThreadPoolExecutor(max_workers=min(8, max(1, len(todo)))) as ex:
        for (src, _, cmd), log in zip(todo, ex.map(lambda u: _run(u[2]), todo)):
            if verbose:
                print("compiled", os.path.basename(src), file=sys.stderr)
                if log.strip():
                    print(log, file=sys.stderr)
    objs = [u[1] for u in units]
    if force or todo or not os.path.exists(LIB):
        tmp = LIB + ".tmp"
        _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + ["-lz"])
        os.replace(tmp, LIB)
    build_cli(force=force)
    return LIB


def build_cli(force: bool = False) -> str:
    """Compile the command-line driver, linked to libqmiv.so beside it through an $ORIGIN rpath."""
    deps = [CLI_SRC, LIB, os.path.join(HERE, "..", "include", "qmiv.h")]
    if force or not os.path.exists(CLI) or any(os.path.getmtime(d) > os.path.getmtime(CLI) for d in deps):
        os.makedirs(os.path.dirname(CLI), exist_ok=True)
        tmp = CLI + ".tmp"
        _run(["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-o", tmp, CLI_SRC, "-L" + HERE, "-lqmiv",
              "-Wl,-rpath,$ORIGIN/.."])
        os.replace(tmp, CLI)
    return CLI


if __name__ == "__main__":
    print(build(force="--force" in sys.argv, verbose=True))
