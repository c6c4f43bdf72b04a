"""Build libsmx_hip.so (gfx950) in-tree with hipcc.

``python -m smx.build`` (with ``fenicsx-micromagnetics_amd`` on sys.path) or
``__graft_entry__.build()`` compiles every ``csrc/*.hip`` file into one shared
library next to this file.  No JIT cache is used: the built ``.so`` lives in the
source tree so that it travels with the repository snapshot.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
CSRC = PKG_DIR.parent / "csrc"
INCLUDE = PKG_DIR.parent.parent / "include"
LIB_NAME = "libsmx_hip.so"
LIB_PATH = PKG_DIR / LIB_NAME
ARCH = os.environ.get("SMX_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared",
    f"--offload-arch={ARCH}",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(exe).exists():
        raise RuntimeError("hipcc not found; the HIP extension cannot be built")
    return exe


def sources():
    return sorted(CSRC.glob("*.hip"))


def build(verbose: bool = True, jobs: int | None = None) -> Path:
    """Compile all kernels into LIB_PATH (always recompiles): every source to
    an object in parallel, then one link."""
    from concurrent.futures import ThreadPoolExecutor
    import tempfile
    srcs = sources()
    if not srcs:
        raise RuntimeError(f"no HIP sources under {CSRC}")
    exe = hipcc()
    objflags = [f for f in HIPCC_FLAGS if f != "-shared"]
    with tempfile.TemporaryDirectory(prefix="smx_build_") as tmpd:
        objs = [Path(tmpd) / (src.stem + ".o") for src in srcs]

        def compile_one(pair):
            src, obj = pair
            cmd = [exe, *objflags, f"-I{INCLUDE}", f"-I{CSRC}", "-c", str(src), "-o", str(obj)]
            return cmd, subprocess.run(cmd, capture_output=True, text=True)

        n = jobs or min(len(srcs), max(1, (os.cpu_count() or 1)))
        with ThreadPoolExecutor(max_workers=n) as pool:
            results = list(pool.map(compile_one, zip(srcs, objs)))
        for cmd, res in results:
            if verbose:
                print("[smx.build]", " ".join(cmd), flush=True)
            if res.returncode != 0:
                sys.stderr.write(res.stdout + res.stderr)
                raise RuntimeError(f"hipcc failed with exit code {res.returncode}")
        tmp = LIB_PATH.with_suffix(".so.tmp")
        cmd = [exe, *HIPCC_FLAGS, "-o", str(tmp), *map(str, objs)]
        if verbose:
            print("[smx.build]", " ".join(cmd), flush=True)
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            raise RuntimeError(f"hipcc link failed with exit code {res.returncode}")
    os.replace(tmp, LIB_PATH)
    if verbose:
        print(f"[smx.build] wrote {LIB_PATH}", flush=True)
    return LIB_PATH


if __name__ == "__main__":
    build()
