"""Build libdeconhip.so in-tree with hipcc for gfx950.

Usage: python deconlib_amd/_build.py   (does not import the package)
Each translation unit compiles in parallel to an object under _lib/obj, then
one hipcc link produces deconlib_amd/_lib/libdeconhip.so.
"""

from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIBDIR = PKG / "_lib"
LIB = LIBDIR / "libdeconhip.so"
SOURCES = ["dh_abi.hip", "dh_noalloc.hip", "dh_fft.hip", "dh_ops.hip", "dh_pdhg.hip", "dh_nlcg.hip", "dh_erdecon.hip", "dh_jetnewton.hip",
           "dh_ct_dispatch.hip",
           "dh_ct_p2a.hip", "dh_ct_p2b.hip", "dh_ct_p2c.hip", "dh_ct_p2d.hip", "dh_ct_p3a.hip",
           "dh_ct_p3b.hip", "dh_ct_p3c.hip", "dh_ct_p5a.hip", "dh_ct_p5b.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CFLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # packed-FP32 instructions off for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    "-I", str(ROOT / "include"),
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    noise = "packed-fp32-ops"
    err = "\n".join(l for l in r.stderr.splitlines() if noise not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed ({r.returncode}):\n{err}")
    if err and verbose:
        print(err, flush=True)


def build(verbose: bool = True) -> Path:
    objdir = LIBDIR / "obj"
    objdir.mkdir(parents=True, exist_ok=True)
    srcs = [s for s in SOURCES if (CSRC / s).exists()]
    objs = [objdir / (Path(s).stem + ".o") for s in srcs]
    cmds = [[HIPCC, *CFLAGS, "-c", str(CSRC / s), "-o", str(o)] for s, o in zip(srcs, objs)]
    with ThreadPoolExecutor(max_workers=min(len(cmds), 8)) as ex:
        list(ex.map(lambda c: _run(c, verbose), cmds))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)], verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
