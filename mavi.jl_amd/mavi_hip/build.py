"""Build recipe for libmavi_hip.so (hipcc, gfx950).

The library is compiled in-tree so that the built ``.so`` sits next to this
package and travels with the source tree.  ``build_library()`` always
recompiles; ``library_path()`` is where the loader looks.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
CSRC = PKG_DIR.parent / "csrc"
INCLUDE = PKG_DIR.parent.parent / "include"
OBJ_DIR = PKG_DIR.parent / "_build"

SOURCES = ["mavi_kernels.hip", "mavi_capi.hip", "mavi_batch.hip", "mavi_mlp.hip", "mavi_learned.hip", "mavi_nlist.hip"]

HIPCC_FLAGS = [
    "-O3",
    "-std=c++17",
    "--offload-arch=gfx950",
    "-fPIC",
    # Julia does not contract a*b+c: keep every unfused expression unfused and
    # fuse only where the kernels call __builtin_fma explicitly.
    "-ffp-contract=off",
    # No packed FP32 instructions in any kernel.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def library_path() -> Path:
    return PKG_DIR / "libmavi_hip.so"


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


def _compile(src: Path, obj: Path) -> None:
    cmd = [hipcc(), *HIPCC_FLAGS, f"-I{INCLUDE}", "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{r.stdout}\n{r.stderr}")


def build_library(verbose: bool = False) -> Path:
    OBJ_DIR.mkdir(exist_ok=True)
    srcs = [CSRC / s for s in SOURCES if (CSRC / s).exists()]
    objs = [OBJ_DIR / (s.stem + ".o") for s in srcs]
    with ThreadPoolExecutor(max_workers=min(4, len(srcs))) as ex:
        list(ex.map(lambda so: _compile(*so), zip(srcs, objs)))
    out = library_path()
    tmp = out.with_suffix(".so.tmp")
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, out)
    if verbose:
        print(f"built {out}", file=sys.stderr)
    return out


if __name__ == "__main__":
    build_library(verbose=True)
