"""Build libmpfhip.so for gfx950 from the HIP sources next to this file.

    python -m mpifield_amd.csrc.build        # or __graft_entry__.build()

Each ``*.hip`` compiles to an object under ``mpifield_amd/csrc/_build`` (in parallel), and
the objects link into ``mpifield_amd/libmpfhip.so``. Flags:

- ``--offload-arch=gfx950``: MI355X only, no other targets;
- ``-ffp-contract=off``: no implicit multiply-add fusion, so the cell selection and the
  FP64 paths keep the reference's rounding; kernels fuse with explicit ``fma()``;
- ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed FP32 instructions.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
PACKAGE = HERE.parent
ROOT = PACKAGE.parent
OUT = PACKAGE / "libmpfhip.so"
BUILD = HERE / "_build"

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

COMMON = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    "--offload-arch=gfx950",
    "-ffp-contract=off",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
    f"-I{ROOT / 'include'}",
]


def sources() -> list[Path]:
    return sorted(HERE.glob("*.hip"))


def _compile(src: Path) -> Path:
    obj = BUILD / (src.stem + ".o")
    cmd = [HIPCC, *COMMON, "-c", str(src), "-o", str(obj)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src.name}:\n{proc.stdout}\n{proc.stderr}")
    if proc.stderr.strip():
        # keep warnings visible but quiet the host pass's note about the packed-FP32 feature
        lines = [ln for ln in proc.stderr.splitlines() if "packed-fp32-ops" not in ln]
        if any(ln.strip() for ln in lines):
            print("\n".join(lines), file=sys.stderr)
    return obj


def build(verbose: bool = True) -> Path:
    """Compile every HIP source and link the shared library; returns its path."""
    BUILD.mkdir(exist_ok=True)
    srcs = sources()
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as pool:
        objs = list(pool.map(_compile, srcs))
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"link failed:\n{proc.stdout}\n{proc.stderr}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT} from {len(srcs)} sources")
    return OUT


if __name__ == "__main__":
    build()
