"""Builds the gfx950 kernel library ``pararealml_amd/libpml.so`` in-tree.

Usage: ``python -m pararealml_amd.build`` (or ``build()`` from Python).
Every HIP source under ``csrc/`` is compiled for gfx950 only, in parallel,
and linked into one shared library exporting the C ABI of ``include/pml.h``.
"""

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD_DIR = CSRC / "_build"
LIB = PKG / "libpml.so"
ARCH = "gfx950"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-variable",
    "-Wno-unused-but-set-variable",
    # No FMA contraction: every kernel variant then rounds identically, so
    # results do not depend on the tuned variant, and FP64 results follow
    # the unfused operation order of the NumPy reference.
    "-ffp-contract=off",
    # Packed FP32 instructions are kept out of every kernel.
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
]


def _compile(src: Path) -> Path:
    obj = BUILD_DIR / (src.stem + ".o")
    cmd = [HIPCC, *COMMON_FLAGS, "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(
            f"hipcc failed for {src.name}:\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}"
        )
    return obj


def _stale(src: Path, headers) -> bool:
    obj = BUILD_DIR / (src.stem + ".o")
    if not obj.exists():
        return True
    t = obj.stat().st_mtime
    return src.stat().st_mtime > t or any(h.stat().st_mtime > t for h in headers)


def build(verbose: bool = True, incremental: bool = False) -> Path:
    """Compiles every ``csrc/*.hip`` and links ``libpml.so``.  Always
    rebuilds every object unless ``incremental`` (then only objects older
    than their source or any header)."""
    if not Path(HIPCC).exists() and shutil.which(HIPCC) is None:
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    BUILD_DIR.mkdir(parents=True, exist_ok=True)
    sources = sorted(CSRC.glob("*.hip"))
    headers = sorted(CSRC.glob("*.h")) + sorted((PKG.parent / "include").glob("*.h"))
    todo = [s for s in sources if not incremental or _stale(s, headers)]
    workers = max(1, min(max(len(todo), 1), os.cpu_count() or 1, 8))
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        list(ex.map(_compile, todo))
    objs = [BUILD_DIR / (s.stem + ".o") for s in sources]
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-o", str(tmp), *map(str, objs)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(sources)} sources", file=sys.stderr)
    return LIB


if __name__ == "__main__":
    build(incremental="--incremental" in sys.argv[1:])
