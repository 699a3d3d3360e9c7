"""Build recipe for ``libmgh.so`` — the HIP/CDNA4 (gfx950) C-ABI library.

Every ``csrc/*.hip`` file is compiled on its own (in parallel) with hipcc and
the objects are linked into ``magnelio_amd/libmgh.so`` in-tree, so the
library travels with the source tree.  The library does not link a HIP
runtime of its own (``-no-hip-rt``): it binds at load time to the
``libamdhip64`` already in the process (PyTorch-ROCm's), so device pointers
and streams handed over from torch tensors are valid in both.

Flags that matter for results:

* ``-ffp-contract=off`` — no implicit FMA, so the kernels reproduce the
  reference's separately rounded multiply/add order bit for bit;
* ``-target-feature -packed-fp32-ops`` — no packed FP32 instructions.

Run ``python -m magnelio_amd.build`` (or ``__graft_entry__.build()``).
"""

from __future__ import annotations

import concurrent.futures as _cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "libmgh.so"
OBJ_DIR = PKG / "_build"

ARCH = os.environ.get("MGH_OFFLOAD_ARCH", "gfx950")


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (set HIPCC or install ROCm under /opt/rocm)")


COMMON_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
    "-Wno-unused-but-set-variable",
]


def _compile(src: Path, hipcc: str) -> Path:
    obj = OBJ_DIR / (src.stem + ".o")
    cmd = [hipcc, *COMMON_FLAGS, "-I", str(CSRC), "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass warns that it does not know the packed-fp32 feature;
    # that line is expected and harmless, everything else is shown.
    noise = "is not a recognized feature for this target"
    err = "\n".join(ln for ln in res.stderr.splitlines() if noise not in ln)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src.name}:\n{err}")
    if err.strip():
        print(err, file=sys.stderr)
    return obj


def build(verbose: bool = True) -> Path:
    """Compile every HIP source and link ``libmgh.so``; returns its path."""
    hipcc = _hipcc()
    OBJ_DIR.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip"))
    if not srcs:
        raise RuntimeError(f"no HIP sources under {CSRC}")
    jobs = min(len(srcs), max(1, min(8, os.cpu_count() or 1)))
    with _cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, hipcc), srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc, f"--offload-arch={ARCH}", "-shared", "-no-hip-rt", "-fPIC", *map(str, objs),
           "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link of libmgh.so failed:\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(srcs)} HIP sources for {ARCH}")
    return LIB


if __name__ == "__main__":
    build()
