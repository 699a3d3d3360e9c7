"""Build the oracle's C restatement (``oracle/libfit_omp.so``) with gcc.

TEST INFRASTRUCTURE ONLY: the library is the CPU checker / CPU baseline, it
is never loaded by the product.  ``load()`` returns a ctypes handle.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
SRC = HERE / "fit_omp.c"
LIB = HERE / "libfit_omp.so"


def build(verbose: bool = True) -> Path:
    tmp = LIB.with_suffix(".so.tmp")
    cmd = ["gcc", "-O3", "-march=x86-64-v2", "-fopenmp", "-ffp-contract=off", "-fPIC",
           "-shared", str(SRC), "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"gcc failed on {SRC.name}:\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


_lib = None


def load() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build(verbose=False)
        lib = ctypes.CDLL(str(LIB))
        fp = ctypes.c_void_p
        i = ctypes.c_int
        for name in ("fit_omp_update_e_f32", "fit_omp_update_h_f32"):
            fn = getattr(lib, name)
            fn.restype = None
            fn.argtypes = [fp, fp, fp, fp, i, i, i]
        lib.fit_omp_max_threads.restype = i
        lib.fit_omp_set_threads.argtypes = [i]
        _lib = lib
    return _lib


def update_e_f32(e, h, ae, be, nx, ny, nz) -> None:
    """In place on float32 C-contiguous numpy arrays (reference layout)."""
    load().fit_omp_update_e_f32(e.ctypes.data, h.ctypes.data, ae.ctypes.data, be.ctypes.data,
                                nx, ny, nz)


def update_h_f32(e, h, ah, bh, nx, ny, nz) -> None:
    load().fit_omp_update_h_f32(e.ctypes.data, h.ctypes.data, ah.ctypes.data, bh.ctypes.data,
                                nx, ny, nz)


if __name__ == "__main__":
    build()
