"""Builds and loads tests/native/adjoint_shim.hip: the library's solver and
adjoint device functions compiled for the host (TEST INFRASTRUCTURE)."""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "native", "adjoint_shim.hip")
OUT = os.path.join(HERE, "_build", "libadjoint_shim.so")
HDRS = [os.path.join(ROOT, "samoa_amd", "csrc", h) for h in ("swe_riemann.hpp", "swe_adjoint.hpp")]

_lib = None
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def load():
    global _lib
    if _lib is None:
        newest = max(os.path.getmtime(p) for p in [SRC, *HDRS])
        if not os.path.exists(OUT) or os.path.getmtime(OUT) < newest:
            os.makedirs(os.path.dirname(OUT), exist_ok=True)
            subprocess.run(["hipcc", "-O2", "-fPIC", "-shared", "-std=c++17", "-ffp-contract=off",
                            "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang",
                            "-packed-fp32-ops", f"-I{os.path.join(ROOT, 'samoa_amd', 'csrc')}",
                            SRC, "-o", f"{OUT}.tmp{os.getpid()}"], check=True,
                           capture_output=True)
            os.replace(f"{OUT}.tmp{os.getpid()}", OUT)  # parallel test workers may race here
        L = C.CDLL(OUT)
        L.shim_edge_flux.argtypes = [C.c_int, C.c_double, C.c_double, _dp, _dp, C.c_double,
                                     C.c_double, _dp, _dp]
        L.shim_edge_flux_adj.argtypes = [C.c_int, C.c_double, C.c_double, _dp, _dp, C.c_double,
                                         C.c_double, _dp, _dp, _dp, _dp]
        _lib = L
    return _lib


def edge_flux(solver, n, QL, QR, drytol, g):
    fL, fR = np.zeros(4), np.zeros(4)
    load().shim_edge_flux(solver, n[0], n[1], np.ascontiguousarray(QL, float),
                          np.ascontiguousarray(QR, float), drytol, g, fL, fR)
    return fL, fR


def edge_flux_adj(solver, n, QL, QR, drytol, g, gfL, gfR):
    gQL, gQR = np.zeros(4), np.zeros(4)
    load().shim_edge_flux_adj(solver, n[0], n[1], np.ascontiguousarray(QL, float),
                              np.ascontiguousarray(QR, float), drytol, g,
                              np.ascontiguousarray(gfL, float), np.ascontiguousarray(gfR, float),
                              gQL, gQR)
    return gQL, gQR
