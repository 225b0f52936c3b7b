"""ctypes binding of liboracle.so, the CPU FP64 restatement.

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, as the checker or the timed CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from typing import Dict, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_HARNESS = os.path.join(HERE, "_ref", "ref_harness")

HIST_BASE = 9
OH = dict(time=0, dt_i=1, dt_i_1=2, dts_min=3, ke=4, ie=5, hg=6, vis=7, wext=8)
OS_NAMES = ["time", "dt_i", "dt_i_1", "dts_min", "ke", "ie", "hg", "vis", "wext", "ke0", "nsteps",
            "step_index", "lmin"]

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.oracle_create.restype = C.c_void_p
        L.oracle_create.argtypes = [C.c_void_p, C.c_int]
        L.oracle_destroy.argtypes = [C.c_void_p]
        L.oracle_init.argtypes = [C.c_void_p]
        L.oracle_run.restype = C.c_int64
        L.oracle_run.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_uint32, C.c_void_p, C.c_int64]
        L.oracle_time_steps.restype = C.c_double
        L.oracle_time_steps.argtypes = [C.c_void_p, C.c_int64]
        L.oracle_get.restype = C.c_int64
        L.oracle_get.argtypes = [C.c_void_p, C.c_char_p, C.c_void_p]
        L.oracle_scalars.argtypes = [C.c_void_p, C.c_void_p]
        L.oracle_threads.restype = C.c_int
        L.oracle_threads.argtypes = [C.c_void_p]
        L.oracle_set_output.argtypes = [C.c_void_p, C.c_double]
        L.oracle_snapshots.restype = C.c_int64
        L.oracle_snapshots.argtypes = [C.c_void_p]
        L.oracle_snapshot.restype = C.c_int64
        L.oracle_snapshot.argtypes = [C.c_void_p, C.c_int64, C.c_char_p, C.c_void_p]
        L.oracle_begin_step.argtypes = [C.c_void_p]
        L.oracle_fsi_setup.restype = C.c_int64
        L.oracle_fsi_setup.argtypes = [C.c_void_p, C.c_int, C.c_double]
        L.oracle_fsi_set_forces.restype = C.c_int
        L.oracle_fsi_set_forces.argtypes = [C.c_void_p, C.c_void_p, C.c_int64]
        L.oracle_fsi_get.restype = C.c_int64
        L.oracle_fsi_get.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        _lib = L
    return _lib


class Oracle:
    """exDyna3D restated on the CPU: init(), then run() the step loop."""

    def __init__(self, model, n_threads: int = 0):
        self.model = model
        self._desc = model.desc()
        self.h = lib().oracle_create(C.byref(self._desc), n_threads)
        self.Nn, self.Ne = model.n_nodes, model.n_elems

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            lib().oracle_destroy(h)
            self.h = None

    def init(self) -> "Oracle":
        lib().oracle_init(self.h)
        return self

    def run(self, max_steps: int = 0, probes: Sequence[int] = (), hist_rows: int = 0):
        pr = np.ascontiguousarray(np.asarray(probes, dtype=np.uint32))
        w = HIST_BASE + 3 * len(pr)
        hist = np.zeros((max(hist_rows, 0), w))
        n = lib().oracle_run(self.h, int(max_steps), pr.ctypes.data if len(pr) else None, len(pr),
                             hist.ctypes.data if hist_rows else None, hist_rows)
        return int(n), hist[: min(n, hist_rows)]

    def time_steps(self, n: int) -> float:
        return float(lib().oracle_time_steps(self.h, int(n)))

    def get(self, name: str) -> np.ndarray:
        shapes = {"u": (self.Nn, 3), "v": (self.Nn, 3), "x": (self.Nn, 3), "a": (self.Nn, 3),
                  "rf": (self.Nn, 3), "sigma": (self.Ne, 6), "qia": (self.Ne, 12), "epsp": (self.Ne,),
                  "dtStable": (self.Ne,), "Length": (self.Ne,), "V0": (self.Ne,), "B0": (self.Ne, 8, 3),
                  "invM": (self.Nn,)}
        out = np.zeros(shapes[name])
        n = lib().oracle_get(self.h, name.encode(), out.ctypes.data)
        if n < 0:
            raise KeyError(name)
        return out

    def set_output(self, interval: float = 0.0) -> "Oracle":
        """Record output_results snapshots (interval <= 0: the reference's rule)."""
        lib().oracle_set_output(self.h, float(interval))
        return self

    def snapshots(self) -> int:
        return int(lib().oracle_snapshots(self.h))

    def snapshot(self, k: int, name: str) -> np.ndarray:
        shapes = {"time": (1,), "x": (self.Nn, 3), "a": (self.Nn, 3), "v": (self.Nn, 3), "u": (self.Nn, 3),
                  "vonMises": (self.Ne,), "Stress": (self.Ne, 6), "Green-Lagrange Strain": (self.Ne, 6)}
        out = np.zeros(shapes[name])
        if lib().oracle_snapshot(self.h, int(k), name.encode(), out.ctypes.data) < 0:
            raise KeyError((k, name))
        return out

    def begin_step(self) -> "Oracle":
        """Enter the next *Step (setCurrentStep) without stepping."""
        lib().oracle_begin_step(self.h)
        return self

    # fluid-structure coupling (exDyna3D.cpp:1281-1758)
    def fsi_setup(self, n_layers: int, dist: float) -> int:
        self._n_fsi = int(lib().oracle_fsi_setup(self.h, int(n_layers), float(dist)))
        return self._n_fsi

    def fsi_set_forces(self, forces: np.ndarray) -> None:
        f = np.ascontiguousarray(forces, dtype=np.float64).reshape(-1, 3)
        if lib().oracle_fsi_set_forces(self.h, f.ctypes.data, f.shape[0]) != 0:
            raise ValueError("force count differs from the particle count")

    def fsi_particles(self):
        n = getattr(self, "_n_fsi", 0)
        x = np.zeros((n, 3))
        v = np.zeros((n, 3))
        lib().oracle_fsi_get(self.h, x.ctypes.data, v.ctypes.data)
        return x, v

    def scalars(self) -> Dict[str, float]:
        o = np.zeros(len(OS_NAMES))
        lib().oracle_scalars(self.h, o.ctypes.data)
        return dict(zip(OS_NAMES, o.tolist()))

    @property
    def threads(self) -> int:
        return int(lib().oracle_threads(self.h))
