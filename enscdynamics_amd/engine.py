"""ctypes binding of libensc_hip.so (include/ensc_engine.h).

``Engine`` mirrors the reference solver class exDyna3D
(Explicit3D/include/exDyna3D.h:36-80): ``init()``, ``set_current_step()``,
``run()`` and a step-loop ``advance()``, with the job's state resident on
the GPU.  There is no CPU fallback: constructing an Engine without the built
library or without a GPU raises.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Dict, Optional, Sequence, Tuple

import numpy as np

from .model import Model

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libensc_hip.so")

FP64 = 0
MIXED = 1

S_NAMES = ["time", "dt_i", "dt_i_1", "dts_min", "ke", "ie", "hg", "vis", "wext", "ke0", "nsteps",
           "step_index", "step_end", "err", "done"]
H_BASE = 9
H_NAMES = ["time", "dt_i", "dt_i_1", "dts_min", "ke", "ie", "hg", "vis", "wext"]

EXPORTS = [
    "ensc_last_error", "ensc_device_count", "ensc_engine_create", "ensc_engine_destroy", "ensc_engine_init",
    "ensc_engine_set_step", "ensc_engine_advance", "ensc_engine_advance_static", "ensc_engine_run",
    "ensc_engine_enqueue_steps",
    "ensc_engine_sync", "ensc_engine_download", "ensc_engine_scalars", "ensc_engine_history",
    "ensc_engine_set_graph", "ensc_copy_bandwidth", "ensc_engine_kernel_times", "ensc_engine_phase_times", "ensc_engine_layout", "ensc_engine_set_timing",
    "ensc_engine_set_output", "ensc_engine_output_pending", "ensc_engine_output_field", "ensc_engine_output_write",
    "ensc_engine_output_pop", "ensc_vtu_write", "ensc_engine_device_bytes", "ensc_engine_autotune",
]


class EnscError(RuntimeError):
    pass


class Options(C.Structure):
    _fields_ = [
        ("precision", C.c_int32),
        ("reorder", C.c_int32),
        ("block_elem", C.c_int32),
        ("block_node", C.c_int32),
        ("device", C.c_int32),
        ("n_probe", C.c_uint32),
        ("probe_nodes", C.POINTER(C.c_uint32)),
        ("hist_cap", C.c_int64),
        ("staging", C.c_int32),
        ("waves", C.c_int32),
        ("fp64_jobs", C.c_int32),
        ("fp64_order", C.c_int32),
    ]


_lib = None


def load_library(path: str = LIB_PATH):
    """Load libensc_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise EnscError(f"{path} is missing: build it with `make -C {HERE}` (or __graft_entry__.build())")
    L = C.CDLL(path)
    L.ensc_last_error.restype = C.c_char_p
    L.ensc_device_count.restype = C.c_int32
    vp = C.c_void_p
    for name, args in [
        ("ensc_engine_create", [vp, vp, C.POINTER(vp)]),
        ("ensc_engine_destroy", [vp]),
        ("ensc_engine_init", [vp]),
        ("ensc_engine_set_step", [vp, C.c_uint32]),
        ("ensc_engine_advance", [vp, C.c_int64, C.POINTER(C.c_int32)]),
        ("ensc_engine_advance_static", [vp, C.c_int64, C.POINTER(C.c_int32)]),
        ("ensc_engine_run", [vp]),
        ("ensc_engine_enqueue_steps", [vp, C.c_int64]),
        ("ensc_engine_sync", [vp]),
        ("ensc_engine_download", [vp, C.c_char_p, vp, C.c_int64]),
        ("ensc_engine_scalars", [vp, vp]),
        ("ensc_engine_history", [vp, vp, C.c_int64, C.POINTER(C.c_int64)]),
        ("ensc_engine_kernel_times", [vp, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
        ("ensc_engine_phase_times", [vp, C.c_void_p]),
        ("ensc_engine_layout", [vp, C.c_void_p]),
        ("ensc_engine_set_output", [vp, C.c_double, C.c_int32, C.c_char_p]),
        ("ensc_engine_output_pending", [vp, C.POINTER(C.c_int32), C.POINTER(C.c_double)]),
        ("ensc_engine_output_field", [vp, C.c_char_p, C.c_void_p, C.c_int64]),
        ("ensc_engine_output_write", [vp, C.c_char_p]),
        ("ensc_engine_output_pop", [vp]),
        ("ensc_engine_set_timing", [vp, C.c_int32]),
        ("ensc_engine_autotune", [vp, C.c_int32, C.POINTER(C.c_int64)]),
        ("ensc_engine_set_graph", [vp, C.c_int32]),
        ("ensc_engine_fsi_setup", [vp, C.c_int32, C.c_double, C.POINTER(C.c_int64)]),
        ("ensc_engine_fsi_set_forces", [vp, vp, C.c_int64]),
        ("ensc_engine_fsi_get", [vp, vp, vp, C.c_int64]),
        ("ensc_engine_fsi_device", [vp, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp), C.POINTER(vp)]),
    ]:
        f = getattr(L, name)
        f.restype = C.c_int
        f.argtypes = args
    L.ensc_engine_device_bytes.restype = C.c_int64
    L.ensc_engine_device_bytes.argtypes = [vp]
    L.ensc_copy_bandwidth.restype = C.c_int
    L.ensc_copy_bandwidth.argtypes = [C.c_int32, C.c_int64, C.c_int32, C.POINTER(C.c_double)]
    _lib = L
    return L


def _check(rc: int) -> None:
    if rc != 0:
        msg = _lib.ensc_last_error().decode(errors="replace") if _lib else "?"
        raise EnscError(f"ensc error {rc}: {msg}")


def device_count() -> int:
    return int(load_library().ensc_device_count())


def copy_bandwidth(nbytes: int = 1 << 30, reps: int = 10, device: int = 0) -> float:
    """Read+write GB/s of a plain device copy of `nbytes` (k_copy_probe)."""
    L = load_library()
    g = C.c_double()
    _check(L.ensc_copy_bandwidth(device, int(nbytes), int(reps), C.byref(g)))
    return g.value


def write_history_csv(path: str, rows: np.ndarray, probes: Sequence[int]) -> int:
    """CSV of history rows laid out as ensc_engine_history returns them."""
    cols = list(H_NAMES) + [f"u{c}_node{p}" for p in probes for c in "xyz"]
    rows = np.asarray(rows, dtype=np.float64).reshape(-1, len(cols))
    with open(path, "w") as f:
        f.write(",".join(["step"] + cols) + "\n")
        for i, r in enumerate(rows):
            f.write(",".join([str(i + 1)] + [repr(float(v)) for v in r]) + "\n")
    return rows.shape[0]



def autotune_handle(h, ticks: int) -> Dict[str, float]:
    """ensc_engine_autotune on an engine or batch handle."""
    out = (C.c_int64 * 4)()
    _check(load_library().ensc_engine_autotune(h, int(ticks), out))
    return {"waves": int(out[0]), "block_node": int(out[1]), "pipe_per_cu": int(out[2]),
            "ms_per_tick": out[3] * 1e-6}

class Engine:
    """One job resident on the GPU, driven like exDyna3D."""

    def __init__(self, model: Model, precision: int = FP64, reorder: bool = False, block_elem: int = 0,
                 block_node: int = 0, probes: Sequence[int] = (), hist_cap: int = 0, device: int = 0,
                 staging: bool = True, waves: int = 0, fp64_jobs: int = 0, fp64_order: int = 0):
        L = load_library()
        self.model = model
        self.precision = precision
        self._desc = model.desc()
        pr = np.ascontiguousarray(np.asarray(probes, dtype=np.uint32))
        self._probes = pr
        o = Options()
        o.precision = precision
        o.reorder = int(reorder)
        o.block_elem = block_elem
        o.block_node = block_node
        o.device = device
        o.n_probe = len(pr)
        o.probe_nodes = pr.ctypes.data_as(C.POINTER(C.c_uint32)) if len(pr) else None
        o.hist_cap = hist_cap
        o.staging = (0 if staging else 2) if isinstance(staging, bool) else int(staging)
        o.waves = waves
        o.fp64_jobs = fp64_jobs
        o.fp64_order = fp64_order
        h = C.c_void_p()
        _check(L.ensc_engine_create(C.byref(self._desc), C.byref(o), C.byref(h)))
        self.h = h
        self.Nn, self.Ne = model.n_nodes, model.n_elems

    def close(self) -> None:
        if getattr(self, "h", None):
            _lib.ensc_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- exDyna3D-shaped interface ---
    def init(self) -> "Engine":
        _check(_lib.ensc_engine_init(self.h))
        return self

    def set_current_step(self, i: int) -> None:
        _check(_lib.ensc_engine_set_step(self.h, i))

    def advance(self, max_steps: int = 0) -> bool:
        done = C.c_int32(0)
        _check(_lib.ensc_engine_advance(self.h, int(max_steps), C.byref(done)))
        return bool(done.value)

    def run(self) -> "Engine":
        _check(_lib.ensc_engine_run(self.h))
        return self

    def advance_static(self, max_steps: int = 0) -> bool:
        """advance() through ensc_engine_advance_static (allocates nothing)."""
        done = C.c_int32(0)
        _check(_lib.ensc_engine_advance_static(self.h, int(max_steps), C.byref(done)))
        return bool(done.value)

    def run_steps(self, max_steps: int) -> int:
        """init + step 1 + up to max_steps step iterations; returns steps taken."""
        self.init()
        self.set_current_step(1)
        self.advance(max_steps)
        return int(self.scalars()["nsteps"])

    # --- low level ---
    def enqueue(self, n: int) -> None:
        _check(_lib.ensc_engine_enqueue_steps(self.h, int(n)))

    def sync(self) -> None:
        _check(_lib.ensc_engine_sync(self.h))

    def set_graph(self, ticks: int) -> None:
        """Replay enqueued steps as a hipGraph of `ticks` step iterations
        (0: plain launches)."""
        _check(_lib.ensc_engine_set_graph(self.h, int(ticks)))

    def autotune(self, ticks: int = 20) -> Dict[str, float]:
        """Time the launch candidates on this engine's jobs and keep the
        fastest (ensc_engine_autotune; call after init and set_current_step,
        the jobs continue from where they were)."""
        return autotune_handle(self.h, ticks)

    def set_timing(self, on: bool) -> None:
        _check(_lib.ensc_engine_set_timing(self.h, int(on)))

    def kernel_times(self) -> Tuple[float, float, int]:
        a, b, n = C.c_double(), C.c_double(), C.c_int64()
        _check(_lib.ensc_engine_kernel_times(self.h, C.byref(a), C.byref(b), C.byref(n)))
        return a.value, b.value, n.value

    def phase_times(self) -> Tuple[float, float, float, float]:
        """Mean ms per launch of the last timed enqueue: element pass, element
        finalize, nodal update, node finalize."""
        out = np.zeros(4)
        _check(_lib.ensc_engine_phase_times(self.h, out.ctypes.data))
        return tuple(float(x) for x in out)

    def layout(self) -> Dict[str, int]:
        """Launch layout of the last step iteration (ensc_engine_layout)."""
        out = np.zeros(7, dtype=np.int64)
        _check(_lib.ensc_engine_layout(self.h, out.ctypes.data))
        keys = ("staged", "pipe_blocks", "elem_blocks", "node_blocks", "lds_bytes", "asm_jobs", "fp64_jobs")
        return {k: int(v) for k, v in zip(keys, out)}

    # --- output snapshots (exDyna3D::output_results) ---
    def set_output(self, interval: float = 0.0, slots: int = 4, directory: Optional[str] = None) -> None:
        """Capture output_results' state at every output time (interval <= 0:
        the reference's rule) into `slots` device slots; with a directory,
        advance/run write <directory>/solution_<n>.vtu as they go."""
        _check(_lib.ensc_engine_set_output(self.h, float(interval), int(slots),
                                           directory.encode() if directory else None))

    def output_pending(self) -> Tuple[int, float]:
        n, t = C.c_int32(), C.c_double()
        _check(_lib.ensc_engine_output_pending(self.h, C.byref(n), C.byref(t)))
        return int(n.value), float(t.value)

    def output_field(self, name: str) -> np.ndarray:
        shapes = {"x": (self.Nn, 3), "a": (self.Nn, 3), "v": (self.Nn, 3), "u": (self.Nn, 3),
                  "vonMises": (self.Ne,), "Stress": (self.Ne, 6), "Green-Lagrange Strain": (self.Ne, 6)}
        out = np.zeros(shapes[name])
        _check(_lib.ensc_engine_output_field(self.h, name.encode(), out.ctypes.data, out.size))
        return out

    def output_write(self, path: str) -> None:
        _check(_lib.ensc_engine_output_write(self.h, path.encode()))

    def output_pop(self) -> None:
        _check(_lib.ensc_engine_output_pop(self.h))

    def get(self, name: str) -> np.ndarray:
        shapes = {"u": (self.Nn, 3), "v": (self.Nn, 3), "x": (self.Nn, 3), "rf": (self.Nn, 3),
                  "sigma": (self.Ne, 6), "qia": (self.Ne, 12), "epsp": (self.Ne,)}
        out = np.zeros(shapes[name])
        _check(_lib.ensc_engine_download(self.h, name.encode(), out.ctypes.data, out.size))
        return out

    def scalars(self) -> Dict[str, float]:
        o = np.zeros(len(S_NAMES))
        _check(_lib.ensc_engine_scalars(self.h, o.ctypes.data))
        return dict(zip(S_NAMES, o.tolist()))

    def history(self, rows: Optional[int] = None) -> np.ndarray:
        w = H_BASE + 3 * len(self._probes)
        rows = int(self.scalars()["nsteps"]) if rows is None else rows
        out = np.zeros((max(rows, 0), w))
        got = C.c_int64()
        _check(_lib.ensc_engine_history(self.h, out.ctypes.data if rows else None, rows, C.byref(got)))
        return out[: got.value]

    def write_history_csv(self, path: str) -> int:
        """Per-step history as CSV (one row per step: time, dt_i, dt_i_1,
        min dtStable, kinetic, internal, hourglass and viscous energy,
        external work, then u of each probe node); returns the rows."""
        return write_history_csv(path, self.history(), [int(p) for p in self._probes])

    def device_bytes(self) -> int:
        return int(_lib.ensc_engine_device_bytes(self.h))

    # --- fluid-structure coupling (exDyna3D.cpp:1281-1758) ---
    def fsi_setup(self, n_layers: int, dist: float) -> int:
        """Virtual particles on the FSI boundary (get_fsiSph_virtualParticles_
        and_vel): n_layers per boundary face, dist apart; returns the count.
        After init()."""
        n = C.c_int64()
        _check(_lib.ensc_engine_fsi_setup(self.h, int(n_layers), float(dist), C.byref(n)))
        self._n_fsi = int(n.value)
        return self._n_fsi

    def fsi_set_forces(self, forces: np.ndarray) -> None:
        """The particles' forces (FSI_virtualParticles_nodeForce), (n, 3)."""
        f = np.ascontiguousarray(forces, dtype=np.float64).reshape(-1, 3)
        _check(_lib.ensc_engine_fsi_set_forces(self.h, f.ctypes.data, f.shape[0]))

    def fsi_particles(self) -> Tuple[np.ndarray, np.ndarray]:
        """The particles' positions and velocities, each (n, 3)."""
        n = getattr(self, "_n_fsi", 0)
        x = np.zeros((n, 3))
        v = np.zeros((n, 3))
        _check(_lib.ensc_engine_fsi_get(self.h, x.ctypes.data, v.ctypes.data, n))
        return x, v
