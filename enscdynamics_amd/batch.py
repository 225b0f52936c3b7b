"""Batched and served execution (ensc_batch_* and ensc_server_* in
include/ensc_engine.h).

``Batch`` runs many independent meshes in one engine: one launch of each
kernel per step iteration advances every unfinished job, each with its own
time step and end time.  ``Server`` is the continuous-batching request loop:
jobs are admitted by device footprint as memory frees up, stepped together
with everything already live, and retired at their end time.
"""
from __future__ import annotations

import ctypes as C
from typing import Dict, List, Sequence, Tuple

import numpy as np

from . import engine as _eng
from .model import Model


def _lib():
    L = _eng.load_library()
    if not getattr(L, "_batch_ready", False):
        vp = C.c_void_p
        L.ensc_batch_create.restype = C.c_int
        L.ensc_batch_create.argtypes = [C.POINTER(C.c_void_p), C.c_uint32, vp, C.POINTER(vp)]
        L.ensc_batch_size.restype = C.c_uint32
        L.ensc_batch_size.argtypes = [vp]
        L.ensc_batch_job_scalars.restype = C.c_int
        L.ensc_batch_job_scalars.argtypes = [vp, C.c_uint32, vp]
        L.ensc_batch_job_download.restype = C.c_int
        L.ensc_batch_job_download.argtypes = [vp, C.c_uint32, C.c_char_p, vp, C.c_int64]
        L.ensc_engine_set_graph.restype = C.c_int
        L.ensc_engine_set_graph.argtypes = [vp, C.c_int32]
        L.ensc_server_create.restype = C.c_int
        L.ensc_server_create.argtypes = [vp, C.c_int64, C.c_int32, C.POINTER(vp)]
        L.ensc_server_destroy.restype = C.c_int
        L.ensc_server_destroy.argtypes = [vp]
        L.ensc_server_set_graph.restype = C.c_int
        L.ensc_server_set_graph.argtypes = [vp, C.c_int32]
        L.ensc_server_submit.restype = C.c_int
        L.ensc_server_submit.argtypes = [vp, vp, C.c_int32, C.c_double, C.POINTER(C.c_uint64)]
        L.ensc_server_pump.restype = C.c_int
        L.ensc_server_pump.argtypes = [vp, C.c_int32, C.POINTER(C.c_uint64), C.c_uint32, C.POINTER(C.c_uint32)]
        L.ensc_server_result.restype = C.c_int
        L.ensc_server_result.argtypes = [vp, C.c_uint64, vp, vp, C.c_int64, vp]
        L.ensc_server_status.restype = C.c_int
        L.ensc_server_status.argtypes = [vp, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_int64)]
        L.ensc_server_clock.restype = C.c_double
        L.ensc_server_clock.argtypes = [vp]
        L._batch_ready = True
    return L


def _options(precision: int, reorder: bool, block_elem: int, staging: bool, waves: int = 0,
             fp64_jobs: int = 0, fp64_order: int = 0) -> _eng.Options:
    o = _eng.Options()
    o.precision = precision
    o.reorder = int(reorder)
    o.block_elem = block_elem
    o.block_node = 0
    o.device = 0
    o.n_probe = 0
    o.probe_nodes = None
    o.hist_cap = 0
    o.staging = (0 if staging else 2) if isinstance(staging, bool) else int(staging)
    o.waves = waves
    o.fp64_jobs = fp64_jobs
    o.fp64_order = fp64_order
    return o


class Batch:
    """Many meshes stepped together, each to its own end time."""

    def __init__(self, models: Sequence[Model], precision: int = _eng.FP64, reorder: bool = True,
                 block_elem: int = 0, graph_ticks: int = 0, staging: bool = True, waves: int = 0,
                 fp64_jobs: int = 0, fp64_order: int = 0):
        L = _lib()
        self.models = list(models)
        self._descs = [m.desc() for m in self.models]
        arr = (C.c_void_p * len(self._descs))(*[C.addressof(d) for d in self._descs])
        self._opts = _options(precision, reorder, block_elem, staging, waves, fp64_jobs, fp64_order)
        h = C.c_void_p()
        _eng._check(L.ensc_batch_create(arr, len(self._descs), C.byref(self._opts), C.byref(h)))
        self.h = h
        if graph_ticks:
            _eng._check(L.ensc_engine_set_graph(self.h, int(graph_ticks)))

    def close(self) -> None:
        if getattr(self, "h", None):
            _lib().ensc_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self) -> "Batch":
        _eng._check(_lib().ensc_engine_run(self.h))
        return self

    # step by step, as Engine: init, then per step set_current_step and advance
    def init(self) -> "Batch":
        _eng._check(_lib().ensc_engine_init(self.h))
        return self

    def set_current_step(self, i: int) -> None:
        _eng._check(_lib().ensc_engine_set_step(self.h, int(i)))

    def advance(self, max_steps: int = 0) -> bool:
        done = C.c_int32()
        _eng._check(_lib().ensc_engine_advance(self.h, int(max_steps), C.byref(done)))
        return bool(done.value)

    def advance_static(self, max_steps: int = 0) -> bool:
        """advance() through ensc_engine_advance_static (allocates nothing)."""
        done = C.c_int32()
        _eng._check(_lib().ensc_engine_advance_static(self.h, int(max_steps), C.byref(done)))
        return bool(done.value)

    def autotune(self, ticks: int = 20) -> Dict[str, float]:
        """ensc_engine_autotune over every job of the batch."""
        return _eng.autotune_handle(self.h, ticks)

    def __len__(self) -> int:
        return int(_lib().ensc_batch_size(self.h))

    def scalars(self, job: int) -> Dict[str, float]:
        o = np.zeros(len(_eng.S_NAMES))
        _eng._check(_lib().ensc_batch_job_scalars(self.h, job, o.ctypes.data))
        return dict(zip(_eng.S_NAMES, o.tolist()))

    def get(self, job: int, name: str) -> np.ndarray:
        m = self.models[job]
        shapes = {"u": (m.n_nodes, 3), "v": (m.n_nodes, 3), "x": (m.n_nodes, 3), "rf": (m.n_nodes, 3),
                  "sigma": (m.n_elems, 6), "qia": (m.n_elems, 12), "epsp": (m.n_elems,)}
        out = np.zeros(shapes[name])
        _eng._check(_lib().ensc_batch_job_download(self.h, job, name.encode(), out.ctypes.data, out.size))
        return out


class Server:
    """Continuous-batching solve server."""

    def __init__(self, precision: int = _eng.FP64, budget_bytes: int = 64 << 30, max_live: int = 1024,
                 reorder: bool = True, graph_ticks: int = 0, block_elem: int = 0, staging: bool = True,
                 waves: int = 0, fp64_jobs: int = 0, fp64_order: int = 0):
        L = _lib()
        self._opts = _options(precision, reorder, block_elem, staging, waves, fp64_jobs, fp64_order)
        h = C.c_void_p()
        _eng._check(L.ensc_server_create(C.byref(self._opts), int(budget_bytes), int(max_live), C.byref(h)))
        self.h = h
        if graph_ticks:
            _eng._check(L.ensc_server_set_graph(self.h, int(graph_ticks)))
        self._nodes: Dict[int, int] = {}

    def close(self) -> None:
        if getattr(self, "h", None):
            _lib().ensc_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def clock(self) -> float:
        return float(_lib().ensc_server_clock(self.h))

    def submit(self, model: Model, keep: bool = False, t_arrival: float = -1.0) -> int:
        d = model.desc()
        t = C.c_uint64()
        _eng._check(_lib().ensc_server_submit(self.h, C.byref(d), int(keep), float(t_arrival), C.byref(t)))
        self._nodes[t.value] = model.n_nodes
        return int(t.value)

    def pump(self, ticks: int, cap: int = 4096) -> List[int]:
        buf = (C.c_uint64 * cap)()
        n = C.c_uint32()
        _eng._check(_lib().ensc_server_pump(self.h, int(ticks), buf, cap, C.byref(n)))
        return [int(buf[i]) for i in range(min(n.value, cap))]

    def status(self) -> Tuple[int, int, int]:
        q, l, b = C.c_uint32(), C.c_uint32(), C.c_int64()
        _eng._check(_lib().ensc_server_status(self.h, C.byref(q), C.byref(l), C.byref(b)))
        return q.value, l.value, b.value

    def result(self, ticket: int, with_u: bool = False):
        sc = np.zeros(len(_eng.S_NAMES))
        times = np.zeros(3)
        u = np.zeros((self._nodes[ticket], 3)) if with_u else None
        _eng._check(_lib().ensc_server_result(self.h, ticket, sc.ctypes.data, u.ctypes.data if with_u else None,
                                              u.size if with_u else 0, times.ctypes.data))
        return dict(zip(_eng.S_NAMES, sc.tolist())), u, times
