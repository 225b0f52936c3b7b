"""Host-side job description: the data exDyna3D's reader fills.

``Model`` holds, as numpy arrays, what the reference's DataIn writes into
exDyna3D (Explicit3D/include/exDyna3D.h:141-220, filled by
Explicit3D/source/dataIn.cpp:1550-1788): node coordinates, C3D8R
connectivity, the material, the per-step boundary conditions with node sets
resolved, the initial velocity, amplitudes and the solver parameters.
``Model.desc()`` flattens it into the C struct ``ensc_model_desc``
(include/ensc_model.h) that the engine's C ABI takes.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np


class BcEntry(C.Structure):
    _fields_ = [
        ("nodes", C.POINTER(C.c_uint32)),
        ("n_nodes", C.c_uint32),
        ("d0", C.c_int32),
        ("d1", C.c_int32),
        ("cond", C.c_int32),
        ("value", C.c_double),
        ("amp", C.c_int32),
        ("pad_", C.c_int32),
    ]


class DsloadEntry(C.Structure):
    _fields_ = [
        ("elems", C.POINTER(C.c_uint32)),
        ("faces", C.POINTER(C.c_int32)),
        ("n_faces", C.c_uint32),
        ("amp", C.c_int32),
        ("value", C.c_double),
    ]


class StepDef(C.Structure):
    _fields_ = [
        ("period", C.c_double),
        ("reset_spc", C.c_int32),
        ("reset_vel", C.c_int32),
        ("reset_dload", C.c_int32),
        ("grav_on", C.c_int32),
        ("grav_value", C.c_double),
        ("grav_dir", C.c_double * 3),
        ("grav_amp", C.c_int32),
        ("n_spc", C.c_uint32),
        ("spc", C.POINTER(BcEntry)),
        ("n_vel", C.c_uint32),
        ("pad_", C.c_int32),
        ("vel", C.POINTER(BcEntry)),
        ("reset_dsload", C.c_int32),
        ("n_dsload", C.c_uint32),
        ("dsload", C.POINTER(DsloadEntry)),
    ]


class Amplitude(C.Structure):
    _fields_ = [
        ("t", C.POINTER(C.c_double)),
        ("a", C.POINTER(C.c_double)),
        ("n", C.c_uint32),
        ("pad_", C.c_int32),
    ]


class ModelDesc(C.Structure):
    _fields_ = [
        ("n_nodes", C.c_uint32),
        ("n_elems", C.c_uint32),
        ("x0", C.POINTER(C.c_double)),
        ("conn", C.POINTER(C.c_uint32)),
        ("E", C.c_double),
        ("nu", C.c_double),
        ("rho", C.c_double),
        ("plastic", C.c_int32),
        ("pad0_", C.c_int32),
        ("yield0", C.c_double),
        ("hard", C.c_double),
        ("c_dr", C.c_double),
        ("c_cr", C.c_double),
        ("factor", C.c_double),
        ("cvisl", C.c_double),
        ("cvisq", C.c_double),
        ("total_time", C.c_double),
        ("v_init", C.POINTER(C.c_double)),
        ("n_steps", C.c_uint32),
        ("n_amps", C.c_uint32),
        ("steps", C.POINTER(StepDef)),
        ("amps", C.POINTER(Amplitude)),
    ]


@dataclass
class Bc:
    """One line of a *Boundary block with its node set resolved.

    ``kind`` is "spc" or "vel"; dofs ``d0..d1`` are 0-based and inclusive
    (BoundaryCondition.h:13-19).  ``cond`` is the SPC line's third numeric
    field: only 1 constrains (exDyna3D.cpp:478).
    """

    kind: str
    nodes: np.ndarray
    d0: int
    d1: int
    cond: int = 1
    value: float = 0.0
    amp: int = -1
    name: str = ""


@dataclass
class Dsload:
    """One *Dsload pressure line with its surface resolved (extension: the
    reference parses *Dsload and never applies it).  Face k of element e is
    C3D8 face S(k+1); a positive value pushes into the element."""

    elems: np.ndarray
    faces: np.ndarray
    value: float
    amp: int = -1


@dataclass
class Step:
    """One *Step (StepData.h:14-34)."""

    name: str
    period: float
    spc: List[Bc] = field(default_factory=list)
    vel: List[Bc] = field(default_factory=list)
    reset_spc: bool = False
    reset_vel: bool = False
    reset_dload: bool = False
    gravity: Optional[Tuple[float, float, float, float, int]] = None  # g, dx, dy, dz, amp
    dsload: List[Dsload] = field(default_factory=list)
    reset_dsload: bool = False


@dataclass
class Model:
    x0: np.ndarray                 # (Nn, 3) float64
    conn: np.ndarray               # (Ne, 8) uint32, 0-based
    E: float = 2.06e11             # Material_elastic.h defaults
    nu: float = 0.3
    rho: float = 7850.0
    plastic: bool = False
    yield0: float = 0.0
    hard: float = 0.0
    c_dr: float = 0.013            # read_project_txt defaults (exDyna3D.cpp:700-702)
    c_cr: float = 0.05
    factor: float = 0.71
    cvisl: float = 0.06            # exDyna3D ctor defaults (exDyna3D.cpp:34-35)
    cvisq: float = 1.2
    total_time: float = 0.0
    v_init: Optional[np.ndarray] = None   # (Nn, 3)
    steps: List[Step] = field(default_factory=list)   # steps[0] = "Initial"
    amplitudes: List[Tuple[str, np.ndarray]] = field(default_factory=list)  # (name, (n,2))
    node_sets: Dict[str, np.ndarray] = field(default_factory=dict)
    # *Initial Conditions, type=VELOCITY lines: (set name, dof 0..2, value)
    ic: List[Tuple[str, int, float]] = field(default_factory=list)
    name: str = "model"

    @property
    def n_nodes(self) -> int:
        return int(self.x0.shape[0])

    @property
    def n_elems(self) -> int:
        return int(self.conn.shape[0])

    def material_consts(self) -> Tuple[float, float, float, float]:
        """G, K, lambda, wave speed as Material_elastic::update computes them."""
        E, v, rho = self.E, self.nu, self.rho
        G = E / ((1.0 + v) * 2.0)
        K = E / (1.0 - 2.0 * v) / 3.0
        lam = 2.0 * G * v / (1.0 - 2.0 * v)
        wos = float(np.sqrt((lam + 2.0 * G) / rho))
        return G, K, lam, wos

    def desc(self) -> ModelDesc:
        """Build the C description.  The returned struct keeps references to
        every buffer it points at (``_keep``)."""
        keep: list = []

        def dptr(a, dt, ct):
            a = np.ascontiguousarray(a, dtype=dt)
            keep.append(a)
            return a.ctypes.data_as(C.POINTER(ct))

        x0 = dptr(self.x0.reshape(-1), np.float64, C.c_double)
        conn = dptr(self.conn.reshape(-1), np.uint32, C.c_uint32)
        vi = self.v_init if self.v_init is not None else np.zeros((self.n_nodes, 3))
        vptr = dptr(np.asarray(vi).reshape(-1), np.float64, C.c_double)

        def entries(bcs: Sequence[Bc]):
            arr = (BcEntry * max(1, len(bcs)))()
            for i, b in enumerate(bcs):
                arr[i].nodes = dptr(np.asarray(b.nodes).reshape(-1), np.uint32, C.c_uint32)
                arr[i].n_nodes = len(b.nodes)
                arr[i].d0 = b.d0
                arr[i].d1 = b.d1
                arr[i].cond = b.cond
                arr[i].value = b.value
                arr[i].amp = b.amp
            keep.append(arr)
            return arr

        steps = (StepDef * max(1, len(self.steps)))()
        for i, s in enumerate(self.steps):
            st = steps[i]
            st.period = s.period
            st.reset_spc = int(s.reset_spc)
            st.reset_vel = int(s.reset_vel)
            st.reset_dload = int(s.reset_dload)
            if s.gravity is not None:
                st.grav_on = 1
                st.grav_value = s.gravity[0]
                st.grav_dir[0], st.grav_dir[1], st.grav_dir[2] = s.gravity[1:4]
                st.grav_amp = int(s.gravity[4])
            else:
                st.grav_on = 0
                st.grav_amp = -1
            sp = entries(s.spc)
            ve = entries(s.vel)
            ds = (DsloadEntry * max(1, len(s.dsload)))()
            for i, d in enumerate(s.dsload):
                ds[i].elems = dptr(np.asarray(d.elems).reshape(-1), np.uint32, C.c_uint32)
                ds[i].faces = dptr(np.asarray(d.faces).reshape(-1), np.int32, C.c_int32)
                ds[i].n_faces = len(d.elems)
                ds[i].amp = d.amp
                ds[i].value = d.value
            keep.append(ds)
            st.reset_dsload = int(s.reset_dsload)
            st.n_dsload = len(s.dsload)
            st.dsload = C.cast(ds, C.POINTER(DsloadEntry))
            st.n_spc = len(s.spc)
            st.spc = C.cast(sp, C.POINTER(BcEntry))
            st.n_vel = len(s.vel)
            st.vel = C.cast(ve, C.POINTER(BcEntry))
        keep.append(steps)
        amps = (Amplitude * max(1, len(self.amplitudes)))()
        for i, (_, tab) in enumerate(self.amplitudes):
            tab = np.asarray(tab, dtype=np.float64).reshape(-1, 2)
            amps[i].t = dptr(tab[:, 0].copy(), np.float64, C.c_double)
            amps[i].a = dptr(tab[:, 1].copy(), np.float64, C.c_double)
            amps[i].n = tab.shape[0]
        keep.append(amps)
        d = ModelDesc()
        d.n_nodes = self.n_nodes
        d.n_elems = self.n_elems
        d.x0 = x0
        d.conn = conn
        d.E, d.nu, d.rho = self.E, self.nu, self.rho
        d.plastic = int(self.plastic)
        d.yield0, d.hard = self.yield0, self.hard
        d.c_dr, d.c_cr, d.factor = self.c_dr, self.c_cr, self.factor
        d.cvisl, d.cvisq = self.cvisl, self.cvisq
        d.total_time = self.total_time
        d.v_init = vptr
        d.n_steps = len(self.steps)
        d.n_amps = len(self.amplitudes)
        d.steps = C.cast(steps, C.POINTER(StepDef))
        d.amps = C.cast(amps, C.POINTER(Amplitude))
        d._keep = keep  # type: ignore[attr-defined]
        return d

    def main_step(self) -> Step:
        return self.steps[1]
