"""Python side of the C++ .inp reader (include/ensc_deck.h).

``read_deck(path)`` parses an Abaqus deck with the engine's reader (the
restatement of the reference's DataIn, Explicit3D/source/dataIn.cpp) and
returns a ``Model``; ``read_project(path)`` reads project.txt parameters
(exDyna3D.cpp:698-772).

By default the reader solves what the reference solves: ``*Plastic``,
``AMPLITUDE=`` on velocity BCs and ``*Dsload`` pressures, which the
reference ignores, are dropped.  ``extensions=PLASTIC | VEL_AMPLITUDE |
DSLOAD`` (or ``ALL``) keeps them.
"""
from __future__ import annotations

import ctypes as C
import json
from typing import Optional

import numpy as np

from . import engine as _eng
from .model import Bc, Dsload, Model, ModelDesc, Step


PLASTIC = 1        # ENSC_DECK_PLASTIC
VEL_AMPLITUDE = 2  # ENSC_DECK_VEL_AMPLITUDE
DSLOAD = 4         # ENSC_DECK_DSLOAD
ALL = PLASTIC | VEL_AMPLITUDE | DSLOAD


class Project(C.Structure):
    _fields_ = [("c_dr", C.c_double), ("c_cr", C.c_double), ("factor", C.c_double), ("deck", C.c_char * 1024)]


def _lib():
    L = _eng.load_library()
    if not getattr(L, "_deck_ready", False):
        vp = C.c_void_p
        L.ensc_project_read.restype = C.c_int
        L.ensc_project_read.argtypes = [C.c_char_p, vp]
        L.ensc_deck_read.restype = C.c_int
        L.ensc_deck_read.argtypes = [C.c_char_p, vp, C.POINTER(vp)]
        L.ensc_deck_parse.restype = C.c_int
        L.ensc_deck_parse.argtypes = [C.c_char_p, C.c_size_t, vp, C.POINTER(vp)]
        L.ensc_deck_read_ex.restype = C.c_int
        L.ensc_deck_read_ex.argtypes = [C.c_char_p, vp, C.c_uint32, C.POINTER(vp)]
        L.ensc_deck_parse_ex.restype = C.c_int
        L.ensc_deck_parse_ex.argtypes = [C.c_char_p, C.c_size_t, vp, C.c_uint32, C.POINTER(vp)]
        L.ensc_deck_model.restype = C.c_int
        L.ensc_deck_model.argtypes = [vp, vp]
        L.ensc_deck_json.restype = C.c_int64
        L.ensc_deck_json.argtypes = [vp, C.c_char_p, C.c_int64]
        L.ensc_deck_free.restype = C.c_int
        L.ensc_deck_free.argtypes = [vp]
        L._deck_ready = True
    return L


def read_project(path: str) -> Project:
    p = Project()
    _eng._check(_lib().ensc_project_read(path.encode(), C.byref(p)))
    return p


def _model_from_handle(h, L, name: str) -> Model:
    d = ModelDesc()
    _eng._check(L.ensc_deck_model(h, C.byref(d)))
    n = L.ensc_deck_json(h, None, 0)
    buf = C.create_string_buffer(int(n) + 1)
    L.ensc_deck_json(h, buf, n + 1)
    info = json.loads(buf.value.decode())
    Nn, Ne = d.n_nodes, d.n_elems
    x0 = np.ctypeslib.as_array(d.x0, shape=(Nn * 3,)).reshape(Nn, 3).copy() if Nn else np.zeros((0, 3))
    conn = np.ctypeslib.as_array(d.conn, shape=(Ne * 8,)).reshape(Ne, 8).copy() if Ne else np.zeros((0, 8), np.uint32)
    v = np.ctypeslib.as_array(d.v_init, shape=(Nn * 3,)).reshape(Nn, 3).copy() if Nn else np.zeros((0, 3))
    m = Model(x0=x0, conn=conn.astype(np.uint32), E=d.E, nu=d.nu, rho=d.rho, plastic=bool(d.plastic),
              yield0=d.yield0, hard=d.hard, c_dr=d.c_dr, c_cr=d.c_cr, factor=d.factor, cvisl=d.cvisl,
              cvisq=d.cvisq, total_time=d.total_time, v_init=v, name=name)
    amps = []
    for i in range(d.n_amps):
        a = d.amps[i]
        t = np.ctypeslib.as_array(a.t, shape=(a.n,)).copy() if a.n else np.zeros(0)
        y = np.ctypeslib.as_array(a.a, shape=(a.n,)).copy() if a.n else np.zeros(0)
        amps.append((list(info["amplitudes"].keys())[i], np.stack([t, y], axis=1)))
    m.amplitudes = amps
    steps = []
    for s in range(d.n_steps):
        sd = d.steps[s]
        js = info["steps"][s]
        st = Step(js["name"], sd.period, reset_spc=bool(sd.reset_spc), reset_vel=bool(sd.reset_vel),
                  reset_dload=bool(sd.reset_dload))
        for i in range(sd.n_spc):
            b = sd.spc[i]
            nodes = np.ctypeslib.as_array(b.nodes, shape=(b.n_nodes,)).copy() if b.n_nodes else np.zeros(0, np.uint32)
            st.spc.append(Bc("spc", nodes.astype(np.uint32), b.d0, b.d1, b.cond, name=js["spc"][i][0]))
        for i in range(sd.n_vel):
            b = sd.vel[i]
            nodes = np.ctypeslib.as_array(b.nodes, shape=(b.n_nodes,)).copy() if b.n_nodes else np.zeros(0, np.uint32)
            st.vel.append(Bc("vel", nodes.astype(np.uint32), b.d0, b.d1, value=b.value, amp=b.amp,
                             name=js["vel"][i][0]))
        if sd.grav_on:
            st.gravity = (sd.grav_value, sd.grav_dir[0], sd.grav_dir[1], sd.grav_dir[2], sd.grav_amp)
        st.reset_dsload = bool(sd.reset_dsload)
        for i in range(sd.n_dsload):
            e = sd.dsload[i]
            n = e.n_faces
            el = np.ctypeslib.as_array(e.elems, shape=(n,)).copy() if n else np.zeros(0, np.uint32)
            fc = np.ctypeslib.as_array(e.faces, shape=(n,)).copy() if n else np.zeros(0, np.int32)
            st.dsload.append(Dsload(el.astype(np.uint32), fc.astype(np.int32), e.value, e.amp))
        steps.append(st)
    m.steps = steps
    m.node_sets = {k: np.asarray(v_, dtype=np.uint32) for k, v_ in info["node_sets"].items()}
    m.parsed = info  # type: ignore[attr-defined]
    return m


def read_deck(path: str, project: Optional[Project] = None, extensions: int = 0) -> Model:
    L = _lib()
    h = C.c_void_p()
    _eng._check(L.ensc_deck_read_ex(path.encode(), C.byref(project) if project else None, int(extensions),
                                    C.byref(h)))
    try:
        return _model_from_handle(h, L, path)
    finally:
        L.ensc_deck_free(h)


def parse_deck(text: str, project: Optional[Project] = None, name: str = "deck", extensions: int = 0) -> Model:
    L = _lib()
    h = C.c_void_p()
    raw = text.encode()
    _eng._check(L.ensc_deck_parse_ex(raw, len(raw), C.byref(project) if project else None, int(extensions),
                                     C.byref(h)))
    try:
        return _model_from_handle(h, L, name)
    finally:
        L.ensc_deck_free(h)
