"""Small models whose energies and driven motion have closed forms.

Used by the oracle's CPU tests (tests/test_oracle.py) and by the GPU tests
that hold the engine to the oracle on the same models.
"""
from __future__ import annotations

import numpy as np

from enscdynamics_amd import decks
from enscdynamics_amd.model import Bc, Dsload, Model, Step

STEEL = dict(E=200e9, nu=0.3, rho=7800.0)


def free_fall(steps: int = 200, g: float = 9.81, v0=(0.0, 0.0, 0.0)) -> Model:
    """A 3x2x2 block with no BCs under gravity g along -z (*Dload GRAV,
    exDyna3D.cpp:800-826) and a uniform initial velocity v0.  Every node
    moves alike, so nothing deforms: IE = HG = VIS = 0 and W_ext = KE - KE0."""
    x0, conn, _ = decks.block_mesh(3, 2, 2, 0.03, 0.02, 0.02)
    m = Model(x0=x0, conn=conn, **STEEL, name="free_fall")
    m.v_init = np.tile(np.asarray(v0, dtype=np.float64), (m.n_nodes, 1))
    dt0 = decks.stable_dt0(m)
    T = steps * dt0
    m.steps = [Step("Initial", 0.0), Step("Step-1", T, gravity=(g, 0.0, 0.0, -1.0, -1))]
    m.total_time = T
    return m


def driven_column(nz: int = 10, strain: float = 1e-3, transits: float = 20.0) -> Model:
    """Uniaxial-strain column of nz elements (h = 0.01 m): x and y fixed on
    every node, z fixed at z = 0, the top face pushed down at a constant
    speed so the nominal strain reaches `strain` after `transits` wave
    transits.  Slow compared with the wave speed, so the work of the top
    reaction approaches the stored energy 1/2 (lambda + 2G) strain^2 V."""
    h = 0.01
    H = nz * h
    x0, conn, ijk = decks.block_mesh(1, 1, nz, h, h, H)
    m = Model(x0=x0, conn=conn, **STEEL, name="driven_column")
    _, _, _, c = m.material_consts()
    T = transits * H / c
    speed = strain * H / T
    alln = np.arange(m.n_nodes, dtype=np.uint32)
    bot = decks.face_nodes(ijk, 2, 0, nz)
    top = decks.face_nodes(ijk, 2, 1, nz)
    m.v_init = np.zeros_like(x0)
    m.steps = [Step("Initial", 0.0, spc=[Bc("spc", alln, 0, 1), Bc("spc", bot, 2, 2)]),
               Step("Step-1", T, vel=[Bc("vel", top, 2, 2, value=-speed)])]
    m.total_time = T
    m.speed = speed  # type: ignore[attr-defined]
    m.height = H  # type: ignore[attr-defined]
    m.top = top  # type: ignore[attr-defined]
    return m


def ramped_block(steps: int = 150, value: float = -2.0) -> Model:
    """A 2x2x3 block, z = 0 fixed, top face driven in z by value * amp(t)
    with a piecewise-linear amplitude (0 -> 1 over the first 40% of the run,
    then down to 0.5): the velocity-BC amplitude extension."""
    x0, conn, ijk = decks.block_mesh(2, 2, 3, 0.02, 0.02, 0.03)
    m = Model(x0=x0, conn=conn, **STEEL, name="ramped_block")
    bot = decks.face_nodes(ijk, 2, 0, 3)
    top = decks.face_nodes(ijk, 2, 1, 3)
    m.v_init = np.zeros_like(x0)
    dt0 = decks.stable_dt0(m)
    T = steps * dt0
    m.amplitudes = [("Amp-1", np.array([[0.0, 0.0], [0.4 * T, 1.0], [T, 0.5], [10 * T, 0.5]]))]
    m.steps = [Step("Initial", 0.0, spc=[Bc("spc", bot, 0, 2)]),
               Step("Step-1", T, vel=[Bc("vel", top, 2, 2, value=value, amp=0)])]
    m.total_time = T
    m.top = top  # type: ignore[attr-defined]
    m.value = value  # type: ignore[attr-defined]
    return m


def fixed_after_release(steps: int = 120) -> Model:
    """Every node starts at 3 m/s along x; step 1 fixes the x = 0 face, so
    the constraint removes that face's kinetic energy at the step switch
    (external work) before the bar starts to ring."""
    x0, conn, ijk = decks.block_mesh(6, 2, 2, 0.06, 0.02, 0.02)
    m = Model(x0=x0, conn=conn, **STEEL, name="fixed_after_release")
    v = np.zeros_like(x0)
    v[:, 0] = 3.0
    m.v_init = v
    left = decks.face_nodes(ijk, 0, 0, 6)
    dt0 = decks.stable_dt0(m)
    T = steps * dt0
    m.steps = [Step("Initial", 0.0), Step("Step-1", T, spc=[Bc("spc", left, 0, 2)])]
    m.total_time = T
    m.left = left  # type: ignore[attr-defined]
    return m


def top_faces(nx: int, ny: int, nz: int) -> np.ndarray:
    """Elements of the top layer (k = nz - 1) of decks.block_mesh."""
    return np.arange(nx * ny * (nz - 1), nx * ny * nz, dtype=np.uint32)


def pressed_block(steps: int = 120, p: float = 1e6, ramp: bool = False) -> Model:
    """A free 2x3x2 block (no BCs) with pressure p on its top faces (face
    S2 = corners 5-8-7-6 of each top element, the *Dsload extension): the
    surface force is p A0 along -z while the faces stay flat, so the
    momentum is -p A0 sum(dt_avg) up to the faces' area change (strain
    p / E ~ 5e-6).  With ramp, the pressure follows an amplitude 0 -> 1
    over the run, so the impulse halves."""
    nx, ny, nz = 2, 3, 2
    x0, conn, _ = decks.block_mesh(nx, ny, nz, 0.02, 0.03, 0.02)
    m = Model(x0=x0, conn=conn, **STEEL, name="pressed_block")
    m.v_init = np.zeros_like(x0)
    dt0 = decks.stable_dt0(m)
    T = steps * dt0
    el = top_faces(nx, ny, nz)
    amp = -1
    if ramp:
        m.amplitudes = [("Ramp", np.array([[0.0, 0.0], [T, 1.0]]))]
        amp = 0
    m.steps = [Step("Initial", 0.0),
               Step("Step-1", T, dsload=[Dsload(el, np.full(len(el), 1, np.int32), p, amp)])]
    m.total_time = T
    m.area = 0.02 * 0.03  # type: ignore[attr-defined]
    m.p = p  # type: ignore[attr-defined]
    return m


def pressure_column(nz: int = 100, p: float = 1e6, transits: float = 0.5) -> Model:
    """Uniaxial-strain column of nz elements (h = 0.01 m, x and y fixed on
    every node, z fixed at z = 0) with a step pressure p on the top face:
    a compression front runs down at c = sqrt((lambda + 2G) / rho) with the
    particle velocity p / (rho c) behind it.  Runs `transits` transits of
    the column; the area stays h^2, so the momentum is exactly -p h^2 t
    while the front has not reached the bottom."""
    h = 0.01
    H = nz * h
    x0, conn, ijk = decks.block_mesh(1, 1, nz, h, h, H)
    m = Model(x0=x0, conn=conn, **STEEL, name="pressure_column")
    _, _, _, c = m.material_consts()
    T = transits * H / c
    alln = np.arange(m.n_nodes, dtype=np.uint32)
    bot = decks.face_nodes(ijk, 2, 0, nz)
    el = top_faces(1, 1, nz)
    m.v_init = np.zeros_like(x0)
    m.steps = [Step("Initial", 0.0, spc=[Bc("spc", alln, 0, 1), Bc("spc", bot, 2, 2)]),
               Step("Step-1", T, dsload=[Dsload(el, np.full(len(el), 1, np.int32), p)])]
    m.total_time = T
    m.height = H  # type: ignore[attr-defined]
    m.c = c  # type: ignore[attr-defined]
    m.p = p  # type: ignore[attr-defined]
    m.area = h * h  # type: ignore[attr-defined]
    return m


def balance(s) -> float:
    """KE + IE + HG + VIS - W_ext - KE0 of a scalar dict (oracle or engine)."""
    return s["ke"] + s["ie"] + s["hg"] + s["vis"] - s["wext"] - s["ke0"]
