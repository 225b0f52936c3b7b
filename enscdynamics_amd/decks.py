"""Synthetic C3D8R decks in the reference's Abaqus dialect.

Builds structured hexahedral blocks and the benchmark configurations, and
writes them as ``.inp`` text the reference's reader accepts
(Explicit3D/source/dataIn.cpp:300-558; node order of
Element_HexN8.cpp:50-89; sets named ``Set-N`` because BC lines refer to a
set only when they contain ``SET-``, dataIn.cpp:928,1144).

Extensions the reference reader ignores (documented where used):
``*Plastic`` (J2 with linear hardening) and ``amplitude=`` on
``*Boundary, type=VELOCITY``.
"""
from __future__ import annotations

import math
from typing import List, Optional, Sequence

import numpy as np

from .model import Bc, Model, Step

# Corner offsets (i, j, k) of the C3D8R node order, Element_HexN8.cpp:50-89.
CORNERS = np.array(
    [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]],
    dtype=np.int64,
)


def block_mesh(nx: int, ny: int, nz: int, lx: float, ly: float, lz: float,
               origin=(0.0, 0.0, 0.0), jitter: float = 0.0, seed: Optional[int] = None):
    """Structured block of nx*ny*nz hexahedra.

    Node (i, j, k) has id i + (nx+1)*(j + (ny+1)*k).  With ``jitter`` > 0
    every node moves by U(-jitter*h, +jitter*h) per component, except that a
    node on a boundary face keeps its coordinate normal to that face.
    Returns (x0 (Nn,3) float64, conn (Ne,8) uint32, ijk (Nn,3) int).
    """
    hx, hy, hz = lx / nx, ly / ny, lz / nz
    i = np.arange(nx + 1)
    j = np.arange(ny + 1)
    k = np.arange(nz + 1)
    K, J, I = np.meshgrid(k, j, i, indexing="ij")
    ijk = np.stack([I.ravel(), J.ravel(), K.ravel()], axis=1)
    x0 = np.empty(ijk.shape, dtype=np.float64)
    x0[:, 0] = origin[0] + ijk[:, 0] * hx
    x0[:, 1] = origin[1] + ijk[:, 1] * hy
    x0[:, 2] = origin[2] + ijk[:, 2] * hz
    if jitter > 0.0:
        rng = np.random.default_rng(seed)
        h = np.array([hx, hy, hz])
        d = rng.uniform(-jitter, jitter, size=x0.shape) * h
        nmax = np.array([nx, ny, nz])
        on_face = (ijk == 0) | (ijk == nmax)
        d[on_face] = 0.0
        x0 = x0 + d
    ei, ej, ek = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    # element id = i + nx*(j + ny*k)
    e_ijk = np.stack([ei.transpose(2, 1, 0).ravel(), ej.transpose(2, 1, 0).ravel(),
                      ek.transpose(2, 1, 0).ravel()], axis=1)
    conn = np.empty((e_ijk.shape[0], 8), dtype=np.uint32)
    for c in range(8):
        p = e_ijk + CORNERS[c]
        conn[:, c] = p[:, 0] + (nx + 1) * (p[:, 1] + (ny + 1) * p[:, 2])
    return x0, conn, ijk


def face_nodes(ijk: np.ndarray, axis: int, side: int, n: int) -> np.ndarray:
    """Node ids on the face axis==0 (side 0) or axis==n (side 1)."""
    target = 0 if side == 0 else n
    return np.nonzero(ijk[:, axis] == target)[0].astype(np.uint32)


def stable_dt0(model: Model) -> float:
    """Initial stable time step of an undeformed mesh at rest:
    factor * min(Lmin/c, (sqrt(1+Cvisl^2)-Cvisl) * L/c) (exDyna3D.cpp:828-834,
    663-679 with VolRate = 0)."""
    _, _, _, c = model.material_consts()
    x = model.x0
    cn = model.conn.astype(np.int64)
    edges = [(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4), (0, 4), (1, 5), (2, 6), (3, 7)]
    lmin2 = np.full(cn.shape[0], np.inf)
    for a, b in edges:
        d = x[cn[:, b]] - x[cn[:, a]]
        lmin2 = np.minimum(lmin2, (d * d).sum(axis=1))
    L = np.sqrt(lmin2)
    xi = model.cvisl
    dts = (math.sqrt(1 + xi * xi) - xi) * L / c
    return model.factor * min(float(L.min()) / c, float(dts.min()))


# --------------------------------------------------------------------------
# benchmark configurations
# --------------------------------------------------------------------------

def config1() -> Model:
    """Cantilever bar, 40x8x8 C3D8R over 0.4x0.08x0.08 m, steel, x=0 face
    ENCASTRE, vz = 5 m/s on every other node, c_cr = 0.1, factor 0.9,
    2000 steps of the initial stable dt."""
    x0, conn, ijk = block_mesh(40, 8, 8, 0.4, 0.08, 0.08)
    fixed = face_nodes(ijk, 0, 0, 40)
    rest = np.setdiff1d(np.arange(x0.shape[0], dtype=np.uint32), fixed)
    m = Model(x0=x0, conn=conn, E=210e9, nu=0.3, rho=7850.0, c_dr=0.013, c_cr=0.1, factor=0.9,
              name="c1_cantilever_40x8x8")
    m.node_sets = {"Set-1": fixed, "Set-2": rest}
    v = np.zeros_like(x0)
    v[rest, 2] = 5.0
    m.v_init = v
    m.ic = [("Set-2", 2, 5.0)]
    dt0 = stable_dt0(m)
    T = 2000 * dt0
    m.steps = [
        Step("Initial", 0.0, spc=[Bc("spc", fixed, 0, 2, 1, name="Set-1")]),
        Step("Step-1", T),
    ]
    m.total_time = T
    m.tip_node = int(np.nonzero((ijk[:, 0] == 40) & (ijk[:, 1] == 4) & (ijk[:, 2] == 4))[0][0])  # type: ignore[attr-defined]
    return m


def config2(nx=200, ny=40, nz=40, steps=5000, seed=2) -> Model:
    """Jittered steel block (h = 0.01 m, nodes moved U(+-0.1h)), z=0 face
    ENCASTRE, top face vz = -2 m/s ramped linearly over the first 10% of the
    run (amplitude on the velocity BC)."""
    h = 0.01
    x0, conn, ijk = block_mesh(nx, ny, nz, nx * h, ny * h, nz * h, jitter=0.1, seed=seed)
    bot = face_nodes(ijk, 2, 0, nz)
    top = face_nodes(ijk, 2, 1, nz)
    m = Model(x0=x0, conn=conn, E=210e9, nu=0.3, rho=7850.0, name=f"c2_block_{nx}x{ny}x{nz}_jitter")
    m.node_sets = {"Set-1": bot, "Set-2": top}
    m.v_init = np.zeros_like(x0)
    dt0 = stable_dt0(m)
    T = steps * dt0
    t_ramp = 0.1 * T
    m.amplitudes = [("Amp-1", np.array([[0.0, 0.0], [t_ramp, 1.0], [10.0 * T, 1.0]]))]
    m.steps = [
        Step("Initial", 0.0, spc=[Bc("spc", bot, 0, 2, 1, name="Set-1")]),
        Step("Step-1", T, vel=[Bc("vel", top, 2, 2, value=-2.0, amp=0, name="Set-2")]),
    ]
    m.total_time = T
    return m


def config3(n=100, steps=3000) -> Model:
    """Unit cube of n^3 C3D8R, E=200 GPa, nu=0.3, rho=7800, J2 yield
    250 MPa with linear hardening 1 GPa; z=0 face ENCASTRE, z=1 face pushed
    at vz = -10 m/s.  Reference defaults for the unstated parameters."""
    x0, conn, ijk = block_mesh(n, n, n, 1.0, 1.0, 1.0)
    bot = face_nodes(ijk, 2, 0, n)
    top = face_nodes(ijk, 2, 1, n)
    m = Model(x0=x0, conn=conn, E=200e9, nu=0.3, rho=7800.0, plastic=True, yield0=250e6, hard=1e9,
              name=f"c3_cube_{n}x{n}x{n}_j2")
    m.node_sets = {"Set-1": bot, "Set-2": top}
    m.v_init = np.zeros_like(x0)
    dt0 = stable_dt0(m)
    T = steps * dt0
    m.steps = [
        Step("Initial", 0.0, spc=[Bc("spc", bot, 0, 2, 1, name="Set-1")]),
        Step("Step-1", T, vel=[Bc("vel", top, 2, 2, value=-10.0, name="Set-2")]),
    ]
    m.total_time = T
    return m


def random_job(rng: np.random.Generator, ne_lo: float = 1e3, ne_hi: float = 5e4,
               plastic_p: float = 0.0, h: float = 0.01, k_lo: int = 2000, k_hi: int = 8000,
               name: str = "job") -> Model:
    """One job of the batched generator: Ne log-uniform in [ne_lo, ne_hi],
    aspect nx:ny:nz each uniform on 1..8, E log-uniform [1, 300] GPa, nu
    uniform [0.2, 0.45], rho uniform [1000, 9000], a random face ENCASTRE,
    initial speed U[1, 10] m/s in a random direction on the other nodes,
    end time K * dt0 with K uniform on k_lo..k_hi."""
    ne = math.exp(rng.uniform(math.log(ne_lo), math.log(ne_hi)))
    asp = rng.integers(1, 9, size=3).astype(np.float64)
    s = (ne / asp.prod()) ** (1.0 / 3.0)
    nx, ny, nz = (max(1, int(round(a * s))) for a in asp)
    E = math.exp(rng.uniform(math.log(1e9), math.log(300e9)))
    nu = rng.uniform(0.2, 0.45)
    rho = rng.uniform(1000.0, 9000.0)
    x0, conn, ijk = block_mesh(nx, ny, nz, nx * h, ny * h, nz * h)
    axis = int(rng.integers(0, 3))
    side = int(rng.integers(0, 2))
    fixed = face_nodes(ijk, axis, side, (nx, ny, nz)[axis])
    rest = np.setdiff1d(np.arange(x0.shape[0], dtype=np.uint32), fixed)
    speed = rng.uniform(1.0, 10.0)
    d = rng.normal(size=3)
    d /= np.linalg.norm(d)
    plastic = bool(rng.uniform() < plastic_p)
    m = Model(x0=x0, conn=conn, E=E, nu=nu, rho=rho, name=f"{name}_{nx}x{ny}x{nz}")
    if plastic:
        m.plastic = True
        m.yield0 = 1.25e-3 * E
        m.hard = 5e-3 * E
    m.node_sets = {"Set-1": fixed, "Set-2": rest}
    v = np.zeros_like(x0)
    v[rest] = speed * d
    m.v_init = v
    m.ic = [("Set-2", q, float(speed * d[q])) for q in range(3)]
    k = int(rng.integers(k_lo, k_hi + 1))
    dt0 = stable_dt0(m)
    T = k * dt0
    m.steps = [Step("Initial", 0.0, spc=[Bc("spc", fixed, 0, 2, 1, name="Set-1")]), Step("Step-1", T)]
    m.total_time = T
    m.target_steps = k  # type: ignore[attr-defined]
    return m


def config4(n_jobs: int = 256, seed: int = 4, k_lo: int = 2000, k_hi: int = 8000,
            ne_lo: float = 1e3, ne_hi: float = 5e4) -> List[Model]:
    rng = np.random.default_rng(seed)
    return [random_job(rng, ne_lo, ne_hi, 0.0, k_lo=k_lo, k_hi=k_hi, name=f"c4j{i}") for i in range(n_jobs)]


def config5(n_jobs: int = 1000, seed: int = 5, k_lo: int = 2000, k_hi: int = 8000,
            ne_lo: float = 1e3, ne_hi: float = 5e5) -> List[Model]:
    """Serving stream requests: the config-4 generator with sizes up to
    5e5 elements, elastic or J2 plastic with probability 1/2 each."""
    rng = np.random.default_rng(seed)
    return [random_job(rng, ne_lo, ne_hi, 0.5, k_lo=k_lo, k_hi=k_hi, name=f"c5j{i}") for i in range(n_jobs)]


# --------------------------------------------------------------------------
# .inp writer
# --------------------------------------------------------------------------

def _id_lines(ids: np.ndarray, per: int = 16) -> List[str]:
    ids = np.asarray(ids, dtype=np.int64) + 1
    out = []
    for s in range(0, len(ids), per):
        out.append(", ".join(str(int(v)) for v in ids[s:s + per]))
    return out


def _set_block(kw: str, name: str, ids: np.ndarray, instance: Optional[str]) -> List[str]:
    ids = np.asarray(ids, dtype=np.int64)
    inst = f", instance={instance}" if instance else ""
    if len(ids) >= 2:
        d = np.diff(ids)
        if np.all(d == d[0]) and d[0] > 0:
            return [f"*{kw}, {kw.lower()}={name}{inst}, generate",
                    f" {ids[0] + 1}, {ids[-1] + 1}, {int(d[0])}"]
    return [f"*{kw}, {kw.lower()}={name}{inst}"] + _id_lines(ids)


def _g(v: float) -> str:
    return repr(float(v))


def write_inp(m: Model, path: str) -> None:
    """Write ``m`` as a single-part, single-instance deck."""
    L: List[str] = ["*Heading", f"** {m.name}", "*Part, name=Part-1", "*Node"]
    for i, p in enumerate(m.x0):
        L.append(f"{i + 1}, {_g(p[0])}, {_g(p[1])}, {_g(p[2])}")
    L.append("*Element, type=C3D8R")
    for e, c in enumerate(m.conn):
        L.append(f"{e + 1}, " + ", ".join(str(int(v) + 1) for v in c))
    L += _set_block("Nset", "Set-1", np.arange(m.n_nodes), None)
    L += _set_block("Elset", "Set-1", np.arange(m.n_elems), None)
    L += ["*Solid Section, elset=Set-1, material=Material-1", ",", "*End Part",
          "*Assembly, name=Assembly", "*Instance, name=Part-1-1, part=Part-1", "*End Instance"]
    for name, ids in m.node_sets.items():
        L += _set_block("Nset", name, ids, "Part-1-1")
    L.append("*End Assembly")
    for name, tab in m.amplitudes:
        tab = np.asarray(tab).reshape(-1, 2)
        L.append(f"*Amplitude, name={name}")
        L.append(", ".join(f"{_g(t)}, {_g(a)}" for t, a in tab))
    L += ["*Material, name=Material-1", "*Density", f"{_g(m.rho)},", "*Elastic", f"{_g(m.E)}, {_g(m.nu)}"]
    if m.plastic:
        L += ["*Plastic", f"{_g(m.yield0)}, 0.", f"{_g(m.yield0 + m.hard)}, 1."]
    amp_names = [n for n, _ in m.amplitudes]

    def bc_lines(bcs: Sequence[Bc], vel: bool) -> List[str]:
        out = []
        for b in bcs:
            if vel:
                kw = "*Boundary, type=VELOCITY"
                if b.amp >= 0:
                    kw += f", amplitude={amp_names[b.amp]}"
                out += [kw, f"{b.name}, {b.d0 + 1}, {b.d1 + 1}, {_g(b.value)}"]
            else:
                if b.d0 == 0 and b.d1 == 2 and b.cond == 1:
                    out += ["*Boundary", f"{b.name}, ENCASTRE"]
                else:
                    out += ["*Boundary", f"{b.name}, {b.d0 + 1}, {b.d1 + 1}, {b.cond}"]
        return out

    init = m.steps[0]
    L += bc_lines(init.spc, False) + bc_lines(init.vel, True)
    for name, dof, val in m.ic:
        L += ["*Initial Conditions, type=VELOCITY", f"{name}, {dof + 1}, {_g(val)}"]
    for s in m.steps[1:]:
        L += [f"*Step, name={s.name}, nlgeom=YES", "*Dynamic, Explicit", f", {_g(s.period)}",
              "*Bulk Viscosity", f"{_g(m.cvisl)}, {_g(m.cvisq)}"]
        if s.reset_spc:
            L += ["*Boundary, op=NEW"]
        if s.reset_vel:
            L += ["*Boundary, op=NEW, type=VELOCITY"]
        L += bc_lines(s.spc, False) + bc_lines(s.vel, True)
        if s.gravity is not None:
            g, dx, dy, dz, amp = s.gravity
            kw = "*Dload" + (", op=NEW" if s.reset_dload else "")
            if amp >= 0:
                kw += f", amplitude={amp_names[amp]}"
            L += [kw, f", GRAV, {_g(g)}, {_g(dx)}, {_g(dy)}, {_g(dz)}"]
        L.append("*End Step")
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")
