"""Write the golden fixtures in tests/golden/ from the real reference.

Runs oracle/_ref/ref_harness (the reference's own solver sources compiled by
oracle/Makefile, target ``ref``) on each deck and packs what it records into
``<case>.npz``: the parsed model (model.json text), setup arrays, a per-step
history and end-of-run fields.  Needs /root/reference; run it here, never on
a machine without the reference:

    make -C oracle ref && python tests/golden/make_golden.py [--only NAME ...]

Decks: synthetic decks written by enscdynamics_amd.decks (configuration 1
and micro decks built to exercise single terms), plus the reference's own
input decks (inp/*.inp), kept gzip'd under tests/golden/decks/ as input data.
"""
from __future__ import annotations

import gzip
import json
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from enscdynamics_amd import decks  # noqa: E402
from enscdynamics_amd.model import Bc, Model, Step  # noqa: E402

REF = "/root/reference"
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
DECKDIR = os.path.join(HERE, "decks")


def run_ref(deck_path: str, max_steps=0, snaps=(), probes=(), params=None, fsi=None):
    work = tempfile.mkdtemp(prefix="gold_")
    shutil.copy(deck_path, os.path.join(work, "deck.inp"))
    args = [HARNESS, work, "deck.inp", work, str(max_steps), ",".join(map(str, snaps)) or "0",
            ",".join(map(str, probes)) or ""]
    if params is not None:
        args += [repr(params[0]), repr(params[1]), repr(params[2])]
    env = dict(os.environ, OMP_NUM_THREADS="8")
    if fsi is not None:
        env["ENSC_FSI"] = ",".join(repr(v) for v in fsi)
    r = subprocess.run(args, capture_output=True, text=True, env=env)
    if r.returncode != 0:
        raise RuntimeError(r.stderr[-2000:])
    out = {}
    for f in os.listdir(work):
        if f.endswith(".npy"):
            out[f[:-4]] = np.load(os.path.join(work, f))
    with open(os.path.join(work, "model.json")) as fh:
        out["model_json"] = np.array(fh.read())
    shutil.rmtree(work)
    return out


def save(name: str, d: dict, keep) -> None:
    sel = {k: v for k, v in d.items() if any(k == p or k.startswith(p) for p in keep)}
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **sel)
    print(f"{name}: {os.path.getsize(path) / 1024:.0f} KiB, {len(sel)} arrays")


def micro_rotation() -> Model:
    """Single unit-ish element spinning about z while stretching along x:
    drives every Jaumann rotation term with a nonzero stress."""
    x0, conn, _ = decks.block_mesh(1, 1, 1, 0.02, 0.015, 0.012)
    w = 400.0  # rad/s
    c = x0.mean(axis=0)
    r = x0 - c
    v = np.zeros_like(x0)
    v[:, 0] = -w * r[:, 1] + 30.0 * r[:, 0] / 0.02 + 2.0 * r[:, 2] / 0.012
    v[:, 1] = w * r[:, 0] + 3.0 * r[:, 2] / 0.012
    v[:, 2] = 0.5 * w * r[:, 1] - 1.5 * r[:, 0] / 0.02
    m = Model(x0=x0, conn=conn, E=210e9, nu=0.3, rho=7850.0, name="micro_rotation")
    m.node_sets = {f"Set-{i + 1}": np.array([i], dtype=np.uint32) for i in range(8)}
    m.v_init = v
    m.ic = [(f"Set-{i + 1}", d, float(v[i, d])) for i in range(8) for d in range(3)]
    T = 400 * decks.stable_dt0(m)
    m.steps = [Step("Initial", 0.0), Step("Step-1", T)]
    m.total_time = T
    return m


def micro_hourglass() -> Model:
    """Two elements in x, the far end fixed, the free nodes given an
    hourglass velocity pattern plus a shear drift: exercises the hourglass
    stiffness (QIA) and viscosity terms and the reaction forces."""
    x0, conn, ijk = decks.block_mesh(2, 1, 1, 0.02, 0.01, 0.01)
    fixed = decks.face_nodes(ijk, 0, 0, 2)
    free = np.setdiff1d(np.arange(x0.shape[0], dtype=np.uint32), fixed)
    g = (-1.0) ** (ijk[:, 0] + ijk[:, 1] + ijk[:, 2])
    v = np.zeros_like(x0)
    v[free, 0] = 0.7 * g[free]
    v[free, 1] = 0.4 * g[free] * (-1.0) ** ijk[free, 2]
    v[free, 2] = 0.25 + 0.3 * g[free]
    m = Model(x0=x0, conn=conn, E=70e9, nu=0.33, rho=2700.0, c_cr=0.1, name="micro_hourglass")
    m.node_sets = {"Set-1": fixed}
    for i, n in enumerate(free):
        m.node_sets[f"Set-{i + 2}"] = np.array([n], dtype=np.uint32)
    m.v_init = v
    m.ic = [(f"Set-{i + 2}", d, float(v[n, d])) for i, n in enumerate(free) for d in range(3)]
    T = 600 * decks.stable_dt0(m)
    m.steps = [Step("Initial", 0.0, spc=[Bc("spc", fixed, 0, 2, 1, name="Set-1")]), Step("Step-1", T)]
    m.total_time = T
    return m


def micro_multistep() -> Model:
    """3x2x2 block over three steps: XSYMM on a face in the Initial step, a
    velocity BC added in step 1, op=NEW resets and gravity with an amplitude
    in step 2, an SPC line whose third field is 0 (ignored), and steps that
    end part-way through a stable increment."""
    x0, conn, ijk = decks.block_mesh(3, 2, 2, 0.03, 0.02, 0.02)
    f_x0 = decks.face_nodes(ijk, 0, 0, 3)
    f_x1 = decks.face_nodes(ijk, 0, 1, 3)
    f_z0 = decks.face_nodes(ijk, 2, 0, 2)
    m = Model(x0=x0, conn=conn, E=200e9, nu=0.28, rho=7800.0, name="micro_multistep")
    m.node_sets = {"Set-1": f_x0, "Set-2": f_x1, "Set-3": f_z0, "Set-4": np.arange(x0.shape[0], dtype=np.uint32)}
    v = np.zeros_like(x0)
    v[:, 1] = 0.8
    m.v_init = v
    m.ic = [("Set-4", 1, 0.8)]
    dt0 = decks.stable_dt0(m)
    m.amplitudes = [("Amp-1", np.array([[0.0, 0.0], [60 * dt0, 1.0], [1e3, 1.0]]))]
    m.steps = [
        Step("Initial", 0.0, spc=[Bc("spc", f_x0, 0, 0, 1, name="Set-1")]),
        Step("Step-1", 80.3 * dt0, vel=[Bc("vel", f_x1, 0, 0, value=0.6, name="Set-2")],
             spc=[Bc("spc", f_z0, 2, 2, 0, name="Set-3")]),
        Step("Step-2", 120.7 * dt0, reset_vel=True,
             vel=[Bc("vel", f_x1, 1, 2, value=-0.3, name="Set-2")],
             gravity=(9000.0, 0.0, 0.0, -1.0, 0)),
        Step("Step-3", 50.2 * dt0, reset_spc=True, spc=[Bc("spc", f_z0, 0, 2, 1, name="Set-3")]),
    ]
    m.total_time = sum(s.period for s in m.steps)
    return m


# FSI coupling cases: (deck builder, layers, dist as a fraction of the
# element size, force scale in N).  The harness generates the particles
# with the reference's own code before the first step and sets force
# scale * sin(0.37 (p + 1) + 1.1 d + 0.05 n) on particle p before step n.
def micro_fsi_block() -> Model:
    """3x2x2 block, x = 0 face fixed: every FSI face is an outer face."""
    x0, conn, ijk = decks.block_mesh(3, 2, 2, 0.03, 0.02, 0.02)
    fixed = decks.face_nodes(ijk, 0, 0, 3)
    m = Model(x0=x0, conn=conn, E=200e9, nu=0.3, rho=7800.0, name="micro_fsi_block")
    m.node_sets = {"Set-1": fixed}
    m.v_init = np.zeros_like(x0)
    T = 150 * decks.stable_dt0(m)
    m.steps = [Step("Initial", 0.0, spc=[Bc("spc", fixed, 0, 2, 1, name="Set-1")]), Step("Step-1", T)]
    m.total_time = T
    m.fsi = (2, 0.25 * 0.01, 50.0)  # type: ignore[attr-defined]
    return m


def micro_fsi_slab() -> Model:
    """3x3x1 jittered slab moving with a linear velocity field, one edge
    fixed: every node is a boundary node, so every face (inner ones too)
    is an FSI face; distorted elements make J n differ from J^T n."""
    x0, conn, ijk = decks.block_mesh(3, 3, 1, 0.03, 0.03, 0.01, jitter=0.15, seed=3)
    edge = np.nonzero((ijk[:, 0] == 0) & (ijk[:, 1] == 0))[0].astype(np.uint32)
    m = Model(x0=x0, conn=conn, E=70e9, nu=0.33, rho=2700.0, name="micro_fsi_slab")
    free = np.setdiff1d(np.arange(x0.shape[0], dtype=np.uint32), edge)
    m.node_sets = {"Set-1": edge}
    v = np.zeros_like(x0)
    v[free, 2] = 2.0 * x0[free, 0] / 0.03 - 0.5 * x0[free, 1] / 0.03
    v[free, 0] = 0.3
    for i, n in enumerate(free):
        m.node_sets[f"Set-{i + 2}"] = np.array([n], dtype=np.uint32)
    m.v_init = v
    m.ic = [(f"Set-{i + 2}", d, float(v[n, d])) for i, n in enumerate(free) for d in range(3) if v[n, d] != 0.0]
    T = 120 * decks.stable_dt0(m)
    m.steps = [Step("Initial", 0.0, spc=[Bc("spc", edge, 0, 2, 1, name="Set-1")]), Step("Step-1", T)]
    m.total_time = T
    m.fsi = (3, 0.2 * 0.01, 20.0)  # type: ignore[attr-defined]
    return m


def main(only=None) -> None:
    if not os.path.exists(HARNESS):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    os.makedirs(DECKDIR, exist_ok=True)
    tmp = tempfile.mkdtemp(prefix="decks_")
    # ---- FSI coupling (virtual particles) ----
    for mk in (micro_fsi_block, micro_fsi_slab):
        mm = mk()
        if only and mm.name not in only:
            continue
        p = os.path.join(tmp, mm.name + ".inp")
        decks.write_inp(mm, p)
        with open(p, "rb") as fh, gzip.open(os.path.join(DECKDIR, mm.name + ".inp.gz"), "wb") as gz:
            gz.write(fh.read())
        d = run_ref(p, 0, (), (), (mm.c_dr, mm.c_cr, mm.factor), fsi=mm.fsi)
        d["fsi"] = np.array(mm.fsi)
        save(mm.name, d, ["model_json", "x0", "conn", "history", "fsi", "final_", "setup_"])
    if only:
        shutil.rmtree(tmp)
        return
    # ---- configuration 1 ----
    m = decks.config1()
    p = os.path.join(tmp, "c1.inp")
    decks.write_inp(m, p)
    d = run_ref(p, 0, (1, 2, 100), (m.tip_node,), (m.c_dr, m.c_cr, m.factor))
    d["tip_node"] = np.array(m.tip_node)
    save("c1", d, ["model_json", "x0", "conn", "V0", "invM", "dtStable0", "history", "tip_node",
                   "final_", "step1_", "step2_"])
    # ---- micro decks (full state) ----
    for mk in (micro_rotation, micro_hourglass, micro_multistep):
        mm = mk()
        p = os.path.join(tmp, mm.name + ".inp")
        decks.write_inp(mm, p)
        with open(p, "rb") as fh, gzip.open(os.path.join(DECKDIR, mm.name + ".inp.gz"), "wb") as gz:
            gz.write(fh.read())
        d = run_ref(p, 0, (1, 2, 3, 50), tuple(range(min(mm.n_nodes, 4))), (mm.c_dr, mm.c_cr, mm.factor))
        save(mm.name, d, ["model_json", "x0", "conn", "B0", "V0", "invM", "dtStable0", "Length0", "v_init",
                          "history", "final_", "step1_", "step2_", "step3_", "step50_"])
    # ---- the reference's own input decks ----
    for name in ("Job-6", "test_single2", "test_step", "test_trans", "test_part1", "GenVirPa"):
        src = os.path.join(REF, "inp", name + ".inp")
        with open(src, "rb") as fh, gzip.open(os.path.join(DECKDIR, name + ".inp.gz"), "wb") as gz:
            gz.write(fh.read())
        d = run_ref(src, 0, (), (0,))
        hist = d["history"]
        stride = max(1, hist.shape[0] // 400)
        d["history_stride"] = np.array(stride)
        d["history_sub"] = hist[::stride]
        d["history_last"] = hist[-5:]
        d["n_steps"] = np.array(hist.shape[0])
        save("ref_" + name.replace("-", "_"), d, ["model_json", "x0", "conn", "history_sub", "history_stride",
                                                   "history_last", "n_steps", "final_u", "final_v", "final_x",
                                                   "final_sigma", "final_rf", "final_scalars"])
    shutil.rmtree(tmp)


if __name__ == "__main__":
    main(sys.argv[sys.argv.index("--only") + 1:] if "--only" in sys.argv else None)
