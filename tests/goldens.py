"""Helpers to read the golden fixtures written by tests/golden/make_golden.py."""
from __future__ import annotations

import gzip
import json
import os
from typing import Dict

import numpy as np

from enscdynamics_amd.model import Bc, Model, Step

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

CASES_MICRO = ["micro_rotation", "micro_hourglass", "micro_multistep"]
CASES_REF = ["ref_Job_6", "ref_test_single2", "ref_test_step", "ref_test_trans", "ref_test_part1",
             "ref_GenVirPa"]


def load(name: str) -> Dict[str, np.ndarray]:
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def deck_path_text(name: str) -> str:
    with gzip.open(os.path.join(GOLDEN, "decks", name + ".inp.gz"), "rt") as f:
        return f.read()


def _resolve(name: str, sets: Dict[str, list]) -> np.ndarray:
    # exDyna3D.cpp:400-419: '#n' is the single node n (1-based), otherwise a
    # set looked up by name; an unknown set is skipped.
    if name.startswith("#"):
        try:
            return np.array([int(name[1:]) - 1], dtype=np.uint32)
        except ValueError:
            return np.zeros(0, dtype=np.uint32)
    return np.asarray(sets.get(name, []), dtype=np.uint32)


def model_from_dump(d: Dict[str, np.ndarray], c_dr=None, c_cr=None, factor=None) -> Model:
    """Rebuild the job the reference parsed, from its dump (model.json +
    x0/conn), applying the reference's own set resolution and initial
    conditions (exDyna3D.cpp:338-377)."""
    j = json.loads(str(d["model_json"]))
    sets = j["node_sets"]
    x0 = d["x0"].astype(np.float64)
    conn = d["conn"].astype(np.uint32)
    m = Model(x0=x0, conn=conn, E=j["E"], nu=j["nu"], rho=j["rho"], c_dr=j["c_dr"], c_cr=j["c_cr"],
              factor=j["factor"], cvisl=j["Cvisl"], cvisq=j["Cvisq"], total_time=j["totalTime"])
    v = np.zeros_like(x0)
    for name, dof, val in j["ini_vel"]:
        nodes = _resolve(name, sets)
        v[nodes, int(dof)] = val
    m.v_init = v
    amp_names = list(j["amplitudes"].keys())
    m.amplitudes = [(n, np.array(j["amplitudes"][n], dtype=np.float64).reshape(-1, 2)) for n in amp_names]
    steps = []
    for s in j["steps"]:
        st = Step(s["name"], s["period"], reset_spc=bool(s["reset_spc"]), reset_vel=bool(s["reset_vel"]),
                  reset_dload=bool(s["reset_dload"]))
        for name, a0, a1, cond in s["spc"]:
            st.spc.append(Bc("spc", _resolve(name, sets), int(a0), int(a1), int(cond), name=name))
        for name, a0, a1, val in s["vel"]:
            st.vel.append(Bc("vel", _resolve(name, sets), int(a0), int(a1), value=float(val), name=name))
        on, amp, g, dx, dy, dz = s["gravity"]
        if on:
            st.gravity = (g, dx, dy, dz, amp_names.index(amp) if amp in amp_names else -1)
        steps.append(st)
    m.steps = steps
    m.node_sets = {k: np.asarray(v_, dtype=np.uint32) for k, v_ in sets.items()}
    return m


def rel_err(a: np.ndarray, b: np.ndarray) -> float:
    """max |a-b| / max |b| (0 when both are all-zero)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    s = float(np.abs(b).max()) if b.size else 0.0
    d = float(np.abs(a - b).max()) if b.size else 0.0
    return d / s if s > 0 else d
