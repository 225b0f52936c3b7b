"""The FSI coupling schedule of the golden fixtures micro_fsi_block and
micro_fsi_slab (tests/golden/make_golden.py), and drivers that replay it on
the oracle and on the engine.

The fixtures come from the reference itself: a subclass of exDyna3D
generates the virtual particles with get_fsiSph_virtualParticles_and_vel
(exDyna3D.cpp:1363-1383) right before the first step and, before step n,
sets force scale * sin(0.37 (p + 1) + 1.1 d + 0.05 n) on particle p,
component d (fsi_share_data, exDyna3D.h:54).
"""
from __future__ import annotations

import numpy as np

import goldens as G

CASES = ("micro_fsi_block", "micro_fsi_slab")


def forces(n_particles: int, step: int, scale: float) -> np.ndarray:
    p = np.arange(n_particles, dtype=np.float64)[:, None]
    d = np.arange(3, dtype=np.float64)[None, :]
    return scale * np.sin(0.37 * (p + 1) + 1.1 * d + 0.05 * step)


def load(case: str):
    d = G.load(case)
    layers, dist, scale = (float(v) for v in d["fsi"])
    return d, G.model_from_dump(d), int(layers), dist, scale


def run_oracle(o, layers: int, dist: float, scale: float):
    """Setup after entering step 1, then the schedule until the step ends;
    returns the particles at setup and the step count."""
    o.begin_step()
    n = o.fsi_setup(layers, dist)
    x0, v0 = o.fsi_particles()
    step = 0
    while True:
        o.fsi_set_forces(forces(n, step, scale))
        taken, _ = o.run(1, [], 0)
        if taken == 0:
            return x0, v0, step
        step += 1


def run_engine(e, layers: int, dist: float, scale: float):
    e.init()
    e.set_current_step(1)
    n = e.fsi_setup(layers, dist)
    x0, v0 = e.fsi_particles()
    step = 0
    done = False
    while not done:
        e.fsi_set_forces(forces(n, step, scale))
        done = e.advance(1)
        step += 1
    return x0, v0, step
