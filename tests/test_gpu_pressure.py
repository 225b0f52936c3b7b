"""Surface pressures (*Dsload P lines; an extension: the reference parses and
inherits them, exDyna3D.cpp:1897-1907, but applies no force) on the GPU
against the oracle, whose pressure forces and work are pinned by the closed
forms in test_oracle.py (impulse, face numbering, wave speed and particle
velocity, inheritance across steps).

FP64 within 1e-10 of the oracle (state) and 1e-9 of the energy scale;
mixed within 1e-4 (displacement) and 1% (external work).  The cases cover
the face kernel alone in a job, with an amplitude, across step switches
(the face arrays and the face block map change), in graphs, in a batch
next to a job without pressures, through the allocation-free entry point,
and the one reference deck that carries a *Dsload (GenVirPa).
"""
import numpy as np
import pytest

import energy_cases as EC
import goldens as G
from enscdynamics_amd import batch
from enscdynamics_amd import deck
from enscdynamics_amd import engine as eng
from enscdynamics_amd.deck import parse_deck
from enscdynamics_amd.model import Step
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-10
TOL_MIXED = 1e-4


def rel(a, b, floor=0.0):
    s = max(float(np.abs(b).max()), floor)
    return float(np.abs(np.asarray(a) - np.asarray(b)).max()) / s if s > 0 else float(np.abs(a - b).max())


def three_steps():
    """Pressed in step 1, the lines inherited by step 2, dropped (op=NEW)
    in step 3."""
    m = EC.pressed_block(steps=60)
    T = m.steps[1].period
    m.steps.append(Step("Step-2", T))
    m.steps.append(Step("Step-3", T, reset_dsload=True))
    m.total_time = 3 * T
    m.name = "pressed_three_steps"
    return m


def genvirpa():
    m = parse_deck(G.deck_path_text("GenVirPa"), extensions=deck.DSLOAD)
    m.name = "GenVirPa+dsload"
    return m


CASES = {"pressed_block": lambda: EC.pressed_block(), "pressed_block_ramp": lambda: EC.pressed_block(ramp=True),
         "pressure_column": lambda: EC.pressure_column(), "three_steps": three_steps, "GenVirPa": genvirpa}


@pytest.fixture(scope="module", autouse=True)
def gpu_present():
    assert eng.device_count() > 0, "GPU tests need a HIP device"


def oracle_of(m):
    o = Oracle(m, 0).init()
    o.run(0, [], 0)
    return o


@pytest.mark.parametrize("case", sorted(CASES))
def test_fp64_pressure_matches_oracle(case):
    m = CASES[case]()
    print(f"[pressure] {case}: fp64 engine vs oracle", flush=True)
    e = eng.Engine(m, eng.FP64, reorder=True).run()
    o = oracle_of(m)
    so, se = o.scalars(), e.scalars()
    assert se["nsteps"] == so["nsteps"]
    # velocities against max(|v|, the particle velocity p / (rho c) of the
    # largest pressure): once a loaded body has come to rest (GenVirPa, after
    # 22k steps) its velocity field is rounding noise (~1e-14 mm/s against
    # p / (rho c) = 21 mm/s), which an engine that groups its sums otherwise
    # than the oracle does not reproduce digit for digit
    c = m.material_consts()[3]
    pmax = max([abs(d.value) for st in m.steps for d in (st.dsload or [])] + [0.0])
    for f, fl in (("u", 1e-12), ("v", pmax / (m.rho * c)), ("sigma", 1e-12)):
        assert rel(e.get(f), o.get(f), fl) <= TOL_FP64, f
    assert so["wext"] > 0.0
    scale = max(abs(so["wext"]), so["ke0"])
    for k in ("ke", "ie", "hg", "vis", "wext"):
        assert abs(se[k] - so[k]) <= 1e-9 * scale, (k, se[k], so[k])
    print(f"[pressure] {case}: W_ext {se['wext']:.6e} (oracle {so['wext']:.6e})", flush=True)


@pytest.mark.parametrize("case", ["pressed_block", "pressure_column", "three_steps", "GenVirPa"])
def test_mixed_pressure_close_to_oracle(case):
    m = CASES[case]()
    print(f"[pressure mixed] {case}", flush=True)
    e = eng.Engine(m, eng.MIXED, reorder=True).run()
    o = oracle_of(m)
    so, se = o.scalars(), e.scalars()
    err = rel(e.get("u"), o.get("u"))
    werr = abs(se["wext"] - so["wext"]) / abs(so["wext"])
    print(f"[pressure mixed] {case}: u rel err {err:.2e}, W_ext rel err {werr:.2e}", flush=True)
    assert err <= TOL_MIXED
    assert werr <= 1e-2


def test_pressure_graph_replay_is_bit_exact():
    # the face arrays change at the step switches: graphs replayed across
    # them equal plain launches
    m = three_steps()
    print("[pressure graph] three_steps, graph of 8 iterations vs plain launches", flush=True)
    a = eng.Engine(m, eng.FP64, reorder=True).run()
    b = eng.Engine(m, eng.FP64, reorder=True)
    b.set_graph(8)
    b.run()
    for f in ("u", "v", "sigma"):
        assert np.array_equal(a.get(f), b.get(f)), f
    assert a.scalars()["wext"] == b.scalars()["wext"]


def test_pressure_in_batch_next_to_unloaded_jobs():
    # a batch of a pressed job, an unloaded one and the three-step job: each
    # equals its own run (the face block map covers only the loaded jobs)
    ms = [EC.pressed_block(), EC.free_fall(steps=150), three_steps()]
    print("[pressure batch] 3 jobs", flush=True)
    bt = batch.Batch(ms, eng.FP64, reorder=True).run()
    for j, m in enumerate(ms):
        solo = eng.Engine(m, eng.FP64, reorder=True).run()
        for f in ("u", "v", "sigma"):
            assert np.array_equal(bt.get(j, f), solo.get(f)), (m.name, f)
        assert bt.scalars(j)["wext"] == solo.scalars()["wext"]


def test_pressure_through_the_allocation_free_entry_point():
    m = three_steps()
    print("[pressure static] three_steps through ensc_engine_advance_static", flush=True)
    a = eng.Engine(m, eng.FP64, reorder=True).run()
    b = eng.Engine(m, eng.FP64, reorder=True).init()
    for s in range(1, len(m.steps)):
        b.set_current_step(s)
        while not b.advance_static(50):
            pass
    for f in ("u", "v", "sigma"):
        assert np.array_equal(a.get(f), b.get(f)), f
