"""FSI virtual particles on the GPU (exDyna3D.cpp:1281-1758) against the
reference's own coupling run (fixtures micro_fsi_block and micro_fsi_slab,
written by the compiled reference, tests/golden/make_golden.py) and the
oracle: FP64 within 1e-10, mixed within 1e-4.

The schedule: particles generated after entering step 1, then before step n
the force scale * sin(0.37 (p + 1) + 1.1 d + 0.05 n) on particle p
(tests/fsi_cases.py).
"""
import ctypes as C

import numpy as np
import pytest

import energy_cases as EC
import fsi_cases as FC
from enscdynamics_amd import batch
from enscdynamics_amd import engine as eng
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def rel(a, b, floor=1e-30):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max()) / max(float(np.abs(b).max()), floor)


@pytest.fixture(scope="module", autouse=True)
def gpu_present():
    assert eng.device_count() > 0, "GPU tests need a HIP device"


@pytest.mark.parametrize("case", FC.CASES)
def test_fsi_fp64_matches_reference_run(case):
    d, m, layers, dist, scale = FC.load(case)
    print(f"[fsi] {case}: fp64 engine vs the reference's coupling run", flush=True)
    e = eng.Engine(m, eng.FP64, reorder=True)
    x0, v0, steps = FC.run_engine(e, layers, dist, scale)
    assert steps == d["history"].shape[0]
    assert rel(x0, d["setup_fsi_x"]) <= 1e-14
    assert rel(v0, d["setup_fsi_v"]) <= 1e-14
    x1, v1 = e.fsi_particles()
    for name, got in (("u", e.get("u")), ("v", e.get("v")), ("x", e.get("x")), ("sigma", e.get("sigma")),
                      ("fsi_x", x1), ("fsi_v", v1)):
        err = rel(got, d["final_" + name])
        print(f"[fsi] {case}: {name} rel err {err:.1e}", flush=True)
        assert err <= 1e-10, name
    # energies against the oracle on the same schedule
    o = Oracle(m, 0).init()
    FC.run_oracle(o, layers, dist, scale)
    so, se = o.scalars(), e.scalars()
    scale_e = max(abs(so["wext"]), so["ke"] + so["ie"])
    for k in ("ke", "ie", "hg", "vis", "wext"):
        assert abs(se[k] - so[k]) <= 1e-9 * scale_e, (k, se[k], so[k])


@pytest.mark.parametrize("case", FC.CASES)
def test_fsi_mixed_close_to_reference_run(case):
    d, m, layers, dist, scale = FC.load(case)
    print(f"[fsi mixed] {case}", flush=True)
    e = eng.Engine(m, eng.MIXED, reorder=True)
    FC.run_engine(e, layers, dist, scale)
    x1, v1 = e.fsi_particles()
    err_u = rel(e.get("u"), d["final_u"])
    err_p = rel(x1 - d["setup_fsi_x"], d["final_fsi_x"] - d["setup_fsi_x"])
    print(f"[fsi mixed] {case}: u rel err {err_u:.1e}, particle displacement {err_p:.1e}", flush=True)
    assert err_u <= 1e-4 and err_p <= 1e-4


def _steady(m, layers, dist, f, graph=0, static=False):
    e = eng.Engine(m, eng.FP64, reorder=True)
    if graph:
        e.set_graph(graph)
    e.init()
    e.set_current_step(1)
    n = e.fsi_setup(layers, dist)
    e.fsi_set_forces(np.tile(f, (n, 1)))
    if static:
        while not e.advance_static(16):
            pass
    else:
        e.advance(0)
    return e


def test_fsi_graph_and_static_entry_point_are_bit_exact():
    # a steady particle load run in one advance: plain launches, graph
    # replay and the allocation-free entry point give the same bits
    d, m, layers, dist, scale = FC.load("micro_fsi_block")
    print("[fsi] steady load: plain vs graph vs static", flush=True)
    f = np.array([0.0, 3.0, -5.0])
    a = _steady(m, layers, dist, f)
    b = _steady(m, layers, dist, f, graph=8)
    c = _steady(m, layers, dist, f, static=True)
    xa, va = a.fsi_particles()
    for other in (b, c):
        for fld in ("u", "v", "sigma"):
            assert np.array_equal(a.get(fld), other.get(fld)), fld
        xo, vo = other.fsi_particles()
        assert np.array_equal(xa, xo) and np.array_equal(va, vo)
    o = Oracle(m, 0).init()
    o.begin_step()
    n = o.fsi_setup(layers, dist)
    o.fsi_set_forces(np.tile(f, (n, 1)))
    o.run(0, [], 0)
    assert rel(a.get("u"), o.get("u")) <= 1e-10
    assert rel(xa, o.fsi_particles()[0]) <= 1e-10


def test_fsi_zero_layers_is_dormant():
    # the reference's own run() calls the setup with 0 layers (exDyna3D.cpp:
    # 67): no particles, the run is unchanged bit for bit
    m = EC.pressed_block()
    print("[fsi] zero layers", flush=True)
    a = eng.Engine(m, eng.FP64, reorder=True).run()
    b = eng.Engine(m, eng.FP64, reorder=True)
    b.init()
    assert b.fsi_setup(0, 0.1) == 0
    for s in range(1, len(m.steps)):
        b.set_current_step(s)
        b.advance(0)
    for fld in ("u", "v", "sigma"):
        assert np.array_equal(a.get(fld), b.get(fld)), fld


def test_fsi_device_pointers_for_a_gpu_coupler():
    # a coupler on the GPU writes the forces through ensc_engine_fsi_device
    # on the engine's stream: the same result as the host path
    d, m, layers, dist, scale = FC.load("micro_fsi_slab")
    print("[fsi] device-pointer coupler", flush=True)
    try:
        hip = C.CDLL("libamdhip64.so")
    except OSError:
        hip = C.CDLL("/opt/rocm/lib/libamdhip64.so")
    hip.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
    hip.hipStreamSynchronize.argtypes = [C.c_void_p]
    a = eng.Engine(m, eng.FP64, reorder=True)
    FC.run_engine(a, layers, dist, scale)
    b = eng.Engine(m, eng.FP64, reorder=True)
    b.init()
    b.set_current_step(1)
    n = b.fsi_setup(layers, dist)
    fp, xp, vp, st = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_void_p()
    eng._check(eng.load_library().ensc_engine_fsi_device(b.h, C.byref(fp), C.byref(xp), C.byref(vp), C.byref(st)))
    assert fp.value and xp.value and vp.value
    step, done = 0, False
    while not done:
        f = np.ascontiguousarray(FC.forces(n, step, scale))
        assert hip.hipMemcpyAsync(fp, f.ctypes.data, f.nbytes, 1, st) == 0  # host to device
        assert hip.hipStreamSynchronize(st) == 0
        done = b.advance(1)
        step += 1
    xa, va = a.fsi_particles()
    xb, vb = b.fsi_particles()
    assert np.array_equal(xa, xb) and np.array_equal(va, vb)
    assert np.array_equal(a.get("u"), b.get("u"))


def test_fsi_rejects_bad_calls():
    m = EC.pressed_block()
    e = eng.Engine(m, eng.FP64)
    with pytest.raises(eng.EnscError):
        e.fsi_setup(2, 0.001)  # before init
    e.init()
    n = e.fsi_setup(2, 0.001)
    assert n > 0
    with pytest.raises(eng.EnscError):
        e.fsi_set_forces(np.zeros((n + 1, 3)))
    with pytest.raises(eng.EnscError):
        e.fsi_setup(2, -1.0)
    bt = batch.Batch([m, m], eng.FP64)
    bt.init()
    with pytest.raises(eng.EnscError):
        eng.load_library()
        eng._check(eng.load_library().ensc_engine_fsi_setup(bt.h, 2, 0.001, None))
