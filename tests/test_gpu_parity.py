"""GPU parity: the HIP engine (through the C ABI) against the reference's
golden vectors and the CPU oracle.

Bars: FP64 path <= 1e-10 relative to the reference / oracle on the same
deck; mixed path <= 1e-4 relative displacement error vs FP64 and an energy
difference <= 1% of the reference energy scale (KE0 or external work).
"""
import numpy as np
import pytest

import energy_cases as EC
import goldens as G
from enscdynamics_amd import decks
from enscdynamics_amd import engine as eng
from oracle.oracle import HIST_BASE, Oracle

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-10
TOL_MIXED = 1e-4


def rel(a, b, floor=0.0):
    s = max(float(np.abs(b).max()) if np.size(b) else 0.0, floor)
    d = float(np.abs(np.asarray(a) - np.asarray(b)).max()) if np.size(b) else 0.0
    return d / s if s > 0 else d


@pytest.fixture(scope="module", autouse=True)
def gpu_present():
    n = eng.device_count()
    print(f"[gpu] devices visible: {n}", flush=True)
    assert n > 0, "GPU tests need a HIP device"


def test_fp64_c1_matches_reference():
    d = G.load("c1")
    m = G.model_from_dump(d)
    tip = int(d["tip_node"])
    e = eng.Engine(m, eng.FP64, probes=[tip], hist_cap=4096)
    e.run()
    h = e.history()
    H = d["history"]
    assert h.shape[0] == H.shape[0] == 2008
    assert rel(h[:, 0], H[:, 0]) <= 1e-14
    assert rel(h[:, 1], H[:, 1]) <= 1e-11
    assert rel(h[:, HIST_BASE:], H[:, 4:]) <= TOL_FP64
    for f in ["u", "v", "x", "sigma", "qia", "rf"]:
        err = rel(e.get(f), d["final_" + f])
        print(f"[c1 fp64] {f}: {err:.2e}", flush=True)
        assert err <= TOL_FP64, f


@pytest.mark.parametrize("case", G.CASES_MICRO)
def test_fp64_micro_matches_reference(case):
    d = G.load(case)
    m = G.model_from_dump(d)
    e = eng.Engine(m, eng.FP64)
    e.init()
    e.set_current_step(1)
    vscale = float(np.abs(d["v_init"]).max()) + 1.0
    taken = 0
    for snap in (1, 2, 3, 50):
        e.advance(snap - taken)
        taken = snap
        assert int(e.scalars()["nsteps"]) == snap
        for f, fl in (("u", 0.0), ("v", 1e-9 * vscale), ("sigma", 0.0), ("rf", 0.0)):
            assert rel(e.get(f), d[f"step{snap}_{f}"], fl) <= TOL_FP64, (snap, f)
    # finish every remaining *Step like exDyna3D::run
    e.advance(0)
    for s in range(2, len(m.steps)):
        e.set_current_step(s)
        e.advance(0)
    assert int(e.scalars()["nsteps"]) == d["history"].shape[0]
    sigmax = float(np.abs(d["final_sigma"]).max())
    for f, fl in (("u", 0.0), ("v", 1e-9 * vscale), ("x", 0.0), ("sigma", 0.0), ("rf", 1e-12 * sigmax)):
        assert rel(e.get(f), d["final_" + f], fl) <= TOL_FP64, f


@pytest.mark.parametrize("case", G.CASES_REF)
def test_fp64_reference_decks(case):
    d = G.load(case)
    m = G.model_from_dump(d)
    n_ref = int(d["n_steps"])
    e = eng.Engine(m, eng.FP64, probes=[0], hist_cap=n_ref + 8)
    e.run()
    h = e.history()
    assert h.shape[0] == n_ref
    st = int(d["history_stride"])
    assert rel(h[::st][:, 0], d["history_sub"][:, 0]) <= 1e-14
    assert rel(h[::st][:, HIST_BASE:], d["history_sub"][:, 4:], 1e-12) <= TOL_FP64
    vmax = max(1.0, float(np.abs(m.v_init).max()))
    for f, fl in (("u", 1e-12), ("v", 1e-9 * vmax), ("x", 0.0), ("sigma", 1.0), ("rf", 1e-12)):
        assert rel(e.get(f), d["final_" + f], fl) <= TOL_FP64, f


def side_centre(m):
    # the node nearest the centre of the x = max face: a free surface point
    # whose motion is computed (the top face's is prescribed)
    x = m.x0
    c = np.array([x[:, 0].max(), 0.5 * (x[:, 1].min() + x[:, 1].max()), 0.5 * (x[:, 2].min() + x[:, 2].max())])
    return int(np.argmin(((x - c) ** 2).sum(axis=1)))


def run_both(m, steps, precision=eng.FP64, reorder=False, probes=()):
    e = eng.Engine(m, precision, reorder=reorder, probes=probes, hist_cap=steps + 8)
    e.run_steps(steps)
    return e


def test_fp64_matches_oracle_c2_small_jittered_ramp():
    # configuration-2 physics (jittered block, velocity ramp by amplitude) at
    # a size the oracle finishes in seconds
    m = decks.config2(nx=24, ny=6, nz=6, steps=400)
    e = eng.Engine(m, eng.FP64, hist_cap=512)
    e.run()
    o = Oracle(m, 0).init()
    o.run(0, [], 0)
    assert int(e.scalars()["nsteps"]) == int(o.scalars()["nsteps"])
    for f in ["u", "v", "x", "sigma", "qia", "rf"]:
        err = rel(e.get(f), o.get(f), 1e-12)
        print(f"[c2s fp64 vs oracle] {f}: {err:.2e}", flush=True)
        assert err <= TOL_FP64, f


def test_fp64_matches_oracle_plastic_cube():
    m = decks.config3(n=16, steps=500)
    e = eng.Engine(m, eng.FP64)
    e.run()
    o = Oracle(m, 0).init()
    o.run(0, [], 0)
    assert (o.get("epsp") > 0).any()
    for f in ["u", "v", "sigma", "qia", "epsp", "rf"]:
        err = rel(e.get(f), o.get(f), 1e-12)
        print(f"[c3s fp64 vs oracle] {f}: {err:.2e}", flush=True)
        assert err <= TOL_FP64, f
    so, se = o.scalars(), e.scalars()
    for k in ("ie", "hg", "vis", "wext", "ke"):
        assert se[k] == pytest.approx(so[k], rel=1e-9, abs=1e-9 * abs(so["wext"])), k


def test_reorder_is_bit_exact():
    # the reference's summation order (fp64_order=1): per node, corner by
    # corner, ascending original element - independent of the numbering
    m = decks.config2(nx=20, ny=7, nz=5, steps=200, seed=11)
    a = eng.Engine(m, eng.FP64, reorder=False, fp64_order=1)
    a.run()
    b = eng.Engine(m, eng.FP64, reorder=True, fp64_order=1)
    b.run()
    for f in ["u", "v", "x", "sigma", "qia", "rf"]:
        assert np.array_equal(a.get(f), b.get(f)), f
    # the oracle sums in the same order; the device's cbrt may differ from
    # glibc's by an ulp, so agreement is to rounding
    o = Oracle(m, 0).init()
    o.run(0, [], 0)
    for f in ["u", "v", "sigma", "qia"]:
        assert rel(a.get(f), o.get(f), 1e-12) <= 1e-12, f


def test_mixed_amplitude_bc_job_in_the_pipelined_pass():
    # a velocity BC ramped by an amplitude: the pipelined pass overrides the
    # staged velocities of the tiles that hold constrained nodes with the
    # value at each step's time.  Against the one-block-per-tile pass
    # (staging 1), which applies the same override while staging (the two
    # differ by contraction only), and both against FP64
    m = decks.config2(nx=40, ny=12, nz=12, steps=600)
    a = eng.Engine(m, eng.FP64, reorder=True).run()
    b = eng.Engine(m, eng.MIXED, reorder=True)
    b.init()
    b.set_current_step(1)
    b.advance(3)
    lay = b.layout()
    b.advance(0)
    assert lay["pipe_blocks"] == (m.n_elems + 255) // 256 and lay["elem_blocks"] == 0, lay
    c = eng.Engine(m, eng.MIXED, reorder=True, staging=1).run()
    assert b.scalars()["done"] == 1.0 and b.scalars()["err"] == 0.0
    err_b, err_c = rel(b.get("u"), a.get("u")), rel(c.get("u"), a.get("u"))
    err_bc = rel(b.get("u"), c.get("u"))
    print(f"[amplitude BC pipelined] mixed vs fp64: pipelined {err_b:.2e}, staged {err_c:.2e}; "
          f"pipelined vs staged {err_bc:.2e}", flush=True)
    assert err_b <= TOL_MIXED and err_c <= TOL_MIXED
    assert err_bc <= 1e-5


def test_fp64_block_assembly_vs_reference_order():
    # default FP64 path (the Walsh-form element body, block partials summed
    # per node) against the reference's order: the same forces regrouped and
    # added in another order, so the results differ by rounding only,
    # renumbered or not
    m = decks.config3(n=14, steps=400)
    a = eng.Engine(m, eng.FP64, reorder=True, fp64_order=1).run()
    for reorder in (True, False):
        b = eng.Engine(m, eng.FP64, reorder=reorder)
        b.init()
        b.set_current_step(1)
        b.advance(3)
        lay = b.layout()
        b.advance(0)
        assert lay["asm_jobs"] == 1 and lay["staged"] == 1, lay
        assert b.scalars()["nsteps"] == a.scalars()["nsteps"]
        for f in ("u", "v", "sigma", "qia", "rf"):
            err = rel(b.get(f), a.get(f), 1e-12)
            assert err <= 1e-12, (reorder, f, err)
    print(f"[fp64 block assembly] within 1e-12 of the reference order", flush=True)


def test_determinism_two_runs():
    m = decks.config3(n=12, steps=300)
    r = []
    for _ in range(2):
        e = eng.Engine(m, eng.MIXED, reorder=True)
        e.run()
        r.append({f: e.get(f) for f in ("u", "v", "sigma", "qia")} | {"s": e.scalars()})
    for f in ("u", "v", "sigma", "qia"):
        assert np.array_equal(r[0][f], r[1][f]), f
    assert r[0]["s"]["ie"] == r[1]["s"]["ie"]


def test_mixed_renumbering_within_tolerance():
    # the mixed engine adds each node's force partials per element block, so
    # a different numbering (different blocks) changes rounding only
    m = decks.config3(n=10, steps=300)
    a = eng.Engine(m, eng.MIXED, reorder=False).run()
    b = eng.Engine(m, eng.MIXED, reorder=True).run()
    for f in ("u", "sigma"):
        err = rel(a.get(f), b.get(f))
        print(f"[mixed renumbering] {f}: {err:.2e}", flush=True)
        assert err <= 1e-5, f
    assert a.scalars()["nsteps"] == b.scalars()["nsteps"]


def rotated_corners(m, every=2):
    """Same mesh with every `every`-th element's corners renumbered by a
    quarter turn about the element's third axis (still right-handed).  Blocks
    then hold elements that share a node at the same corner index, which
    rules out block assembly for the job."""
    q = m.copy() if hasattr(m, "copy") else m
    c = np.array(q.conn, copy=True)
    rot = [1, 2, 3, 0, 5, 6, 7, 4]
    c[::every] = c[::every][:, rot]
    q.conn = c
    return q


def test_mixed_gathered_assembly_fallback_vs_fp64():
    m = rotated_corners(decks.config3(n=8, steps=400))
    a = eng.Engine(m, eng.FP64, reorder=True).run()
    b = eng.Engine(m, eng.MIXED, reorder=True).run()
    err = rel(b.get("u"), a.get("u"))
    print(f"[mixed gathered fallback] u vs fp64: {err:.2e}", flush=True)
    assert err <= TOL_MIXED
    o = Oracle(m, 0).init()
    o.run(0, [], 0)
    assert rel(a.get("u"), o.get("u"), 1e-12) <= TOL_FP64


def test_mixed_vs_fp64_c1():
    m = decks.config1()
    tip = m.tip_node
    a = eng.Engine(m, eng.FP64, probes=[tip], hist_cap=4096).run()
    b = eng.Engine(m, eng.MIXED, probes=[tip], hist_cap=4096).run()
    ha, hb = a.history(), b.history()
    n = min(len(ha), len(hb))
    err = rel(hb[:n, HIST_BASE:], ha[:n, HIST_BASE:])
    print(f"[c1 mixed vs fp64] tip u history rel err {err:.2e}", flush=True)
    assert err <= TOL_MIXED
    assert rel(b.get("u"), a.get("u")) <= TOL_MIXED
    sa, sb = a.scalars(), b.scalars()
    ea = sa["ke"] + sa["ie"] + sa["hg"] + sa["vis"]
    eb = sb["ke"] + sb["ie"] + sb["hg"] + sb["vis"]
    assert abs(ea - eb) <= 0.01 * sa["ke0"]


def test_mixed_vs_fp64_plastic_cube():
    m = decks.config3(n=20, steps=800)
    a = eng.Engine(m, eng.FP64).run()
    b = eng.Engine(m, eng.MIXED).run()
    err_u = rel(b.get("u"), a.get("u"))
    err_s = rel(b.get("sigma"), a.get("sigma"))
    sa, sb = a.scalars(), b.scalars()
    e_err = abs((sb["ke"] + sb["ie"] + sb["hg"] + sb["vis"]) - (sa["ke"] + sa["ie"] + sa["hg"] + sa["vis"]))
    print(f"[c3s mixed vs fp64] u {err_u:.2e} sigma {err_s:.2e} energy {e_err / sa['wext']:.2e} of W_ext",
          flush=True)
    assert err_u <= TOL_MIXED
    assert e_err <= 0.01 * sa["wext"]


def test_mixed_pipelined_pass_vs_fp64_plastic_cube():
    # the persistent element pass that copies the next tile's nodes into
    # LDS while the current tile computes (default for such jobs)
    m = decks.config3(n=20, steps=800)
    a = eng.Engine(m, eng.FP64).run()
    b = eng.Engine(m, eng.MIXED, reorder=True)
    b.init()
    b.set_current_step(1)
    b.advance(5)
    lay = b.layout()
    assert lay["staged"] == 1 and lay["asm_jobs"] == 1
    assert lay["pipe_blocks"] == (m.n_elems + 255) // 256 and lay["elem_blocks"] == 0, lay
    b.advance(0)
    sb = b.scalars()
    assert sb["done"] == 1.0 and sb["err"] == 0.0
    err_u = rel(b.get("u"), a.get("u"))
    sa = a.scalars()
    e_err = abs((sb["ke"] + sb["ie"] + sb["hg"] + sb["vis"]) - (sa["ke"] + sa["ie"] + sa["hg"] + sa["vis"]))
    print(f"[c3s mixed pipelined vs fp64] u {err_u:.2e} energy {e_err / sa['wext']:.2e} of W_ext", flush=True)
    assert err_u <= TOL_MIXED
    assert e_err <= 0.01 * sa["wext"]
    # staging option 1 on the same job: same layout without the pipeline
    c = eng.Engine(m, eng.MIXED, reorder=True, staging=1)
    c.init()
    c.set_current_step(1)
    c.advance(5)
    lay = c.layout()
    assert lay["pipe_blocks"] == 0 and lay["elem_blocks"] == (m.n_elems + 255) // 256, lay


def test_full_size_cube_fp64_vs_oracle_few_steps():
    # configuration 3 at its full 1M-element size: a few steps against the
    # oracle (size-independent check of the same arithmetic at scale)
    m = decks.config3(n=100, steps=3000)
    e = eng.Engine(m, eng.FP64, reorder=True)
    e.run_steps(6)
    o = Oracle(m, 0).init()
    o.run(6, [], 0)
    for f in ["u", "v", "sigma", "qia"]:
        err = rel(e.get(f), o.get(f), 1e-12)
        print(f"[c3 full fp64 vs oracle, 6 steps] {f}: {err:.2e}", flush=True)
        assert err <= TOL_FP64, f


def test_fp64_full_size_cube_split_finalize_is_deterministic():
    # the 1M-element cube in FP64 with 128-element blocks has 7813 element
    # blocks, so its element finalize folds the partials over 8 blocks and
    # the last block to count itself decides; which block that is varies
    # from run to run, the result must not: two runs agree bit for bit,
    # scalars included, and dt and the energies agree with the oracle over
    # the same steps
    m = decks.config3(n=100, steps=3000)
    runs = []
    for _ in range(2):
        e = eng.Engine(m, eng.FP64, reorder=True, block_elem=128)
        e.run_steps(40)
        runs.append(e)
    a, b = runs
    assert a.layout()["elem_blocks"] > 4096
    for f in ("u", "v", "sigma", "qia", "epsp"):
        assert np.array_equal(a.get(f), b.get(f)), f
    sa, sb = a.scalars(), b.scalars()
    for k in ("time", "dt_i", "dt_i_1", "dts_min", "ke", "ie", "hg", "vis", "wext"):
        assert sa[k] == sb[k], k
    o = Oracle(m, 0).init()
    o.run(40, [], 0)
    so = o.scalars()
    for k in ("time", "dt_i"):
        assert sa[k] == pytest.approx(so[k], rel=1e-12), k
    for k in ("ie", "hg", "vis", "wext", "ke"):
        assert sa[k] == pytest.approx(so[k], rel=1e-9, abs=1e-9 * max(abs(so["wext"]), so["ke0"])), k
    print(f"[c3 fp64 split finalize] 40 steps, two runs bit-identical, dt {sa['dt_i']:.6e} "
          f"(oracle {so['dt_i']:.6e})", flush=True)


def test_mixed_vs_fp64_full_plastic_cube_3000_steps():
    # configuration 3 as specified: 1M elements, J2, run to the end time of
    # 3000 initial stable steps (the step shrinks as the cube deforms, so a
    # few more steps are taken).  The mixed engine against the FP64 engine
    # (itself pinned to the oracle above): displacement within 1e-4, total
    # energy within 1% of the external work.
    m = decks.config3(n=100, steps=3000)
    tip = side_centre(m)
    a = eng.Engine(m, eng.FP64, reorder=True, probes=[tip], hist_cap=4000).run()
    b = eng.Engine(m, eng.MIXED, reorder=True, probes=[tip], hist_cap=4000).run()
    sa, sb = a.scalars(), b.scalars()
    assert sa["nsteps"] >= 3000 and abs(sb["nsteps"] - sa["nsteps"]) <= 1
    assert sa["done"] == 1.0 and sb["done"] == 1.0
    err_u = rel(b.get("u"), a.get("u"))
    err_s = rel(b.get("sigma"), a.get("sigma"))
    ha, hb = a.history(), b.history()
    n = min(len(ha), len(hb))
    err_tip = rel(hb[:n, HIST_BASE:], ha[:n, HIST_BASE:])
    ea = sa["ke"] + sa["ie"] + sa["hg"] + sa["vis"]
    eb = sb["ke"] + sb["ie"] + sb["hg"] + sb["vis"]
    e_err = abs(eb - ea) / sa["wext"]
    plastic = float((a.get("epsp") > 0).mean())
    print(f"[c3 full mixed vs fp64, {int(sa['nsteps'])} steps] u {err_u:.2e}, tip history {err_tip:.2e}, "
          f"sigma {err_s:.2e}, energy {e_err:.2e} of W_ext, {100 * plastic:.1f}% of elements yielded", flush=True)
    assert plastic > 0.0
    assert err_u <= TOL_MIXED and err_tip <= TOL_MIXED
    assert err_s <= 1e-3
    assert e_err <= 0.01


def test_mixed_vs_fp64_full_jittered_block_c2():
    # configuration 2 as specified: 200x40x40 jittered block (320k elements),
    # top face driven at -2 m/s ramped over the first 10% by an amplitude,
    # 5000 steps; the FP64 GPU engine against the mixed one
    m = decks.config2()
    tip = side_centre(m)
    a = eng.Engine(m, eng.FP64, reorder=True, probes=[tip], hist_cap=6000).run()
    b = eng.Engine(m, eng.MIXED, reorder=True, probes=[tip], hist_cap=6000).run()
    sa, sb = a.scalars(), b.scalars()
    assert sa["done"] == 1.0 and sb["done"] == 1.0 and abs(sb["nsteps"] - sa["nsteps"]) <= 1
    err_u = rel(b.get("u"), a.get("u"))
    err_s = rel(b.get("sigma"), a.get("sigma"))
    ha, hb = a.history(), b.history()
    n = min(len(ha), len(hb))
    err_tip = rel(hb[:n, HIST_BASE:], ha[:n, HIST_BASE:])
    e_err = abs((sb["ke"] + sb["ie"] + sb["hg"] + sb["vis"]) - (sa["ke"] + sa["ie"] + sa["hg"] + sa["vis"]))
    print(f"[c2 full mixed vs fp64, {int(sa['nsteps'])} steps] u {err_u:.2e} tip history {err_tip:.2e} "
          f"sigma {err_s:.2e} energy {e_err / sa['wext']:.2e} of W_ext", flush=True)
    assert err_u <= TOL_MIXED and err_tip <= TOL_MIXED
    assert err_s <= 1e-3
    assert e_err <= 0.01 * sa["wext"]


def edge_block(nx, ny, nz, plastic=False):
    # a clamped block with a transverse initial velocity: x = 0 face fixed
    x0, conn, ijk = decks.block_mesh(nx, ny, nz, 0.01 * nx, 0.01 * ny, 0.01 * nz)
    fixed = decks.face_nodes(ijk, 0, 0, nx)
    rest = np.setdiff1d(np.arange(x0.shape[0], dtype=np.uint32), fixed)
    m = decks.Model(x0=x0, conn=conn, E=200e9, nu=0.3, rho=7800.0, name=f"edge_{nx}x{ny}x{nz}")
    if plastic:
        m.plastic, m.yield0, m.hard = True, 2e6, 1e9
    m.node_sets = {"Set-1": fixed, "Set-2": rest}
    v = np.zeros_like(x0)
    v[rest, 2] = 3.0
    m.v_init = v
    m.ic = [("Set-2", 2, 3.0)]
    T = 150 * decks.stable_dt0(m)
    m.steps = [decks.Step("Initial", 0.0, spc=[decks.Bc("spc", fixed, 0, 2, 1, name="Set-1")]),
               decks.Step("Step-1", T)]
    m.total_time = T
    return m


@pytest.mark.parametrize("dims", [(1, 1, 1), (7, 1, 1), (5, 5, 10), (8, 8, 4), (257, 1, 1), (9, 9, 4)])
def test_edge_sizes_match_oracle(dims):
    # one element, a partial element block, exactly one block, one block plus
    # one element along a line (gathered: its blocks do not fit LDS), and two
    # tiles of the pipelined pass with a partial second tile
    m = edge_block(*dims, plastic=dims == (9, 9, 4))
    o = Oracle(m, 0).init()
    o.run(0, [], 0)
    a = eng.Engine(m, eng.FP64, reorder=True).run()
    for f in ("u", "v", "sigma"):
        assert rel(a.get(f), o.get(f), 1e-12) <= TOL_FP64, (dims, f)
    b = eng.Engine(m, eng.MIXED, reorder=True)
    b.init()
    b.set_current_step(1)
    b.advance(3)
    lay = b.layout()
    b.advance(0)
    # relative to the displacement scale of the driven motion (v0 T): a single
    # clamped element barely moves, and its max |u| is no scale for rounding
    err = rel(b.get("u"), a.get("u"), 3.0 * m.total_time)
    print(f"[edge {dims}] layout {lay}, mixed u vs fp64 {err:.2e}", flush=True)
    assert b.scalars()["done"] == 1.0 and b.scalars()["err"] == 0.0
    assert err <= TOL_MIXED


def test_bad_models_fail_loudly():
    m = decks.config3(n=4, steps=10)
    bad = decks.config3(n=4, steps=10)
    bad.conn = bad.conn.copy()
    bad.conn[0] = bad.conn[0][[4, 5, 6, 7, 0, 1, 2, 3]]  # top and bottom faces swapped: inverted element
    with pytest.raises(eng.EnscError):
        eng.Engine(bad)
    nan = decks.config3(n=4, steps=10)
    nan.v_init = nan.v_init.copy()
    nan.v_init[3, 1] = np.nan
    with pytest.raises(eng.EnscError):
        eng.Engine(nan)
    eng.Engine(m).run()


def test_mixed_rigid_translation_u_is_the_compensated_sum():
    # A free cube in uniform motion: every corner velocity is the same, so
    # the Walsh differences, the stress, the hourglass and viscous terms and
    # hence the accelerations are exactly zero, v stays v0, and after N steps
    # u = sum_i fl(dt_i * v0) (update_displacement, exDyna3D.cpp:651-654)
    # with dt_i the step's increment from the history.  The mixed engine adds
    # the increments with Kahan compensation; a plain sum of these equal
    # increments drifts by hundreds of ulp.
    from fractions import Fraction
    from math import ulp

    from enscdynamics_amd.model import Model, Step

    x0, conn, _ = decks.block_mesh(4, 4, 4, 0.04, 0.04, 0.04)
    v0 = np.array([3.1, -1.7, 0.9])
    m = Model(x0=x0, conn=conn, E=210e9, nu=0.3, rho=7850.0, name="rigid_translation_4x4x4")
    m.node_sets = {"Set-1": np.arange(x0.shape[0], dtype=np.uint32)}
    m.v_init = np.tile(v0, (x0.shape[0], 1))
    m.ic = [("Set-1", d, float(v0[d])) for d in range(3)]
    T = 5000 * decks.stable_dt0(m)
    m.steps = [Step("Initial", 0.0), Step("Step-1", T)]
    m.total_time = T
    # 5000 steps on a 4-element cube span ~830 wave transits: the automatic
    # rule would run this quasi-static job in FP64, so mixed is forced
    e = eng.Engine(m, eng.MIXED, reorder=True, hist_cap=8000, fp64_jobs=1).run()
    assert e.layout()["fp64_jobs"] == 0
    h = e.history()
    n = int(e.scalars()["nsteps"])
    assert h.shape[0] == n and n >= 5000
    v, u = e.get("v"), e.get("u")
    assert np.array_equal(v, np.tile(v0, (x0.shape[0], 1)))
    worst, plain_worst = 0.0, 0.0
    for d in range(3):
        incs = [float(dt) * float(v0[d]) for dt in h[:, 1]]
        exact = sum((Fraction(y) for y in incs), Fraction(0))
        plain = 0.0
        for y in incs:
            plain += y
        un = ulp(float(exact))
        assert np.all(u[:, d] == u[0, d])  # every node moved alike
        worst = max(worst, float(abs(Fraction(float(u[0, d])) - exact)) / un)
        plain_worst = max(plain_worst, float(abs(Fraction(plain) - exact)) / un)
    print(f"[rigid translation, {n} steps] compensated u within {worst:.2f} ulp of the exact sum; "
          f"a plain sum is {plain_worst:.1f} ulp off", flush=True)
    assert plain_worst >= 10.0  # the case does expose summation error
    assert worst <= 2.0


ENERGY_CASES = {"free_fall": lambda: EC.free_fall(steps=200), "driven_column": lambda: EC.driven_column(),
                "ramped_block": lambda: EC.ramped_block(), "fixed_after_release": lambda: EC.fixed_after_release()}


@pytest.mark.parametrize("case", sorted(ENERGY_CASES))
def test_fp64_energies_match_oracle(case):
    # the external work (reactions, gravity, constraint jumps) and the other
    # energy terms of the engine against the oracle, whose energy account is
    # pinned by closed forms in test_oracle.py
    m = ENERGY_CASES[case]()
    print(f"[energy] {case}: fp64 engine vs oracle", flush=True)
    e = eng.Engine(m, eng.FP64, reorder=True).run()
    o = Oracle(m, 0).init()
    o.run(0, [], 0)
    so, se = o.scalars(), e.scalars()
    assert se["nsteps"] == so["nsteps"]
    for f in ("u", "v"):
        assert rel(e.get(f), o.get(f), 1e-12) <= TOL_FP64, f
    # stress against max(|sigma|, E * 1e-10): in free fall the exact stress
    # is zero and the oracle's is rounding noise (~3e-11 Pa: the corner sum
    # of v_k b_k with sum_k b_k not exactly 0), which the engine's Walsh-form
    # FP64 body cancels exactly
    assert rel(e.get("sigma"), o.get("sigma"), 1e-10 * m.E) <= TOL_FP64, "sigma"
    scale = max(abs(so["wext"]), so["ke0"])
    for k in ("ke", "ie", "hg", "vis", "wext", "ke0"):
        assert abs(se[k] - so[k]) <= 1e-9 * scale, (k, se[k], so[k])
    print(f"[energy] {case}: W_ext {se['wext']:.6e} (oracle {so['wext']:.6e}), balance "
          f"{EC.balance(se) / scale:.2e} of the energy scale", flush=True)


@pytest.mark.parametrize("case", sorted(ENERGY_CASES))
def test_fp64_history_rows_match_oracle(case):
    # every column of every history row against the oracle's rows: the time
    # columns come from the element finalize's deciding block, the energy
    # columns from the finalize blocks that own those scalars (the last row
    # from the node finalize that ends the run)
    m = ENERGY_CASES[case]()
    probe = m.x0.shape[0] - 1
    e = eng.Engine(m, eng.FP64, probes=[probe], hist_cap=100000).run()
    n = int(e.scalars()["nsteps"])
    o = Oracle(m, 0).init()
    no, H = o.run(0, [probe], n)
    h = e.history()
    assert no == n and h.shape == H.shape == (n, HIST_BASE + 3)
    so = o.scalars()
    scale = max(abs(so["wext"]), so["ke0"])
    assert rel(h[:, 0], H[:, 0]) <= 1e-13
    assert rel(h[:, 1:4], H[:, 1:4]) <= 1e-11
    de = float(np.abs(h[:, 4:HIST_BASE] - H[:, 4:HIST_BASE]).max())
    assert de <= 1e-9 * scale, de
    assert rel(h[:, HIST_BASE:], H[:, HIST_BASE:], 1e-12) <= TOL_FP64
    print(f"[history] {case}: {n} rows, energy columns within {de / scale:.1e} of the energy scale", flush=True)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_history_alike_one_step_per_launch(prec):
    # a continuous run's rows are written by the element finalize's blocks,
    # a run of one-step launches' rows by the node finalize that ends each
    # launch: the two histories, and the final scalars, agree bit for bit
    p = eng.FP64 if prec == "fp64" else eng.MIXED
    m = EC.driven_column()
    a = eng.Engine(m, p, probes=[0], hist_cap=100000).run()
    b = eng.Engine(m, p, probes=[0], hist_cap=100000).init()
    for st in range(1, len(m.steps)):
        b.set_current_step(st)
        while not b.advance(1):
            pass
    ha, hb = a.history(), b.history()
    print(f"[history chunked] {prec}: {ha.shape[0]} rows", flush=True)
    assert ha.shape == hb.shape and ha.shape[0] > 100
    assert np.array_equal(ha, hb)
    sa, sb = a.scalars(), b.scalars()
    for k in ("time", "dt_i", "ke", "ie", "hg", "vis", "wext"):
        assert sa[k] == sb[k], k


def test_history_rows_with_four_finalize_blocks():
    # a job of more than 1024 element blocks: its element finalize runs as
    # four blocks (the decision, the nodal update's close, two work blocks),
    # each writing its own history columns.  60 steps of the J2 cube of 70^3
    # elements: the FP64 rows against the oracle's, and in both precisions a
    # continuous run against one-step launches (whose rows the node finalize
    # writes), bit for bit
    m = decks.config3(n=70, steps=3000)
    probe = m.x0.shape[0] - 1
    n = 60

    def run(p, chunked):
        e = eng.Engine(m, p, probes=[probe], hist_cap=n + 8)
        e.init()
        e.set_current_step(1)
        if chunked:
            for _ in range(n):
                e.advance(1)
        else:
            e.advance(n)
        return e

    a = run(eng.FP64, False)
    assert a.layout()["elem_blocks"] > 1024
    o = Oracle(m, 0).init()
    no, H = o.run(n, [probe], n)
    h = a.history()
    assert no == n and h.shape == H.shape == (n, HIST_BASE + 3)
    so = o.scalars()
    scale = max(abs(so["wext"]), so["ke0"])
    assert rel(h[:, 0], H[:, 0]) <= 1e-13
    assert rel(h[:, 1:4], H[:, 1:4]) <= 1e-11
    de = float(np.abs(h[:, 4:HIST_BASE] - H[:, 4:HIST_BASE]).max())
    assert de <= 1e-9 * scale, de
    assert rel(h[:, HIST_BASE:], H[:, HIST_BASE:], 1e-12) <= TOL_FP64
    print(f"[history 4-block finalize] fp64 vs oracle: energy columns within {de / scale:.1e} of the scale",
          flush=True)
    for p, cont in ((eng.FP64, a), (eng.MIXED, run(eng.MIXED, False))):
        ch = run(p, True)
        assert np.array_equal(cont.history(), ch.history()), p
        sa, sb = cont.scalars(), ch.scalars()
        for k in ("time", "dt_i", "ke", "ie", "hg", "vis", "wext"):
            assert sa[k] == sb[k], (p, k)


def test_mixed_energies_close_to_fp64_ramp_and_gravity():
    # mixed engine: the same external work within 1% (the mixed bar on energy)
    for m in (EC.ramped_block(), EC.free_fall(steps=200)):
        a = eng.Engine(m, eng.FP64).run()
        b = eng.Engine(m, eng.MIXED).run()
        sa, sb = a.scalars(), b.scalars()
        err = abs(sb["wext"] - sa["wext"]) / abs(sa["wext"])
        print(f"[energy mixed] {m.name}: W_ext {sb['wext']:.6e} vs fp64 {sa['wext']:.6e} ({err:.1e})", flush=True)
        assert err <= 1e-2
        assert rel(b.get("u"), a.get("u")) <= TOL_MIXED
