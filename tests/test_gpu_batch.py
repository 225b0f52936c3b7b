"""Batched and served execution on the GPU.

* alone == in-mix: each job's result in a batch (or in the serving loop,
  admitted while others run) equals the same job run alone in its own
  engine, bit for bit (per-job reductions only, fixed order);
* a hipGraph-captured step replays to the same state as eager launches;
* every job reaches its own end time, and FP64 batch results match the
  CPU oracle.
"""
import numpy as np
import pytest

import goldens as G
from enscdynamics_amd import decks
from enscdynamics_amd import engine as eng
from enscdynamics_amd.batch import Batch, Server
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def small_jobs(n, seed=7, plastic_p=0.0):
    rng = np.random.default_rng(seed)
    return [decks.random_job(rng, 300, 4000, plastic_p, k_lo=150, k_hi=400, name=f"t{i}") for i in range(n)]


def thin_slab(steps=300):
    """One element thick: its 256-element blocks hold 578 nodes, more than
    the pipelined pass stages, so its element blocks run the
    one-block-per-tile pass."""
    from enscdynamics_amd.model import Bc, Model, Step
    h = 0.01
    x0, conn, ijk = decks.block_mesh(32, 32, 1, 32 * h, 32 * h, h)
    fixed = decks.face_nodes(ijk, 0, 0, 32)
    rest = np.setdiff1d(np.arange(x0.shape[0], dtype=np.uint32), fixed)
    m = Model(x0=x0, conn=conn, E=70e9, nu=0.33, rho=2700.0, name="thin_slab_32x32x1")
    m.node_sets = {"Set-1": fixed, "Set-2": rest}
    v = np.zeros_like(x0)
    v[rest, 2] = 3.0
    m.v_init = v
    m.ic = [("Set-2", 2, 3.0)]
    T = steps * decks.stable_dt0(m)
    m.steps = [Step("Initial", 0.0, spc=[Bc("spc", fixed, 0, 2, 1, name="Set-1")]), Step("Step-1", T)]
    m.total_time = T
    return m


@pytest.mark.parametrize("precision,staging", [(eng.FP64, 0), (eng.MIXED, 0), (eng.MIXED, 1)])
def test_batch_alone_equals_in_mix(precision, staging):
    # staging 0: element blocks that fit the pipelined pass run it, the
    # others the one-block-per-tile pass, decided from each job alone;
    # staging 1: every job the one-block-per-tile pass
    jobs = small_jobs(10, seed=3, plastic_p=0.5)
    # a ramped velocity BC (amplitude): the pipelined pass overrides the
    # staged velocities of its constrained tiles; the thin slab's blocks run
    # the one-block-per-tile pass, so a mixed batch at staging 0 runs both
    # passes per step
    jobs.append(decks.config2(nx=12, ny=6, nz=6, steps=300))
    jobs.append(thin_slab())
    b = Batch(jobs, precision, reorder=True, staging=staging).run()
    for i, m in enumerate(jobs):
        s = b.scalars(i)
        assert s["done"] == 1.0 and s["err"] == 0.0
        assert s["time"] == pytest.approx(m.total_time, rel=1e-12)
        a = eng.Engine(m, precision, reorder=True, staging=staging).run()
        for f in ("u", "v", "sigma", "qia"):
            assert np.array_equal(b.get(i, f), a.get(f)), (i, f)
        sa = a.scalars()
        for k in ("nsteps", "ie", "hg", "vis", "ke"):
            assert s[k] == sa[k], (i, k)
    print(f"[batch] {len(jobs)} jobs equal alone and in the mix", flush=True)


@pytest.mark.parametrize("precision", [eng.FP64, eng.MIXED])
def test_batch_of_large_and_small_jobs_alone_equals_in_mix(precision):
    # a job of more than 1024 element blocks (its element finalize runs as
    # four blocks, one per group of scalars) beside small jobs (one block
    # each): every job equals its own run, bit for bit, energies included
    big = decks.config3(n=70, steps=600)
    jobs = [big] + small_jobs(6, seed=5, plastic_p=0.5)
    b = Batch(jobs, precision, reorder=True).run()
    for i, m in enumerate(jobs):
        s = b.scalars(i)
        assert s["done"] == 1.0 and s["err"] == 0.0
        a = eng.Engine(m, precision, reorder=True).run()
        for f in ("u", "v", "sigma", "qia"):
            assert np.array_equal(b.get(i, f), a.get(f)), (i, f)
        sa = a.scalars()
        for k in ("nsteps", "time", "dt_i", "ke", "ie", "hg", "vis", "wext"):
            assert s[k] == sa[k], (i, k)
    print(f"[batch] a {big.n_elems}-element job and {len(jobs) - 1} small ones equal alone and in the mix",
          flush=True)


def test_mixed_batch_of_staged_and_gathered_jobs_alone_equals_in_mix():
    # without renumbering, a job with shuffled elements has element blocks
    # whose node lists overflow the LDS budget and gathers its corners from
    # HBM, while the others stage theirs: each job's kernels and arithmetic
    # are chosen from the job alone
    import dataclasses
    jobs = small_jobs(4, seed=11, plastic_p=0.5)
    rng = np.random.default_rng(12)
    m = decks.random_job(rng, 3000, 4000, 0.0, k_lo=150, k_hi=300, name="shuffled")
    m = dataclasses.replace(m, conn=np.ascontiguousarray(m.conn[rng.permutation(m.n_elems)]))
    jobs.append(m)
    b = Batch(jobs, eng.MIXED, reorder=False)
    b.run()
    staged = []
    for i, mi in enumerate(jobs):
        a = eng.Engine(mi, eng.MIXED, reorder=False)
        a.init()
        a.set_current_step(1)
        a.advance(3)
        lay = a.layout()
        print(f"[mix] job {i}: {lay}", flush=True)
        staged.append(lay["staged"])
        a.advance(0)
        s = b.scalars(i)
        assert s["done"] == 1.0 and s["err"] == 0.0
        for f in ("u", "v", "sigma", "qia"):
            assert np.array_equal(b.get(i, f), a.get(f)), (i, f)
    assert staged[-1] == 0 and any(staged[:-1]), staged


def test_batch_fp64_matches_oracle():
    jobs = small_jobs(4, seed=5, plastic_p=0.5)
    b = Batch(jobs, eng.FP64).run()
    for i, m in enumerate(jobs):
        o = Oracle(m, 0).init()
        o.run(0, [], 0)
        for f in ("u", "sigma"):
            ref = o.get(f)
            err = np.abs(b.get(i, f) - ref).max() / max(np.abs(ref).max(), 1e-30)
            assert err <= 1e-10, (i, f, err)


def test_graph_replay_matches_eager():
    # Both element passes run in the captured step (the thin slab's blocks
    # take the staged pass on the side stream beside the pipelined one; the
    # ramped-BC job's constrained tiles override their staged velocities).
    # The graph is replayed many times, with host reads and syncs between
    # the replays, and every checkpoint equals the eager run.
    jobs = small_jobs(6, seed=9)
    jobs.append(decks.config2(nx=12, ny=6, nz=6, steps=300))
    jobs.append(thin_slab())
    a = Batch(jobs, eng.MIXED, graph_ticks=0).init()
    g = Batch(jobs, eng.MIXED, graph_ticks=16).init()
    a.set_current_step(1)
    g.set_current_step(1)
    for chunk in range(6):
        a.advance(64)
        g.advance(64)
        for i in range(len(jobs)):
            for f in ("u", "v", "sigma"):
                assert np.array_equal(a.get(i, f), g.get(i, f)), (chunk, i, f)
            assert a.scalars(i)["nsteps"] == g.scalars(i)["nsteps"], (chunk, i)
    a.advance(0)
    g.advance(0)
    for i in range(len(jobs)):
        assert np.array_equal(a.get(i, "u"), g.get(i, "u")), i
        assert g.scalars(i)["done"] == 1.0


def test_server_alone_equals_in_mix_and_latency():
    jobs = small_jobs(8, seed=21, plastic_p=0.5)
    srv = Server(eng.MIXED, budget_bytes=64 << 20, max_live=4, graph_ticks=8)
    tickets = []
    done = []
    # staggered arrivals: submit two jobs, pump, submit more while they run
    for i, m in enumerate(jobs):
        tickets.append(srv.submit(m, keep=True))
        if i % 2 == 1:
            done += srv.pump(24)
    guard = 0
    while len(done) < len(jobs):
        done += srv.pump(64)
        guard += 1
        assert guard < 10000
    assert sorted(done) == sorted(tickets)
    q, live, used = srv.status()
    assert q == 0 and live == 0 and used == 0
    for t, m in zip(tickets, jobs):
        s, u, times = srv.result(t, with_u=True)
        assert times[0] <= times[1] <= times[2]
        assert s["time"] == pytest.approx(m.total_time, rel=1e-12)
        a = eng.Engine(m, eng.MIXED, reorder=True).run()
        assert np.array_equal(u, a.get("u")), t
        assert s["nsteps"] == a.scalars()["nsteps"]


def test_server_reports_a_job_that_cannot_be_set_up():
    jobs = small_jobs(3, seed=31)
    bad = jobs[1]
    bad.conn = np.array(bad.conn, copy=True)
    bad.conn[0] = bad.conn[0][[4, 5, 6, 7, 0, 1, 2, 3]]  # inverted element
    srv = Server(eng.MIXED, budget_bytes=64 << 20, max_live=4, graph_ticks=8)
    tickets = [srv.submit(m, keep=False) for m in jobs]
    done = []
    guard = 0
    while len(done) < len(jobs):
        done += srv.pump(64)
        guard += 1
        assert guard < 10000
    assert sorted(done) == sorted(tickets)
    with pytest.raises(eng.EnscError, match="set up"):
        srv.result(tickets[1])
    for t in (tickets[0], tickets[2]):
        s, _, _ = srv.result(t)
        assert s["done"] == 1.0 and s["err"] == 0.0


@pytest.mark.parametrize("precision", [eng.MIXED, eng.FP64])
def test_autotune_leaves_the_solution_unchanged(precision):
    # ensc_engine_autotune runs trial steps from the current state and puts
    # it back; its candidates change launch shapes only.  A batch tuned after
    # step 1 is set must finish with the same fields as an untuned one, bit
    # for bit; kinetic energy and external work may differ in their last
    # bits (the nodal update's block size regroups their block sums).
    jobs = small_jobs(6, seed=31, plastic_p=0.5)
    jobs.append(decks.config2(nx=12, ny=6, nz=6, steps=300))  # amplitude BC
    jobs.append(thin_slab())  # staged pass
    ref = Batch(jobs, precision, reorder=True).run()
    b = Batch(jobs, precision, reorder=True, graph_ticks=4)
    b.init()
    b.set_current_step(1)
    cfg = b.autotune(6)
    print(f"[autotune] chose {cfg}", flush=True)
    assert cfg["waves"] in (1, 2, 3) and cfg["block_node"] in (64, 128, 256) and cfg["pipe_per_cu"] in (1, 2)
    assert cfg["ms_per_tick"] > 0.0
    b.advance(0)
    for si in range(2, max(len(m.steps) for m in jobs)):
        b.set_current_step(si)
        b.advance(0)
    for i in range(len(jobs)):
        s, sr = b.scalars(i), ref.scalars(i)
        assert s["done"] == 1.0 and s["nsteps"] == sr["nsteps"], i
        for f in ("u", "v", "sigma", "qia", "epsp"):
            assert np.array_equal(b.get(i, f), ref.get(i, f)), (i, f)
        for k in ("ie", "hg", "vis", "time"):
            assert s[k] == sr[k], (i, k)
        for k in ("ke", "wext"):
            assert s[k] == pytest.approx(sr[k], rel=1e-12, abs=1e-300), (i, k)


def test_mixed_batch_within_tolerance_of_fp64_batch():
    # every job of a mixed batch (uneven sizes, half J2) against the same job
    # in an FP64 batch: displacement within 1e-4 relative, total energy
    # within 1% of the external work (or of the initial kinetic energy when
    # nothing does work on the job)
    jobs = small_jobs(8, seed=41, plastic_p=0.5)
    a = Batch(jobs, eng.FP64, reorder=True).run()
    b = Batch(jobs, eng.MIXED, reorder=True).run()
    worst_u, worst_e = 0.0, 0.0
    for i in range(len(jobs)):
        sa, sb = a.scalars(i), b.scalars(i)
        assert sa["done"] == 1.0 and sb["done"] == 1.0 and abs(sa["nsteps"] - sb["nsteps"]) <= 1, i
        ua, ub = a.get(i, "u"), b.get(i, "u")
        du = float(np.abs(ub - ua).max() / max(np.abs(ua).max(), 1e-30))
        ea = sa["ke"] + sa["ie"] + sa["hg"] + sa["vis"]
        eb = sb["ke"] + sb["ie"] + sb["hg"] + sb["vis"]
        scale = max(abs(sa["wext"]), sa["ke0"], abs(ea))
        de = abs(eb - ea) / scale
        worst_u, worst_e = max(worst_u, du), max(worst_e, de)
        assert du <= 1e-4, (i, du)
        assert de <= 0.01, (i, de)
    print(f"[mixed batch vs fp64 batch] {len(jobs)} jobs: u {worst_u:.2e}, energy {worst_e:.2e}", flush=True)


def serve_all(srv, jobs, keep=True, first=2, ticks=48):
    """Submit jobs two at a time with pumps between, so later jobs arrive
    while earlier ones run; pump until every ticket is retired."""
    tickets, done = [], []
    for i, m in enumerate(jobs):
        tickets.append(srv.submit(m, keep=keep))
        if i + 1 >= first and i % 2 == 1:
            done += srv.pump(ticks)
    guard = 0
    while len(done) < len(jobs):
        done += srv.pump(64)
        guard += 1
        assert guard < 100000
    assert sorted(done) == sorted(tickets)
    return tickets


def test_server_fp64_matches_oracle():
    # FP64 serving with staggered arrivals: every retired job equals the CPU
    # oracle (the restatement of the reference's step loop) within 1e-10
    jobs = small_jobs(8, seed=51, plastic_p=0.5)
    srv = Server(eng.FP64, budget_bytes=48 << 20, max_live=3, graph_ticks=8)
    tickets = serve_all(srv, jobs)
    for t, m in zip(tickets, jobs):
        s, u, times = srv.result(t, with_u=True)
        o = Oracle(m, 0).init()
        o.run(0, [], 0)
        so = o.scalars()
        err = float(np.abs(u - o.get("u")).max() / max(np.abs(o.get("u")).max(), 1e-30))
        assert err <= 1e-10, (t, err)
        assert s["nsteps"] == so["nsteps"] and s["time"] == so["time"]
        scale = max(abs(so["wext"]), so["ke0"])
        for k in ("ke", "ie", "hg", "vis", "wext"):
            assert abs(s[k] - so[k]) <= 1e-9 * scale, (t, k)
    print(f"[server fp64] {len(jobs)} staggered jobs equal the oracle", flush=True)


def test_server_runs_every_step_of_a_multi_step_deck():
    # test_step (a reference deck: 3 *Steps, op=NEW, gravity) served among
    # other jobs: the server switches it from step to step like
    # exDyna3D::run, and the result equals Engine.run() on the same deck
    d = G.load("ref_test_step")
    deck = G.model_from_dump(d)
    assert len(deck.steps) > 2
    jobs = small_jobs(3, seed=61)
    jobs.insert(1, deck)
    srv = Server(eng.FP64, budget_bytes=64 << 20, max_live=4, graph_ticks=8)
    tickets = serve_all(srv, jobs)
    s, u, _ = srv.result(tickets[1], with_u=True)
    a = eng.Engine(deck, eng.FP64, reorder=True).run()
    sa = a.scalars()
    assert s["nsteps"] == sa["nsteps"] == int(d["n_steps"])
    assert s["step_index"] == len(deck.steps) - 1
    assert np.array_equal(u, a.get("u"))
    assert float(np.abs(u - d["final_u"]).max() / np.abs(d["final_u"]).max()) <= 1e-10


def test_server_mixed_config5_jobs_vs_fp64():
    # requests from the configuration-5 generator (sizes log-uniform up to
    # 5e5 elements, half J2) served in mixed precision, each against the
    # same job on the FP64 engine: displacement within 1e-4, total energy
    # within 1% of the energy scale (initial kinetic energy: nothing does
    # external work on these jobs)
    jobs = decks.config5(n_jobs=16)
    print(f"[c5 server] {len(jobs)} jobs, {sum(m.n_elems for m in jobs)} elements, "
          f"largest {max(m.n_elems for m in jobs)}, J2 {sum(m.plastic for m in jobs)}", flush=True)
    srv = Server(eng.MIXED, budget_bytes=8 << 30, max_live=8, graph_ticks=20)
    tickets = serve_all(srv, jobs, ticks=100)
    worst_u, worst_e = 0.0, 0.0
    for t, m in zip(tickets, jobs):
        s, u, _ = srv.result(t, with_u=True)
        a = eng.Engine(m, eng.FP64, reorder=True).run()
        sa = a.scalars()
        assert s["done"] == 1.0 and s["err"] == 0.0 and abs(s["nsteps"] - sa["nsteps"]) <= 1
        ua = a.get("u")
        du = float(np.abs(u - ua).max() / max(np.abs(ua).max(), 1e-30))
        ea = sa["ke"] + sa["ie"] + sa["hg"] + sa["vis"] - sa["wext"]
        eb = s["ke"] + s["ie"] + s["hg"] + s["vis"] - s["wext"]
        de = abs(eb - ea) / max(abs(sa["wext"]), sa["ke0"])
        worst_u, worst_e = max(worst_u, du), max(worst_e, de)
        assert du <= 1e-4, (t, du)
        assert de <= 0.01, (t, de)
    print(f"[c5 server mixed vs fp64] u {worst_u:.2e}, energy {worst_e:.2e}", flush=True)


def test_mixed_engine_runs_quasi_static_jobs_in_fp64():
    # per-job precision: in a mixed batch, the two quasi-static reference
    # decks (thousands of wave transits) run the FP64 kernels and equal an
    # FP64 engine with the same element blocks (256 elements, the mixed
    # engine's) bit for bit, while the other jobs run mixed and equal
    # themselves alone
    quasi = [G.model_from_dump(G.load(c)) for c in ("ref_test_part1", "ref_Job_6")]
    jobs = small_jobs(4, seed=71, plastic_p=0.5)
    jobs = [jobs[0], quasi[0], jobs[1], jobs[2], quasi[1], jobs[3]]
    b = Batch(jobs, eng.MIXED, reorder=True).run()
    for i, m in enumerate(jobs):
        s = b.scalars(i)
        assert s["done"] == 1.0 and s["err"] == 0.0, i
        if any(m is q for q in quasi):
            a = eng.Engine(m, eng.FP64, reorder=True, block_elem=256).run()
            assert a.layout()["fp64_jobs"] == 1
        else:
            a = eng.Engine(m, eng.MIXED, reorder=True).run()
            assert a.layout()["fp64_jobs"] == 0
        for f in ("u", "v", "sigma", "qia"):
            assert np.array_equal(b.get(i, f), a.get(f)), (i, f)
        assert s["nsteps"] == a.scalars()["nsteps"], i
    forced = eng.Engine(quasi[0], eng.MIXED, reorder=True, fp64_jobs=1)
    forced.run_steps(3)
    assert forced.layout()["fp64_jobs"] == 0
    print(f"[per-job precision] {len(jobs)} jobs, 2 quasi-static in FP64", flush=True)
