#!/usr/bin/env python3
"""Benchmark of the MI355X step engine.

Headline (configuration 3): the 1M-element plastic cube (100^3 C3D8R, J2
yield 250 MPa, hardening 1 GPa, z=0 fixed, z=1 pushed at 10 m/s), inputs
resident in HBM.  One "step" = one iteration of exDyna3D's step loop
(element pass + nodal update + dt reduction) over the whole mesh.  The
headline value is the mixed-precision element-steps/s; the FP64 engine is
timed in the same run on the same steps, and the relative displacement error
of mixed vs FP64 after those steps is reported beside it.

Also in the same JSON line:
  * batched  (configuration 4): 256 independent meshes in one engine, each
    run to its own end time; solves/s and element-steps/s.
  * serving  (configuration 5): continuous-batching server fed with
    requests drawn from the config-4 generator (up to 5e5 elements, half
    J2); capacity (all submitted at once), Poisson arrivals at 80% of that
    capacity with p50/p99 latency, and the same jobs run one at a time.
  * cpu_baseline: the reference solver's own step loop (oracle/_ref/ref_timer,
    built from the reference sources) timed on the host's cores on a bounded
    sample of configuration 3, with the FP64 restatement (oracle/) beside it
    as cpu_restatement; the restatement alone where the reference build is
    absent.

    python bench.py [--gpus 1] [--steps 200] [--warmup 20] [--sections c3,out,c4,c5]
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from enscdynamics_amd import decks  # noqa: E402
from enscdynamics_amd import engine as eng  # noqa: E402
from enscdynamics_amd.batch import Batch, Server  # noqa: E402

METRIC = "elem-steps/s mixed vs FP64 at ≤1e-4 rel disp err; batched solves/s + p99"
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pmc_traffic_c3_cube_1M.json")


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def survey_bytes(precision: int, plastic: bool, r: float):
    """Algorithmic bytes per element-step as the hot-path contract fixes
    them: the once-per-step read and write of the persistent state at the
    stored precision, element part and node part.
      FP64:  conn 32 + B0 192 + V0 8 read; sigma 48 + QIA 96 read and
             written (+ plastic strain 8 read and written) = 520 (+16);
             nodes: x, v, u read and written, 1/m read = 152 per node.
      mixed: conn 32 + B0 96 + V0 4 read; sigma 24 + QIA 48 read and
             written (+ plastic strain 4 read and written) = 276 (+8);
             nodes: x, v, u FP64 read and written, 1/m, a compensation
             residual for x (24) read and written = 200 per node.
    Returns (element part, node part) per element (r = nodes / elements)."""
    if precision == eng.MIXED:
        return 276.0 + (8.0 if plastic else 0.0), 200.0 * r
    return 520.0 + (16.0 if plastic else 0.0), 152.0 * r


def alg_bytes(precision: int, plastic: bool, r: float):
    """Algorithmic bytes per element-step: the compulsory once-per-step read
    and write of persistent state as this engine stores it (force exchange,
    staging redundancy and block partials are overhead and not counted).
    Returns (element pass, nodal update) per element."""
    mixed = precision == eng.MIXED
    s = 4 if mixed else 8
    # 8 corner ids (16-bit block-local), Gi (9) and V0 (FP64 engine; the
    # mixed body rebuilds them from X), sigma r+w, QIA r+w, eps_p r+w
    elem = 16 + (0 if mixed else 10 * s) + 12 * s + 24 * s + (2 * s if plastic else 0)
    # node values each element pass reads once: X, v, u
    elem += 72 * r
    if precision == eng.MIXED:
        # v, u read+write, invM, compensation of u (one packed word) read+write
        node = (48 + 48 + 8 + 8) * r
    else:
        # x, v, u read+write, invM
        node = (144 + 8) * r
    return elem, node


# The engines of the configuration-3 leg: mixed precision; FP64 (the
# Walsh-form FP64 element body with block partials, the default); FP64 in the
# reference's operation and summation order (fp64_order=1), which reproduces
# the reference bit for bit up to libm rounding.
C3_ENGINES = (("mixed", eng.MIXED, 0), ("fp64", eng.FP64, 0), ("fp64_reference_order", eng.FP64, 1))


def c3_section(args, m):
    r = m.n_nodes / m.n_elems
    res = {}
    for name, prec, order in C3_ENGINES:
        e = eng.Engine(m, prec, reorder=True, staging=not args.no_staging, fp64_order=order)
        if args.graph:
            e.set_graph(args.graph)
        e.init()
        e.set_current_step(1)
        tuned = e.autotune(args.c3_autotune) if args.c3_autotune else None
        if tuned:
            log(f"{name}: autotuned {tuned}")
        e.enqueue(args.warmup)
        e.sync()
        n0 = int(e.scalars()["nsteps"])
        t0 = time.perf_counter()
        e.enqueue(args.steps)
        e.sync()
        t1 = time.perf_counter()
        n1 = int(e.scalars()["nsteps"])
        assert n1 - n0 == args.steps, f"only {n1 - n0} of {args.steps} ticks were full steps"
        # per-launch durations, events on the engine's stream around each launch
        e.set_timing(True)
        e.enqueue(args.kernel_steps)
        e.sync()
        ph = e.phase_times()
        e.set_timing(False)
        res[name] = dict(engine=e, secs=t1 - t0, phases=ph, layout=e.layout(), tuned=tuned)
        log(f"{name}: {args.steps} steps in {t1 - t0:.4f} s; per launch (ms): element {ph[0]:.4f}, "
            f"element finalize {ph[1]:.4f}, node {ph[2]:.4f}, node finalize {ph[3]:.4f}")
    u_mix = res["mixed"]["engine"].get("u")
    u_64 = res["fp64"]["engine"].get("u")
    rel_err = float(np.abs(u_mix - u_64).max() / max(np.abs(u_64).max(), 1e-300))
    out = {}
    for name, prec, _ in C3_ENGINES:
        d = res[name]
        be, bn = alg_bytes(prec, m.plastic, r)
        se, sn = survey_bytes(prec, m.plastic, r)
        ph = d["phases"]
        out[name] = {
            "value": m.n_elems * args.steps / d["secs"],
            "ms_per_step": 1e3 * d["secs"] / args.steps,
            "launch_ms": {"element": ph[0], "element_finalize": ph[1], "node": ph[2], "node_finalize": ph[3]},
            # contract bytes (state at the stored precision) and this engine's
            # own stored bytes (no stored B0/V0 in mixed; node reads of the
            # element pass counted there)
            "contract_bytes_per_elem": {"element": se, "node": sn, "step": se + sn},
            "alg_bytes_per_elem": {"element": be, "node": bn},
            "element_gbs": se * m.n_elems / (ph[0] * 1e-3) / 1e9,
            "node_gbs": sn * m.n_elems / (ph[2] * 1e-3) / 1e9,
            "step_gbs": (se + sn) * m.n_elems * args.steps / d["secs"] / 1e9,
            "engine_bytes_gbs": {"element": be * m.n_elems / (ph[0] * 1e-3) / 1e9,
                                 "node": bn * m.n_elems / (ph[2] * 1e-3) / 1e9},
            "layout": d["layout"],
            "autotuned": d["tuned"],
        }
        d["engine"].close()
    out["rel_disp_err_mixed_vs_fp64"] = rel_err
    return out


def out_section(args, m, ms_per_step):
    """Cost of output_results on the device engine (mixed, configuration 3):
    snapshot capture inside the step, post-processing (von Mises, stress in
    VTK order, Green-Lagrange strain) with the copy to the host, and writing
    one .vtu file."""
    import tempfile

    res = {}
    e = eng.Engine(m, eng.MIXED, reorder=True)
    if args.graph:
        e.set_graph(args.graph)
    e.init()
    e.set_current_step(1)
    e.enqueue(args.warmup)
    e.sync()
    dt = e.scalars()["dt_i"]
    e.set_output(10.5 * dt, slots=4)  # a snapshot every 10-11 steps
    e.enqueue(2)  # the layout and graph pick up the output ring; one snapshot
    e.sync()
    while e.output_pending()[0]:
        e.output_pop()
    n = 40
    t0 = time.perf_counter()
    e.enqueue(n)
    e.sync()
    t_out = time.perf_counter() - t0
    k, _ = e.output_pending()
    res["steps"] = n
    res["snapshots"] = k
    res["ms_per_step_with_output"] = 1e3 * t_out / n
    res["capture_ms_per_snapshot"] = (1e3 * t_out - n * ms_per_step) / max(k, 1)
    t0 = time.perf_counter()
    for f in ("vonMises", "Stress", "Green-Lagrange Strain", "x", "a", "v", "u"):
        e.output_field(f)
    res["fields_to_host_ms"] = 1e3 * (time.perf_counter() - t0)
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "solution_0.vtu")
        t0 = time.perf_counter()
        e.output_write(p)
        res["vtu_write_ms"] = 1e3 * (time.perf_counter() - t0)
        res["vtu_bytes"] = os.path.getsize(p)
    e.close()
    log(f"out: {k} snapshots in {n} steps, capture {res['capture_ms_per_snapshot']:.3f} ms each, "
        f"fields to host {res['fields_to_host_ms']:.1f} ms, .vtu {res['vtu_bytes'] / 1e6:.0f} MB in "
        f"{res['vtu_write_ms']:.0f} ms")
    return res


def traffic_entry(kernel_key: str):
    """HBM bytes per launch of a kernel from the committed counter summary
    (rocprofv3 FETCH_SIZE and WRITE_SIZE passes, scaled by the same counters
    on k_copy_probe), or None when the summary is absent."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    k = t.get("kernels", {}).get(kernel_key)
    return None if k is None else k.get("hbm_bytes_per_launch")


def c4_section(args):
    models = decks.config4(args.c4_jobs)
    ne = np.array([m.n_elems for m in models], dtype=np.float64)
    log(f"c4: {len(models)} meshes, {int(ne.sum())} elements")
    t0 = time.perf_counter()
    b = Batch(models, eng.MIXED, reorder=True, graph_ticks=args.graph)
    t1 = time.perf_counter()
    b.run()
    t2 = time.perf_counter()
    steps = np.array([b.scalars(i)["nsteps"] for i in range(len(models))])
    errs = [b.scalars(i)["err"] for i in range(len(models))]
    assert all(x == 0.0 for x in errs), "a batched job raised an error flag"
    b.close()
    secs = t2 - t1
    log(f"c4: built in {t1 - t0:.2f} s, ran in {secs:.3f} s")
    out = {"workload": f"config 4: {len(models)} hex-block meshes, Ne log-uniform [1e3, 5e4], "
                       f"{int(ne.sum())} elements, own end times (2000-8000 CFL steps), elastic, mixed",
           "jobs": len(models), "solves_per_s": len(models) / secs, "seconds": secs,
           "elem_steps_per_s": float((ne * steps).sum() / secs), "build_seconds": t1 - t0,
           "graph_ticks": args.graph}
    if args.c4_autotune:
        # the same batch with its launch configuration autotuned after step 1
        # is set (the tuning's trial steps are undone; its time is reported)
        b = Batch(models, eng.MIXED, reorder=True, graph_ticks=args.graph)
        t3 = time.perf_counter()
        b.init()
        b.set_current_step(1)
        t4 = time.perf_counter()
        cfg = b.autotune(args.c4_autotune)
        t5 = time.perf_counter()
        b.advance(0)
        t6 = time.perf_counter()
        steps_t = np.array([b.scalars(i)["nsteps"] for i in range(len(models))])
        assert np.array_equal(steps_t, steps), "the tuned batch took other step counts"
        b.close()
        run_t = (t4 - t3) + (t6 - t5)
        log(f"c4 autotuned: {cfg}; tuning {t5 - t4:.3f} s, ran in {run_t:.3f} s")
        out["autotuned"] = {"config": cfg, "tune_seconds": t5 - t4, "seconds": run_t,
                            "solves_per_s": len(models) / run_t}
    return out


def _pump_until(srv, pending, t0, done_set, ticks):
    """Submit arrivals whose time has come, pump, collect completions."""
    now = time.perf_counter() - t0
    while pending and pending[0][0] <= now:
        ta, m, idx = pending.pop(0)
        tk = srv.submit(m, keep=False, t_arrival=srv.clock() - (now - ta))
        done_set[tk] = idx
    return srv.pump(ticks)


def c5_section(args):
    models = decks.config5(args.c5_jobs)
    ne = np.array([m.n_elems for m in models], dtype=np.float64)
    log(f"c5: {len(models)} requests, {int(ne.sum())} elements, largest {int(ne.max())}")
    budget = 64 << 30

    def serve(arrivals):
        srv = Server(eng.MIXED, budget_bytes=budget, max_live=args.c5_live, graph_ticks=args.graph)
        pending = [(a, m, i) for i, (a, m) in enumerate(zip(arrivals, models))]
        pending.sort(key=lambda x: x[0])
        tick2idx = {}
        finished = []
        t0 = time.perf_counter()
        guard = 0
        while len(finished) < len(models):
            q, live, _ = srv.status()
            if not live and not q and pending:
                wait = pending[0][0] - (time.perf_counter() - t0)
                if wait > 0:
                    time.sleep(wait)
            finished += _pump_until(srv, pending, t0, tick2idx, args.c5_ticks)
            guard += 1
            assert guard < 10_000_000
        t1 = time.perf_counter()
        lat = []
        for tk in finished:
            s, _, times = srv.result(tk)
            assert s["err"] == 0.0 and s["done"] == 1.0
            lat.append(times[2] - times[0])
        srv.close()
        return t1 - t0, np.array(lat)

    secs_sat, lat_sat = serve([0.0] * len(models))
    cap = len(models) / secs_sat
    log(f"c5: saturated {secs_sat:.3f} s ({cap:.2f} solves/s)")
    # the same jobs as one batch, all resident from the start (no admission,
    # no pumps): the device-side bound of the saturated server
    t0 = time.perf_counter()
    b = Batch(models, eng.MIXED, reorder=True, graph_ticks=args.graph)
    t1 = time.perf_counter()
    b.run()
    secs_batch = time.perf_counter() - t1
    b.close()
    log(f"c5: as one batch {secs_batch:.3f} s (built in {t1 - t0:.2f} s)")
    rate = 0.8 * cap
    rng = np.random.default_rng(55)
    arr = np.cumsum(rng.exponential(1.0 / rate, size=len(models)))
    arr -= arr[0]
    secs_p, lat_p = serve(list(arr))
    log(f"c5: Poisson {rate:.2f}/s run {secs_p:.3f} s, p50 {np.percentile(lat_p, 50):.3f} s, "
        f"p99 {np.percentile(lat_p, 99):.3f} s")
    # the same jobs one at a time, one engine each: (a) build, run and read
    # the results inside the clock, as a caller without a server would;
    # (b) engines built before the clock (in groups, to bound memory), so the
    # ratio compares device work alone
    t0 = time.perf_counter()
    lat_seq = []
    for i, m in enumerate(models):
        a = time.perf_counter()
        e = eng.Engine(m, eng.MIXED, reorder=True)
        if args.graph:
            e.set_graph(args.graph)
        e.run()
        s = e.scalars()
        assert s["err"] == 0.0
        e.close()
        lat_seq.append(time.perf_counter() - a)
        if (i + 1) % 200 == 0 and i + 1 < len(models):
            log(f"c5: one at a time (build inside the clock): {i + 1} of {len(models)} done")
    secs_seq = time.perf_counter() - t0
    log(f"c5: one at a time {secs_seq:.3f} s (engine build inside the clock)")
    secs_pre, lat_pre = 0.0, []
    for g0 in range(0, len(models), 40):
        group = []
        for m in models[g0:g0 + 40]:
            e = eng.Engine(m, eng.MIXED, reorder=True)
            if args.graph:
                e.set_graph(args.graph)
            group.append(e)
        for e in group:
            a = time.perf_counter()
            e.run()
            s = e.scalars()
            assert s["err"] == 0.0
            dt = time.perf_counter() - a
            secs_pre += dt
            lat_pre.append(dt)
        for e in group:
            e.close()
        if (g0 + 40) % 200 == 0 and g0 + 40 < len(models):
            log(f"c5: one at a time (prebuilt): {g0 + 40} of {len(models)} done")
    log(f"c5: one at a time {secs_pre:.3f} s (engines built before the clock)")
    return {"workload": f"config 5: {len(models)} requests from the config-4 generator, Ne log-uniform "
                        f"[1e3, 5e5], half J2, mixed, admitted by HBM footprint (budget {budget >> 30} GiB, "
                        f"<= {args.c5_live} live)",
            "jobs": len(models),
            "jobs_note": f"{len(models)} completions per leg keep the whole bench within a few minutes "
                         f"(p99 of {len(models)} latencies); --c5-jobs 1000 runs the full stream",
            "capacity_solves_per_s": cap,
            "one_batch_solves_per_s": len(models) / secs_batch,
            "poisson_rate_per_s": rate,
            "poisson_solves_per_s": len(models) / secs_p,
            "latency_s": {"p50": float(np.percentile(lat_p, 50)), "p99": float(np.percentile(lat_p, 99)),
                          "saturated_p50": float(np.percentile(lat_sat, 50)),
                          "saturated_p99": float(np.percentile(lat_sat, 99))},
            "one_at_a_time_solves_per_s": len(models) / secs_seq,
            "one_at_a_time_latency_s": {"p50": float(np.percentile(lat_seq, 50)),
                                        "p99": float(np.percentile(lat_seq, 99))},
            "one_at_a_time_prebuilt_solves_per_s": len(models) / secs_pre,
            "one_at_a_time_prebuilt_latency_s": {"p50": float(np.percentile(lat_pre, 50)),
                                                 "p99": float(np.percentile(lat_pre, 99))},
            "speedup_vs_one_at_a_time": cap / (len(models) / secs_seq),
            "speedup_vs_one_at_a_time_prebuilt": cap / (len(models) / secs_pre),
            "ticks_per_pump": args.c5_ticks}


def gpu_rate(m, precision, steps, warm, graph):
    """elem-steps/s of one job on the GPU: `steps` full step iterations
    timed after `warm` (steps <= 0: the whole run, setup excluded)."""
    e = eng.Engine(m, precision, reorder=True)
    if graph:
        e.set_graph(graph)
    e.init()
    e.set_current_step(1)
    if steps <= 0:
        t0 = time.perf_counter()
        e.advance(0)
        secs = time.perf_counter() - t0
        n = int(e.scalars()["nsteps"])
    else:
        e.enqueue(warm)
        e.sync()
        n0 = int(e.scalars()["nsteps"])
        t0 = time.perf_counter()
        e.enqueue(steps)
        e.sync()
        secs = time.perf_counter() - t0
        n = int(e.scalars()["nsteps"]) - n0
    e.close()
    return m.n_elems * n / secs, n


def configs_section(args):
    """Configurations 1, 2 and 4 beside the headline: GPU element-steps/s
    (FP64 and mixed) and CPU baselines on bounded samples."""
    from oracle.oracle import Oracle  # CPU baseline leg only

    out = {}
    c1 = decks.config1()
    c2 = decks.config2()
    for name, m, steps in (("c1", c1, 0), ("c2", c2, 200)):
        r = {}
        for prec, pn in ((eng.FP64, "fp64"), (eng.MIXED, "mixed")):
            v, n = gpu_rate(m, prec, steps, 20, args.graph)
            r[f"gpu_{pn}_elem_steps_per_s"] = v
            r["gpu_steps"] = n
        log(f"{name}: gpu fp64 {r['gpu_fp64_elem_steps_per_s']:.3e}, mixed {r['gpu_mixed_elem_steps_per_s']:.3e} "
            f"elem-steps/s ({r['gpu_steps']} steps)")
        out[name] = r
    if not args.no_cpu:
        threads = os.cpu_count() or 1
        # C1: all 2000 steps; C2: 10 warm-up + 100 timed steps
        for name, m, steps, warm, ref_steps in (("c1", c1, 2000, 0, 2000), ("c2", c2, 100, 10, 10)):
            o = Oracle(m, threads).init()
            if warm:
                o.time_steps(warm)
            secs = o.time_steps(steps)
            out[name]["cpu_restatement"] = {"value": m.n_elems * steps / secs, "cores": o.threads,
                                            "sample": f"{warm} warm-up + {steps} timed steps"}
            if os.path.exists(REF_TIMER):
                try:
                    out[name]["cpu_reference"] = cpu_reference(m, ref_steps, warm=2)
                except (RuntimeError, OSError, subprocess.SubprocessError) as ex:
                    log(f"{name}: cpu reference not timed: {ex}")
            log(f"{name}: cpu restatement {out[name]['cpu_restatement']['value']:.3e} elem-steps/s")
        # C4: a fixed 100 steps per mesh over all 256 meshes, the
        # restatement (one reference process per mesh would time its deck
        # parsing instead)
        models = decks.config4(args.c4_jobs)
        tot, secs = 0, 0.0
        for m in models:
            o = Oracle(m, threads).init()
            secs += o.time_steps(100)
            tot += m.n_elems * 100
        out["c4"] = {"cpu_restatement": {"value": tot / secs, "cores": threads,
                                         "sample": f"100 steps of each of the {len(models)} meshes"}}
        log(f"c4: cpu restatement {tot / secs:.3e} elem-steps/s")
    return out


REF_TIMER = os.path.join(ROOT, "oracle", "_ref", "ref_timer")


def _cpu_model() -> str:
    cpu = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return cpu


def cpu_port(m, steps=100, warm=10):
    """The FP64 restatement (oracle/restate.cpp) on every host core."""
    from oracle.oracle import Oracle  # CPU baseline leg only

    threads = os.cpu_count() or 1
    o = Oracle(m, threads).init()
    o.time_steps(warm)
    secs = o.time_steps(steps)
    return {"value": m.n_elems * steps / secs, "unit": "elem-steps/s", "cores": o.threads, "kind": "port",
            "sample": f"{m.name}: {warm} warm-up + {steps} timed steps of the FP64 restatement, "
                      f"{o.threads} OpenMP threads on {_cpu_model()}"}


def cpu_reference(m, steps=20, warm=2):
    """The reference solver itself (oracle/_ref/ref_timer, built from its
    sources): its reader parses the deck of the same mesh, init() runs, and
    `steps` iterations of its step loop are timed after `warm` (dt update,
    computeSate with output_results, its VTK writer stubbed)."""
    import tempfile

    with tempfile.TemporaryDirectory() as d:
        decks.write_inp(m, os.path.join(d, "deck.inp"))
        r = subprocess.run([REF_TIMER, d, "deck.inp", str(warm), str(steps), repr(m.c_dr), repr(m.c_cr),
                            repr(m.factor)], capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        raise RuntimeError(f"ref_timer failed: {r.stderr[-500:]}")
    f = r.stdout.split()
    kv = {f[i]: f[i + 1] for i in range(len(f) - 1) if f[i] in ("elements", "steps", "seconds", "threads")}
    ne, n, secs, thr = int(kv["elements"]), int(kv["steps"]), float(kv["seconds"]), int(kv["threads"])
    return {"value": ne * n / secs, "unit": "elem-steps/s", "cores": thr, "kind": "reference",
            "sample": f"{m.name}: the reference's own reader, init and step loop (oracle/_ref/ref_timer), "
                      f"{warm} warm-up + {n} timed steps ({secs:.2f} s), {thr} OpenMP threads on {_cpu_model()}"}


def cpu_baseline(m, args):
    """The reference's CPU step loop when its build travelled with the tree,
    with the restatement beside it; else the restatement alone."""
    port = cpu_port(m, args.cpu_steps)
    log(f"cpu restatement: {port['value']:.3e} elem-steps/s")
    if os.path.exists(REF_TIMER):
        try:
            ref = cpu_reference(m, args.cpu_ref_steps)
            log(f"cpu reference: {ref['value']:.3e} elem-steps/s")
            return ref, port
        except (RuntimeError, OSError, subprocess.SubprocessError) as ex:
            log(f"cpu reference not timed: {ex}")
    return port, None


# (substrings of the demangled kernel name, key); the first row whose
# substrings all occur names the kernel
KERNEL_KEYS = ((("k_element_pipe",), "k_element_pipe_mixed"), (("k_element<true",), "k_element_mixed"),
               (("k_element<false", ", true>("), "k_element_fp64"),
               (("k_element<false",), "k_element_fp64_reference_order"),
               (("k_node<true, 256, true>",), "k_node_mixed_partials"), (("k_node<true",), "k_node_mixed"),
               (("k_node<false, 256, true>",), "k_node_fp64"), (("k_node<false",), "k_node_fp64_reference_order"),
               (("k_fin_element",), "k_fin_element"), (("k_fin_node",), "k_fin_node"),
               (("k_copy_probe",), "k_copy_probe"))


def _pmc_rows(d: str, counter: str):
    import csv
    import glob

    rows = {}
    files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise SystemExit(f"no counter_collection.csv under {d}")
    for fn in files:
        with open(fn, newline="") as f:
            for row in csv.DictReader(f):
                if row.get("Counter_Name") != counter:
                    continue
                name = row.get("Kernel_Name", "")
                for pats, key in KERNEL_KEYS:
                    if all(p in name for p in pats):
                        rows.setdefault(key, []).append(float(row["Counter_Value"]))
                        break
    return rows


def pmc_summary(fetch_dir: str, write_dir: str, copy_bytes: int, command: str) -> dict:
    """HBM bytes per launch from two counter passes (FETCH_SIZE, WRITE_SIZE,
    KiB per dispatch), each scaled by what the same counter reads for
    k_copy_probe, which moves copy_bytes in and copy_bytes out."""
    fe, wr = _pmc_rows(fetch_dir, "FETCH_SIZE"), _pmc_rows(write_dir, "WRITE_SIZE")
    cf = float(np.median(fe["k_copy_probe"]))
    cw = float(np.median(wr["k_copy_probe"]))
    ff, fw = copy_bytes / 1024.0 / cf, copy_bytes / 1024.0 / cw
    out = {"command": command,
           "copy_probe": {"bytes_each_way": copy_bytes, "fetch_kib": cf, "write_kib": cw,
                          "fetch_factor": ff, "write_factor": fw},
           "kernels": {}}
    for _, key in KERNEL_KEYS:
        if key == "k_copy_probe" or key not in fe or key not in wr:
            continue
        f_med, w_med = float(np.median(fe[key])), float(np.median(wr[key]))
        out["kernels"][key] = {"dispatches": len(fe[key]), "fetch_kib_median": f_med, "write_kib_median": w_med,
                               "hbm_bytes_per_launch": f_med * 1024.0 * ff + w_med * 1024.0 * fw}
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--kernel-steps", type=int, default=50, help="steps timed per launch with events")
    ap.add_argument("--sections", default="c3,out,c4,c5,cfg", help="comma list of c3, out, c4, c5, cfg")
    ap.add_argument("--graph", type=int, default=20, help="step iterations per captured hipGraph (0: none)")
    ap.add_argument("--c3-autotune", type=int, default=0,
                    help="step iterations per autotuner trial before the c3 timing (0: the default launch shapes)")
    ap.add_argument("--c4-jobs", type=int, default=256)
    ap.add_argument("--c4-autotune", type=int, default=20,
                    help="step iterations per autotuner trial for a second, tuned c4 run (0: none)")
    ap.add_argument("--c5-jobs", type=int, default=240)
    ap.add_argument("--c5-live", type=int, default=64)
    ap.add_argument("--c5-ticks", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=100, help="timed steps of the CPU restatement")
    ap.add_argument("--cpu-ref-steps", type=int, default=20, help="timed steps of the reference solver")
    ap.add_argument("--no-staging", action="store_true", help="gather element corners from HBM instead of LDS")
    ap.add_argument("--copy-bytes", type=int, default=1 << 30, help="size of the copy-bandwidth probe")
    ap.add_argument("--pmc-summary", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="summarize two rocprofv3 counter passes (FETCH_SIZE, WRITE_SIZE) of this script "
                         "into HBM bytes per launch (JSON on stdout) and exit")
    ap.add_argument("--pmc-command", default="", help="command line recorded in the --pmc-summary output")
    args = ap.parse_args()
    if args.pmc_summary:
        print(json.dumps(pmc_summary(args.pmc_summary[0], args.pmc_summary[1], args.copy_bytes, args.pmc_command),
                         indent=1))
        return 0
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: every job of this engine fits one MI355X and the "
              "multi-GPU layout is design-only; run with --gpus 1", file=sys.stderr)
        return 2
    if eng.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        return 1
    sections = set(args.sections.split(","))
    copy_gbs = eng.copy_bandwidth(args.copy_bytes, 10)
    log(f"device copy bandwidth (read+write): {copy_gbs:.1f} GB/s")
    m = decks.config3(n=100, steps=3000)
    log(f"workload {m.name}: {m.n_elems} elements, {m.n_nodes} nodes")
    c3 = c3_section(args, m)
    mixed, fp64 = c3["mixed"], c3["fp64"]
    be = mixed["contract_bytes_per_elem"]["element"]
    achieved = mixed["element_gbs"]
    piped = mixed["layout"]["pipe_blocks"] > 0
    ekey = "k_element_pipe_mixed" if piped else "k_element_mixed"
    ename = ("k_element_pipe (mixed, pipelined LDS staging, block assembly)" if piped
             else "k_element (mixed, staged, block assembly)")
    out = {
        "metric": METRIC,
        "value": mixed["value"],
        "unit": "elem-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mixed["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"config 3, {m.name}: {m.n_elems} C3D8R, {m.n_nodes} nodes, "
                               f"{'J2 plastic' if m.plastic else 'elastic'}, mixed precision (FP32 element and "
                               f"constitutive update, FP64 nodal state, dt and energies)",
                   "elements": m.n_elems, "nodes": m.n_nodes, "graph_ticks": args.graph},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic_entry(ekey),
                     "kernel": ename,
                     "alg_bytes_per_launch": be * m.n_elems, "avg_launch_ms": mixed["launch_ms"]["element"],
                     "bytes_basis": "contract state bytes: mixed 276 + 8 (J2) per element; whole step "
                                    "(276 + 8 + 200 r) per element-step in mixed.step_gbs",
                     "step_frac": mixed["step_gbs"] / HBM_PEAK_GBS},
        "mixed": mixed,
        "fp64": fp64,
        "mixed_over_fp64": mixed["value"] / fp64["value"],
        "fp64_reference_order": c3["fp64_reference_order"],
        "mixed_over_fp64_reference_order": mixed["value"] / c3["fp64_reference_order"]["value"],
        "rel_disp_err_mixed_vs_fp64": c3["rel_disp_err_mixed_vs_fp64"],
        "copy_probe_gbs": copy_gbs,
    }
    if "out" in sections:
        out["output"] = out_section(args, m, mixed["ms_per_step"])
    if "c4" in sections:
        out["batched"] = c4_section(args)
    if "c5" in sections:
        out["serving"] = c5_section(args)
    if "cfg" in sections:
        out["configs"] = configs_section(args)
    if not args.no_cpu:
        base, port = cpu_baseline(m, args)
        out["cpu_baseline"] = base
        if port is not None:
            out["cpu_restatement"] = port
        log(f"cpu baseline: {out['cpu_baseline']['value']:.3e} elem-steps/s")
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
