#!/usr/bin/env python3
"""Benchmark of the MI355X step engine.

Default workload: the 1M-element plastic cube (100^3 C3D8R, J2 yield 250 MPa,
hardening 1 GPa, z=0 fixed, z=1 pushed at 10 m/s), inputs resident in HBM.
One "step" = one iteration of exDyna3D's step loop (element pass + nodal
update + dt reduction) over the whole mesh.  The headline value is the
mixed-precision element-steps/s; the FP64 engine is timed in the same run on
the same steps, and the relative displacement error of mixed vs FP64 after
those steps is reported beside it.  The CPU baseline is the FP64
restatement (oracle/) timed on the host's cores on a bounded sample.

    python bench.py [--gpus 1] [--steps 200] [--warmup 20] [--workload c3|c2|c1]
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from enscdynamics_amd import decks  # noqa: E402
from enscdynamics_amd import engine as eng  # noqa: E402

METRIC = "elem-steps/s mixed vs FP64 at ≤1e-4 rel disp err; batched solves/s + p99"
HBM_PEAK_GBS = 8000.0


def bytes_per_elem_step(precision: int, plastic: bool, r: float):
    """Algorithmic bytes per element-step: compulsory once-per-step reads and
    writes of persistent state at the stored precision.  Returns
    (whole step, element kernel share)."""
    if precision == eng.FP64:
        elem = 520 + (16 if plastic else 0)
        node = 152 * r
    else:
        elem = 276 + (8 if plastic else 0)
        node = 200 * r
    return elem + node, elem + 72 * r


def workload(name: str):
    if name == "c3":
        return decks.config3(n=100, steps=3000)
    if name == "c2":
        return decks.config2()
    if name == "c1":
        return decks.config1()
    raise SystemExit(f"unknown workload {name}")


def time_engine(m, precision, warmup, steps, timed_kernel_steps, staging=True):
    e = eng.Engine(m, precision, reorder=True, staging=staging)
    e.init()
    e.set_current_step(1)
    e.enqueue(warmup)
    e.sync()
    n0 = int(e.scalars()["nsteps"])
    t0 = time.perf_counter()
    e.enqueue(steps)
    e.sync()
    t1 = time.perf_counter()
    n1 = int(e.scalars()["nsteps"])
    assert n1 - n0 == steps, f"only {n1 - n0} of {steps} ticks were full steps"
    e.set_timing(True)
    e.enqueue(timed_kernel_steps)
    e.sync()
    elem_ms, node_ms, launches = e.kernel_times()
    e.set_timing(False)
    s = e.scalars()
    return e, (t1 - t0), elem_ms, node_ms, s


def cpu_baseline(m, steps=100, warm=10):
    from oracle.oracle import Oracle  # CPU baseline leg only

    threads = os.cpu_count() or 1
    o = Oracle(m, threads).init()
    o.time_steps(warm)
    secs = o.time_steps(steps)
    cpu = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": m.n_elems * steps / secs, "unit": "elem-steps/s", "cores": o.threads, "kind": "port",
            "sample": f"{m.name}: {warm} warm-up + {steps} timed steps of the FP64 restatement, "
                      f"{o.threads} OpenMP threads on {cpu}"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="c3")
    ap.add_argument("--kernel-steps", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=100)
    ap.add_argument("--no-staging", action="store_true", help="gather element corners from HBM instead of LDS")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: every job of this engine fits one MI355X and the "
              "multi-GPU layout is design-only (see DESIGN.md); run with --gpus 1", file=sys.stderr)
        return 2
    if eng.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        return 1
    m = workload(args.workload)
    print(f"[bench] workload {m.name}: {m.n_elems} elements, {m.n_nodes} nodes", file=sys.stderr, flush=True)
    r = m.n_nodes / m.n_elems
    res = {}
    for prec, name in ((eng.MIXED, "mixed"), (eng.FP64, "fp64")):
        e, secs, kms, nms, s = time_engine(m, prec, args.warmup, args.steps, args.kernel_steps,
                                           staging=not args.no_staging)
        res[name] = dict(engine=e, secs=secs, elem_ms=kms, node_ms=nms, scalars=s)
        print(f"[bench] {name}: {args.steps} steps in {secs:.4f} s; element kernel {kms:.4f} ms, "
              f"node kernel {nms:.4f} ms", file=sys.stderr, flush=True)
    u_mix = res["mixed"]["engine"].get("u")
    u_64 = res["fp64"]["engine"].get("u")
    rel_err = float(np.abs(u_mix - u_64).max() / max(np.abs(u_64).max(), 1e-300))
    mixed = res["mixed"]
    fp64 = res["fp64"]
    ms_per_step = 1e3 * mixed["secs"] / args.steps
    value = m.n_elems * args.steps / mixed["secs"]
    v64 = m.n_elems * args.steps / fp64["secs"]
    b_step, b_k1 = bytes_per_elem_step(eng.MIXED, m.plastic, r)
    b_step64, b_k1_64 = bytes_per_elem_step(eng.FP64, m.plastic, r)
    achieved = b_k1 * m.n_elems / (mixed["elem_ms"] * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "elem-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"{m.name}: {m.n_elems} C3D8R, {m.n_nodes} nodes, "
                               f"{'J2 plastic' if m.plastic else 'elastic'}, mixed precision (FP32 element, "
                               f"FP64 nodal/dt/energy)",
                   "elements": m.n_elems, "nodes": m.n_nodes, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_element (mixed)", "bytes_per_elem": b_k1,
                     "avg_launch_ms": mixed["elem_ms"]},
        "fp64": {"value": v64, "ms_per_step": 1e3 * fp64["secs"] / args.steps, "elem_kernel_ms": fp64["elem_ms"],
                 "node_kernel_ms": fp64["node_ms"],
                 "elem_kernel_gbs": b_k1_64 * m.n_elems / (fp64["elem_ms"] * 1e-3) / 1e9},
        "mixed_over_fp64": value / v64,
        "rel_disp_err_mixed_vs_fp64": rel_err,
        "step_roofline": {"mixed_gbs": b_step * value / 1e9, "fp64_gbs": b_step64 * v64 / 1e9,
                          "mixed_frac": b_step * value / 1e9 / HBM_PEAK_GBS,
                          "fp64_frac": b_step64 * v64 / 1e9 / HBM_PEAK_GBS,
                          "bytes_per_elem_step": {"mixed": b_step, "fp64": b_step64}},
        "node_kernel_ms": mixed["node_ms"],
    }
    for k in ("engine",):
        mixed.pop(k)
        fp64.pop(k)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(m, args.cpu_steps)
        print(f"[bench] cpu baseline: {out['cpu_baseline']['value']:.3e} elem-steps/s", file=sys.stderr, flush=True)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
