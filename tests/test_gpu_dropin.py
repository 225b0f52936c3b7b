"""Drop-in check: the reference's own reader and solver class, with the
MI355X binding (enscdynamics_amd/binding/exDyna3D_mi355x.h, the subclass a
maintainer adds to the reference) replacing its CPU step loop.

`make -C oracle shim` links the reference's sources (where /root/reference
exists) with the binding and the engine library into oracle/_ref/ref_dropin;
the binary travels with the tree.  The reference's DataIn parses each deck,
exDyna3D_MI355X::run executes every *Step on the GPU through the C ABI, and
the final state must match the CPU reference's golden run of the same deck:
<= 1e-10 in FP64, <= 1e-4 (displacement) in mixed precision.
"""
import gzip
import json
import os
import subprocess

import numpy as np
import pytest

import goldens as G

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DROPIN = os.path.join(ROOT, "oracle", "_ref", "ref_dropin")
DECKS = {"ref_Job_6": "Job-6", "ref_test_single2": "test_single2", "ref_test_step": "test_step",
         "ref_test_trans": "test_trans", "ref_test_part1": "test_part1", "ref_GenVirPa": "GenVirPa"}


def rel(a, b, floor=0.0):
    s = max(float(np.abs(b).max()), floor)
    return float(np.abs(a - b).max()) / s if s > 0 else float(np.abs(a - b).max())


def run_dropin(case, precision, tmp_path, stepwise=False, fsi=None, deck=None):
    if not os.path.exists(DROPIN):
        pytest.skip("oracle/_ref/ref_dropin not built (needs the reference sources: make -C oracle shim)")
    d = G.load(case)
    j = json.loads(str(d["model_json"]))
    tmp_path.mkdir(parents=True, exist_ok=True)
    with gzip.open(os.path.join(G.GOLDEN, "decks", (deck or DECKS[case]) + ".inp.gz"), "rt") as f:
        (tmp_path / "deck.inp").write_text(f.read())
    args = [DROPIN, str(tmp_path), "deck.inp", str(tmp_path), precision,
            repr(j["c_dr"]), repr(j["c_cr"]), repr(j["factor"])] + (["stepwise"] if stepwise else [])
    print(f"[dropin] {case} {precision}{' stepwise' if stepwise else ''}{' fsi' if fsi else ''}", flush=True)
    env = dict(os.environ)
    if fsi is not None:
        env["ENSC_FSI"] = ",".join(repr(float(v)) if i else str(int(v)) for i, v in enumerate(fsi))
    r = subprocess.run(args, capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stderr[-2000:]
    sc = open(tmp_path / "scalars.txt").read().split()
    steps = int(sc[sc.index("steps") + 1])
    if stepwise:
        # the subclass's computeSate() hook saw every step iteration
        assert int(sc[sc.index("hooked") + 1]) == steps
    return d, np.load(tmp_path / "u.npy"), np.load(tmp_path / "sigma.npy"), np.load(tmp_path / "x.npy"), steps


@pytest.mark.parametrize("case", sorted(DECKS))
def test_reference_reader_with_gpu_step_loop_fp64(case, tmp_path):
    d, u, s, x, steps = run_dropin(case, "fp64", tmp_path)
    assert steps == int(d["n_steps"])
    assert rel(u, d["final_u"], 1e-12) <= 1e-10
    assert rel(x, d["final_x"]) <= 1e-10
    assert rel(s, d["final_sigma"], 1.0) <= 1e-10


@pytest.mark.parametrize("case", sorted(DECKS))
def test_reference_reader_with_gpu_step_loop_mixed(case, tmp_path):
    # every reference deck is quasi-static (810-7500 wave transits in its
    # run), so the mixed engine accumulates sigma and QIA with FP32
    # residuals there (ensc_options.state_comp, automatic): without them
    # Job-6 and test_part1 drift to 2e-4 and 2e-3
    d, u, s, x, steps = run_dropin(case, "mixed", tmp_path)
    assert steps == int(d["n_steps"])
    err = rel(u, d["final_u"], 1e-12)
    err_s = rel(s, d["final_sigma"], 1.0)
    print(f"[dropin mixed] {case}: u rel err {err:.2e}, sigma {err_s:.2e}", flush=True)
    assert err <= 1e-4


def test_dropin_writes_the_reference_result_files(tmp_path):
    # exDyna3D_MI355X::run also writes output_results' files,
    # ResultData/solution_<n>.vtu, from device snapshots: as many as the
    # reference's cadence gives, each holding the reference's state of that
    # step (the oracle's output snapshots), in Float32 as VTK stores it
    from enscdynamics_amd.vtu import read_vtu
    from oracle.oracle import Oracle
    case = "ref_test_step"
    d, u, s, x, steps = run_dropin(case, "fp64", tmp_path)
    o = Oracle(G.model_from_dump(d), 0).set_output(0.0).init()
    o.run(0, [], 0)
    files = os.listdir(tmp_path / "ResultData")
    print(f"[dropin output] {case}: {len(files)} files, oracle {o.snapshots()}", flush=True)
    assert sorted(files) == sorted(f"solution_{k}.vtu" for k in range(o.snapshots()))
    for k in (0, o.snapshots() // 2, o.snapshots() - 1):
        r = read_vtu(str(tmp_path / "ResultData" / f"solution_{k}.vtu"))
        assert r["time"] == pytest.approx(o.snapshot(k, "time")[0], rel=1e-7)
        for f in ("u", "v", "a"):
            ref = o.snapshot(k, f)
            assert np.allclose(r["point_data"][f], ref, rtol=1e-6, atol=1e-6 * max(np.abs(ref).max(), 1e-30)), f
        for f in ("vonMises", "Stress", "Green-Lagrange Strain"):
            ref = o.snapshot(k, f).reshape(r["cell_data"][f].shape)
            assert np.allclose(r["cell_data"][f], ref, rtol=1e-5, atol=1e-6 * max(np.abs(ref).max(), 1e-30)), f


def test_stepwise_binding_hooks_every_step(tmp_path):
    # the binding's per-step mode: run() calls computeSate() once per step
    # iteration, the subclass hooks it, and the result equals the whole-run
    # mode bit for bit (the same device steps) and the reference's golden
    case = "ref_test_step"
    d, u, s, x, steps = run_dropin(case, "fp64", tmp_path / "whole")
    d, u2, s2, x2, steps2 = run_dropin(case, "fp64", tmp_path / "stepwise", stepwise=True)
    assert steps2 == steps == int(d["n_steps"])
    assert np.array_equal(u2, u) and np.array_equal(s2, s) and np.array_equal(x2, x)
    assert rel(u2, d["final_u"], 1e-12) <= 1e-10


@pytest.mark.parametrize("case", ["micro_fsi_block", "micro_fsi_slab"])
def test_dropin_fsi_coupler_matches_the_reference(case, tmp_path):
    # an FSI coupler written against the reference's interface (a subclass
    # that calls the particle setup at its first step and writes
    # fsi_share_data's forces before every step) runs unchanged on the
    # binding: the state and the particles equal the reference's own
    # coupling run of the same subclass
    d, u, s, x, steps = run_dropin(case, "fp64", tmp_path, stepwise=True, fsi=tuple(d for d in G.load(case)["fsi"]),
                                   deck=case)
    assert steps == d["history"].shape[0]
    assert rel(u, d["final_u"], 1e-12) <= 1e-10
    assert rel(x, d["final_x"]) <= 1e-10
    assert rel(s, d["final_sigma"], 1.0) <= 1e-10
    for f in ("fsi_x", "fsi_v"):
        got = np.load(tmp_path / (f + ".npy"))
        assert got.shape == d["final_" + f].shape
        assert rel(got, d["final_" + f]) <= 1e-10, f
