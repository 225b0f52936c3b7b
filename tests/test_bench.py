"""The benchmark's measurement plumbing, on the CPU: the HBM-traffic summary
of two rocprofv3 counter passes (FETCH_SIZE, WRITE_SIZE, KiB per dispatch,
each scaled by what the same counter reads for a plain copy of known size)
and the one-GPU contract of bench.py."""
import csv
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BENCH = os.path.join(ROOT, "bench.py")


def write_counters(d, counter, rows):
    os.makedirs(os.path.join(d, "run"), exist_ok=True)
    with open(os.path.join(d, "run", "run_counter_collection.csv"), "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"])
        w.writeheader()
        for i, (name, value) in enumerate(rows):
            w.writerow({"Dispatch_Id": i, "Kernel_Name": name, "Counter_Name": counter, "Counter_Value": value})


def test_pmc_summary_scales_each_counter_by_the_copy_probe(tmp_path):
    copy_bytes = 1 << 28
    probe = "(anonymous namespace)::k_copy_probe(HIP_vector_type<double, 2u> const*, HIP_vector_type<double, 2u>*, "\
            "unsigned long)"
    pipe = "void ensc::k_element_pipe<2, false>(ensc::JobDev const*, HIP_vector_type<unsigned int, 4u> const*, "\
           "unsigned int)"
    w64 = "void ensc::k_element<false, 256, true, 2, true>(ensc::JobDev const*, HIP_vector_type<unsigned int, 4u> "\
          "const*)"
    ref = "void ensc::k_element<false, 128, true, 1, false>(ensc::JobDev const*, HIP_vector_type<unsigned int, 4u> "\
          "const*)"
    kib = copy_bytes / 1024.0
    # the fetch counter reads half the bytes a copy reads (factor 2), the
    # write counter all of them (factor 1)
    fetch = [(probe, kib / 2)] * 3 + [(pipe, 1000.0), (pipe, 1200.0), (pipe, 1100.0), (w64, 4000.0), (ref, 5000.0)]
    write = [(probe, kib)] * 3 + [(pipe, 500.0)] * 3 + [(w64, 2000.0), (ref, 3000.0)]
    fd, wd = str(tmp_path / "fetch"), str(tmp_path / "write")
    write_counters(fd, "FETCH_SIZE", fetch)
    write_counters(wd, "WRITE_SIZE", write)
    r = subprocess.run([sys.executable, BENCH, "--pmc-summary", fd, wd, "--copy-bytes", str(copy_bytes),
                        "--pmc-command", "unit"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout)
    cp = out["copy_probe"]
    assert cp["fetch_factor"] == pytest.approx(2.0) and cp["write_factor"] == pytest.approx(1.0)
    k = out["kernels"]
    # medians over the dispatches: fetch 1100 KiB x 2, write 500 KiB x 1
    assert k["k_element_pipe_mixed"]["dispatches"] == 3
    assert k["k_element_pipe_mixed"]["hbm_bytes_per_launch"] == pytest.approx((1100.0 * 2 + 500.0) * 1024)
    # the two FP64 element kernels are told apart by their template arguments
    assert k["k_element_fp64"]["hbm_bytes_per_launch"] == pytest.approx((4000.0 * 2 + 2000.0) * 1024)
    assert k["k_element_fp64_reference_order"]["hbm_bytes_per_launch"] == pytest.approx((5000.0 * 2 + 3000.0) * 1024)
    assert "k_copy_probe" not in k
    assert out["command"] == "unit"


def test_bench_runs_on_one_gpu_only():
    r = subprocess.run([sys.executable, BENCH, "--gpus", "2"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
