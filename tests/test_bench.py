"""bench.py end to end with small settings: one JSON line carrying every
section (roofline, per-shape lines, host drivers), all measured."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.gpu
@pytest.mark.timeout(400)
def test_bench_line_is_complete():
    print("running bench.py --no-cpu --steps 2 --warmup 1 --serve-requests 500", flush=True)
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--no-cpu", "--steps", "2",
                        "--warmup", "1", "--serve-requests", "500"],
                       capture_output=True, text=True, timeout=380)
    assert r.returncode == 0, r.stderr[-3000:]
    lines = [ln for ln in r.stdout.splitlines() if ln.strip()]
    assert len(lines) == 1, r.stdout[-2000:]
    d = json.loads(lines[0])
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "dtype",
              "config", "roofline", "cpu_baseline"):
        assert k in d, k
    assert d["n_gpus"] == 1 and d["value"] > 0
    rf = d["roofline"]
    assert rf["bound"] == "hbm" and 0 < rf["frac"] < 1 and rf["achieved"] > 0
    cc = d["copy_ceiling"]   # the same buffers copied on the same box
    assert cc["achieved"] > 0 and 0 < cc["partition_vs_copy"] < 1.5
    rc = d["read_ceiling"]   # a streaming read of the same input on the same box
    assert rc["achieved"] > 0 and 0 < rc["count_vs_read"] < 1.5
    assert 0 < rc["step_vs_read_plus_copy"] < 1.5
    for m in ("m16", "m32", "m64"):
        assert d["soa_records_2p26"][m]["partition_frac"] > 0
    for name, line in d["general"].items():
        assert line["partition_frac"] and line["partition_frac"] > 0, name
    for m in ("m2", "m4", "m16", "m32", "m64", "mod7", "mod33"):
        assert 0 < d["plain_across_m"][m]["partition_frac"] < 1, m
    for k in ("n100", "n4096"):
        assert d["small_calls"][k]["us_per_call"] > 0
    assert 0 < d["hbm_frac_whole_path_12B"] < 1
    assert "traffic_source" in rf
    assert d["batched_segments"]["identical"] is True
    assert d["serving"]["identical_spot_checks"] is True


def test_bench_refuses_several_gpus():
    """Runs on the CPU: the refusal comes before anything touches a GPU."""
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_cpu_baseline_script_quick():
    """oracle/cpu_baseline.py on small samples (CPU): every configuration's
    line, the host description and the reference method."""
    env = dict(os.environ, OMP_PROC_BIND="close", OMP_NUM_THREADS="2")
    r = subprocess.run([sys.executable, "-m", "oracle.cpu_baseline", "--quick", "--runs", "3"],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    d = json.loads(r.stdout.strip().splitlines()[-1])
    for k in ("config1_seq", "config2_seq", "config2_openmp", "config4_seq"):
        assert d[k]["value"] > 0 and len(d[k]["runs_ms"]) == 3, k
    assert d["config2_openmp"]["cores"] == 2
    for m in ("m16", "m32", "m64"):
        assert d["config3_seq"][m]["value"] > 0
    assert d["host"]["nproc"] >= 1 and "run.py" in d["method"]
