"""The C and C++ host programs over the C ABI, run as the reference's test
executables are: the reference-shaped C++ driver validates against the oracle
and prints VALID RESULT; the C benchmark drivers check batched against plain
outputs bit for bit."""
import json
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_DRIVER = os.path.join(ROOT, "tests", "c", "mdiscr_ref_driver")
BIN = os.path.join(ROOT, "mdiscriminator_amd", "bin")


def run(args, timeout=120):
    print("running", " ".join(str(a) for a in args), flush=True)
    return subprocess.run([str(a) for a in args], capture_output=True, text=True, timeout=timeout)


@pytest.mark.parametrize("n,m,mode", [(1000003, 4, ""), (1000003, 8, ""), (77777, 3, ""),
                                      (4096, 64, ""), (1, 5, ""), (300000, 16, "segm"),
                                      (123457, 4, "segm"), (100003, 4, "arrays"),
                                      (77777, 5, "arrays"), (4097, 64, "arrays"),
                                      (1000003, 4, "tuples"), (65536, 8, "tuples"),
                                      (300007, 64, "select"), (100003, 5, "select")])
def test_reference_shaped_driver(n, m, mode):
    r = run([REF_DRIVER, n, m] + ([mode] if mode else []))
    assert r.returncode == 0, r.stdout + r.stderr
    assert "VALID RESULT!" in r.stdout


def test_segbench_driver_small():
    r = run([os.path.join(BIN, "mdiscr_segbench"), 300, 2000, 1.5, 8, 2])
    assert r.returncode == 0, r.stderr
    res = json.loads(r.stdout.strip().splitlines()[-1])
    assert res["identical"] is True


def test_servebench_driver_small():
    r = run([os.path.join(BIN, "mdiscr_servebench"), 400, 0.8, 4, 2, 1], timeout=200)
    assert r.returncode == 0, r.stderr
    res = json.loads(r.stdout.strip().splitlines()[-1])
    assert res["identical_spot_checks"] is True
