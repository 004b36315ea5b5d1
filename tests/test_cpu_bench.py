"""The CPU baseline driver (oracle/cpu_bench.py) runs as a child interpreter
and reports the summed rate of its worker processes (no GPU)."""
import json
import os
import subprocess
import sys

from conftest import ROOT


def test_cpu_bench_child_reports_rate():
    env = dict(os.environ, OMP_NUM_THREADS="1", OPENBLAS_NUM_THREADS="1", MKL_NUM_THREADS="1")
    r = subprocess.run([sys.executable, "-m", "oracle.cpu_bench", "--n", "16", "--seconds", "0.5", "--workers", "2",
                        "--batch", "256"], cwd=ROOT, env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["cores"] == 2 and out["kind"] == "port" and out["value"] > 0


def test_cpu_bench_hashgrid_child_reports_rate():
    env = dict(os.environ, OMP_NUM_THREADS="1", OPENBLAS_NUM_THREADS="1", MKL_NUM_THREADS="1")
    r = subprocess.run([sys.executable, "-m", "oracle.cpu_bench", "--workload", "hashgrid", "--n", "16", "--seconds",
                        "0.5", "--workers", "1", "--batch", "16"], cwd=ROOT, env=env, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["cores"] == 1 and out["kind"] == "port" and out["value"] > 0 and "hashgrid" in out["sample"]
