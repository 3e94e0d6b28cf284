"""The proxy executable (C++ host over the C ABI) on the GPU."""
import json
import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

EXE = Path(__file__).resolve().parents[1] / "seissol_amd" / "_lib" / "seissol_amd_proxy"


@pytest.mark.parametrize("arith", ["fp64", "bf16x6"])
@pytest.mark.parametrize("kernel", ["all", "local", "neigh", "ader", "localwoader", "ader,neigh"])
def test_proxy_runs_every_kernel(gpu, arith, kernel):
    print(f"proxy {arith} {kernel}", flush=True)
    r = subprocess.run([str(EXE), "--order", "5", "--arithmetic", arith, "-f", "json", "2000", "3", kernel],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["name"] == kernel
    assert out["time"] > 0 and out["gflop-nz"] > 0 and out["gib"] > 0
    assert "element-updates/s" in r.stderr


def test_proxy_plain_summary_periodic(gpu):
    r = subprocess.run([str(EXE), "--order", "6", "--arithmetic", "fp32", "--mesh", "periodic", "4000", "2", "all"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "PERFORMANCE SUMMARY" in r.stdout and "GFLOPS (non-zero) for seissol proxy" in r.stdout


@pytest.mark.parametrize("arith", ["fp64", "fp32"])
@pytest.mark.parametrize("kernel", ["neigh_dr", "godunov_dr", "all_dr", "local,godunov_dr,neigh_dr"])
def test_proxy_runs_the_dr_kernels(gpu, arith, kernel):
    print(f"proxy {arith} {kernel}", flush=True)
    r = subprocess.run([str(EXE), "--order", "4", "--arithmetic", arith, "-f", "json", "1000", "3", kernel],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["name"] == kernel and out["time"] > 0 and out["gflop-nz"] > 0


def test_proxy_dr_kernels_reject_matrix_core_forms(gpu):
    r = subprocess.run([str(EXE), "--order", "4", "--arithmetic", "f16x3", "1000", "1", "all_dr"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 255 and "fp64 or fp32" in r.stderr
