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


@pytest.mark.parametrize("arith", ["fp64", "fp32", "f16x3"])
@pytest.mark.parametrize("kernel", ["neigh_dr", "godunov_dr", "all_dr", "local,godunov_dr,neigh_dr"])
def test_proxy_runs_the_dr_kernels(gpu, arith, kernel):
    print(f"proxy {arith} {kernel}", flush=True)
    r = subprocess.run([str(EXE), "--order", "4", "--arithmetic", arith, "-f", "json", "1000", "3", kernel],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["name"] == kernel and out["time"] > 0 and out["gflop-nz"] > 0


@pytest.mark.parametrize("arith,order,mesh,kernel,tol", [
    ("fp64", 4, "proxy", "all", 1e-12),
    ("fp64", 6, "periodic", "all", 1e-12),
    ("fp64", 5, "proxy", "local", 1e-12),
    ("fp32", 6, "proxy", "all", 2e-5),
    ("f16x3", 6, "periodic", "all", 5e-6),
    ("f16x3", 4, "proxy", "local", 5e-6),
])
def test_proxy_dump_matches_oracle(gpu, tmp_path, arith, order, mesh, kernel, tol):
    """The C++ host path end to end (argument parsing, synthetic data, device
    set-up, the kernel classes of seissol_amd.hpp) against the FP64 oracle on
    the same seeded data: one untimed pass plus `timesteps` passes."""
    import numpy as np

    import oracle
    from helpers import rel_l2
    from seissol_amd.kernels import global_matrices, monomial_integrate
    from seissol_amd.proxy import make_setup

    cells, timesteps, seed = 600, 2, 77
    dump = tmp_path / "dofs.bin"
    print(f"proxy dump {arith} order={order} mesh={mesh} kernel={kernel}", flush=True)
    r = subprocess.run([str(EXE), "--order", str(order), "--arithmetic", arith, "--mesh", mesh, "--seed", str(seed),
                        "--dump", str(dump), str(cells), str(timesteps), kernel],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    oracle.build()
    setup = make_setup(order, n=cells, mesh=mesh, seed=seed)
    got = np.fromfile(dump, dtype=np.float64).reshape(setup.dofs.shape)
    gm = global_matrices(order)["flat"]
    coeffs = monomial_integrate(order, 0.0, setup.dt)
    star, ap, am = oracle.cell_local_matrices(setup.vertices, setup.material, setup.neighbors)
    steps = 1 + timesteps
    if kernel == "all":
        ref, _ = oracle.proxy_all(order, gm, steps, coeffs, star, ap, am, setup.neighbors, setup.face_relations,
                                  setup.dofs)
    else:
        ref = setup.dofs.copy()
        for _ in range(steps):
            ref = oracle.local(order, gm, oracle.ader(order, gm, coeffs, ref, star), star, ap, ref)
    # FP64: the update itself; the FP32-data forms: Q, as the other parity tests
    err = rel_l2(got - setup.dofs, ref - setup.dofs) if arith == "fp64" else rel_l2(got, ref)
    print(f"rel L2 vs oracle: {err:.2e}")
    assert err < tol
