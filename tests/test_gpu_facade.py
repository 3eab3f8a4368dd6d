"""The BioFVM-compatible C++ façade, driven like the reference's examples.

tests/cpp/facade_driver builds a Microenvironment through the reference's
API (set_density/add_density, resize_space_uniform, density_vector,
add_dirichlet_node, bulk callbacks, create_basic_agent/assign_position/
set_internal_uptake_constants) and loops bulk -> cells -> diffusion; its
output must match the compiled reference's fixtures (FP64 <= 1e-12) and the
oracle (FP32 store <= 1e-5).  In its timing mode (--time) the same loop runs
without reading densities in between: the Thomas tables are built once for
the whole loop, and the bulk fields kept on the device (the default) give the
same bits as the callbacks evaluated at every step (--bulk-every-step).
"""
import os
import subprocess

import numpy as np
import pytest

import gen_golden
import oracle
from biofvm_amd import FP32, FP64
from biofvm_amd import workloads as W

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "facade_driver")
TOL = {FP64: 1e-12, FP32: 1e-5}


def run_driver(tmp_path, case, precision):
    cin, cout = tmp_path / "case.bin", tmp_path / "out.bin"
    n = W.write_driver_case(cin, case, precision)
    r = subprocess.run([DRIVER, str(cin), str(cout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    out = np.fromfile(cout, dtype="<f8")
    assert out.size == n
    return out


@pytest.mark.parametrize("precision", [FP64, FP32], ids=["fp64", "fp32"])
@pytest.mark.parametrize("name", ["g1_lod", "g2_dirichlet", "g3_bulk", "g4_agents", "g8_shell",
                                  "g5_degenerate_1x6x5"])
def test_facade_matches_reference_fixture(tmp_path, name, precision):
    case, expected, _ = gen_golden.load(name)
    got = run_driver(tmp_path, case, precision)
    err = float(np.max(np.abs(got - expected)) / np.max(np.abs(expected)))
    assert err <= TOL[precision], err


def test_facade_many_agents_per_voxel(tmp_path):
    rng = np.random.default_rng(21)
    n = 12
    case = W.make_case(n, n, n, 2, 10.0, [1e4, 1e2], [0.1, 0.01], 0.01, 6, W.ALL,
                       rng.uniform(0, 1, n ** 3 * 2))
    case.update(W.random_agents(rng, 500, 2, n, n, n, 10.0, cluster_sigma=8.0, share_voxel=40))
    ref = oracle.oracle_run(case)
    got = run_driver(tmp_path, case, FP64)
    assert float(np.max(np.abs(got - ref)) / np.max(np.abs(ref))) <= 1e-12


@pytest.mark.parametrize("precision", [FP64, FP32], ids=["fp64", "fp32"])
def test_facade_gradients_and_matlab_snapshot(tmp_path, precision):
    # compute_all_gradient_vectors and write_to_matlab through the façade,
    # after the steps of a fixture with every phase
    import oracle
    from biofvm_amd.matlab import read_matlab
    case, _, _ = gen_golden.load("g8_shell")
    cin, cout, cgrad, cmat = (tmp_path / n for n in ("case.bin", "out.bin", "grad.bin", "snap.mat"))
    W.write_driver_case(cin, case, precision)
    r = subprocess.run([DRIVER, str(cin), str(cout), str(cgrad), str(cmat)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    rho = np.fromfile(cout, dtype="<f8")
    nx, ny, nz, S, dx = case["nx"], case["ny"], case["nz"], case["S"], case["dx"]
    grad = np.fromfile(cgrad, dtype="<f8").reshape(-1, S, 3)
    ref = oracle.gradients(nx, ny, nz, S, dx, rho)
    if precision == FP64:
        np.testing.assert_array_equal(grad, ref)
    else:
        assert np.max(np.abs(grad - ref)) <= 1e-6 * np.max(np.abs(ref))
    name, M = read_matlab(cmat)
    assert name == "multiscale_microenvironment" and M.shape == (4 + S, nx * ny * nz)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    np.testing.assert_array_equal(M[0], ((i + 0.5) * dx).reshape(-1))
    np.testing.assert_array_equal(M[1], ((j + 0.5) * dx).reshape(-1))
    np.testing.assert_array_equal(M[2], ((k + 0.5) * dx).reshape(-1))
    assert np.all(M[3] == dx * dx * dx)
    np.testing.assert_array_equal(M[4:].T.reshape(-1), rho)


@pytest.mark.parametrize("precision", [FP64, FP32], ids=["fp64", "fp32"])
def test_facade_lod2d_solver(tmp_path, precision):
    # diffusion_decay_solver__constant_coefficients_LOD_2D assigned as the examples do
    rng = np.random.default_rng(23)
    nx, ny, S = 17, 11, 2
    case = W.make_case(nx, ny, 1, S, 20.0, [3e4, 8e2], [0.2, 0.02], 0.01, 5, W.ALL, rng.uniform(0, 1, nx * ny * S))
    case.update(W.random_bulk(rng, nx * ny, S))
    case.update(W.random_agents(rng, 25, S, nx, ny, 1, 20.0, share_voxel=3))
    j, i = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    idx = np.flatnonzero(((i == 0) | (i == nx - 1) | (j == 0)).reshape(-1)).astype(np.int32)  # 3 edges
    case.update(dir_idx=idx, dir_val=rng.uniform(0, 2, idx.size * S))
    case["lod_dims"] = 2
    ref = oracle.oracle_run(case)
    got = run_driver(tmp_path, case, precision)
    assert float(np.max(np.abs(got - ref)) / np.max(np.abs(ref))) <= TOL[precision]


def shell_case(steps):
    case, _, _ = gen_golden.load("g8_shell")  # every phase: Dirichlet, bulk, agents
    case = dict(case)
    case["steps"] = steps
    return case


@pytest.mark.parametrize("precision", [FP64, FP32], ids=["fp64", "fp32"])
def test_facade_steady_loop_keeps_its_tables_and_bulk_fields(tmp_path, precision):
    import json
    warm, steps = 3, 12
    case = shell_case(steps)
    cin = tmp_path / "case.bin"
    W.write_driver_case(cin, case, precision)
    outs, lines = [], []
    for mode in ([], ["--bulk-every-step"]):
        cout = tmp_path / f"out{len(outs)}.bin"
        r = subprocess.run([DRIVER, str(cin), str(cout), "--time", str(warm), *mode], capture_output=True,
                           text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        lines.append(json.loads(r.stdout.strip().splitlines()[-1]))
        outs.append(np.fromfile(cout, dtype="<f8"))
    assert lines[0]["coefficient_builds"] == 1 and lines[1]["coefficient_builds"] == 1, lines
    assert lines[0]["bulk_rates_host_seconds"] == 0.0  # fields evaluated once, during the warm-up
    assert lines[1]["bulk_rates_host_seconds"] > 0.0
    np.testing.assert_array_equal(outs[0], outs[1])
    ref_case = dict(case)
    ref_case["steps"] = warm + steps
    ref = oracle.oracle_run(ref_case)
    assert float(np.max(np.abs(outs[0] - ref)) / np.max(np.abs(ref))) <= TOL[precision]
