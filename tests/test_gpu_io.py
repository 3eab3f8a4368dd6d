"""GPU MAT v4 snapshots (bfvm_write_matlab) against the reference's own
write_to_matlab: byte-identical files for the FP64 store, and the snapshot
of the asynchronous form equals the state at the call while later steps run."""
import numpy as np
import pytest

import oracle
from biofvm_amd import ALL, FP32, FP64, setup_case
from biofvm_amd import workloads as W
from biofvm_amd.matlab import read_matlab

pytestmark = pytest.mark.gpu


@pytest.mark.skipif(not oracle.have_ref(), reason="compiled reference not built")
@pytest.mark.parametrize("dims", [(9, 7, 5, 3), (1, 6, 4, 2), (4, 3, 2, 1)])
def test_matlab_file_is_byte_identical_to_reference(tmp_path, dims):
    nx, ny, nz, S = dims
    dx = 20.0
    rho = np.random.default_rng(sum(dims)).uniform(0, 1, nx * ny * nz * S)
    oracle.ref_write_matlab(nx, ny, nz, S, dx, rho, tmp_path / "ref.mat")
    case = W.make_case(nx, ny, nz, S, dx, np.ones(S), np.ones(S), 0.01, 0, ALL, rho)
    with setup_case(case, FP64) as m:
        m.write_matlab(tmp_path / "gpu.mat")
        m.write_matlab(tmp_path / "gpu_async.mat", asynchronous=True)
        m.io_wait()
    ref = (tmp_path / "ref.mat").read_bytes()
    assert (tmp_path / "gpu.mat").read_bytes() == ref
    assert (tmp_path / "gpu_async.mat").read_bytes() == ref


@pytest.mark.parametrize("precision", [FP64, FP32], ids=["fp64", "fp32"])
def test_async_snapshot_captures_the_state_at_the_call(tmp_path, precision):
    case = W.ensemble_mixed(steps=3)[3]  # 20x14x40x5, bulk, agents, Dirichlet
    with setup_case(case, precision) as m:
        m.run(case["dt"], 2, case["phases"])
        before = m.download()
        m.write_matlab(tmp_path / "a.mat", origin=(-10.0, 5.0, 2.5), asynchronous=True)
        m.run(case["dt"], 5, case["phases"])  # runs while the snapshot is copied and written
        m.io_wait()
        after = m.download()
    _, M = read_matlab(tmp_path / "a.mat")
    np.testing.assert_array_equal(M[4:].T.reshape(-1), before)
    assert not np.array_equal(before, after)
    nx, dx = case["nx"], case["dx"]
    np.testing.assert_array_equal(M[0, :nx], -10.0 + (np.arange(nx) + 0.5) * dx)
