"""Device reductions and failure paths, through the C ABI.

- bfvm_mass_balance: per-substrate FP64 sums of the store against numpy's
  sums of the downloaded store (shapes with 1..300 substrates, both store
  precisions), the same bits on repeated calls, and the closed form of the
  LOD step's mass balance: every column of each sweep matrix sums to 1 + c2
  (BioFVM_solvers.cpp:121-135), so with no-flux boundaries and no sources
  each step multiplies each substrate's sum by (1 + lambda dt/3)^-3.
- Non-finite values: NaN and infinities injected into the store are counted,
  a clean store counts none, and a NaN spreads along its lines in a step.
- Device memory exhaustion: creating handles until the device refuses one
  gives BFVM_E_DEVICE with a message, and a handle created before keeps its
  densities and steps as before.
"""
import ctypes

import numpy as np
import pytest

from biofvm_amd import DIFFUSION, FP32, FP64, BfvmError, DeviceMicroenvironment, load
from biofvm_amd import workloads as W

pytestmark = pytest.mark.gpu

BFVM_E_DEVICE = -2


def random_handle(rng, shape, S, precision, dx=10.0):
    m = DeviceMicroenvironment(*shape, S, dx, precision)
    m.set_coefficients(10 ** rng.uniform(2, 5, S), 10 ** rng.uniform(-3, 0, S))
    m.upload(rng.uniform(0.0, 1.0, int(np.prod(shape)) * S))
    return m


@pytest.mark.parametrize("precision", [FP64, FP32], ids=["fp64", "fp32"])
@pytest.mark.parametrize("shape,S", [((1, 1, 1), 1), ((7, 5, 3), 3), ((40, 33, 21), 8), ((64, 64, 64), 6),
                                     ((9, 8, 7), 300)])
def test_mass_balance_equals_host_sums(shape, S, precision):
    rng = np.random.default_rng(sum(shape) + S)
    with random_handle(rng, shape, S, precision) as m:
        sums, bad = m.mass_balance()
        host = m.download().reshape(-1, S)
        again, _ = m.mass_balance()
    assert bad == 0
    expect = np.array([np.sum(host[:, s], dtype=np.float64) for s in range(S)])
    assert np.max(np.abs(sums - expect) / np.abs(expect)) <= 1e-13
    np.testing.assert_array_equal(sums, again)  # fixed summation order


@pytest.mark.parametrize("precision,tol", [(FP64, 1e-12), (FP32, 1e-6)], ids=["fp64", "fp32"])
def test_mass_decays_by_the_exact_factor_small(precision, tol):
    rng = np.random.default_rng(5)
    S, dt, steps = 5, 0.01, 4
    lam = 10 ** rng.uniform(-3, 0, S)
    with DeviceMicroenvironment(30, 20, 25, S, 10.0, precision) as m:
        m.set_coefficients(10 ** rng.uniform(2, 5, S), lam)
        m.upload(rng.uniform(0.0, 1.0, 30 * 20 * 25 * S))
        start, _ = m.mass_balance()
        m.run(dt, steps, DIFFUSION)
        end, bad = m.mass_balance()
    assert bad == 0
    expect = start * (1.0 + lam * dt / 3.0) ** (-3 * steps)
    assert np.max(np.abs(end - expect) / expect) <= tol


def test_nonfinite_values_are_counted():
    rng = np.random.default_rng(8)
    S, n = 3, 24
    rho = rng.uniform(0.0, 1.0, n ** 3 * S)
    with DeviceMicroenvironment(n, n, n, S, 10.0, FP64) as m:
        m.set_coefficients([1e3, 1e2, 1e4], [0.1, 0.01, 1.0])
        m.upload(rho)
        assert m.mass_balance()[1] == 0
        bad = rho.copy()
        bad[17] = np.nan
        bad[5 * S + 1] = np.inf
        bad[-1] = -np.inf
        m.upload(bad)
        assert m.mass_balance()[1] == 3
        # one step: the NaN (voxel 5, substrate 2) spreads along its x line
        # in the x sweep, then over its plane and the grid in y and z
        m.upload(np.where(np.isnan(bad), np.nan, rho))
        m.run(0.01, 1, DIFFUSION)
        sums, count = m.mass_balance()
        assert count >= n ** 3 and np.isnan(sums[17 % S])
        assert np.all(np.isfinite(np.delete(sums, 17 % S)))
    with DeviceMicroenvironment(n, n, n, 2, 10.0, FP32) as m:
        x = np.zeros(n ** 3 * 2)
        x[100] = np.inf
        m.upload(x)
        assert m.mass_balance()[1] == 1


def test_device_out_of_memory_is_an_error_and_leaves_handles_intact():
    lib = load()
    rng = np.random.default_rng(9)
    keep = DeviceMicroenvironment(20, 20, 20, 2, 10.0, FP64)
    keep.set_coefficients([1e3, 1e4], [0.1, 0.5])
    rho = rng.uniform(0.0, 1.0, 20 ** 3 * 2)
    keep.upload(rho)
    before, _ = keep.mass_balance()
    big = []
    # stores of 2^31 - 2^24 FP64 values (16 GiB each) until the device refuses one
    nx = 2048
    nz = ((1 << 31) - (1 << 24)) // (nx * nx)
    rc = 0
    try:
        for _ in range(40):
            h = ctypes.c_void_p()
            rc = lib.bfvm_create(nx, nx, nz, 1, 1.0, FP64, ctypes.byref(h))
            if rc != 0:
                break
            big.append(h)
        assert rc == BFVM_E_DEVICE, f"{len(big)} stores of 16 GiB created, last rc {rc}"
        msg = lib.bfvm_last_error(None).decode()
        assert "bfvm_create" in msg and msg.strip(), msg
    finally:
        for h in big:
            assert lib.bfvm_destroy(h) == 0
    after, bad = keep.mass_balance()
    np.testing.assert_array_equal(after, before)
    np.testing.assert_array_equal(keep.download(), rho)
    keep.run(0.01, 2, DIFFUSION)
    end, bad = keep.mass_balance()
    assert bad == 0
    expect = before * (1.0 + np.array([0.1, 0.5]) * 0.01 / 3.0) ** (-6)
    assert np.max(np.abs(end - expect) / expect) <= 1e-12
    keep.close()
