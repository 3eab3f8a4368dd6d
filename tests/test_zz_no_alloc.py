"""bfvm_mass_balance_noalloc (no device allocation, static workspace) against
bfvm_mass_balance, through the C ABI: the same bits, on two calls in a row,
the second on a smaller store, and the substrate limit of its workspace."""
import ctypes

import numpy as np
import pytest

from biofvm_amd import DIFFUSION, FP32, FP64, DeviceMicroenvironment, load

pytestmark = pytest.mark.gpu

BFVM_E_ARG = -1


def noalloc(m):
    import torch
    sums = torch.full((m.S,), -1.0, dtype=torch.float64, device="cuda")
    bad = torch.full((1,), -1, dtype=torch.int64, device="cuda")
    torch.cuda.synchronize()
    rc = m.lib.bfvm_mass_balance_noalloc(m.h, ctypes.c_void_p(sums.data_ptr()), ctypes.c_void_p(bad.data_ptr()))
    assert rc == 0, m.lib.bfvm_last_error(m.h)
    assert m.lib.bfvm_sync(m.h) == 0
    return sums.cpu().numpy(), int(bad.cpu().item())


def make(shape, S, precision, seed, nan_at=None):
    rng = np.random.default_rng(seed)
    m = DeviceMicroenvironment(*shape, S, 10.0, precision)
    m.set_coefficients(10 ** rng.uniform(2, 5, S), 10 ** rng.uniform(-3, 0, S))
    rho = rng.uniform(0.0, 1.0, int(np.prod(shape)) * S)
    if nan_at is not None:
        rho[nan_at] = np.nan
    m.upload(rho)
    m.run(0.01, 2, DIFFUSION)
    return m


def test_noalloc_mass_balance_equals_allocating_one():
    load()
    big = make((96, 80, 64), 8, FP32, 1)
    small = make((13, 11, 7), 3, FP64, 2, nan_at=5)
    try:
        for m in (big, small):  # the second call on a smaller store
            got, got_bad = noalloc(m)
            ref, ref_bad = m.mass_balance()
            np.testing.assert_array_equal(got[~np.isnan(ref)], ref[~np.isnan(ref)])
            assert np.array_equal(np.isnan(got), np.isnan(ref))
            assert got_bad == ref_bad
        assert noalloc(small)[1] > 0 and noalloc(big)[1] == 0
    finally:
        big.close()
        small.close()


def test_noalloc_rejects_more_substrates_than_its_workspace_holds():
    with DeviceMicroenvironment(4, 4, 4, 65, 10.0, FP64) as m:
        rc = m.lib.bfvm_mass_balance_noalloc(m.h, None, None)
        assert rc == BFVM_E_ARG
        assert b"64" in m.lib.bfvm_last_error(m.h)
