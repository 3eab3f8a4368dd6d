"""GPU parity at the measured configurations' full sizes.

- The FP64 drop-in at scale (200^3 x 4, Dirichlet faces, bulk fields, 1e5
  agents): with the reference's whole-vector Dirichlet semantics, against the
  compiled reference itself (one thread: its agent loop races with more);
  with the per-substrate mask (substrate 0 only), against the oracle.  Two
  steps each, FP64 store, 1e-12 of max|ref|.
- 400^3 x 8 (the precision form's grid), diffusion only: every sweep's column
  sums are 1 + c2, so each substrate's mass is multiplied by exactly
  (1 + lambda dt / 3)^-3 per step (no-flux boundaries).  FP64 store within
  1e-12, FP32 store within 1e-6 (its sum is taken in FP64 over FP32 values).
"""
import numpy as np
import pytest

import oracle
from biofvm_amd import FP32, FP64, setup_case
from biofvm_amd import workloads as W

pytestmark = pytest.mark.gpu


def rel_err(got, ref):
    return float(np.max(np.abs(got - ref)) / max(np.max(np.abs(ref)), 1e-300))


@pytest.mark.skipif(not oracle.have_ref(), reason="compiled reference not built (oracle/Makefile)")
def test_drop_in_200_whole_vector_dirichlet_against_compiled_reference():
    case = W.drop_in_200(steps=2, mask_substrate0=False)
    ref = oracle.ref_run(case, nthreads=1)
    with setup_case(case, FP64) as m:
        m.run(case["dt"], case["steps"], case["phases"])
        got = m.download()
    assert rel_err(got, ref) <= 1e-12


def test_drop_in_200_masked_dirichlet_against_oracle():
    case = W.drop_in_200(steps=2)
    ref = oracle.oracle_run(case)
    with setup_case(case, FP64) as m:
        m.run(case["dt"], case["steps"], case["phases"])
        got = m.download()
    assert rel_err(got, ref) <= 1e-12


@pytest.mark.parametrize("precision,tol", [(FP64, 1e-12), (FP32, 1e-6)], ids=["fp64", "fp32"])
def test_400_cubed_mass_decays_by_the_exact_factor(precision, tol):
    n, S, steps, dt = 400, 8, 2, 0.01
    rng = np.random.default_rng(3)
    D = 10.0 ** rng.uniform(2.0, 5.0, S)
    lam = 10.0 ** rng.uniform(-3.0, 0.0, S)
    rho0 = rng.uniform(0.0, 1.0, n ** 3 * S)
    case = W.make_case(n, n, n, S, 10.0, D, lam, dt, steps, W.DIFFUSION, rho0)
    with setup_case(case, precision) as m:
        start = m.download().reshape(-1, S).sum(axis=0)  # the store's own values (FP32-rounded)
        m.run(dt, steps, W.DIFFUSION)
        end = m.download().reshape(-1, S).sum(axis=0)
    c2 = lam * dt / 3.0
    expect = start * (1.0 + c2) ** (-3 * steps)
    assert np.max(np.abs(end - expect) / expect) <= tol
