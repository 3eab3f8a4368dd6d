"""GPU parity at the measured configurations' full sizes.

- The FP64 drop-in at scale (200^3 x 4, Dirichlet faces, bulk fields, 1e5
  agents): with the reference's whole-vector Dirichlet semantics, against the
  compiled reference itself (one thread: its agent loop races with more);
  with the per-substrate mask (substrate 0 only), against the oracle.  Two
  steps each, FP64 store, 1e-12 of max|ref|.
- 400^3 x 8 (the precision form's grid), diffusion only: every sweep's column
  sums are 1 + c2, so each substrate's mass is multiplied by exactly
  (1 + lambda dt / 3)^-3 per step (no-flux boundaries); the sums are the
  device's FP64 reduction (bfvm_mass_balance).  FP64 store within 1e-12,
  FP32 store within 1e-6 (its sum is taken in FP64 over FP32 values).
- A store above 4 GiB (416 x 416 x 400 x 8, FP64, 4.43 GB): the x sweep runs
  in two runs of z planes and the z sweep with one segment per wave (its
  buffers start at each segment), both past the 32-bit offsets of one
  buffer.  A separable cosine mode decays by the closed-form factor of each
  axis, 1 / (1 + c2 + 4 c1 sin^2(pi m / 2n)), at every value.
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
        start, bad0 = m.mass_balance()  # FP64 sums on the device of the store's own values
        m.run(dt, steps, W.DIFFUSION)
        end, bad = m.mass_balance()
    assert bad0 == 0 and bad == 0
    c2 = lam * dt / 3.0
    expect = start * (1.0 + c2) ** (-3 * steps)
    assert np.max(np.abs(end - expect) / expect) <= tol


def test_store_above_4_gib_cosine_mode_decays_at_its_eigenvalue():
    nx, ny, nz, S = 416, 416, 400, 8
    modes = (1, 2, 3)
    D = np.array([1e5, 3e4, 1e4, 3e3, 1e3, 300.0, 100.0, 30.0])
    lam = np.array([0.1, 0.01, 0.5, 0.001, 1.0, 0.05, 0.2, 0.003])
    dt, dx = 0.01, 10.0
    cx, cy, cz = [np.cos(np.pi * m * (np.arange(n) + 0.5) / n) for m, n in zip(modes, (nx, ny, nz))]
    amp = np.arange(1.0, S + 1.0)
    rho0 = (cz[:, None, None, None] * cy[None, :, None, None] * cx[None, None, :, None]
            * amp[None, None, None, :]).reshape(-1)
    assert rho0.size * 8 > 2 ** 32
    case = W.make_case(nx, ny, nz, S, dx, D, lam, dt, 1, W.DIFFUSION, rho0)
    with setup_case(case, FP64) as m:
        plans = m.sweep_plans()
        assert plans[2][0] == 64, plans  # segmented z, one segment per wave
        m.run(dt, 1, W.DIFFUSION)
        got = m.download().reshape(-1, S)
    c1 = D * dt / dx ** 2
    c2 = lam * dt / 3.0
    f = np.ones(S)
    for mo, n in zip(modes, (nx, ny, nz)):
        f /= 1.0 + c2 + 4.0 * c1 * np.sin(np.pi * mo / (2 * n)) ** 2
    expect = rho0.reshape(-1, S) * f
    err = np.max(np.abs(got - expect), axis=0) / amp
    assert np.all(err < 1e-13), err
