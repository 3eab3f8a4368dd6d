"""The measured configurations at their own sizes, through the C ABI.

- 400^3 x 8 with 1e6 agents in a Gaussian cluster (the precision form's
  workload, biofvm_amd.workloads.precision_400):
  * the agent phase alone on the FP64 store, two steps, against the oracle's
    agent step (BioFVM_basic_agent.cpp:229-238 in list order) on the agent
    voxels, bit for bit, every other voxel unchanged;
  * the FP32 store against the FP64 store over the whole 1000-step run
    (agents + LOD), checked every 250 steps: max relative error of the values
    above 1e-3 of the largest, max normwise error and the per-substrate mass
    (bfvm_mass_balance) all within 1e-5.
- 64 uneven microenvironments (edges from {32, ..., 128}, 1-6 substrates,
  0-2e4 agents; workloads.ensemble_64) advanced together with one HIP graph
  replay per step: every member against the oracle after 3 steps (FP64 store
  1e-12, FP32 1e-5), and every member after 100 steps against the same
  member advanced alone (bfvm_run), bit for bit.
- 128^3 x 2 with 2e4 agents, 500 steps, checkpoints every 50
  (workloads.differentiation_128): the FP32 and FP64 adjoint gradients
  (initial field, D, lambda, secretion and uptake rates) on 8 random joint
  directions against 4th-order central differences of the FP64 loss
  (h = 1e-4 relative, Richardson: (8 (J(h) - J(-h)) - (J(2h) - J(-2h))) / 12h,
  truncation O(h^4); J summed with math.fsum): FP64 adjoint within 1e-6 of
  |FD|, FP32 adjoint within 1e-4 of |FD|.
"""
import math

import numpy as np
import pytest

import oracle
from biofvm_amd import CELLS, DIFFUSION, FP32, FP64, Ensemble, setup_case
from biofvm_amd import workloads as W

pytestmark = pytest.mark.gpu


def rel_err(got, ref):
    return float(np.max(np.abs(got - ref)) / max(np.max(np.abs(ref)), 1e-300))


def test_c3_agent_phase_at_full_size_is_bit_exact():
    case = W.precision_400(steps=2)
    S, dt, dx = case["S"], case["dt"], case["dx"]
    with setup_case(case, FP64) as m:
        vox = m.set_agent_rates(case["agent_pos"], case["agent_volume"], case["agent_secretion"],
                                case["agent_saturation"], case["agent_uptake"], dt)
        m.run(dt, case["steps"], CELLS)
        got = m.download().reshape(-1, S)
    rho0 = case["rho0"].reshape(-1, S)
    used = np.unique(vox)
    t1, t2 = oracle.agent_constants(S, dt, dx ** 3, case["agent_volume"], case["agent_secretion"],
                                    case["agent_saturation"], case["agent_uptake"])
    compact = np.ascontiguousarray(rho0[used]).reshape(-1)
    local = np.searchsorted(used, vox).astype(np.int32)
    for _ in range(case["steps"]):
        oracle.agent_step(S, compact, local, t1, t2)
    np.testing.assert_array_equal(got[used], compact.reshape(-1, S))
    others = np.ones(rho0.shape[0], dtype=bool)
    others[used] = False
    assert np.array_equal(got[others], rho0[others])
    assert used.size > 500_000  # the cluster's voxels with agents


def test_c3_fp32_store_tracks_the_fp64_store_for_1000_steps():
    case = W.precision_400()
    S, dt, ph = case["S"], case["dt"], case["phases"]
    m32, m64 = setup_case(case, FP32), setup_case(case, FP64)
    try:
        for n in range(250, 1001, 250):
            m32.run(dt, 250, ph)
            m64.run(dt, 250, ph)
            mass32, bad32 = m32.mass_balance()
            mass64, bad64 = m64.mass_balance()
            assert bad32 == 0 and bad64 == 0
            a, b = m32.download(), m64.download()
            diff = np.abs(a - b)
            scale = np.abs(b)
            big = scale > 1e-3 * scale.max()
            elem = float((diff[big] / scale[big]).max())
            norm = float(diff.max() / scale.max())
            mass = float(np.max(np.abs(mass32 - mass64) / np.abs(mass64)))
            print(f"step {n}: elementwise {elem:.3e} normwise {norm:.3e} mass {mass:.3e}", flush=True)
            assert elem <= 1e-5 and norm <= 1e-5 and mass <= 1e-5, (n, elem, norm, mass)
            del a, b, diff, scale, big
    finally:
        m32.close()
        m64.close()


@pytest.fixture(scope="module")
def c4():
    cases = W.ensemble_64(steps=3)
    return cases, [oracle.oracle_run(c) for c in cases]


def _ensemble(cases, precision):
    members = [setup_case(c, precision) for c in cases]
    return members, Ensemble(members)


@pytest.mark.parametrize("precision,tol", [(FP64, 1e-12), (FP32, 1e-5)], ids=["fp64", "fp32"])
def test_c4_all_64_members_in_one_graph_against_oracle(c4, precision, tol):
    cases, refs = c4
    assert len(cases) == 64
    members, ens = _ensemble(cases, precision)
    try:
        ens.run(cases[0]["dt"], cases[0]["steps"], cases[0]["phases"], graph=True)
        for i, (m, ref) in enumerate(zip(members, refs)):
            err = rel_err(m.download(), ref)
            assert err <= tol, (i, err)
    finally:
        ens.close()
        for m in members:
            m.close()


@pytest.mark.parametrize("precision", [FP64, FP32], ids=["fp64", "fp32"])
def test_c4_ensemble_equals_each_member_alone_over_100_steps(c4, precision):
    cases, _ = c4
    steps, dt, ph = 100, cases[0]["dt"], cases[0]["phases"]
    members, ens = _ensemble(cases, precision)
    try:
        ens.run(dt, steps, ph, graph=True)
        together = [m.download() for m in members]
    finally:
        ens.close()
        for m in members:
            m.close()
    for i, c in enumerate(cases):
        with setup_case(c, precision) as m:
            m.run(dt, steps, ph)
            np.testing.assert_array_equal(together[i], m.download(), err_msg=f"member {i}")


KEYS = ["rho0", "D", "lam", "agent_secretion", "agent_uptake"]
GRAD = {"rho0": "grad_rho0", "D": "grad_D", "lam": "grad_lambda", "agent_secretion": "grad_secretion",
        "agent_uptake": "grad_uptake"}


def test_c5_adjoint_at_full_size_against_fourth_order_differences():
    hidden = W.differentiation_128()
    dt, ph, steps = hidden["dt"], hidden["phases"], hidden["steps"]
    assert (hidden["nx"], hidden["S"], len(hidden["agent_volume"]), steps) == (128, 2, 20_000, 500)
    rng = np.random.default_rng(12)
    with setup_case(hidden, FP64) as m:
        m.run(dt, steps, ph)
        target = m.download()
    case = dict(hidden)  # the start of a fit: D, lambda and secretion off by up to 35 %
    for k, shape in (("D", hidden["S"]), ("lam", hidden["S"]), ("agent_secretion", hidden["agent_secretion"].shape)):
        case[k] = np.asarray(hidden[k]) * np.exp(rng.uniform(-0.3, 0.3, shape))

    grads = {}
    for precision in (FP32, FP64):
        with setup_case(case, precision) as m:
            m.tape_record(dt, steps, ph, 50)
            m.tape_seed_l2(target)
            grads[precision] = m.tape_backward(len(case["agent_volume"]))

    def J(ds, h):
        c = dict(case)
        for k, d in ds.items():
            c[k] = np.asarray(case[k]) + h * d
        with setup_case(c, FP64) as mm:
            mm.run(dt, steps, ph)
            r = mm.download() - target
        return 0.5 * math.fsum(r * r)

    h = 1e-4
    worst = {FP32: 0.0, FP64: 0.0}
    for _ in range(8):
        ds = {k: rng.normal(size=np.shape(case[k])) * np.asarray(case[k]) for k in KEYS}
        fd = (8 * (J(ds, h) - J(ds, -h)) - (J(ds, 2 * h) - J(ds, -2 * h))) / (12 * h)
        for precision, g in grads.items():
            ad = math.fsum(np.concatenate([np.asarray(g[GRAD[k]]).reshape(-1) * ds[k].reshape(-1) for k in KEYS]))
            worst[precision] = max(worst[precision], abs(ad - fd) / abs(fd))
    print(f"max |ad - fd| / |fd|: fp32 {worst[FP32]:.3e}, fp64 {worst[FP64]:.3e}", flush=True)
    assert worst[FP64] <= 1e-6, worst
    assert worst[FP32] <= 1e-4, worst
