"""Steps replayed from a HIP graph (bfvm_set_graph_steps) against the same
steps launched kernel by kernel.

Two handles on the same inputs, one replaying groups of 4 steps from a graph,
one launching every kernel on its own, go through the same sequence of runs
with downloads (and so synchronizations) between them; every run's result
must be the same bits.  Between runs the inputs a graph bakes in change — the
coefficients (new tables), the bulk fields, the agents, the x launch shape,
the phases, dt — so the graph is captured again, and reused where nothing
changed.  The first run is also checked against the oracle.
"""
import numpy as np
import pytest

import oracle
from biofvm_amd import FP32, FP64, BfvmError, setup_case
from biofvm_amd import workloads as W

pytestmark = pytest.mark.gpu

TOL = {FP64: 1e-12, FP32: 1e-5}


def rel_err(got, ref):
    return float(np.max(np.abs(got - ref)) / max(np.max(np.abs(ref)), 1e-300))


def case_with_sources(rng, nx=30, ny=20, nz=17, S=3, steps=11):
    nvox = nx * ny * nz
    case = W.make_case(nx, ny, nz, S, 10.0, 10.0 ** rng.uniform(2, 5, S), 10.0 ** rng.uniform(-3, 0, S), 0.01,
                       steps, W.ALL, rng.uniform(0, 1, nvox * S))
    case.update(W.random_bulk(rng, nvox, S))
    case.update(W.random_agents(rng, nvox // 4, S, nx, ny, nz, 10.0, share_voxel=4))
    idx = W.outer_face_voxels(nx, ny, nz).astype(np.int32)
    case.update(dir_idx=idx, dir_val=rng.uniform(0, 2, idx.size * S), dir_mask=(np.arange(S) % 2 == 0).astype(np.uint8))
    return case


@pytest.mark.parametrize("precision", [FP64, FP32], ids=["fp64", "fp32"])
def test_graph_replays_equal_kernel_by_kernel_steps(precision):
    rng = np.random.default_rng(811)
    case = case_with_sources(rng)
    dt, S, nvox = case["dt"], case["S"], case["nx"] * case["ny"] * case["nz"]
    g, e = setup_case(case, precision), setup_case(case, precision)
    try:
        g.graph_steps = 4
        e.graph_steps = 0
        assert (g.graph_steps, e.graph_steps) == (4, 0)

        def both(fn):
            for m in (g, e):
                fn(m)

        def same(label):
            a, b = g.download(), e.download()
            np.testing.assert_array_equal(a, b, err_msg=label)
            return a

        both(lambda m: m.run(dt, 11, W.ALL))  # 1 eager + 2 replays of 4 + 2 eager
        assert rel_err(same("first run"), oracle.oracle_run(case)) <= TOL[precision]
        both(lambda m: m.run(dt, 6, W.ALL))  # the same graph again
        same("graph reused")
        D2 = np.asarray(case["D"]) * 1.5
        both(lambda m: m.set_coefficients(D2, case["lam"]))
        both(lambda m: m.run(dt, 9, W.ALL))
        same("new coefficients")
        bulk = W.random_bulk(rng, nvox, S)
        both(lambda m: m.set_bulk(bulk["bulk_supply"], bulk["bulk_target"], bulk["bulk_uptake"]))
        both(lambda m: m.run(dt, 7, W.ALL))
        same("new bulk fields")
        ag = W.random_agents(rng, nvox // 5, S, case["nx"], case["ny"], case["nz"], 10.0)
        both(lambda m: m.set_agent_rates(ag["agent_pos"], ag["agent_volume"], ag["agent_secretion"],
                                         ag["agent_saturation"], ag["agent_uptake"], dt))
        both(lambda m: m.run(dt, 8, W.ALL))
        same("new agents")
        both(lambda m: m.set_sweep_plan(0, 16, 8, 0))
        both(lambda m: m.run(dt, 10, W.ALL))
        same("new x launch shape")
        both(lambda m: m.run(dt, 13, W.DIFFUSION))
        same("diffusion only")
        both(lambda m: m.run(2 * dt, 5, W.ALL))
        same("new dt")
        both(lambda m: m.timed_run(dt, 12, W.ALL))
        same("timed run")
        both(lambda m: m.run(dt, 3, W.ALL))  # too few steps for a graph
        same("short run")
    finally:
        g.close()
        e.close()


def test_graph_steps_argument_checks():
    rng = np.random.default_rng(5)
    case = case_with_sources(rng, 12, 10, 8, 2, 3)
    with setup_case(case, FP64) as m:
        assert m.graph_steps == 8
        with pytest.raises(BfvmError):
            m.graph_steps = -1
        m.graph_steps = 1  # a graph of one step is never worth it: kernel by kernel
        m.run(case["dt"], 5, W.ALL)
        assert np.all(np.isfinite(m.download()))
