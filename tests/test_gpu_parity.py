"""GPU parity: the HIP path through the C ABI against the golden fixtures and the oracle.

Tolerances (stated per the precision of the device store):
  FP64  max|gpu - ref| <= 1e-12 * max|ref|  (FMA and reciprocal rounding only)
  FP32  max|gpu - ref| <= 1e-5  * max|ref|  (FP32 store rounded once per sweep; FP64 arithmetic)
"""
import numpy as np
import pytest

import gen_golden
import oracle
from biofvm_amd import FP32, FP64, BfvmError, DeviceMicroenvironment, run_case, setup_case
from biofvm_amd import workloads as W

pytestmark = pytest.mark.gpu

TOL = {FP64: 1e-12, FP32: 1e-5}
FIXTURES = gen_golden.names()


def rel_err(got, ref):
    scale = max(np.max(np.abs(ref)), 1e-300)
    return float(np.max(np.abs(got - ref)) / scale)


@pytest.mark.parametrize("precision", [FP64, FP32], ids=["fp64", "fp32"])
@pytest.mark.parametrize("name", FIXTURES)
def test_fixture_parity(name, precision):
    case, expected, voxels = gen_golden.load(name)
    if voxels is not None:
        with DeviceMicroenvironment(case["nx"], case["ny"], case["nz"], case["S"], case["dx"],
                                    precision) as m:
            got_vox = m.set_agent_rates(case["agent_pos"], case["agent_volume"],
                                        case["agent_secretion"], case["agent_saturation"],
                                        case["agent_uptake"], case["dt"])
        np.testing.assert_array_equal(got_vox, voxels)
    got = run_case(case, precision)
    err = rel_err(got, expected)
    assert err <= TOL[precision], f"{name}: rel err {err:.3e}"


def _random_case(rng, sources):
    nx, ny, nz = (int(v) for v in rng.integers(1, 40, size=3))
    S = int(rng.integers(1, 9))
    D = 10.0 ** rng.uniform(2, 5, S)
    lam = 10.0 ** rng.uniform(-3, 0, S)
    case = W.make_case(nx, ny, nz, S, 10.0, D, lam, 0.01, int(rng.integers(1, 5)),
                       W.ALL if sources else W.DIFFUSION, rng.uniform(0, 1, nx * ny * nz * S))
    if sources:
        case.update(W.random_bulk(rng, nx * ny * nz, S))
        case.update(W.random_agents(rng, int(rng.integers(0, 300)), S, nx, ny, nz, 10.0,
                                    share_voxel=int(rng.integers(0, 20))))
        idx = W.outer_face_voxels(nx, ny, nz)
        case.update(dir_idx=idx, dir_val=rng.uniform(0, 2, idx.size * S))
        if rng.uniform() < 0.5:
            case["dir_mask"] = (rng.uniform(size=S) < 0.5).astype(np.uint8)
    return case


@pytest.mark.parametrize("seed", range(16))
def test_random_shapes_against_oracle(seed):
    rng = np.random.default_rng(1000 + seed)
    case = _random_case(rng, sources=seed % 2 == 1)
    ref = oracle.oracle_run(case)
    for precision in (FP64, FP32):
        err = rel_err(run_case(case, precision), ref)
        assert err <= TOL[precision], (seed, precision, err)


@pytest.mark.parametrize("n", [1, 2, 3, 63, 64, 65, 127, 200, 401])
def test_axis_lengths_each_axis(n):
    # a long line along each axis in turn, with a short cross-section
    rng = np.random.default_rng(n)
    for dims in [(n, 3, 2), (3, n, 2), (2, 3, n)]:
        case = W.make_case(*dims, 5, 10.0, [1e5, 1e2, 3e4, 7e3, 1e3], [0.1, 1.0, 0.01, 0.3, 0.001],
                           0.01, 2, W.DIFFUSION, rng.uniform(0, 1, int(np.prod(dims)) * 5))
        ref = oracle.oracle_run(case)
        for precision in (FP64, FP32):
            err = rel_err(run_case(case, precision), ref)
            assert err <= TOL[precision], (dims, precision, err)


def test_dirichlet_mask_parity():
    rng = np.random.default_rng(77)
    nx, ny, nz, S = 20, 17, 12, 4
    idx = W.outer_face_voxels(nx, ny, nz)
    val = np.zeros((idx.size, S))
    val[:, 0] = 1.0
    case = W.make_case(nx, ny, nz, S, 10.0, [1e5, 1e3, 5e4, 1e4], [0.1, 0.01, 1, 0.5], 0.01, 5,
                       W.DIFFUSION, rng.uniform(0, 1, nx * ny * nz * S), dir_idx=idx,
                       dir_val=val.reshape(-1), dir_mask=np.array([1, 0, 0, 0], np.uint8))
    ref = oracle.oracle_run(case)
    assert rel_err(run_case(case, FP64), ref) <= TOL[FP64]


def test_agents_given_as_constants_match_rates():
    case, expected, voxels = gen_golden.load("g4_agents")
    dV = case["dx"] ** 3
    t1, t2 = oracle.agent_constants(case["S"], case["dt"], dV, case["agent_volume"],
                                    case["agent_secretion"], case["agent_saturation"],
                                    case["agent_uptake"])
    with setup_case({k: v for k, v in case.items() if not k.startswith("agent_")}, FP64) as m:
        m.set_agents(voxels, t1, t2)
        m.run(case["dt"], case["steps"], case["phases"])
        got = m.download()
    assert rel_err(got, expected) <= TOL[FP64]


def test_non_uniform_mesh_is_refused_and_state_kept():
    rng = np.random.default_rng(5)
    rho = rng.uniform(0, 1, 6 * 5 * 4)
    with DeviceMicroenvironment(6, 5, 4, 1, 10.0, FP64, dy=12.0) as m:
        m.set_coefficients([1e3], [0.1])
        m.upload(rho)
        with pytest.raises(BfvmError) as ei:
            m.step(0.01, W.DIFFUSION)
        assert ei.value.code == -4
        np.testing.assert_array_equal(m.download(), rho)


def test_argument_errors():
    with DeviceMicroenvironment(4, 4, 4, 2, 10.0, FP64) as m:
        with pytest.raises(BfvmError):
            m.step(0.01, W.DIFFUSION)  # coefficients not set
        with pytest.raises(BfvmError):
            m.set_dirichlet([64], [1.0, 1.0])
        with pytest.raises(BfvmError):
            m.set_agent_rates([1.0, 1.0, 40.5], [1.0], [1, 1], [1, 1], [1, 1], 0.01)
        with pytest.raises(BfvmError):
            m.set_agents([-1], [0.0, 0.0], [1.0, 1.0])


def test_clearing_sources():
    rng = np.random.default_rng(8)
    case = W.make_case(6, 6, 6, 2, 10.0, [1e3, 1e4], [0.1, 0.2], 0.01, 2, W.ALL,
                       rng.uniform(0, 1, 6 ** 3 * 2))
    ref = oracle.oracle_run(case)
    full = dict(case, **W.random_bulk(rng, 6 ** 3, 2), **W.random_agents(rng, 30, 2, 6, 6, 6, 10.0))
    with setup_case(full, FP64) as m:
        m.set_bulk(None, None, None)
        m.set_agents([], [], [])
        m.run(case["dt"], case["steps"], case["phases"])
        assert rel_err(m.download(), ref) <= TOL[FP64]


def test_step_by_step_equals_run_and_timed_run():
    case, expected, _ = gen_golden.load("g8_shell")
    with setup_case(case, FP64) as m:
        for _ in range(case["steps"]):
            m.step(case["dt"], case["phases"])
        a = m.download()
    with setup_case(case, FP64) as m:
        ms, per = m.timed_run(case["dt"], case["steps"], case["phases"], per_kernel=True)
        b = m.download()
    # the bulk update and the Dirichlet overwrites before the x and y sweeps run inside the x sweep
    assert ms > 0 and per["sweep_x"][1] == case["steps"] and per["dirichlet"][1] == case["steps"]
    assert per["bulk"][1] == 0 and per["agents"][1] == case["steps"]
    np.testing.assert_array_equal(a, b)
    assert rel_err(a, expected) <= TOL[FP64]
    with setup_case(case, FP64) as m:
        m.fused_sources = False
        ms, per = m.timed_run(case["dt"], case["steps"], case["phases"], per_kernel=True)
        assert per["dirichlet"][1] == 3 * case["steps"] and per["bulk"][1] == case["steps"]
        np.testing.assert_array_equal(m.download(), a)


@pytest.mark.parametrize("precision", [FP64, FP32], ids=["fp64", "fp32"])
def test_autotune_keeps_state_and_parity(precision):
    case, expected, _ = gen_golden.load("g6_ensemble")
    with setup_case(case, precision) as m:
        before = m.download()
        m.autotune(case["dt"])
        np.testing.assert_array_equal(m.download(), before)
        for axis, (cols, L, streaming) in enumerate(m.sweep_plans()):
            assert 4 <= L <= 32 and streaming in (0, 1, 2)
            if streaming == 2:
                assert axis == 0 and cols % case["S"] == 0
            else:
                assert cols in (16, 32, 64)
        m.run(case["dt"], case["steps"], case["phases"])
        assert rel_err(m.download(), expected) <= TOL[precision]


@pytest.mark.parametrize("precision", [FP64, FP32], ids=["fp64", "fp32"])
def test_many_substrates_keep_the_segmented_kernels(precision):
    # 24 substrates, one diffusing fast (its end corrections run the whole
    # line): the x and y tables take 69 and 57 KiB of LDS per block
    rng = np.random.default_rng(29)
    nx, ny, nz, S = 128, 96, 24, 24
    D = np.exp(rng.uniform(np.log(1e2), np.log(1e5), S))
    D[3] = 1e5
    lam = np.exp(rng.uniform(np.log(1e-3), np.log(1.0), S))
    case = W.make_case(nx, ny, nz, S, 10.0, D, lam, 0.01, 2, W.DIFFUSION, rng.uniform(0, 1, nx * ny * nz * S))
    ref = oracle.oracle_run(case)
    with setup_case(case, precision) as m:
        plans = m.sweep_plans()
        assert plans[0][0] > 0 and plans[1][0] > 0, plans  # not the one-thread-per-line fallback
        m.run(case["dt"], case["steps"], case["phases"])
        assert rel_err(m.download(), ref) <= TOL[precision]


@pytest.fixture(scope="module")
def plan_case():
    # large enough that resident (streaming) blocks loop over several tiles on
    # every axis: x and y have 80 planes of 3*128/cols or 3*112/cols tiles
    rng = np.random.default_rng(77)
    S = 3
    case = W.make_case(112, 128, 80, S, 10.0, [1e5, 2e3, 4e4], [0.1, 0.01, 1.0], 0.01, 2,
                       W.DIFFUSION, rng.uniform(0, 1, 112 * 128 * 80 * S))
    return case, oracle.oracle_run(case)


def _plans(n, precision):
    for cols in (16, 32, 64):
        for L in (4, 8, 12, 16, 20, 24, 32):
            if cols * -(-n // L) > 1024 or (L > 8 and L >= 2 * n):
                continue
            yield cols, L, False
            if L in (8, 12, 16) or (L == 24 and precision == FP32):
                yield cols, L, True


@pytest.mark.parametrize("precision", [FP64, FP32], ids=["fp64", "fp32"])
def test_every_launch_shape_against_oracle(plan_case, precision):
    case, ref = plan_case
    dims = (case["nx"], case["ny"], case["nz"])
    with setup_case(case, precision) as m:
        start = case["rho0"]
        for axis in range(3):
            for cols, L, streaming in _plans(dims[axis], precision):
                m.upload(start)
                m.set_sweep_plan(axis, cols, L, streaming)
                assert m.sweep_plans()[axis] == (cols, L, int(streaming))
                m.run(case["dt"], case["steps"], case["phases"])
                err = rel_err(m.download(), ref)
                assert err <= TOL[precision], (axis, cols, L, streaming, err)
            m.set_sweep_plan(axis, 0, 0)


@pytest.mark.parametrize("precision", [FP64, FP32], ids=["fp64", "fp32"])
def test_xcd_spread_never_changes_results(plan_case, precision):
    # the spread only renumbers which workgroup solves which tile
    case, ref = plan_case
    with setup_case(case, precision) as m:
        outs = []
        for spread in (0, 1, 3, 16, 64, 1000, -1):
            for axis in range(3):
                m.set_sweep_spread(axis, spread)
                assert m.sweep_spread(axis) == spread
            m.upload(case["rho0"])
            m.run(case["dt"], case["steps"], case["phases"])
            outs.append(m.download())
        for o in outs[1:]:
            np.testing.assert_array_equal(o, outs[0])
        assert rel_err(outs[0], ref) <= TOL[precision]
        with pytest.raises(BfvmError):
            m.set_sweep_spread(0, -2)


ROW_SHAPES = [(112, 128, 80, 3), (37, 21, 5, 2), (400, 6, 3, 8), (2, 9, 7, 1), (70, 3, 2, 5)]


@pytest.mark.parametrize("precision", [FP64, FP32], ids=["fp64", "fp32"])
def test_row_tile_x_sweep_equals_column_tiles(precision):
    # the LDS row-tile x sweep runs the same arithmetic as the column-tile
    # kernel with the same segment length: bit-identical stores, for every
    # row-tile shape the library accepts (partial last tiles and segments)
    rng = np.random.default_rng(91)
    for nx, ny, nz, S in ROW_SHAPES:
        case = W.make_case(nx, ny, nz, S, 10.0, 10.0 ** rng.uniform(2, 5, S), 10.0 ** rng.uniform(-3, 0, S),
                           0.01, 2, W.DIFFUSION, rng.uniform(0, 1, nx * ny * nz * S))
        ref = oracle.oracle_run(case)
        tried = 0
        with setup_case(case, precision) as m:
            for L in (4, 8, 12, 16, 24, 32):
                base = None
                for cols in (16, 32, 64):
                    try:
                        m.set_sweep_plan(0, cols, L, 0)
                    except BfvmError:
                        continue
                    m.upload(case["rho0"])
                    m.run(case["dt"], case["steps"], case["phases"])
                    base = m.download()
                    break
                for R in (1, 2, 4, 8, 16, 32, 64):
                    try:
                        m.set_sweep_plan(0, R * S, L, 2)
                    except BfvmError:
                        continue
                    assert m.sweep_plans()[0] == (R * S, L, 2)
                    m.upload(case["rho0"])
                    m.run(case["dt"], case["steps"], case["phases"])
                    got = m.download()
                    tried += 1
                    assert rel_err(got, ref) <= TOL[precision], (nx, ny, nz, S, R, L)
                    if base is not None:
                        np.testing.assert_array_equal(got, base, err_msg=str((nx, ny, nz, S, R, L)))
        assert tried > 0, (nx, ny, nz, S)


# substrate counts whose nodes split into whole 16-byte pieces (S in {4, 8,
# 16}; line lengths below, at and past the segment lengths, partial last
# segments and tiles), and FP32 rows of a multiple of 4 nodes with any S (a
# lane's 16 bytes spanning two nodes, lanes past a wave's whole segments)
WHOLE_LINE_SHAPES = [(400, 6, 3, 8), (37, 5, 4, 8), (61, 7, 3, 4), (20, 9, 5, 16), (3, 11, 2, 8),
                     (130, 4, 3, 2), (48, 13, 2, 4), (64, 5, 3, 3), (32, 7, 2, 1), (48, 6, 4, 5),
                     (128, 3, 2, 6), (36, 4, 3, 2), (40, 5, 3, 12), (4, 9, 2, 7)]


@pytest.mark.parametrize("precision", [FP64, FP32], ids=["fp64", "fp32"])
def test_whole_line_x_sweep_equals_column_tiles(precision):
    # the x sweep with whole-cache-line loads (kernel 3) runs the column
    # kernel's arithmetic with the same segment length: bit-identical stores
    # for every shape the library accepts, and within tolerance of the oracle
    rng = np.random.default_rng(93)
    tried = 0
    for nx, ny, nz, S in WHOLE_LINE_SHAPES:
        case = W.make_case(nx, ny, nz, S, 10.0, 10.0 ** rng.uniform(2, 5, S), 10.0 ** rng.uniform(-3, 0, S),
                           0.01, 3, W.DIFFUSION, rng.uniform(0, 1, nx * ny * nz * S))
        ref = oracle.oracle_run(case)
        with setup_case(case, precision) as m:
            for L in (8, 16, 20):
                base = None
                for cols in (16, 32, 64):
                    try:
                        m.set_sweep_plan(0, cols, L, 0)
                    except BfvmError:
                        continue
                    m.upload(case["rho0"])
                    m.run(case["dt"], case["steps"], case["phases"])
                    base = m.download()
                    break
                for R in (1, 2, 3, 4, 5, 6, 7, 8, 16, 32, 64, 128):
                    try:
                        m.set_sweep_plan(0, R * S, L, 3)
                    except BfvmError:
                        continue
                    assert m.sweep_plans()[0] == (R * S, L, 3)
                    for spread in (0, -1):
                        m.set_sweep_spread(0, spread)
                        m.upload(case["rho0"])
                        m.run(case["dt"], case["steps"], case["phases"])
                        got = m.download()
                        tried += 1
                        assert rel_err(got, ref) <= TOL[precision], (nx, ny, nz, S, R, L)
                        if base is not None:
                            np.testing.assert_array_equal(got, base, err_msg=str((nx, ny, nz, S, R, L, spread)))
    assert tried > 0


def test_whole_line_x_sweep_refuses_other_substrate_counts():
    # S = 3 does not split into 16-byte pieces: FP32 rows of 42 nodes are not
    # whole chunks, and the FP64 form needs S in {4, 8, 16}; y and z take no
    # such kernel
    with DeviceMicroenvironment(42, 8, 6, 3, 10.0, FP32) as m:
        m.set_coefficients([1e3, 1e4, 1e5], [0.1, 0.2, 0.3])
        with pytest.raises(BfvmError):
            m.set_sweep_plan(0, 12, 16, 3)
    with DeviceMicroenvironment(40, 8, 6, 3, 10.0, FP64) as m:
        m.set_coefficients([1e3, 1e4, 1e5], [0.1, 0.2, 0.3])
        with pytest.raises(BfvmError):
            m.set_sweep_plan(0, 12, 16, 3)
    with DeviceMicroenvironment(40, 8, 6, 3, 10.0, FP32) as m:
        m.set_coefficients([1e3, 1e4, 1e5], [0.1, 0.2, 0.3])
        m.set_sweep_plan(0, 12, 16, 3)
    with DeviceMicroenvironment(40, 40, 40, 8, 10.0, FP32) as m:
        m.set_coefficients([1e3] * 8, [0.1] * 8)
        m.set_sweep_plan(0, 32, 16, 3)
        for axis in (1, 2):
            with pytest.raises(BfvmError):
                m.set_sweep_plan(axis, 32, 16, 3)


@pytest.mark.parametrize("precision", [FP64, FP32], ids=["fp64", "fp32"])
def test_cache_policy_never_changes_results(precision):
    rng = np.random.default_rng(41)
    nx, ny, nz, S = 40, 36, 30, 4
    case = W.make_case(nx, ny, nz, S, 10.0, [1e5, 2e3, 3e4, 500.0], [0.1, 0.3, 0.01, 1.0], 0.01, 3,
                       W.DIFFUSION, rng.uniform(0, 1, nx * ny * nz * S))
    outs = []
    for nt in (0, 1):
        with setup_case(case, precision) as m:
            m.set_sweep_plan(0, 4 * S, 8, 3)  # the whole-cache-line x kernel, 4 rows per block
            for axis in (1, 2):
                m.set_sweep_plan(axis, 32, 16)  # row pieces of 128 or 256 bytes
            for axis in (0, 1, 2):
                m.set_sweep_cache_policy(axis, nt)
            assert m.sweep_cache_policies() == [nt, nt, nt]
            m.run(case["dt"], case["steps"], case["phases"])
            outs.append(m.download())
        with pytest.raises(BfvmError):
            with setup_case(case, precision) as m:
                m.set_sweep_plan(0, 32, 8)  # column tiles: x keeps the default policy
                m.set_sweep_cache_policy(0, 1)
    np.testing.assert_array_equal(outs[1], outs[0])
    assert rel_err(outs[0], oracle.oracle_run(case)) <= TOL[precision]


@pytest.mark.parametrize("precision", [FP64, FP32], ids=["fp64", "fp32"])
def test_cache_policy_of_every_shape_never_changes_results(precision):
    # every y and z shape whose tiles' row pieces are whole 128-byte lines,
    # one block per tile and streaming: non-temporal = default, bit for bit
    rng = np.random.default_rng(43)
    nx, ny, nz, S = 24, 44, 37, 4
    case = W.make_case(nx, ny, nz, S, 10.0, [1e5, 2e3, 3e4, 500.0], [0.1, 0.3, 0.01, 1.0], 0.01, 2,
                       W.DIFFUSION, rng.uniform(0, 1, nx * ny * nz * S))
    elem = 8 if precision == FP64 else 4
    tried = 0
    with setup_case(case, precision) as m:
        for axis, n in ((1, ny), (2, nz)):
            for cols, L, streaming in _plans(n, precision):
                if cols * elem < 128:
                    continue
                outs = []
                for nt in (0, 1):
                    m.upload(case["rho0"])
                    m.set_sweep_plan(axis, cols, L, streaming)
                    m.set_sweep_cache_policy(axis, nt)
                    assert m.sweep_cache_policies()[axis] == nt
                    m.run(case["dt"], case["steps"], case["phases"])
                    outs.append(m.download())
                np.testing.assert_array_equal(outs[1], outs[0], err_msg=str((axis, cols, L, streaming)))
                tried += 1
            m.set_sweep_plan(axis, 0, 0)
        # the whole-cache-line x kernel, 1-8 rows per block
        for rows in (1, 2, 3, 5, 8):
            for L in (8, 16):
                outs = []
                for nt in (0, 1):
                    m.upload(case["rho0"])
                    m.set_sweep_plan(0, rows * S, L, 3)
                    m.set_sweep_cache_policy(0, nt)
                    assert m.sweep_cache_policies()[0] == nt
                    m.run(case["dt"], case["steps"], case["phases"])
                    outs.append(m.download())
                np.testing.assert_array_equal(outs[1], outs[0], err_msg=str((0, rows, L)))
                tried += 1
    assert tried >= 30


@pytest.mark.parametrize("precision", [FP64, FP32], ids=["fp64", "fp32"])
def test_autotuned_results_do_not_depend_on_timing(precision):
    # the autotuner keeps the segment length fixed, so two handles tuned
    # separately (whatever shapes their timings pick) give the same bits as
    # the untuned default shapes
    rng = np.random.default_rng(45)
    nx, ny, nz, S = 64, 48, 40, 3
    case = W.make_case(nx, ny, nz, S, 10.0, [1e5, 2e3, 3e4], [0.1, 0.3, 0.01], 0.01, 3, W.ALL,
                       rng.uniform(0, 1, nx * ny * nz * S))
    case.update(W.random_bulk(rng, nx * ny * nz, S))
    case.update(W.random_agents(rng, 400, S, nx, ny, nz, 10.0, share_voxel=6))
    idx = W.outer_face_voxels(nx, ny, nz)
    case.update(dir_idx=idx, dir_val=np.tile([1.0, 0.0, 0.5], idx.size))
    outs = []
    for tune in (False, True, True):
        with setup_case(case, precision) as m:
            if tune:
                m.autotune(case["dt"])
            m.run(case["dt"], case["steps"], case["phases"])
            outs.append(m.download())
    np.testing.assert_array_equal(outs[1], outs[0])
    np.testing.assert_array_equal(outs[2], outs[0])
    assert rel_err(outs[0], oracle.oracle_run(case)) <= TOL[precision]


def test_unrunnable_launch_shapes_are_refused():
    with DeviceMicroenvironment(40, 40, 300, 2, 10.0, FP64) as m:
        with pytest.raises(BfvmError):
            m.set_sweep_plan(2, 64, 4)  # 64 lines x 75 segments > 1024 threads
        with pytest.raises(BfvmError):
            m.set_sweep_plan(0, 32, 24, True)  # no FP64 streaming kernel with 24 nodes
        with pytest.raises(BfvmError):
            m.set_sweep_plan(0, 48, 8)
        with pytest.raises(BfvmError):
            m.set_sweep_plan(1, 2, 8, 2)  # row tiles are for the x sweep only
        with pytest.raises(BfvmError):
            m.set_sweep_plan(0, 3, 8, 2)  # lines per row tile must be a multiple of S


@pytest.mark.parametrize("precision", [FP64, FP32], ids=["fp64", "fp32"])
def test_plane_chunked_xy_sweeps_are_bit_identical(precision):
    # x then y per chunk of z planes, with Dirichlet faces (masked), bulk and agents
    rng = np.random.default_rng(61)
    nx, ny, nz, S = 30, 20, 17, 3
    case = W.make_case(nx, ny, nz, S, 10.0, [1e5, 2e3, 3e4], [0.1, 0.3, 0.01], 0.01, 3, W.ALL,
                       rng.uniform(0, 1, nx * ny * nz * S))
    case.update(W.random_bulk(rng, nx * ny * nz, S))
    case.update(W.random_agents(rng, 200, S, nx, ny, nz, 10.0, share_voxel=5))
    idx = W.outer_face_voxels(nx, ny, nz)
    case.update(dir_idx=idx, dir_val=rng.uniform(0, 2, idx.size * S), dir_mask=np.array([1, 0, 1], np.uint8))
    ref = oracle.oracle_run(case)
    with setup_case(case, precision) as m:
        m.fused_sources = False  # chunked sweeps run the separate source kernels
        m.run(case["dt"], case["steps"], case["phases"])
        whole = m.download()
    assert rel_err(whole, ref) <= TOL[precision]
    for chunk in (1, 4, 16, 17, 40):
        with setup_case(case, precision) as m:
            m.sweep_chunk = chunk
            assert m.sweep_chunk == chunk
            m.run(case["dt"], case["steps"], case["phases"])
            np.testing.assert_array_equal(m.download(), whole, err_msg=f"chunk {chunk}")


FUSED_SHAPES = [(30, 20, 17, 3, 3), (37, 21, 5, 2, 3), (64, 9, 7, 1, 3), (70, 3, 2, 5, 3), (45, 33, 1, 2, 2),
                (2, 9, 7, 1, 3)]


@pytest.mark.parametrize("precision", [FP64, FP32], ids=["fp64", "fp32"])
def test_fused_sources_equal_separate_kernels(precision):
    # bulk, agents (some stacked on one voxel) and masked Dirichlet faces, with
    # a voxel listed twice (its last entry wins): the x sweep with the bulk
    # update and Dirichlet overwrites fused in gives the bits of the separate
    # kernels, for every x launch shape
    rng = np.random.default_rng(303)
    for nx, ny, nz, S, dims in FUSED_SHAPES:
        nvox = nx * ny * nz
        case = W.make_case(nx, ny, nz, S, 10.0, 10.0 ** rng.uniform(2, 5, S), 10.0 ** rng.uniform(-3, 0, S), 0.01,
                           3, W.ALL, rng.uniform(0, 1, nvox * S))
        case.update(W.random_bulk(rng, nvox, S))
        case.update(W.random_agents(rng, 3 * nvox // 10 + 5, S, nx, ny, nz, 10.0, share_voxel=4))
        idx = np.concatenate([W.outer_face_voxels(nx, ny, nz), [nvox // 2, 0]]).astype(np.int32)
        mask = (np.arange(S) % 2 == 0).astype(np.uint8)
        case.update(dir_idx=idx, dir_val=rng.uniform(0, 2, idx.size * S), dir_mask=mask)
        if dims == 2:
            case["lod_dims"] = 2
        ref = oracle.oracle_run(case)
        for cols, L in [(0, 0), (16, 4), (32, 8), (16, 12), (64, 16), (32, 24), (32, 32)]:
            out = {}
            for fused in (False, True):
                with setup_case(case, precision) as m:
                    if cols:
                        try:
                            m.set_sweep_plan(0, cols, L, 0)
                        except BfvmError:
                            break
                    m.fused_sources = fused
                    ms, per = m.timed_run(case["dt"], case["steps"], case["phases"], per_kernel=True)
                    if fused:  # bulk and two of the Dirichlet overwrites inside the x sweep
                        steps = case["steps"]
                        assert per["bulk"][1] == 0 and per["agents"][1] == steps, (nx, ny, nz, S, cols, L)
                        assert per["dirichlet"][1] == (steps if dims == 3 else 0), (nx, ny, nz, S, cols, L)
                    out[fused] = m.download()
            if not out:
                continue
            assert rel_err(out[False], ref) <= TOL[precision], (nx, ny, nz, S, cols, L)
            np.testing.assert_array_equal(out[True], out[False], err_msg=str((nx, ny, nz, S, dims, cols, L)))


@pytest.mark.parametrize("precision", [FP64, FP32], ids=["fp64", "fp32"])
def test_gradients_against_oracle(precision):
    # gradient vectors of the device's densities after a few steps, against the
    # oracle's gradients of the same densities (FP64: bit-identical formula)
    rng = np.random.default_rng(70)
    for dims in [(17, 13, 9, 3), (1, 12, 7, 2), (40, 2, 3, 1)]:
        nx, ny, nz, S = dims
        case = W.make_case(nx, ny, nz, S, 10.0, 10.0 ** rng.uniform(2, 5, S), 10.0 ** rng.uniform(-3, 0, S),
                           0.01, 2, W.DIFFUSION, rng.uniform(0, 1, nx * ny * nz * S))
        with setup_case(case, precision) as m:
            m.run(case["dt"], case["steps"], case["phases"])
            rho = m.download()
            g = m.gradients()
        ref = oracle.gradients(nx, ny, nz, S, 10.0, rho)
        if precision == FP64:
            np.testing.assert_array_equal(g, ref, err_msg=str(dims))
        else:  # FP32 output of the FP64 difference of FP32 values
            assert np.max(np.abs(g - ref)) <= 1e-6 * np.max(np.abs(ref)) + 1e-12, dims


@pytest.mark.parametrize("precision", [FP64, FP32], ids=["fp64", "fp32"])
def test_lod2d_against_oracle(precision):
    rng = np.random.default_rng(81)
    for nx, ny, S in [(40, 33, 3), (1, 50, 2), (70, 1, 1), (300, 200, 2)]:
        case = W.make_case(nx, ny, 1, S, 10.0, 10.0 ** rng.uniform(2, 5, S), 10.0 ** rng.uniform(-3, 0, S),
                           0.01, 3, W.ALL, rng.uniform(0, 1, nx * ny * S))
        case.update(W.random_bulk(rng, nx * ny, S))
        case.update(W.random_agents(rng, 50, S, nx, ny, 1, 10.0, share_voxel=4))
        idx = W.outer_face_voxels(nx, ny, 1)
        case.update(dir_idx=idx, dir_val=rng.uniform(0, 1, idx.size * S))
        case["lod_dims"] = 2
        err = rel_err(run_case(case, precision), oracle.oracle_run(case))
        assert err <= TOL[precision], (nx, ny, S, err)


def test_lod2d_needs_a_flat_mesh():
    with DeviceMicroenvironment(5, 4, 3, 1, 10.0, FP64) as m:
        with pytest.raises(BfvmError):
            m.set_lod_dimensions(2)
        with pytest.raises(BfvmError):
            m.set_lod_dimensions(1)


@pytest.mark.parametrize("precision", [FP64, FP32], ids=["fp64", "fp32"])
def test_device_agent_setup_equals_host_setup(precision):
    # voxels, constants and grouping computed on the GPU from device arrays,
    # against the host path; then the agents move and are re-binned
    import torch
    rng = np.random.default_rng(91)
    nx, ny, nz, S = 23, 17, 11, 3
    case = W.make_case(nx, ny, nz, S, 10.0, [3e4, 1e3, 2e2], [0.1, 0.5, 0.01], 0.01, 3, W.CELLS | W.DIFFUSION,
                       rng.uniform(0, 1, nx * ny * nz * S))
    case.update(W.random_agents(rng, 3000, S, nx, ny, nz, 10.0, share_voxel=40))
    n = len(case["agent_volume"])
    dev = {k: torch.tensor(case[k], dtype=torch.float64, device="cuda")
           for k in ("agent_pos", "agent_volume", "agent_secretion", "agent_saturation", "agent_uptake")}
    with setup_case(case, precision) as host, setup_case(W.make_case(nx, ny, nz, S, 10.0, case["D"], case["lam"],
                                                                     0.01, 3, case["phases"], case["rho0"]),
                                                         precision) as m:
        vox = torch.empty(n, dtype=torch.int32, device="cuda")
        m.set_agent_rates_device(dev["agent_pos"], dev["agent_volume"], dev["agent_secretion"],
                                 dev["agent_saturation"], dev["agent_uptake"], case["dt"], out_voxel=vox)
        torch.cuda.synchronize()
        ref_vox = oracle.oracle_run(case, return_voxels=True)[1]
        np.testing.assert_array_equal(vox.cpu().numpy(), ref_vox)
        host.run(case["dt"], 3, case["phases"])
        m.run(case["dt"], 3, case["phases"])
        np.testing.assert_array_equal(m.download(), host.download())
        # move every agent by up to two voxels, re-bin both ways
        moved = np.clip(case["agent_pos"] + rng.uniform(-20, 20, case["agent_pos"].size), 0.0,
                        np.repeat([[nx * 10.0, ny * 10.0, nz * 10.0]], n, axis=0).reshape(-1))
        host.set_agent_rates(moved, case["agent_volume"], case["agent_secretion"], case["agent_saturation"],
                             case["agent_uptake"], case["dt"])
        dev["agent_pos"].copy_(torch.tensor(moved, dtype=torch.float64))
        m.set_agent_rates_device(dev["agent_pos"], dev["agent_volume"], dev["agent_secretion"],
                                 dev["agent_saturation"], dev["agent_uptake"], case["dt"])
        host.run(case["dt"], 2, case["phases"])
        m.run(case["dt"], 2, case["phases"])
        np.testing.assert_array_equal(m.download(), host.download())
        # an agent outside the domain is refused
        dev["agent_pos"][0] = -1.0
        with pytest.raises(BfvmError):
            m.set_agent_rates_device(dev["agent_pos"], dev["agent_volume"], dev["agent_secretion"],
                                     dev["agent_saturation"], dev["agent_uptake"], case["dt"])


@pytest.mark.parametrize("precision", [FP64, FP32], ids=["fp64", "fp32"])
def test_device_agent_setup_at_scale_equals_host_setup(precision):
    # 400^3 voxels (26-bit voxel keys: four radix passes over many blocks) and
    # 3e5 agents, 1e5 of them stacked 2000 deep on 50 voxels: the device
    # grouping gives the host grouping's step, bit for bit, and the voxels of
    # BioFVM_mesh.cpp:930-948
    import torch
    rng = np.random.default_rng(92)
    n, S, dx, dt = 400, 1, 10.0, 0.01
    ag = W.random_agents(rng, 200_000, S, n, n, n, dx, cluster_sigma=300.0)
    stack_vox = rng.integers(0, n, (50, 3))
    stacked = np.repeat((stack_vox + 0.5) * dx, 2000, axis=0)
    pos = np.concatenate([ag["agent_pos"].reshape(-1, 3), stacked]).reshape(-1)
    na = pos.size // 3
    vol = np.full(na, 2494.0)
    sec, upt = rng.uniform(0, 10, na * S), rng.uniform(0, 10, na * S)
    sat = np.ones(na * S)
    rho0 = rng.uniform(0, 1, n ** 3 * S)
    p3 = pos.reshape(-1, 3)
    ijk = np.clip(np.floor(p3 / dx).astype(np.int64), 0, n - 1)
    want_vox = (ijk[:, 2] * n + ijk[:, 1]) * n + ijk[:, 0]
    with DeviceMicroenvironment(n, n, n, S, dx, precision) as host, \
            DeviceMicroenvironment(n, n, n, S, dx, precision) as dev:
        for m in (host, dev):
            m.set_coefficients([1e3], [0.1])
            m.upload(rho0)
        got_host = host.set_agent_rates(pos, vol, sec, sat, upt, dt)
        t = {k: torch.tensor(v, dtype=torch.float64, device="cuda") for k, v in
             (("pos", pos), ("vol", vol), ("sec", sec), ("sat", sat), ("upt", upt))}
        vox = torch.empty(na, dtype=torch.int32, device="cuda")
        dev.set_agent_rates_device(t["pos"], t["vol"], t["sec"], t["sat"], t["upt"], dt, out_voxel=vox)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(vox.cpu().numpy(), want_vox)
        np.testing.assert_array_equal(np.asarray(got_host), want_vox)
        for m in (host, dev):
            m.run(dt, 2, W.CELLS)
        np.testing.assert_array_equal(dev.download(), host.download())


def test_sources_only_steps_are_bit_exact():
    # bulk and agent updates (FP64 store) have no reassociation: bit for bit the oracle's
    rng = np.random.default_rng(101)
    nx, ny, nz, S = 19, 13, 7, 3
    case = W.make_case(nx, ny, nz, S, 20.0, [1e4] * S, [0.1] * S, 0.01, 4, W.BULK | W.CELLS,
                       rng.uniform(0, 1, nx * ny * nz * S))
    case.update(W.random_bulk(rng, nx * ny * nz, S))
    case.update(W.random_agents(rng, 500, S, nx, ny, nz, 20.0, share_voxel=25))
    np.testing.assert_array_equal(run_case(case, FP64), oracle.oracle_run(case))
