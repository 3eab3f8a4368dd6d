"""GPU parity of ensembles: uneven grids advanced together by batched launches
(and HIP graph replays) against the oracle run member by member.

Tolerances as for single grids: FP64 1e-12, FP32 1e-5 of max|ref| per member.
"""
import numpy as np
import pytest

import oracle
from biofvm_amd import ALL, DIFFUSION, FP32, FP64, BfvmError, DeviceMicroenvironment, Ensemble, setup_case
from biofvm_amd import workloads as W

pytestmark = pytest.mark.gpu

TOL = {FP64: 1e-12, FP32: 1e-5}


def rel_err(got, ref):
    scale = max(np.max(np.abs(ref)), 1e-300)
    return float(np.max(np.abs(got - ref)) / scale)


def _members(cases, precision):
    return [setup_case(c, precision) for c in cases]


def _close(ens, members):
    ens.close()
    for m in members:
        m.close()


@pytest.fixture(scope="module")
def mixed():
    cases = W.ensemble_mixed(steps=3)
    return cases, [oracle.oracle_run(c) for c in cases]


@pytest.mark.parametrize("graph", [False, True], ids=["eager", "graph"])
@pytest.mark.parametrize("precision", [FP64, FP32], ids=["fp64", "fp32"])
def test_mixed_ensemble_against_oracle(mixed, precision, graph):
    cases, refs = mixed
    members = _members(cases, precision)
    ens = Ensemble(members)
    try:
        ens.run(cases[0]["dt"], cases[0]["steps"], ALL, graph=graph)
        ens.sync()
        for i, (m, ref) in enumerate(zip(members, refs)):
            err = rel_err(m.download(), ref)
            assert err <= TOL[precision], (i, cases[i]["nx"], cases[i]["ny"], cases[i]["nz"], err)
    finally:
        _close(ens, members)


@pytest.mark.parametrize("L", [4, 8, 16, 32])
def test_every_segment_length(mixed, L):
    cases, refs = mixed
    members = _members(cases, FP64)
    ens = Ensemble(members)
    try:
        ens.segment = L
        assert ens.segment == L
        ens.run(cases[0]["dt"], cases[0]["steps"], ALL, graph=False)
        for i, (m, ref) in enumerate(zip(members, refs)):
            assert rel_err(m.download(), ref) <= TOL[FP64], (L, i)
    finally:
        _close(ens, members)


def test_graph_replays_match_eager():
    # two copies of one ensemble: one replays a graph step by step, the other
    # launches eagerly; every replay is checked after eager work on the host
    # and on the device (downloads, a member's own Dirichlet change)
    cases = W.ensemble_mixed(seed=7, steps=1)
    a, b = _members(cases, FP32), _members(cases, FP32)
    ea, eb = Ensemble(a), Ensemble(b)
    try:
        dt = cases[0]["dt"]
        for rep in range(5):
            ea.run(dt, 1, ALL, graph=True)
            eb.run(dt, 1, ALL, graph=False)
            for i, (ma, mb) in enumerate(zip(a, b)):
                np.testing.assert_array_equal(ma.download(), mb.download(), err_msg=f"replay {rep} member {i}")
            if rep == 2:  # a changed source must reach the next replay
                for m in (a[0], b[0]):
                    idx = W.outer_face_voxels(m.nx, m.ny, m.nz)
                    m.set_dirichlet(idx, np.full(idx.size * m.S, 0.25))
    finally:
        _close(ea, a)
        _close(eb, b)


def test_ensemble_equals_single_runs_after_changes(mixed):
    cases, _ = mixed
    members = _members(cases, FP64)
    ens = Ensemble(members)
    try:
        ens.run(cases[0]["dt"], 2, ALL, graph=True)
        # change coefficients of one member and agents of another, then go on
        members[1].set_coefficients(cases[1]["D"] * 0.5, cases[1]["lam"] * 2.0)
        members[3].set_agents(np.array([0, 0, 5], dtype=np.int32), np.full(3 * cases[3]["S"], 0.01),
                              np.full(3 * cases[3]["S"], 1.02))
        ens.run(cases[0]["dt"], 2, ALL, graph=True)
        got = [m.download() for m in members]
    finally:
        _close(ens, members)
    # the same sequence member by member through the single-grid entry points
    for i, c in enumerate(cases):
        with setup_case(c, FP64) as m:
            m.run(c["dt"], 2, ALL)
            if i == 1:
                m.set_coefficients(c["D"] * 0.5, c["lam"] * 2.0)
            if i == 3:
                m.set_agents(np.array([0, 0, 5], dtype=np.int32), np.full(3 * c["S"], 0.01),
                             np.full(3 * c["S"], 1.02))
            m.run(c["dt"], 2, ALL)
            ref = m.download()
        assert rel_err(got[i], ref) <= TOL[FP64], i


def test_multi_step_graphs_equal_eager_steps():
    # runs of graph_steps steps replay one graph holding them, the rest
    # one-step graphs: the bits of launching every step eagerly
    cases = W.ensemble_mixed(seed=8, steps=1)
    out = {}
    for mode in ("eager", "graph3", "graph1"):
        members = _members(cases, FP32)
        ens = Ensemble(members)
        try:
            if mode == "graph3":
                ens.graph_steps = 3
                assert ens.graph_steps == 3
            if mode == "graph1":
                ens.graph_steps = 1
            with pytest.raises(BfvmError):
                ens.graph_steps = 0
            for steps in (7, 2, 3):  # 2 x 3 + 1, one-step graphs, one 3-step graph
                ens.run(cases[0]["dt"], steps, ALL, graph=mode != "eager")
            out[mode] = [m.download() for m in members]
        finally:
            _close(ens, members)
    for mode in ("graph3", "graph1"):
        for i, (a, b) in enumerate(zip(out[mode], out["eager"])):
            np.testing.assert_array_equal(a, b, err_msg=f"{mode} member {i}")


def test_members_at_their_own_segment_lengths_equal_solo_runs():
    # by default each axis is batched at the length most members take on
    # their own (FP32 lines of 32 nodes or more: 16, for x if the rows are
    # whole 4-node chunks; shorter ones: 8) and members with another length
    # sweep on their own: every member's bits equal its solo run's
    cases = W.ensemble_mixed(seed=5, steps=3)
    members = _members(cases, FP32)
    ens = Ensemble(members)
    try:
        assert ens.segment == 0
        # x: the 64-node member (12288 values) against the rest; z: one
        # 40-node axis holds most of the values
        assert ens.axis_segments() == [8, 8, 16]
        ens.run(cases[0]["dt"], 3, ALL, graph=True)
        got = [m.download() for m in members]
        ens.segment = 8
        assert ens.axis_segments() == [8, 8, 8]
    finally:
        _close(ens, members)
    for i, c in enumerate(cases):
        with setup_case(c, FP32) as m:
            m.run(c["dt"], 3, ALL)
            np.testing.assert_array_equal(got[i], m.download(), err_msg=f"member {i}")


def test_autotune_keeps_densities(mixed):
    cases, refs = mixed
    members = _members(cases, FP32)
    ens = Ensemble(members)
    try:
        before = [m.download() for m in members]
        L = ens.autotune(cases[0]["dt"])
        assert L in (0, 4, 8, 16, 32)  # 0: the members' own lengths
        for m, b in zip(members, before):
            np.testing.assert_array_equal(m.download(), b)
        ens.run(cases[0]["dt"], cases[0]["steps"], ALL, graph=True)
        for i, (m, ref) in enumerate(zip(members, refs)):
            assert rel_err(m.download(), ref) <= TOL[FP32], i
    finally:
        _close(ens, members)


def test_large_uneven_ensemble_sample():
    # a few members of the 64-member workload at full size, 3 steps, FP32
    cases = W.ensemble_64(steps=3, members=64)[:6]
    members = _members(cases, FP32)
    ens = Ensemble(members)
    try:
        ens.autotune(0.01)
        ens.run(0.01, 3, cases[0]["phases"], graph=True)
        for i, (m, c) in enumerate(zip(members, cases)):
            err = rel_err(m.download(), oracle.oracle_run(c))
            assert err <= TOL[FP32], (i, c["nx"], c["ny"], c["nz"], c["S"], err)
    finally:
        _close(ens, members)


def test_ensemble_errors():
    a = DeviceMicroenvironment(8, 8, 8, 2, 10.0, FP64)
    b = DeviceMicroenvironment(8, 8, 8, 2, 10.0, FP32)
    c = DeviceMicroenvironment(6, 5, 4, 1, 10.0, FP64)
    try:
        with pytest.raises(BfvmError):
            Ensemble([a, b])  # mixed precision
        with pytest.raises(BfvmError):
            Ensemble([a, a])  # repeated member
        ens = Ensemble([a, c])
        with pytest.raises(BfvmError):
            Ensemble([c])  # already in an ensemble
        assert a.lib.bfvm_destroy(a.h) != 0  # members outlive their ensemble
        with pytest.raises(BfvmError):
            ens.run(0.01, 1, DIFFUSION)  # no coefficients
        with pytest.raises(BfvmError):
            ens.segment = 12
        ens.close()
        Ensemble([c]).close()
    finally:
        for m in (a, b, c):
            m.close()


def test_ensemble_with_lod2d_members():
    rng = np.random.default_rng(31)
    cases = W.ensemble_mixed(seed=9, steps=2)[:3]
    for nx, ny, S in [(24, 16, 2), (1, 30, 1)]:
        c = W.make_case(nx, ny, 1, S, 20.0, 10.0 ** rng.uniform(2, 5, S), 10.0 ** rng.uniform(-3, 0, S), 0.01, 2,
                        ALL, rng.uniform(0, 1, nx * ny * S))
        c.update(W.random_agents(rng, 20, S, nx, ny, 1, 20.0))
        c["lod_dims"] = 2
        cases.append(c)
    members = _members(cases, FP64)
    ens = Ensemble(members)
    try:
        ens.run(0.01, 2, ALL, graph=True)
        for i, (m, c) in enumerate(zip(members, cases)):
            assert rel_err(m.download(), oracle.oracle_run(c)) <= TOL[FP64], i
    finally:
        _close(ens, members)


@pytest.mark.parametrize("precision", [FP64, FP32], ids=["fp64", "fp32"])
@pytest.mark.parametrize("L", [8, 16])
def test_row_tile_x_sweeps_equal_column_tiles(mixed, precision, L):
    # members' x sweeps on row tiles or whole-cache-line chunks (one batched
    # launch each) give the bits of the column-tile batch with the same
    # segment length
    cases, refs = mixed
    out = {}
    for chunks, rows in ((False, False), (False, True), (True, False), (True, True)):
        members = _members(cases, precision)
        ens = Ensemble(members)
        try:
            ens.segment = L
            ens.row_tiles = rows
            ens.x_chunks = chunks
            assert ens.row_tiles == rows and ens.x_chunks == chunks
            ens.run(cases[0]["dt"], cases[0]["steps"], ALL, graph=rows)
            out[chunks, rows] = [m.download() for m in members]
        finally:
            _close(ens, members)
    for key in ((False, True), (True, False), (True, True)):
        for i, (a, b, ref) in enumerate(zip(out[key], out[False, False], refs)):
            np.testing.assert_array_equal(a, b, err_msg=f"member {i} (chunks, rows) = {key}")
            assert rel_err(a, ref) <= TOL[precision], (i, key)
        assert rel_err(a, ref) <= TOL[precision], i


@pytest.mark.parametrize("precision", [FP64, FP32], ids=["fp64", "fp32"])
def test_cache_policy_never_changes_ensemble_results(mixed, precision):
    # the batched y and z sweeps with non-temporal loads and stores give the
    # default policy's bits
    cases, refs = mixed
    out = {}
    for nt in (False, True):
        members = _members(cases, precision)
        ens = Ensemble(members)
        try:
            ens.non_temporal = nt
            assert ens.non_temporal == nt
            ens.run(cases[0]["dt"], cases[0]["steps"], ALL, graph=True)
            out[nt] = [m.download() for m in members]
        finally:
            _close(ens, members)
    for i, (a, b, ref) in enumerate(zip(out[True], out[False], refs)):
        np.testing.assert_array_equal(a, b, err_msg=f"member {i}")
        assert rel_err(a, ref) <= TOL[precision], i
