"""The CPU oracle against the golden fixtures and the compiled reference.

Fixtures (tests/golden/*.npz) were written by tests/golden/gen_golden.py from
BioFVM 1.0.4 compiled from its own sources; the oracle must reproduce them
bit for bit.  Where oracle/_ref is built (this container), random cases are
also run through both.
"""
import numpy as np
import pytest

import gen_golden
import oracle
from biofvm_amd import workloads as W

FIXTURES = gen_golden.names()


@pytest.mark.parametrize("name", FIXTURES)
def test_oracle_matches_fixture_bitwise(name):
    case, expected, voxels = gen_golden.load(name)
    if voxels is not None:
        got, got_vox = oracle.oracle_run(case, return_voxels=True)
        np.testing.assert_array_equal(got_vox, voxels)
    else:
        got = oracle.oracle_run(case)
    assert got.dtype == np.float64 and got.shape == expected.shape
    assert np.array_equal(got, expected), f"max diff {np.max(np.abs(got - expected))}"


def test_fixture_manifest_hashes():
    import hashlib
    import json
    import os
    with open(os.path.join(gen_golden.HERE, "manifest.json")) as f:
        man = json.load(f)["fixtures"]
    for name in FIXTURES:
        _, expected, _ = gen_golden.load(name)
        assert hashlib.sha256(expected.tobytes()).hexdigest() == man[name]["sha256_expected"]


needs_ref = pytest.mark.skipif(not oracle.have_ref(), reason="oracle/_ref not built (no reference tree)")


def _random_case(rng, with_sources):
    nx, ny, nz = (int(v) for v in rng.integers(1, 9, size=3))
    S = int(rng.integers(1, 5))
    D = 10.0 ** rng.uniform(2, 5, S)
    lam = 10.0 ** rng.uniform(-3, 0, S)
    dx = float(rng.choice([10.0, 20.0]))
    case = W.make_case(nx, ny, nz, S, dx, D, lam, 0.01, int(rng.integers(1, 4)),
                       W.ALL if with_sources else W.DIFFUSION,
                       rng.uniform(0, 1, nx * ny * nz * S))
    if with_sources:
        case.update(W.random_bulk(rng, nx * ny * nz, S))
        case.update(W.random_agents(rng, int(rng.integers(0, 40)), S, nx, ny, nz, dx,
                                    share_voxel=int(rng.integers(0, 5))))
        idx = rng.choice(nx * ny * nz, size=min(5, nx * ny * nz), replace=False).astype(np.int32)
        case.update(dir_idx=idx, dir_val=rng.uniform(0, 2, idx.size * S))
    return case


@needs_ref
@pytest.mark.parametrize("seed", range(12))
def test_oracle_matches_compiled_reference_random(seed):
    rng = np.random.default_rng(100 + seed)
    case = _random_case(rng, with_sources=seed % 2 == 1)
    a = oracle.oracle_run(case)
    b = oracle.ref_run(case, nthreads=1)
    assert np.array_equal(a, b)


@needs_ref
def test_reference_agent_voxels_match_oracle():
    rng = np.random.default_rng(7)
    nx, ny, nz, S, dx = 6, 5, 4, 1, 20.0
    case = W.make_case(nx, ny, nz, S, dx, [1e3], [0.1], 0.01, 1, W.CELLS, rng.uniform(0, 1, nx * ny * nz))
    # positions on voxel faces and on the domain boundary exercise floor + clamp
    pos = np.array([[0, 0, 0], [nx * dx, ny * dx, nz * dx], [dx, 2 * dx, 3 * dx], [59.999, 0.0, 80.0]],
                   dtype=np.float64)
    case.update(agent_pos=pos.reshape(-1), agent_volume=np.full(4, 100.0),
                agent_secretion=np.ones(4), agent_saturation=np.ones(4), agent_uptake=np.ones(4))
    _, va = oracle.oracle_run(case, return_voxels=True)
    _, vb = oracle.ref_run(case, return_voxels=True)
    np.testing.assert_array_equal(va, vb)
