"""Writes the golden fixtures of the diffusion-decay path from the compiled reference.

    make -C oracle && python tests/golden/gen_golden.py

Each fixture ``<name>.npz`` holds one case's inputs and the compiled
reference's output (oracle/_ref/libbfvm_ref.so, BioFVM 1.0.4 built from its
own sources, run on one OpenMP thread so the agent loop is serial).  The
manifest ``manifest.json`` lists the parameters, seed and the sha256 of each
output.  The reference has no in-tree golden data, so these are the pins.

Cases
  g1_lod          9x7x5, S=3, LOD only, 1 and 5 steps
  g2_dirichlet    g1 + Dirichlet on every outer face, values {1, 0.5, 0}
  g3_bulk         g1 + bulk fields (S, U ~ U[0,0.1), T = 1), all phases
  g4_agents       g3 + 200 random agents + 12 stacked on one voxel
  g5_degenerate_* nx=1; nz=1; 1x1xn; 1x1x1; 2x3x2 axes
  g6_ensemble     12x20x28, S=6, log-uniform D/lambda, 3 steps (uneven ensemble member shape)
  g7_lod50        50^3 x 1, dx=20, D=1e5, lambda=0.1, 100 steps (the CPU-runnable case)
  g8_shell        28^3 x 2, Dirichlet + bulk + 400 agents in a shell, 20 steps
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from biofvm_amd import workloads as W  # noqa: E402

ARRAY_KEYS = ["D", "lam", "rho0", "dir_idx", "dir_val", "dir_mask", "bulk_supply", "bulk_target",
              "bulk_uptake", "agent_pos", "agent_volume", "agent_secretion", "agent_saturation",
              "agent_uptake"]
SCALAR_KEYS = ["nx", "ny", "nz", "S", "dx", "dt", "steps", "phases"]


def cases():
    out = {}
    rng = np.random.default_rng(1)
    nx, ny, nz, S = 9, 7, 5, 3
    g1 = W.make_case(nx, ny, nz, S, 20.0, [1e5, 1e3, 5e4], [0.1, 0.01, 1.0], 0.01, 5, W.DIFFUSION,
                     rng.uniform(0.0, 1.0, nx * ny * nz * S))
    out["g1_lod_1step"] = dict(g1, steps=1)
    out["g1_lod"] = g1
    idx = W.outer_face_voxels(nx, ny, nz)
    out["g2_dirichlet"] = dict(g1, dir_idx=idx, dir_val=np.tile([1.0, 0.5, 0.0], idx.size))
    g3 = dict(g1, phases=W.ALL, **W.random_bulk(rng, nx * ny * nz, S))
    out["g3_bulk"] = g3
    out["g4_agents"] = dict(g3, **W.random_agents(rng, 200, S, nx, ny, nz, 20.0, share_voxel=12))
    for dims in [(1, 6, 5), (7, 6, 1), (1, 1, 9), (1, 1, 1), (2, 3, 2)]:
        name = "g5_degenerate_" + "x".join(map(str, dims))
        out[name] = W.make_case(*dims, 2, 20.0, [1e5, 1e3], [0.1, 0.01], 0.01, 3, W.DIFFUSION,
                                rng.uniform(0.0, 1.0, int(np.prod(dims)) * 2))
    rng6 = np.random.default_rng(6)
    D6 = 10.0 ** rng6.uniform(2.0, 5.0, 6)
    l6 = 10.0 ** rng6.uniform(-3.0, 0.0, 6)
    out["g6_ensemble"] = W.make_case(12, 20, 28, 6, 10.0, D6, l6, 0.01, 3, W.DIFFUSION,
                                     rng6.uniform(0.0, 1.0, 12 * 20 * 28 * 6))
    out["g7_lod50"] = W.lod_50()
    rng8 = np.random.default_rng(8)
    n8 = 28
    g8 = W.make_case(n8, n8, n8, 2, 10.0, [1e5, 5e3], [0.1, 0.5], 0.01, 20, W.ALL,
                     rng8.uniform(0.0, 1.0, n8 ** 3 * 2))
    g8.update(W.random_bulk(rng8, n8 ** 3, 2))
    ag = W.random_agents(rng8, 400, 2, n8, n8, n8, 10.0, cluster_sigma=60.0)
    g8.update(ag)
    i8 = W.outer_face_voxels(n8, n8, n8)
    g8.update(dir_idx=i8, dir_val=np.tile([1.0, 0.0], i8.size))
    out["g8_shell"] = g8
    return out


def save(name, case, expected, voxels):
    arrays = {k: np.asarray(case[k]) for k in ARRAY_KEYS if case.get(k) is not None}
    arrays["expected"] = expected
    if voxels is not None:
        arrays["agent_voxel"] = voxels
    for k in SCALAR_KEYS:
        arrays[k] = np.asarray(case[k])
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **arrays)
    return hashlib.sha256(expected.tobytes()).hexdigest()


def load(name):
    """Fixture -> (case dict, expected output, agent voxels or None)."""
    z = np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False)
    case = {}
    for k in SCALAR_KEYS:
        v = z[k]
        case[k] = float(v) if k in ("dx", "dt") else int(v)
    for k in ARRAY_KEYS:
        if k in z.files:
            case[k] = z[k]
    return case, z["expected"], (z["agent_voxel"] if "agent_voxel" in z.files else None)


def names():
    with open(os.path.join(HERE, "manifest.json")) as f:
        return sorted(json.load(f)["fixtures"].keys())


def main():
    if not oracle.have_ref():
        sys.exit("oracle/_ref/libbfvm_ref.so is missing: run `make -C oracle` where the reference exists")
    manifest = {"generator": "tests/golden/gen_golden.py",
                "reference": "BioFVM 1.0.4 compiled from source (oracle/Makefile), 1 OpenMP thread",
                "rng": "numpy.random.default_rng(seed), densities then bulk then agents",
                "fixtures": {}}
    for name, case in cases().items():
        n_agents = 0 if case.get("agent_volume") is None else len(case["agent_volume"])
        if n_agents:
            expected, vox = oracle.ref_run(case, nthreads=1, return_voxels=True)
        else:
            expected, vox = oracle.ref_run(case, nthreads=1), None
        restated = oracle.oracle_run(case)
        if not np.array_equal(restated, expected):
            sys.exit(f"{name}: the C restatement differs from the compiled reference")
        sha = save(name, case, expected, vox)
        manifest["fixtures"][name] = {
            "grid": [case["nx"], case["ny"], case["nz"]], "S": case["S"], "dx": case["dx"],
            "dt": case["dt"], "steps": case["steps"], "phases": case["phases"],
            "n_dirichlet": 0 if case.get("dir_idx") is None else int(len(case["dir_idx"])),
            "n_agents": n_agents, "bulk": case.get("bulk_supply") is not None,
            "sha256_expected": sha}
        print(f"{name}: ok ({expected.size} values)")
    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
