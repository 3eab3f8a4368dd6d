"""Seeded synthetic inputs for the diffusion-decay step.

Every generator returns a *case* dict (see ``make_case``) whose arrays are
float64, voxel-major / substrate-innermost, so the same host arrays can feed
the GPU library, the CPU oracle and the compiled reference.  Randomness comes
from ``numpy.random.default_rng(seed)`` drawn in a fixed order: densities,
then bulk fields, then agents.

The named workloads follow the five benchmark configurations of this path:

* ``lod_50``       50^3 x 1 substrate, dx=20, D=1e5, lambda=0.1, dt=0.01, LOD only, 100 steps
* ``drop_in_200``  200^3 x 4, dx=10, Dirichlet faces, bulk fields, 1e5 agents, 200 steps
* ``precision_400`` 400^3 x 8, dx=10, log-uniform D/lambda, 1e6 agents (Gaussian cluster), 1000 steps
* ``ensemble_64``  64 independent grids, edges from {32,48,64,96,128} per axis, 1-6 substrates,
                   log-uniform D/lambda, 0-2e4 uniform agents, 500 steps
"""
import numpy as np

BULK, CELLS, DIFFUSION = 1, 2, 4
ALL = BULK | CELLS | DIFFUSION


def make_case(nx, ny, nz, S, dx, D, lam, dt, steps, phases=DIFFUSION, rho0=None, **extra):
    case = dict(nx=int(nx), ny=int(ny), nz=int(nz), S=int(S), dx=float(dx),
                D=np.asarray(D, dtype=np.float64), lam=np.asarray(lam, dtype=np.float64),
                dt=float(dt), steps=int(steps), phases=int(phases))
    nvox = nx * ny * nz
    case["rho0"] = np.zeros(nvox * S) if rho0 is None else np.asarray(rho0, np.float64).reshape(-1)
    case.update(extra)
    return case


def outer_face_voxels(nx, ny, nz):
    """Indices of every voxel on the outer faces, in ascending voxel order."""
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    on = (i == 0) | (i == nx - 1) | (j == 0) | (j == ny - 1) | (k == 0) | (k == nz - 1)
    return np.flatnonzero(on.reshape(-1)).astype(np.int32)


def random_agents(rng, n, S, nx, ny, nz, dx, volume=2494.0, rate_hi=10.0,
                  cluster_sigma=None, share_voxel=0):
    """n agents: positions uniform in the box (or a Gaussian cluster of the
    given sigma at the centre, rejection-sampled to the box), rates U[0,rate_hi),
    saturation 1.  `share_voxel` extra agents are stacked on one voxel centre to
    exercise same-voxel ordering."""
    L = np.array([nx * dx, ny * dx, nz * dx])
    if cluster_sigma is None:
        pos = rng.uniform(0.0, 1.0, size=(n, 3)) * L
    else:
        pos = np.empty((0, 3))
        while pos.shape[0] < n:
            p = rng.normal(L / 2, cluster_sigma, size=(n, 3))
            ok = np.all((p >= 0) & (p <= L), axis=1)
            pos = np.concatenate([pos, p[ok]])
        pos = pos[:n]
    if share_voxel:
        c = (np.floor(L / dx / 2) + 0.5) * dx
        pos = np.concatenate([pos, np.tile(c, (share_voxel, 1))])
        n = pos.shape[0]
    sec = rng.uniform(0.0, rate_hi, size=(n, S))
    upt = rng.uniform(0.0, rate_hi, size=(n, S))
    sat = np.ones((n, S))
    vol = np.full(n, float(volume))
    return dict(agent_pos=pos.reshape(-1), agent_volume=vol, agent_secretion=sec.reshape(-1),
                agent_saturation=sat.reshape(-1), agent_uptake=upt.reshape(-1))


def random_bulk(rng, nvox, S, hi=0.1, target=1.0):
    """Bulk supply/uptake rates U[0,hi) per voxel-substrate, supply target constant."""
    return dict(bulk_supply=rng.uniform(0.0, hi, nvox * S),
                bulk_target=np.full(nvox * S, float(target)),
                bulk_uptake=rng.uniform(0.0, hi, nvox * S))


def lod_50(steps=100, seed=1):
    """50^3 voxels, 1 substrate, LOD only — the reference's CPU-runnable case."""
    rng = np.random.default_rng(seed)
    n = 50
    return make_case(n, n, n, 1, 20.0, [1e5], [0.1], 0.01, steps, DIFFUSION,
                     rng.uniform(0.0, 1.0, n ** 3))


def drop_in_200(steps=200, seed=2, n=200, n_agents=100_000, mask_substrate0=True):
    """200^3 voxels x 4 substrates with Dirichlet faces, bulk fields and agents."""
    rng = np.random.default_rng(seed)
    S = 4
    rho0 = rng.uniform(0.0, 1.0, n ** 3 * S)
    case = make_case(n, n, n, S, 10.0, [1e5, 1e3, 5e4, 1e4], [0.1, 0.01, 1.0, 0.5], 0.01,
                     steps, ALL, rho0)
    case.update(random_bulk(rng, n ** 3, S))
    case.update(random_agents(rng, n_agents, S, n, n, n, 10.0))
    idx = outer_face_voxels(n, n, n)
    val = np.zeros((idx.size, S))
    val[:, 0] = 1.0
    case.update(dir_idx=idx, dir_val=val.reshape(-1))
    if mask_substrate0:
        case["dir_mask"] = np.array([1, 0, 0, 0], dtype=np.uint8)
    return case


def precision_400(steps=1000, seed=3, n=400, S=8, n_agents=1_000_000):
    """400^3 x 8 substrates, log-uniform D and lambda, 1e6 agents in a
    Gaussian cluster (sigma=600 um) at the centre; agents + LOD each step."""
    rng = np.random.default_rng(seed)
    D = 10.0 ** rng.uniform(2.0, 5.0, S)
    lam = 10.0 ** rng.uniform(-3.0, 0.0, S)
    rho0 = rng.uniform(0.0, 1.0, n ** 3 * S)
    case = make_case(n, n, n, S, 10.0, D, lam, 0.01, steps, CELLS | DIFFUSION, rho0)
    case.update(random_agents(rng, n_agents, S, n, n, n, 10.0, cluster_sigma=600.0))
    return case


def ensemble_64(steps=500, seed=4, members=64, edges=(32, 48, 64, 96, 128), max_agents=20_000,
                max_substrates=6):
    """A list of `members` independent cases: each axis length drawn from
    `edges` (non-cubic allowed), 1..max_substrates substrates, D and lambda
    log-uniform as in ``precision_400``, 0..max_agents uniform agents; dx=10,
    dt=0.01, agents + LOD each step."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(members):
        nx, ny, nz = (int(v) for v in rng.choice(edges, size=3))
        S = int(rng.integers(1, max_substrates + 1))
        D = 10.0 ** rng.uniform(2.0, 5.0, S)
        lam = 10.0 ** rng.uniform(-3.0, 0.0, S)
        n_agents = int(rng.integers(0, max_agents + 1))
        rho0 = rng.uniform(0.0, 1.0, nx * ny * nz * S)
        case = make_case(nx, ny, nz, S, 10.0, D, lam, 0.01, steps, CELLS | DIFFUSION, rho0)
        case.update(random_agents(rng, n_agents, S, nx, ny, nz, 10.0))
        out.append(case)
    return out


def ensemble_mixed(seed=5, steps=3):
    """A small uneven ensemble for parity tests: members with Dirichlet faces
    (some masked), bulk fields, agents (some sharing a voxel), single-node and
    long axes, and 1-5 substrates; all phases, `steps` steps."""
    rng = np.random.default_rng(seed)
    shapes = [(12, 9, 7, 3), (1, 16, 5, 2), (33, 1, 1, 1), (20, 14, 40, 5), (7, 30, 2, 4),
              (64, 24, 8, 1), (9, 9, 9, 2)]
    out = []
    for m, (nx, ny, nz, S) in enumerate(shapes):
        D = 10.0 ** rng.uniform(2.0, 5.0, S)
        lam = 10.0 ** rng.uniform(-3.0, 0.0, S)
        rho0 = rng.uniform(0.0, 1.0, nx * ny * nz * S)
        case = make_case(nx, ny, nz, S, 20.0, D, lam, 0.01, steps, ALL, rho0)
        if m % 2 == 0:
            case.update(random_bulk(rng, nx * ny * nz, S))
        if m % 3 != 2:
            case.update(random_agents(rng, 40 + 10 * m, S, nx, ny, nz, 20.0, share_voxel=3 * (m % 2)))
        if m in (0, 3, 5):
            idx = outer_face_voxels(nx, ny, nz)
            val = rng.uniform(0.0, 1.0, (idx.size, S))
            case.update(dir_idx=idx, dir_val=val.reshape(-1))
            if m == 3:
                case["dir_mask"] = (np.arange(S) % 2 == 0).astype(np.uint8)
        out.append(case)
    return out


def differentiation_128(steps=500, seed=6, n=128, S=2, n_agents=20_000):
    """128^3 x 2 substrates, 2e4 uniform agents, dt=0.01, agents + LOD each
    step.  The case holds the hidden parameters (per-agent secretion, D,
    lambda drawn from the seed); ``target`` is produced by running it."""
    rng = np.random.default_rng(seed)
    D = 10.0 ** rng.uniform(2.0, 5.0, S)
    lam = 10.0 ** rng.uniform(-3.0, 0.0, S)
    rho0 = rng.uniform(0.0, 1.0, n ** 3 * S)
    case = make_case(n, n, n, S, 10.0, D, lam, 0.01, steps, CELLS | DIFFUSION, rho0)
    case.update(random_agents(rng, n_agents, S, n, n, n, 10.0))
    return case


def adjoint_small(seed=7, steps=6):
    """A small case with every phase (bulk, agents sharing voxels, masked
    Dirichlet faces) for gradient checks against finite differences."""
    rng = np.random.default_rng(seed)
    nx, ny, nz, S = 9, 7, 5, 2
    case = make_case(nx, ny, nz, S, 20.0, [3e4, 2e3], [0.1, 0.5], 0.01, steps, ALL,
                     rng.uniform(0.0, 1.0, nx * ny * nz * S))
    case.update(random_bulk(rng, nx * ny * nz, S))
    case.update(random_agents(rng, 30, S, nx, ny, nz, 20.0, share_voxel=4))
    idx = outer_face_voxels(nx, ny, nz)[::3]
    case.update(dir_idx=idx, dir_val=rng.uniform(0.0, 1.0, idx.size * S),
                dir_mask=np.array([1, 0], dtype=np.uint8))
    return case


def write_driver_case(path, case, precision):
    """Writes `case` in the input format of the façade driver
    (tests/cpp/facade_driver.cpp, which reads it back through the reference's
    API) for a store of `precision`; returns the number of values of the store."""
    S = case["S"]
    nvox = case["nx"] * case["ny"] * case["nz"]
    n_dir = 0 if case.get("dir_idx") is None else len(case["dir_idx"])
    n_ag = 0 if case.get("agent_volume") is None else len(case["agent_volume"])
    has_bulk = case.get("bulk_supply") is not None
    with open(path, "wb") as f:
        phases = case["phases"] | (8 if case.get("lod_dims", 3) == 2 else 0)  # 8: the LOD_2D solver
        np.array([case["nx"], case["ny"], case["nz"], S, case["steps"], phases, n_dir, n_ag,
                  int(has_bulk), precision], dtype="<i4").tofile(f)
        np.array([case["dx"], case["dt"]], dtype="<f8").tofile(f)
        for k in ("D", "lam", "rho0"):
            np.asarray(case[k], dtype="<f8").tofile(f)
        if n_dir:
            np.asarray(case["dir_idx"], dtype="<i4").tofile(f)
            np.asarray(case["dir_val"], dtype="<f8").tofile(f)
        if has_bulk:
            for k in ("bulk_supply", "bulk_target", "bulk_uptake"):
                np.asarray(case[k], dtype="<f8").tofile(f)
        if n_ag:
            for k in ("agent_pos", "agent_volume", "agent_secretion", "agent_saturation",
                      "agent_uptake"):
                np.asarray(case[k], dtype="<f8").tofile(f)
    return nvox * S
