"""The CPU oracle against what the mathematics fixes, independent of the reference.

Each check is chosen so that a plausible slip in the restatement (wrong end
coefficient, a decay split other than thirds, a transposed axis stride,
agents applied out of order, a wrong implicit form) fails it:

* LOD sweep matrices are (1+c2) I + c1 L with L the Neumann Laplacian, whose
  eigenvectors are cos(pi m (i+1/2)/n) with eigenvalues 4 sin^2(pi m / 2n):
  a separable cosine mode is scaled by prod_a 1/(1+c2+4 c1 sin^2(pi m_a/2n_a)).
* every column of those matrices sums to 1+c2, so without sources the total
  of each substrate decays by exactly (1+c2)^-3 per step;
* a step equals three dense solves A_z^-1 A_y^-1 A_x^-1 on a tiny grid, and
  the Thomas coefficients are the LU factors of A;
* bulk and agent updates are implicit Euler steps with closed-form solutions;
* the 1-D analytic solution of examples/convergence_test1_analytical.cpp is
  approached at O(dx^2) + O(dt).
"""
import numpy as np
import pytest

import oracle
from biofvm_amd import workloads as W


def consts(D, lam, dt, dx):
    c1 = np.asarray(D) * dt / dx / dx
    c2 = np.asarray(lam) * dt / 3.0
    return c1, c2


def axis_matrix(n, c1, c2):
    """The tridiagonal matrix one Thomas sweep inverts (BioFVM_solvers.cpp:121-135)."""
    if n == 1:
        return np.array([[1.0 + c2]])
    A = np.diag(np.full(n, 1 + 2 * c1 + c2)) - c1 * np.eye(n, k=1) - c1 * np.eye(n, k=-1)
    A[0, 0] = A[-1, -1] = 1 + c1 + c2
    return A


@pytest.mark.parametrize("dims,modes", [((8, 6, 5), (1, 2, 3)), ((11, 4, 7), (0, 3, 6)),
                                        ((1, 9, 6), (0, 4, 1)), ((6, 1, 1), (5, 0, 0))])
def test_cosine_modes_decay_at_their_eigenvalue(dims, modes):
    nx, ny, nz = dims
    D, lam, dt, dx = [1e5, 2e3], [0.1, 0.7], 0.01, 20.0
    c1, c2 = consts(D, lam, dt, dx)
    i = (np.arange(nx) + 0.5) / nx
    j = (np.arange(ny) + 0.5) / ny
    k = (np.arange(nz) + 0.5) / nz
    mode = (np.cos(np.pi * modes[2] * k)[:, None, None] * np.cos(np.pi * modes[1] * j)[None, :, None]
            * np.cos(np.pi * modes[0] * i)[None, None, :])
    rho0 = np.stack([mode, 2 * mode], axis=-1).reshape(-1)
    steps = 3
    case = W.make_case(nx, ny, nz, 2, dx, D, lam, dt, steps, W.DIFFUSION, rho0)
    got = oracle.oracle_run(case).reshape(-1, 2)
    for s in range(2):
        f = 1.0
        for m, n in zip(modes, dims):
            f /= 1 + c2[s] + 4 * c1[s] * np.sin(np.pi * m / (2 * n)) ** 2
        expect = rho0.reshape(-1, 2)[:, s] * f ** steps
        diff = np.max(np.abs(got[:, s] - expect))
        # rounding is relative to the initial amplitude; damped modes shrink it
        assert diff / np.max(np.abs(rho0.reshape(-1, 2)[:, s])) < 1e-14, diff
        assert diff / np.max(np.abs(expect)) < 1e-9, diff


def test_mass_decays_by_exact_factor():
    rng = np.random.default_rng(3)
    nx, ny, nz, S = 10, 7, 9, 3
    D, lam, dt, dx = [1e5, 1e3, 3e4], [0.1, 0.01, 1.0], 0.01, 10.0
    c1, c2 = consts(D, lam, dt, dx)
    rho0 = rng.uniform(0, 1, nx * ny * nz * S)
    steps = 4
    got = oracle.oracle_run(W.make_case(nx, ny, nz, S, dx, D, lam, dt, steps, W.DIFFUSION, rho0))
    m0 = rho0.reshape(-1, S).sum(0)
    m1 = got.reshape(-1, S).sum(0)
    np.testing.assert_allclose(m1, m0 * (1 + c2) ** (-3 * steps), rtol=1e-13)


def test_constant_field_only_decays():
    nx, ny, nz = 5, 4, 3
    got = oracle.oracle_run(W.make_case(nx, ny, nz, 1, 20.0, [1e5], [0.3], 0.01, 2, W.DIFFUSION,
                                        np.full(nx * ny * nz, 2.5)))
    c1, c2 = consts([1e5], [0.3], 0.01, 20.0)
    np.testing.assert_allclose(got, 2.5 * (1 + c2[0]) ** -6, rtol=1e-14)


def test_step_equals_dense_axis_solves():
    rng = np.random.default_rng(4)
    nx, ny, nz, S = 5, 4, 6, 2
    D, lam, dt, dx = [1e5, 5e2], [0.2, 0.05], 0.01, 20.0
    c1, c2 = consts(D, lam, dt, dx)
    rho0 = rng.uniform(0, 1, nx * ny * nz * S)
    got = oracle.oracle_run(W.make_case(nx, ny, nz, S, dx, D, lam, dt, 1, W.DIFFUSION, rho0))
    r = rho0.reshape(nz, ny, nx, S).copy()
    for s in range(S):
        f = r[..., s]
        f = np.linalg.solve(axis_matrix(nx, c1[s], c2[s]), f.reshape(-1, nx).T).T.reshape(nz, ny, nx)
        f = np.moveaxis(np.linalg.solve(axis_matrix(ny, c1[s], c2[s]),
                                        np.moveaxis(f, 1, 0).reshape(ny, -1)).reshape(ny, nz, nx), 0, 1)
        f = np.linalg.solve(axis_matrix(nz, c1[s], c2[s]), f.reshape(nz, -1)).reshape(nz, ny, nx)
        r[..., s] = f
    np.testing.assert_allclose(got, r.reshape(-1), rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize("n", [1, 2, 3, 17])
def test_thomas_coefficients_are_lu_factors(n):
    D, lam, dt, dx = [1e5, 10.0], [0.1, 0.9], 0.01, 10.0
    c1, c2 = consts(D, lam, dt, dx)
    denom, c = oracle.thomas_coefficients(n, 2, dx, dt, D, lam)
    for s in range(2):
        A = axis_matrix(n, c1[s], c2[s])
        # A = L U with unit upper U: U[i,i+1] = c_i, L diagonal = denom_i, L[i+1,i] = -c1
        L = np.diag(denom[:, s]) - c1[s] * np.eye(n, k=-1)
        U = np.eye(n) + np.diag(c[:-1, s], k=1)
        np.testing.assert_allclose(L @ U, A, rtol=1e-14, atol=1e-14)


def test_bulk_is_implicit_euler_with_closed_form():
    rng = np.random.default_rng(5)
    n, S, dt, steps = 3 * 2 * 2, 2, 0.05, 7
    sup = rng.uniform(0.1, 2.0, n * S)
    tgt = rng.uniform(0.5, 3.0, n * S)
    upt = rng.uniform(0.0, 2.0, n * S)
    rho0 = rng.uniform(0, 1, n * S)
    case = W.make_case(3, 2, 2, S, 20.0, [0, 0], [0, 0], dt, steps, W.BULK, rho0,
                       bulk_supply=sup, bulk_target=tgt, bulk_uptake=upt)
    got = oracle.oracle_run(case)
    star = sup * tgt / (sup + upt)
    expect = star + (rho0 - star) / (1 + dt * (sup + upt)) ** steps
    np.testing.assert_allclose(got, expect, rtol=1e-13)


def test_agent_constants_steady_state():
    rng = np.random.default_rng(6)
    n, S, dt, dV = 50, 3, 0.01, 8000.0
    vol = rng.uniform(100, 4000, n)
    sec, sat, upt = rng.uniform(0, 10, (n, S)), rng.uniform(0, 2, (n, S)), rng.uniform(0, 10, (n, S))
    t1, t2 = oracle.agent_constants(S, dt, dV, vol, sec, sat, upt)
    c = dt * vol / dV
    np.testing.assert_allclose(t2 - 1.0, c[:, None] * (sec + upt), rtol=1e-12)
    np.testing.assert_allclose(t1, c[:, None] * sec * sat, rtol=1e-14)


def test_stacked_agents_compose_in_list_order():
    # k copies of one agent on a voxel: closed form of k implicit updates
    nx, ny, nz, dx, dt = 3, 3, 3, 20.0, 0.01
    k = 9
    rho0 = np.linspace(0.1, 0.9, nx * ny * nz)
    centre = np.array([1.5, 1.5, 1.5]) * dx
    sec, sat, upt, vol = 7.0, 0.6, 3.0, 900.0
    case = W.make_case(nx, ny, nz, 1, dx, [0], [0], dt, 1, W.CELLS, rho0,
                       agent_pos=np.tile(centre, k), agent_volume=np.full(k, vol),
                       agent_secretion=np.full(k, sec), agent_saturation=np.full(k, sat),
                       agent_uptake=np.full(k, upt))
    got, vox = oracle.oracle_run(case, return_voxels=True)
    v = 13
    assert np.all(vox == v)
    c = dt * vol / dx ** 3
    t1, t2 = c * sec * sat, 1 + c * (sec + upt)
    star = t1 / (t2 - 1)
    np.testing.assert_allclose(got[v], star + (rho0[v] - star) / t2 ** k, rtol=1e-13)
    others = np.delete(np.arange(rho0.size), v)
    np.testing.assert_array_equal(got[others], rho0[others])
    # order matters when agents differ: swapping two unlike agents changes the result
    case2 = dict(case, agent_secretion=np.r_[np.full(k - 1, sec), 0.0],
                 agent_uptake=np.r_[np.full(k - 1, upt), 50.0])
    a = oracle.oracle_run(case2)
    rev = {key: case2[key][::-1].copy() if key in ("agent_secretion", "agent_uptake") else case2[key]
           for key in case2}
    b = oracle.oracle_run(rev)
    assert a[v] != b[v]


def test_nearest_voxel_brute_force():
    nx, ny, nz, dx = 4, 3, 5, 10.0
    rng = np.random.default_rng(9)
    pts = np.concatenate([rng.uniform(0, 1, (200, 3)) * [nx * dx, ny * dx, nz * dx],
                          [[0, 0, 0], [40, 30, 50], [10, 10, 10], [39.999, 0, 49.999]]])
    for x, y, z in pts:
        i = min(int(x // dx), nx - 1)
        j = min(int(y // dx), ny - 1)
        k = min(int(z // dx), nz - 1)
        assert oracle.nearest_voxel(nx, ny, nz, dx, x, y, z) == (k * ny + j) * nx + i
    assert oracle.nearest_voxel(nx, ny, nz, dx, -1e-9, 5, 5) == -1
    assert oracle.nearest_voxel(nx, ny, nz, dx, 5, 30.000001, 5) == -1


def test_dirichlet_mask_leaves_other_substrates_alone():
    rng = np.random.default_rng(10)
    nx, ny, nz, S = 4, 4, 4, 3
    idx = W.outer_face_voxels(nx, ny, nz)
    rho0 = rng.uniform(0, 1, nx * ny * nz * S)
    base = W.make_case(nx, ny, nz, S, 20.0, [1e4] * 3, [0.1] * 3, 0.01, 2, W.DIFFUSION, rho0)
    val = np.tile([1.0, 5.0, 7.0], idx.size)
    masked = oracle.oracle_run(dict(base, dir_idx=idx, dir_val=val, dir_mask=np.array([1, 0, 0], np.uint8)))
    whole = oracle.oracle_run(dict(base, dir_idx=idx, dir_val=val))
    free = oracle.oracle_run(base)
    # masked substrates 1, 2 evolve as if there were no Dirichlet nodes
    np.testing.assert_array_equal(masked.reshape(-1, S)[:, 1:], free.reshape(-1, S)[:, 1:])
    np.testing.assert_array_equal(masked.reshape(-1, S)[:, 0], whole.reshape(-1, S)[:, 0])


def test_convergence_to_1d_analytic_solution():
    # examples/convergence_test1.cpp / _analytical.cpp: 1 + cos(pi x / L0) exp(-D pi^2 t / L0^2)
    D, L0, t = 1e5, 500.0, 0.5
    errs = []
    for dx, dt in [(20.0, 0.01), (10.0, 0.0025)]:
        n = int(round(2 * L0 / dx))
        x = -L0 + (np.arange(n) + 0.5) * dx
        steps = int(round(t / dt))
        case = W.make_case(n, 1, 1, 1, dx, [D], [0.0], dt, steps, W.DIFFUSION, 1 + np.cos(np.pi * x / L0))
        got = oracle.oracle_run(case)
        exact = 1 + np.cos(np.pi * x / L0) * np.exp(-D * np.pi ** 2 * t / L0 ** 2)
        errs.append(np.max(np.abs(got - exact)))
    assert errs[1] < errs[0] / 3.0, errs
    assert errs[0] < 1e-2


# ---- gradient vectors (compute_all_gradient_vectors) ----------------------

def test_gradient_of_quadratic_fields_is_exact_inside_and_zero_on_boundary_layers():
    # central differences are exact for polynomials of degree <= 2
    nx, ny, nz, S, dx = 7, 6, 5, 2, 20.0
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    x, y, z = (i + 0.5) * dx, (j + 0.5) * dx, (k + 0.5) * dx
    f0 = 3.0 * x - 2.0 * y + 0.5 * z
    f1 = 0.01 * x * x + 0.02 * y * y - 0.03 * z * z
    rho = np.stack([f0.reshape(-1), f1.reshape(-1)], axis=1).reshape(-1)
    g = oracle.gradients(nx, ny, nz, S, dx, rho)
    coords = [i.reshape(-1), j.reshape(-1), k.reshape(-1)]
    n_axis = [nx, ny, nz]
    expect0 = [3.0, -2.0, 0.5]
    expect1 = [0.02 * x.reshape(-1), 0.04 * y.reshape(-1), -0.06 * z.reshape(-1)]
    for a in range(3):
        inside = (coords[a] > 0) & (coords[a] < n_axis[a] - 1)
        np.testing.assert_allclose(g[inside, 0, a], expect0[a], rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(g[inside, 1, a], expect1[a][inside], rtol=1e-12, atol=1e-12)
        assert np.all(g[~inside, :, a] == 0.0)


@pytest.mark.skipif(not oracle.have_ref(), reason="compiled reference not built")
@pytest.mark.parametrize("shape", [(9, 7, 5, 3), (1, 6, 4, 2), (5, 1, 1, 1), (2, 3, 2, 2)])
def test_gradient_matches_compiled_reference(shape):
    # one spacing in this process: the reference keeps 2*dx in a function-static
    nx, ny, nz, S = shape
    rho = np.random.default_rng(sum(shape)).uniform(0, 1, nx * ny * nz * S)
    np.testing.assert_array_equal(oracle.gradients(nx, ny, nz, S, 20.0, rho),
                                  oracle.ref_gradients(nx, ny, nz, S, 20.0, rho))


# ---- LOD_2D (diffusion_decay_solver__constant_coefficients_LOD_2D) --------

def _lod2d_case(nx, ny, S, steps, rho0, D, lam, phases=4):
    from biofvm_amd import workloads as W
    case = W.make_case(nx, ny, 1, S, 20.0, D, lam, 0.01, steps, phases, rho0)
    case["lod_dims"] = 2
    return case


def test_lod2d_cosine_modes_and_mass_decay():
    nx, ny, dx, dt, D, lam = 12, 9, 20.0, 0.01, 4e4, 0.3
    c1, c2 = dt * D / dx ** 2, dt * lam * 0.5
    j, i = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    mx, my = 3, 2
    mode = (np.cos(np.pi * mx * (i + 0.5) / nx) * np.cos(np.pi * my * (j + 0.5) / ny)).reshape(-1)
    out = oracle.oracle_run(_lod2d_case(nx, ny, 1, 3, mode, [D], [lam]))
    f = 1.0
    for m, n in ((mx, nx), (my, ny)):
        f /= 1 + c2 + 4 * c1 * np.sin(np.pi * m / (2 * n)) ** 2
    np.testing.assert_allclose(out, mode * f ** 3, rtol=1e-11, atol=1e-13)
    rho = np.random.default_rng(1).uniform(0, 1, nx * ny)
    out = oracle.oracle_run(_lod2d_case(nx, ny, 1, 4, rho, [D], [lam]))
    np.testing.assert_allclose(out.sum(), rho.sum() * (1 + c2) ** -8, rtol=1e-12)


@pytest.mark.skipif(not oracle.have_ref(), reason="compiled reference not built")
@pytest.mark.parametrize("dims", [(13, 9, 3), (1, 7, 2), (6, 1, 1)])
def test_lod2d_matches_compiled_reference(dims):
    from biofvm_amd import workloads as W
    nx, ny, S = dims
    rng = np.random.default_rng(sum(dims))
    case = _lod2d_case(nx, ny, S, 4, rng.uniform(0, 1, nx * ny * S), 10 ** rng.uniform(2, 5, S),
                       10 ** rng.uniform(-3, 0, S), phases=7)
    case.update(W.random_bulk(rng, nx * ny, S))
    case.update(W.random_agents(rng, 20, S, nx, ny, 1, 20.0, share_voxel=3))
    idx = W.outer_face_voxels(nx, ny, 1)[::2]
    case.update(dir_idx=idx, dir_val=rng.uniform(0, 1, idx.size * S))
    np.testing.assert_array_equal(oracle.oracle_run(case), oracle.ref_run(case))
