"""Python handles over the C ABI.

``DeviceMicroenvironment`` owns a ``bfvm_handle`` (one device-resident
microenvironment); ``Ensemble`` advances several of them together with
batched launches; ``run_case`` drives one synthetic case (see
``biofvm_amd.workloads``) end to end through the C ABI.
"""
import ctypes

import numpy as np

from . import _capi
from ._capi import ALL, BULK, CELLS, DIFFUSION, FP32, FP64, check, check_ensemble


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _iptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class DeviceMicroenvironment:
    """A uniform nx*ny*nz grid with S substrates, densities resident in HBM."""

    def __init__(self, nx, ny, nz, S, dx, precision=FP64, dy=None, dz=None):
        self.lib = _capi.load()
        self.h = ctypes.c_void_p()
        dy = dx if dy is None else dy
        dz = dx if dz is None else dz
        rc = self.lib.bfvm_create_mesh(nx, ny, nz, S, dx, dy, dz, precision, ctypes.byref(self.h))
        check(rc, None)
        self.nx, self.ny, self.nz, self.S, self.dx = nx, ny, nz, S, dx
        self.precision = precision
        self.nvox = nx * ny * nz

    def close(self):
        if self.h:
            self.lib.bfvm_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        check(rc, self.h)

    def set_coefficients(self, D, lam):
        D, lam = _f64(D), _f64(lam)
        assert D.size == self.S and lam.size == self.S
        self._check(self.lib.bfvm_set_coefficients(self.h, _dptr(D), _dptr(lam)))

    def upload(self, rho):
        rho = _f64(rho).reshape(-1)
        assert rho.size == self.nvox * self.S
        self._check(self.lib.bfvm_upload_density(self.h, _dptr(rho)))

    def download(self):
        out = np.empty(self.nvox * self.S)
        self._check(self.lib.bfvm_download_density(self.h, _dptr(out)))
        return out

    def set_dirichlet(self, idx, values, mask=None):
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        values = _f64(values).reshape(-1)
        assert values.size == idx.size * self.S
        mp = None
        if mask is not None:
            m = np.ascontiguousarray(mask, dtype=np.uint8)
            mp = m.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte))
        self._check(self.lib.bfvm_set_dirichlet(self.h, int(idx.size), _iptr(idx), _dptr(values), mp))

    def clear_dirichlet(self):
        self._check(self.lib.bfvm_set_dirichlet(self.h, 0, None, None, None))

    def set_bulk(self, supply, target, uptake):
        if supply is None:
            self._check(self.lib.bfvm_set_bulk(self.h, None, None, None))
            return
        f = [_f64(x).reshape(-1) for x in (supply, target, uptake)]
        assert all(x.size == self.nvox * self.S for x in f)
        self._check(self.lib.bfvm_set_bulk(self.h, _dptr(f[0]), _dptr(f[1]), _dptr(f[2])))

    def set_agents(self, voxel, temp1, temp2):
        voxel = np.ascontiguousarray(voxel, dtype=np.int32)
        t1, t2 = _f64(temp1).reshape(-1), _f64(temp2).reshape(-1)
        self._check(self.lib.bfvm_set_agents(self.h, int(voxel.size), _iptr(voxel), _dptr(t1), _dptr(t2)))

    def set_agent_rates(self, pos, volume, secretion, saturation, uptake, dt):
        volume = _f64(volume).reshape(-1)
        n = int(volume.size)
        pos = _f64(pos).reshape(-1)
        sec, sat, upt = (_f64(x).reshape(-1) for x in (secretion, saturation, uptake))
        assert pos.size == 3 * n and sec.size == n * self.S
        vox = np.zeros(max(n, 1), dtype=np.int32)
        self._check(self.lib.bfvm_set_agent_rates(self.h, n, _dptr(pos), _dptr(volume), _dptr(sec),
                                                  _dptr(sat), _dptr(upt), float(dt), _iptr(vox)))
        return vox[:n]

    def set_agent_rates_device(self, pos, volume, secretion, saturation, uptake, dt, out_voxel=None):
        """Agents from device arrays: tensors (or objects with data_ptr()) of float64 in
        device memory, contiguous; out_voxel an int32 device tensor of n or None."""
        n = int(volume.numel())
        ptrs = [t.data_ptr() for t in (pos, volume, secretion, saturation, uptake)]
        self._check(self.lib.bfvm_set_agent_rates_device(self.h, n, *ptrs, float(dt),
                                                         out_voxel.data_ptr() if out_voxel is not None else None))

    def step(self, dt, flags=ALL):
        self._check(self.lib.bfvm_step(self.h, float(dt), int(flags)))

    def run(self, dt, steps, flags=ALL):
        self._check(self.lib.bfvm_run(self.h, float(dt), int(steps), int(flags)))

    def sync(self):
        self._check(self.lib.bfvm_sync(self.h))

    def autotune(self, dt):
        """Picks the fastest launch shape of each sweep (on a scratch copy)."""
        self._check(self.lib.bfvm_autotune(self.h, float(dt)))

    def set_autotune_segments(self, on):
        """Let autotune() choose the nodes per segment too (changes the rounding)."""
        self._check(self.lib.bfvm_set_autotune_segments(self.h, 1 if on else 0))

    def coefficient_builds(self):
        """Times the Thomas/Woodbury tables were built for this handle."""
        n = ctypes.c_uint(0)
        self._check(self.lib.bfvm_get_coefficient_builds(self.h, ctypes.byref(n)))
        return int(n.value)

    def set_sweep_plan(self, axis, lines_per_block, nodes_per_segment, streaming=0):
        """streaming (the kernel): 0 one block per column tile, 1 resident
        streaming blocks, 2 row tiles through LDS (x only; lines = rows x S)."""
        self._check(self.lib.bfvm_set_sweep_plan(self.h, int(axis), int(lines_per_block),
                                                 int(nodes_per_segment), int(streaming)))

    @property
    def fused_sources(self):
        """Bulk, agents and the first two Dirichlet overwrites inside the x sweep."""
        v = ctypes.c_int()
        self._check(self.lib.bfvm_get_fused_sources(self.h, ctypes.byref(v)))
        return bool(v.value)

    @fused_sources.setter
    def fused_sources(self, on):
        self._check(self.lib.bfvm_set_fused_sources(self.h, int(bool(on))))

    def sweep_spread(self, axis):
        """Tiles per XCD chunk of one axis (0 as dispatched, -1 contiguous eighths)."""
        v = ctypes.c_int()
        self._check(self.lib.bfvm_get_sweep_spread(self.h, int(axis), ctypes.byref(v)))
        return v.value

    def set_sweep_spread(self, axis, tiles_per_chunk):
        self._check(self.lib.bfvm_set_sweep_spread(self.h, int(axis), int(tiles_per_chunk)))

    @property
    def graph_steps(self):
        """Steps per HIP graph replay in run/timed_run (0 = every kernel launched on its own)."""
        v = ctypes.c_int()
        self._check(self.lib.bfvm_get_graph_steps(self.h, ctypes.byref(v)))
        return v.value

    @graph_steps.setter
    def graph_steps(self, steps):
        self._check(self.lib.bfvm_set_graph_steps(self.h, int(steps)))

    def set_lod_dimensions(self, dims):
        """3 = LOD_3D (default), 2 = LOD_2D (nz must be 1)."""
        self._check(self.lib.bfvm_set_lod_dimensions(self.h, int(dims)))

    def write_matlab(self, filename, origin=(0.0, 0.0, 0.0), asynchronous=False):
        """MAT v4 snapshot in write_to_matlab's layout (see biofvm_amd.matlab.read_matlab)."""
        fn = self.lib.bfvm_write_matlab_async if asynchronous else self.lib.bfvm_write_matlab
        self._check(fn(self.h, str(filename).encode(), float(origin[0]), float(origin[1]), float(origin[2])))

    def io_wait(self):
        self._check(self.lib.bfvm_io_wait(self.h))

    def mass_balance(self):
        """(per-substrate FP64 sums of the store, number of non-finite values),
        reduced on the device (bfvm_mass_balance)."""
        sums = np.zeros(self.S)
        bad = ctypes.c_longlong(0)
        self._check(self.lib.bfvm_mass_balance(self.h, _dptr(sums), ctypes.byref(bad)))
        return sums, int(bad.value)

    def gradients(self):
        """Gradient vectors of the current densities, shape (nvox, S, 3)."""
        self._check(self.lib.bfvm_compute_gradients(self.h))
        out = np.empty(self.nvox * self.S * 3)
        self._check(self.lib.bfvm_download_gradients(self.h, _dptr(out)))
        return out.reshape(self.nvox, self.S, 3)

    @property
    def sweep_chunk(self):
        """z planes per x/y chunk (0 = whole-store sweeps)."""
        v = ctypes.c_int(0)
        self._check(self.lib.bfvm_get_sweep_chunk(self.h, ctypes.byref(v)))
        return v.value

    @sweep_chunk.setter
    def sweep_chunk(self, planes):
        self._check(self.lib.bfvm_set_sweep_chunk(self.h, int(planes)))

    def sweep_plans(self):
        """[(lines_per_block, nodes_per_segment, streaming)] for the x, y and z sweeps."""
        out = []
        for axis in range(3):
            a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
            self._check(self.lib.bfvm_get_sweep_plan(self.h, axis, ctypes.byref(a), ctypes.byref(b),
                                                     ctypes.byref(c)))
            out.append((a.value, b.value, c.value))
        return out

    def sweep_cache_policies(self):
        """[non_temporal] for the x, y and z sweeps (1: streaming loads and stores)."""
        out = []
        for axis in range(3):
            v = ctypes.c_int()
            self._check(self.lib.bfvm_get_sweep_cache_policy(self.h, axis, ctypes.byref(v)))
            out.append(v.value)
        return out

    def set_sweep_cache_policy(self, axis, non_temporal):
        self._check(self.lib.bfvm_set_sweep_cache_policy(self.h, int(axis), int(bool(non_temporal))))

    # ---- differentiation (adjoint) -------------------------------------
    def tape_record(self, dt, steps, flags=ALL, checkpoint_every=50):
        """Runs `steps` steps from the current store, recording checkpoints."""
        self._check(self.lib.bfvm_tape_record(self.h, float(dt), int(steps), int(flags),
                                              int(checkpoint_every)))

    def tape_seed(self, seed):
        """Sets dJ/drho_N (nvox*S values)."""
        seed = _f64(seed)
        assert seed.size == self.nvox * self.S
        self._check(self.lib.bfvm_tape_seed(self.h, _dptr(seed)))

    def tape_seed_l2(self, target):
        """J = sum (rho_N - target)^2 / 2 (FP64); the seed becomes rho_N - target. Returns J."""
        target = _f64(target)
        assert target.size == self.nvox * self.S
        loss = ctypes.c_double(0)
        self._check(self.lib.bfvm_tape_seed_l2(self.h, _dptr(target), ctypes.byref(loss)))
        return loss.value

    def tape_backward(self, n_agents=0, rates=True):
        """Runs the adjoint; returns dict of gradients (grad_rho0, grad_D,
        grad_lambda and, with agents set from rates, grad_secretion /
        grad_uptake of shape (n_agents, S))."""
        g0 = np.empty(self.nvox * self.S)
        gD = np.empty(self.S)
        gL = np.empty(self.S)
        gs = np.empty(max(n_agents, 1) * self.S)
        gu = np.empty(max(n_agents, 1) * self.S)
        want = rates and n_agents > 0
        self._check(self.lib.bfvm_tape_backward(self.h, _dptr(g0), _dptr(gD), _dptr(gL),
                                                _dptr(gs) if want else None,
                                                _dptr(gu) if want else None))
        out = dict(grad_rho0=g0, grad_D=gD, grad_lambda=gL)
        if want:
            out["grad_secretion"] = gs.reshape(n_agents, self.S)
            out["grad_uptake"] = gu.reshape(n_agents, self.S)
        return out

    def tape_free(self):
        self._check(self.lib.bfvm_tape_free(self.h))

    @property
    def tape_keep_states(self):
        """The backward pass keeps each step's states where memory allows (True)."""
        v = ctypes.c_int(0)
        self._check(self.lib.bfvm_get_tape_keep_states(self.h, ctypes.byref(v)))
        return bool(v.value)

    @tape_keep_states.setter
    def tape_keep_states(self, on):
        self._check(self.lib.bfvm_set_tape_keep_states(self.h, 1 if on else 0))

    def tape_states_kept(self):
        """Whether the last backward pass kept every step's states."""
        v = ctypes.c_int(0)
        self._check(self.lib.bfvm_tape_states_kept(self.h, ctypes.byref(v)))
        return bool(v.value)

    def timed_run(self, dt, steps, flags=ALL, per_kernel=False):
        """Returns (ms_total, {kernel: (ms_sum, launches)} or None)."""
        ms = ctypes.c_double(0)
        pk = np.zeros(_capi.KERNEL_KINDS)
        nl = np.zeros(_capi.KERNEL_KINDS, dtype=np.int32)
        self._check(self.lib.bfvm_timed_run(self.h, float(dt), int(steps), int(flags), ctypes.byref(ms),
                                            _dptr(pk) if per_kernel else None,
                                            _iptr(nl) if per_kernel else None))
        if not per_kernel:
            return ms.value, None
        return ms.value, {n: (float(pk[k]), int(nl[k])) for k, n in enumerate(_capi.KERNEL_NAMES)}


def setup_case(case, precision=FP64):
    """Creates a DeviceMicroenvironment holding every input of `case`."""
    m = DeviceMicroenvironment(case["nx"], case["ny"], case["nz"], case["S"], case["dx"], precision)
    if case.get("lod_dims", 3) != 3:
        m.set_lod_dimensions(case["lod_dims"])
    m.set_coefficients(case["D"], case["lam"])
    m.upload(case["rho0"])
    if case.get("dir_idx") is not None and len(case["dir_idx"]) > 0:
        m.set_dirichlet(case["dir_idx"], case["dir_val"], case.get("dir_mask"))
    if case.get("bulk_supply") is not None:
        m.set_bulk(case["bulk_supply"], case["bulk_target"], case["bulk_uptake"])
    if case.get("agent_pos") is not None and len(case["agent_volume"]) > 0:
        m.set_agent_rates(case["agent_pos"], case["agent_volume"], case["agent_secretion"],
                          case["agent_saturation"], case["agent_uptake"], case["dt"])
    return m


def run_case(case, precision=FP64):
    """Runs `case` for its step count on the device; returns final densities (float64)."""
    with setup_case(case, precision) as m:
        m.run(case["dt"], case["steps"], case["phases"])
        return m.download()


class Ensemble:
    """Several DeviceMicroenvironments (one precision) advanced together: one
    batched launch per phase and sweep axis, optionally replayed from a HIP
    graph.  The members stay usable; their entry points wait for the
    ensemble's work.  Close the ensemble before closing its members."""

    def __init__(self, members):
        self.lib = _capi.load()
        self.members = list(members)
        arr = (ctypes.c_void_p * len(self.members))(*[m.h.value for m in self.members])
        self.e = ctypes.c_void_p()
        rc = self.lib.bfvm_ensemble_create(len(self.members), arr, ctypes.byref(self.e))
        if rc:
            raise _capi.BfvmError(rc, "bfvm_ensemble_create: members must be distinct live handles "
                                      "of one precision, in no other ensemble")

    def close(self):
        if self.e:
            self.lib.bfvm_ensemble_destroy(self.e)
            self.e = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        check_ensemble(rc, self.e)

    def run(self, dt, steps, flags=ALL, graph=True):
        self._check(self.lib.bfvm_ensemble_run(self.e, float(dt), int(steps), int(flags), int(graph)))

    def sync(self):
        self._check(self.lib.bfvm_ensemble_sync(self.e))

    def timed_run(self, dt, steps, flags=ALL, graph=True):
        ms = ctypes.c_double(0)
        self._check(self.lib.bfvm_ensemble_timed_run(self.e, float(dt), int(steps), int(flags),
                                                     int(graph), ctypes.byref(ms)))
        return ms.value

    def autotune(self, dt):
        self._check(self.lib.bfvm_ensemble_autotune(self.e, float(dt)))
        return self.segment

    @property
    def segment(self):
        v = ctypes.c_int(0)
        self._check(self.lib.bfvm_ensemble_get_segment(self.e, ctypes.byref(v)))
        return v.value

    @segment.setter
    def segment(self, nodes):
        self._check(self.lib.bfvm_ensemble_set_segment(self.e, int(nodes)))

    @property
    def graph_steps(self):
        """Steps per graph replay of run(..., graph=True)."""
        v = ctypes.c_int(0)
        self._check(self.lib.bfvm_ensemble_get_graph_steps(self.e, ctypes.byref(v)))
        return v.value

    @graph_steps.setter
    def graph_steps(self, steps):
        self._check(self.lib.bfvm_ensemble_set_graph_steps(self.e, int(steps)))

    def axis_segments(self):
        """Nodes per segment the batched x, y and z sweeps run at."""
        out = []
        for axis in range(3):
            v = ctypes.c_int(0)
            self._check(self.lib.bfvm_ensemble_get_axis_segment(self.e, axis, ctypes.byref(v)))
            out.append(v.value)
        return out

    @property
    def x_chunks(self):
        """x sweeps with whole-cache-line chunk loads where a member can (True)."""
        v = ctypes.c_int(0)
        self._check(self.lib.bfvm_ensemble_get_x_chunks(self.e, ctypes.byref(v)))
        return bool(v.value)

    @x_chunks.setter
    def x_chunks(self, on):
        self._check(self.lib.bfvm_ensemble_set_x_chunks(self.e, 1 if on else 0))

    @property
    def row_tiles(self):
        """x sweeps on LDS-staged row tiles (True) or column tiles (False)."""
        v = ctypes.c_int(0)
        self._check(self.lib.bfvm_ensemble_get_row_tiles(self.e, ctypes.byref(v)))
        return bool(v.value)

    @row_tiles.setter
    def row_tiles(self, on):
        self._check(self.lib.bfvm_ensemble_set_row_tiles(self.e, int(bool(on))))

    @property
    def non_temporal(self):
        """Batched y and z sweeps with non-temporal loads and stores."""
        v = ctypes.c_int(0)
        self._check(self.lib.bfvm_ensemble_get_cache_policy(self.e, ctypes.byref(v)))
        return bool(v.value)

    @non_temporal.setter
    def non_temporal(self, on):
        self._check(self.lib.bfvm_ensemble_set_cache_policy(self.e, int(bool(on))))
