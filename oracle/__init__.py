"""TEST INFRASTRUCTURE ONLY: the CPU oracle for the diffusion-decay path.

Two CPU implementations of BioFVM 1.0.4's step, loaded through ctypes:

* ``oracle_run`` — the plain-C restatement in ``oracle/bfvm_oracle.c``
  (built to ``oracle/libbfvm_oracle.so``), which follows the reference's
  arithmetic order (BioFVM_solvers.cpp:77-276, BioFVM_microenvironment.cpp:
  192-197, 584-633, BioFVM_basic_agent.cpp:114-148, 229-238).
* ``ref_run`` / ``RefSession`` — the compiled reference itself
  (``oracle/_ref/libbfvm_ref.so``, built from the unmodified reference sources
  by ``oracle/Makefile``), driven through ``oracle/ref_driver.cpp``.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg import this package, and only as the checker.  The product library under
``biofvm_amd/`` never loads it.

A *case* is a dict with keys
``nx, ny, nz, S, dx, D, lam, dt, steps, phases, rho0`` and optionally
``dir_idx, dir_val, dir_mask, bulk_supply, bulk_target, bulk_uptake,
agent_pos, agent_volume, agent_secretion, agent_saturation, agent_uptake``.
Arrays are float64 in voxel-major / substrate-innermost order.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "libbfvm_oracle.so")
REF_SO = os.path.join(HERE, "_ref", "libbfvm_ref.so")

BULK, CELLS, DIFFUSION = 1, 2, 4
ALL = BULK | CELLS | DIFFUSION

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_u8p = ctypes.POINTER(ctypes.c_ubyte)

_oracle = None
_ref = None


def _load_oracle():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            raise FileNotFoundError(f"{ORACLE_SO} missing: run `make -C oracle`")
        lib = ctypes.CDLL(ORACLE_SO)
        lib.oracle_run.restype = ctypes.c_int
        lib.oracle_thomas_coefficients.restype = None
        lib.oracle_nearest_voxel.restype = ctypes.c_int
        lib.oracle_nearest_voxel.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 7
        _oracle = lib
    return _oracle


def have_ref():
    return os.path.exists(REF_SO)


def _load_ref():
    global _ref
    if _ref is None:
        if not have_ref():
            raise FileNotFoundError(f"{REF_SO} missing: run `make -C oracle ref` where the reference tree exists")
        lib = ctypes.CDLL(REF_SO)
        lib.ref_run.restype = ctypes.c_int
        lib.ref_create.restype = ctypes.c_void_p
        lib.ref_steps.restype = ctypes.c_int
        lib.ref_steps.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  _dp]
        lib.ref_read.restype = None
        lib.ref_read.argtypes = [ctypes.c_void_p, _dp]
        lib.ref_destroy.restype = None
        lib.ref_destroy.argtypes = [ctypes.c_void_p]
        _ref = lib
    return _ref


def _ptr(a, kind=_dp):
    if a is None:
        return ctypes.cast(None, kind)
    return a.ctypes.data_as(kind)


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.int32)


def _phases(case):
    """Phase flags plus 8 (the 2-D solver) for a case with lod_dims == 2."""
    return int(case["phases"]) | (8 if case.get("lod_dims", 3) == 2 else 0)


def _common(case):
    S = int(case["S"])
    D = _f64(case["D"])
    lam = _f64(case["lam"])
    assert D.shape == (S,) and lam.shape == (S,)
    rho = _f64(case["rho0"]).copy()
    assert rho.size == case["nx"] * case["ny"] * case["nz"] * S
    dir_idx = _i32(case.get("dir_idx"))
    dir_val = _f64(case.get("dir_val"))
    n_dir = 0 if dir_idx is None else int(dir_idx.size)
    bulk = [_f64(case.get(k)) for k in ("bulk_supply", "bulk_target", "bulk_uptake")]
    ag = [_f64(case.get(k)) for k in ("agent_pos", "agent_volume", "agent_secretion",
                                        "agent_saturation", "agent_uptake")]
    n_agents = 0 if ag[0] is None else int(ag[1].size)
    return S, D, lam, rho, n_dir, dir_idx, dir_val, bulk, n_agents, ag


def oracle_run(case, return_voxels=False):
    """Runs the C restatement on `case`; returns the final densities (flat)."""
    lib = _load_oracle()
    S, D, lam, rho, n_dir, dir_idx, dir_val, bulk, n_agents, ag = _common(case)
    mask = case.get("dir_mask")
    mask = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
    vox = np.zeros(max(n_agents, 1), dtype=np.int32)
    rc = lib.oracle_run(
        ctypes.c_int(case["nx"]), ctypes.c_int(case["ny"]), ctypes.c_int(case["nz"]),
        ctypes.c_int(S), ctypes.c_double(case["dx"]), _ptr(D), _ptr(lam),
        ctypes.c_double(case["dt"]), ctypes.c_int(case["steps"]), ctypes.c_int(_phases(case)),
        _ptr(rho), ctypes.c_int(n_dir), _ptr(dir_idx, _ip), _ptr(dir_val), _ptr(mask, _u8p),
        _ptr(bulk[0]), _ptr(bulk[1]), _ptr(bulk[2]),
        ctypes.c_int(n_agents), _ptr(ag[0]), _ptr(ag[1]), _ptr(ag[2]), _ptr(ag[3]), _ptr(ag[4]),
        _ptr(vox, _ip))
    if rc != 0:
        raise ValueError(f"oracle_run rejected the case (code {rc})")
    return (rho, vox[:n_agents]) if return_voxels else rho


def ref_run(case, nthreads=1, return_voxels=False, return_seconds=False):
    """Runs the compiled reference on `case` (whole-vector Dirichlet only)."""
    lib = _load_ref()
    S, D, lam, rho, n_dir, dir_idx, dir_val, bulk, n_agents, ag = _common(case)
    if case.get("dir_mask") is not None and not np.all(np.asarray(case["dir_mask"]) != 0):
        raise ValueError("the reference has no per-substrate Dirichlet mask")
    vox = np.zeros(max(n_agents, 1), dtype=np.int32)
    secs = ctypes.c_double(0.0)
    rc = lib.ref_run(
        ctypes.c_int(case["nx"]), ctypes.c_int(case["ny"]), ctypes.c_int(case["nz"]),
        ctypes.c_int(S), ctypes.c_double(case["dx"]), _ptr(D), _ptr(lam),
        ctypes.c_double(case["dt"]), ctypes.c_int(case["steps"]), ctypes.c_int(_phases(case)),
        _ptr(rho), ctypes.c_int(n_dir), _ptr(dir_idx, _ip), _ptr(dir_val),
        _ptr(bulk[0]), _ptr(bulk[1]), _ptr(bulk[2]),
        ctypes.c_int(n_agents), _ptr(ag[0]), _ptr(ag[1]), _ptr(ag[2]), _ptr(ag[3]), _ptr(ag[4]),
        _ptr(vox, _ip), ctypes.c_int(nthreads), ctypes.byref(secs))
    if rc != 0:
        raise ValueError(f"ref_run rejected the case (code {rc})")
    out = [rho]
    if return_voxels:
        out.append(vox[:n_agents])
    if return_seconds:
        out.append(secs.value)
    return out[0] if len(out) == 1 else tuple(out)


class RefSession:
    """The compiled reference set up once on `case` and stepped on demand
    (ref_create / ref_steps / ref_read / ref_destroy in ref_driver.cpp), so
    that its set-up (tens of seconds at 400^3) stays out of a timing.  One
    session at a time: the reference keeps its agents in globals."""

    def __init__(self, case):
        lib = self.lib = _load_ref()
        if case.get("dir_mask") is not None and not np.all(np.asarray(case["dir_mask"]) != 0):
            raise ValueError("the reference has no per-substrate Dirichlet mask")
        S, D, lam, rho, n_dir, dir_idx, dir_val, bulk, n_agents, ag = _common(case)
        self.case, self.nvs = case, rho.size
        self._keep = (bulk, ag, dir_idx, dir_val)  # the bulk callbacks read these at every step
        self.voxels = np.zeros(max(n_agents, 1), dtype=np.int32)
        rc = ctypes.c_int(0)
        self.h = lib.ref_create(
            ctypes.c_int(case["nx"]), ctypes.c_int(case["ny"]), ctypes.c_int(case["nz"]),
            ctypes.c_int(S), ctypes.c_double(case["dx"]), _ptr(D), _ptr(lam), ctypes.c_double(case["dt"]),
            ctypes.c_int(_phases(case)), _ptr(rho), ctypes.c_int(n_dir), _ptr(dir_idx, _ip), _ptr(dir_val),
            _ptr(bulk[0]), _ptr(bulk[1]), _ptr(bulk[2]),
            ctypes.c_int(n_agents), _ptr(ag[0]), _ptr(ag[1]), _ptr(ag[2]), _ptr(ag[3]), _ptr(ag[4]),
            _ptr(self.voxels, _ip), ctypes.byref(rc))
        if not self.h:
            raise ValueError(f"ref_create rejected the case (code {rc.value})")
        self.voxels = self.voxels[:n_agents]

    def steps(self, steps, phases=None, nthreads=0):
        """Runs `steps` steps; returns the wall-clock seconds of the step loop."""
        secs = ctypes.c_double(0.0)
        ph = _phases(self.case) if phases is None else int(phases)
        self.lib.ref_steps(self.h, ctypes.c_double(self.case["dt"]), int(steps), ph, int(nthreads),
                           ctypes.byref(secs))
        return secs.value

    def read(self):
        rho = np.empty(self.nvs)
        self.lib.ref_read(self.h, _ptr(rho))
        return rho

    def close(self):
        if self.h:
            self.lib.ref_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def thomas_coefficients(n, S, dx, dt, D, lam):
    """(denom, c), each shaped (n, S), in the reference's order (BioFVM_solvers.cpp:104-142)."""
    lib = _load_oracle()
    D = _f64(D)
    lam = _f64(lam)
    denom = np.zeros((n, S))
    c = np.zeros((n, S))
    lib.oracle_thomas_coefficients(ctypes.c_int(n), ctypes.c_int(S), ctypes.c_double(dx),
                                   ctypes.c_double(dt), _ptr(D), _ptr(lam), _ptr(denom), _ptr(c))
    return denom, c


def agent_constants(S, dt, voxel_volume, volume, secretion, saturation, uptake):
    """(temp1, temp2), each (n, S) (BioFVM_basic_agent.cpp:114-148)."""
    lib = _load_oracle()
    volume = _f64(volume)
    n = int(volume.size)
    sec, sat, upt = _f64(secretion), _f64(saturation), _f64(uptake)
    t1 = np.zeros((n, S))
    t2 = np.zeros((n, S))
    lib.oracle_agent_constants(ctypes.c_int(n), ctypes.c_int(S), ctypes.c_double(dt),
                               ctypes.c_double(voxel_volume), _ptr(volume), _ptr(sec), _ptr(sat),
                               _ptr(upt), _ptr(t1), _ptr(t2))
    return t1, t2


def agent_step(S, rho, voxel, temp1, temp2):
    """One agent phase on `rho` (flat, S values per voxel) in place: every
    agent in list order applies rho[v] = (rho[v] + temp1) / temp2
    (oracle_agent_step; BioFVM_basic_agent.cpp:229-238)."""
    lib = _load_oracle()
    assert rho.dtype == np.float64 and rho.flags.c_contiguous
    voxel, t1, t2 = _i32(voxel), _f64(temp1), _f64(temp2)
    lib.oracle_agent_step(ctypes.c_int(S), _ptr(rho), ctypes.c_int(int(voxel.size)), _ptr(voxel, _ip),
                          _ptr(t1), _ptr(t2))
    return rho


def nearest_voxel(nx, ny, nz, dx, x, y, z):
    """Voxel index of a point, or -1 outside the closed box (BioFVM_mesh.cpp:930-948)."""
    return _load_oracle().oracle_nearest_voxel(nx, ny, nz, dx, 0.0, 0.0, 0.0, x, y, z)


def gradients(nx, ny, nz, S, dx, rho, dy=None, dz=None):
    """Gradient vectors of every voxel, shape (nvox, S, 3) (oracle_gradients)."""
    lib = _load_oracle()
    rho = _f64(rho)
    out = np.empty(nx * ny * nz * S * 3)
    lib.oracle_gradients(nx, ny, nz, S, ctypes.c_double(dx), ctypes.c_double(dx if dy is None else dy),
                         ctypes.c_double(dx if dz is None else dz), _ptr(rho), _ptr(out))
    return out.reshape(nx * ny * nz, S, 3)


def ref_gradients(nx, ny, nz, S, dx, rho):
    """The compiled reference's gradient vectors, shape (nvox, S, 3).  The
    reference fixes 2*dx on its first call in a process: use one dx per process."""
    lib = _load_ref()
    rho = _f64(rho)
    out = np.empty(nx * ny * nz * S * 3)
    lib.ref_gradients.restype = ctypes.c_int
    rc = lib.ref_gradients(nx, ny, nz, S, ctypes.c_double(dx), _ptr(rho), _ptr(out))
    if rc != 0:
        raise RuntimeError(f"ref_gradients failed ({rc})")
    return out.reshape(nx * ny * nz, S, 3)


def ref_write_multicellds(nx, ny, nz, S, dx, D, lam, rho, filename_base, current_time=0.0,
                          current_runtime=0.0, mesh_as_matlab=False, data_as_matlab=True, fresh=True):
    """The compiled reference's save_microenvironment_to_MultiCellDS_xml_pugi
    (writes filename_base.xml and the MAT files it names); fresh=False
    updates the document of an earlier call, as the reference's later calls do."""
    lib = _load_ref()
    rho, D, lam = _f64(rho), _f64(D), _f64(lam)
    lib.ref_write_multicellds.restype = ctypes.c_int
    rc = lib.ref_write_multicellds(nx, ny, nz, S, ctypes.c_double(dx), _ptr(D), _ptr(lam), _ptr(rho),
                                   str(filename_base).encode(), ctypes.c_double(current_time),
                                   ctypes.c_double(current_runtime), int(bool(mesh_as_matlab)),
                                   int(bool(data_as_matlab)), int(bool(fresh)))
    if rc != 0:
        raise RuntimeError(f"ref_write_multicellds failed ({rc})")


def ref_write_matlab(nx, ny, nz, S, dx, rho, filename):
    """The compiled reference's write_to_matlab of densities rho (origin 0)."""
    lib = _load_ref()
    rho = _f64(rho)
    lib.ref_write_matlab.restype = ctypes.c_int
    rc = lib.ref_write_matlab(nx, ny, nz, S, ctypes.c_double(dx), _ptr(rho), str(filename).encode())
    if rc != 0:
        raise RuntimeError(f"ref_write_matlab failed ({rc})")
