"""ctypes binding of the C ABI declared in include/biofvm_amd.h.

The shared library is built in-tree (``python -m biofvm_amd.build``).  There
is no fallback: if ``libbiofvm_amd.so`` is missing or fails to load, every
entry point raises.
"""
import ctypes
import os

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libbiofvm_amd.so")

FP64, FP32 = 0, 1
BULK, CELLS, DIFFUSION, ALL = 1, 2, 4, 7
K_DIRICHLET, K_BULK, K_AGENTS, K_SWEEP_X, K_SWEEP_Y, K_SWEEP_Z, KERNEL_KINDS = range(7)
KERNEL_NAMES = ["dirichlet", "bulk", "agents", "sweep_x", "sweep_y", "sweep_z"]

# Every exported symbol with its ctypes signature: (restype, argtypes).
_h = ctypes.c_void_p
_i = ctypes.c_int
_d = ctypes.c_double
_vp = ctypes.c_void_p
_ip = ctypes.POINTER(ctypes.c_int)
_dp = ctypes.POINTER(ctypes.c_double)
_up = ctypes.POINTER(ctypes.c_ubyte)
_sp = ctypes.POINTER(ctypes.c_size_t)

SIGNATURES = {
    "bfvm_create": (_i, [_i, _i, _i, _i, _d, _i, ctypes.POINTER(_h)]),
    "bfvm_create_mesh": (_i, [_i, _i, _i, _i, _d, _d, _d, _i, ctypes.POINTER(_h)]),
    "bfvm_destroy": (_i, [_h]),
    "bfvm_set_coefficients": (_i, [_h, _dp, _dp]),
    "bfvm_upload_density": (_i, [_h, _dp]),
    "bfvm_download_density": (_i, [_h, _dp]),
    "bfvm_set_dirichlet": (_i, [_h, _i, _ip, _dp, _up]),
    "bfvm_set_bulk": (_i, [_h, _dp, _dp, _dp]),
    "bfvm_set_agents": (_i, [_h, _i, _ip, _dp, _dp]),
    "bfvm_set_agent_rates": (_i, [_h, _i, _dp, _dp, _dp, _dp, _dp, _d, _ip]),
    "bfvm_set_agent_rates_device": (_i, [_h, _i, _vp, _vp, _vp, _vp, _vp, _d, _vp]),
    "bfvm_step": (_i, [_h, _d, _i]),
    "bfvm_run": (_i, [_h, _d, _i, _i]),
    "bfvm_sync": (_i, [_h]),
    "bfvm_timed_run": (_i, [_h, _d, _i, _i, _dp, _dp, _ip]),
    "bfvm_autotune": (_i, [_h, _d]),
    "bfvm_set_autotune_segments": (_i, [_h, _i]),
    "bfvm_get_autotune_segments": (_i, [_h, _ip]),
    "bfvm_get_coefficient_builds": (_i, [_h, ctypes.POINTER(ctypes.c_uint)]),
    "bfvm_get_sweep_plan": (_i, [_h, _i, _ip, _ip, _ip]),
    "bfvm_set_sweep_plan": (_i, [_h, _i, _i, _i, _i]),
    "bfvm_get_sweep_spread": (_i, [_h, _i, _ip]),
    "bfvm_get_sweep_cache_policy": (_i, [_h, _i, _ip]),
    "bfvm_set_sweep_cache_policy": (_i, [_h, _i, _i]),
    "bfvm_set_fused_sources": (_i, [_h, _i]),
    "bfvm_get_fused_sources": (_i, [_h, _ip]),
    "bfvm_set_sweep_spread": (_i, [_h, _i, _i]),
    "bfvm_set_graph_steps": (_i, [_h, _i]),
    "bfvm_get_graph_steps": (_i, [_h, _ip]),
    "bfvm_mass_balance": (_i, [_h, _dp, ctypes.POINTER(ctypes.c_longlong)]),
    "bfvm_mass_balance_noalloc": (_i, [_h, _vp, _vp]),
    "bfvm_compute_gradients": (_i, [_h]),
    "bfvm_download_gradients": (_i, [_h, _dp]),
    "bfvm_gradient_device_ptr": (_i, [_h, ctypes.POINTER(_vp)]),
    "bfvm_write_matlab": (_i, [_h, ctypes.c_char_p, _d, _d, _d]),
    "bfvm_write_matlab_async": (_i, [_h, ctypes.c_char_p, _d, _d, _d]),
    "bfvm_io_wait": (_i, [_h]),
    "bfvm_set_lod_dimensions": (_i, [_h, _i]),
    "bfvm_get_sweep_chunk": (_i, [_h, _ip]),
    "bfvm_set_sweep_chunk": (_i, [_h, _i]),
    "bfvm_get_info": (_i, [_h, _ip, _ip, _ip, _ip, _ip, _sp]),
    "bfvm_density_device_ptr": (_i, [_h, ctypes.POINTER(_vp)]),
    "bfvm_last_error": (ctypes.c_char_p, [_h]),
    "bfvm_ensemble_create": (_i, [_i, ctypes.POINTER(_h), ctypes.POINTER(_vp)]),
    "bfvm_ensemble_destroy": (_i, [_vp]),
    "bfvm_ensemble_run": (_i, [_vp, _d, _i, _i, _i]),
    "bfvm_ensemble_sync": (_i, [_vp]),
    "bfvm_ensemble_timed_run": (_i, [_vp, _d, _i, _i, _i, _dp]),
    "bfvm_ensemble_autotune": (_i, [_vp, _d]),
    "bfvm_ensemble_get_segment": (_i, [_vp, _ip]),
    "bfvm_ensemble_set_segment": (_i, [_vp, _i]),
    "bfvm_ensemble_get_axis_segment": (_i, [_vp, _i, _ip]),
    "bfvm_ensemble_set_graph_steps": (_i, [_vp, _i]),
    "bfvm_ensemble_set_x_chunks": (_i, [_vp, _i]),
    "bfvm_ensemble_get_x_chunks": (_i, [_vp, _ip]),
    "bfvm_ensemble_get_graph_steps": (_i, [_vp, _ip]),
    "bfvm_ensemble_set_autotune_segments": (_i, [_vp, _i]),
    "bfvm_ensemble_set_row_tiles": (_i, [_vp, _i]),
    "bfvm_ensemble_get_row_tiles": (_i, [_vp, _ip]),
    "bfvm_ensemble_set_cache_policy": (_i, [_vp, _i]),
    "bfvm_ensemble_get_cache_policy": (_i, [_vp, _ip]),
    "bfvm_ensemble_last_error": (ctypes.c_char_p, [_vp]),
    "bfvm_tape_record": (_i, [_h, _d, _i, _i, _i]),
    "bfvm_tape_seed": (_i, [_h, _dp]),
    "bfvm_tape_seed_l2": (_i, [_h, _dp, _dp]),
    "bfvm_tape_backward": (_i, [_h, _dp, _dp, _dp, _dp, _dp]),
    "bfvm_tape_free": (_i, [_h]),
    "bfvm_set_tape_keep_states": (_i, [_vp, _i]),
    "bfvm_get_tape_keep_states": (_i, [_vp, _ip]),
    "bfvm_tape_states_kept": (_i, [_vp, _ip]),
}

_lib = None


class BfvmError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"biofvm_amd error {code}: {message}")
        self.code = code


def load():
    """Loads libbiofvm_amd.so once and binds every signature."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m biofvm_amd.build`")
        # One HIP runtime per process: PyTorch's HIP libraries ask for
        # libamdhip64 under another name than this library does, so if this
        # library loaded the system runtime first, torch would load its own
        # copy beside it and find no GPU.  Loaded after torch, this library
        # binds to the runtime torch brought in (same soname).
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(rc, handle=None):
    if rc != 0:
        msg = load().bfvm_last_error(handle)
        raise BfvmError(rc, msg.decode() if msg else "")


def check_ensemble(rc, ens):
    if rc != 0:
        msg = load().bfvm_ensemble_last_error(ens)
        raise BfvmError(rc, msg.decode() if msg else "")
