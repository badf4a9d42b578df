"""ctypes binding of ``liblbm_hip.so`` (the C ABI of ``include/lbm_hip.h``).

The library is built in-tree by ``lbm_isca_amd.build`` (``hipcc
--offload-arch=gfx950``).  There is no fallback: if the library is missing
or no HIP device is visible, every entry point raises.
"""

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblbm_hip.so")

LBM_OK = 0
LBM_ERR_VALUE = -1
LBM_ERR_RUNTIME = -2
LBM_ERR_NONFINITE = -3
LBM_ERR_STATE = -4

LBM_FP64 = 0
LBM_FP32 = 1
LBM_BF16X3 = 2

# LBM_ABI_VERSION of include/lbm_hip.h these bindings are written for
ABI_VERSION = 5

# Every symbol include/lbm_hip.h declares.
EXPORTS = (
    "lbm_plan_create", "lbm_plan_destroy", "lbm_plan_get",
    "lbm_grid_to_wave", "lbm_wave_to_grid", "lbm_uv_to_vordiv", "lbm_psichi_to_uv",
    "lbm_hgrad", "lbm_set_basic_state", "lbm_set_forcing", "lbm_set_state",
    "lbm_set_lengths", "lbm_run", "lbm_get_state", "lbm_get_spectral",
    "lbm_steps_taken", "lbm_profile", "lbm_profile_read", "lbm_last_error",
    "lbm_set_spectral_state", "lbm_set_response", "lbm_get_response", "lbm_adjoint_run",
    "lbm_grid_to_wave_adj", "lbm_wave_to_grid_adj", "lbm_uv_to_vordiv_adj",
    "lbm_psichi_to_uv_adj", "lbm_hgrad_adj",
    "lbm_run_to", "lbm_set_history", "lbm_get_history", "lbm_get_diagnostics",
    "lbm_set_mjo_forcing", "lbm_run_segment_done", "lbm_spec_resize", "lbm_abi_version",
    "lbm_set_steady_forcing", "lbm_set_state_keep", "lbm_get_prev_state", "lbm_tendencies",
    "lbm_prepare_dt", "lbm_step", "lbm_robert_asselin", "lbm_lnps", "lbm_grid_stats",
    "lbm_grid_stats_noalloc",
    "lbm_spec_scale", "lbm_hydrostatic", "lbm_run_steps", "lbm_run_abort",
    "lbm_autotune", "lbm_set_tuning", "lbm_get_tuning", "lbm_set_checkpoints",
)

PROFILE_SLOTS = ("syn_level", "syn_surf", "ifft", "tend", "fft", "ana_level", "ana_surf", "step")

_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p


class LbmDesc(ctypes.Structure):
    """Mirror of ``lbm_desc``."""
    _fields_ = [
        ("nlon", ctypes.c_int32), ("nlat", ctypes.c_int32), ("nmax", ctypes.c_int32),
        ("nlev", ctypes.c_int32), ("nmem", ctypes.c_int32), ("nbasic", ctypes.c_int32),
        ("precision", ctypes.c_int32), ("strictly_linear", ctypes.c_int32),
        ("inistp", ctypes.c_int32), ("skip_low_n_vdiff", ctypes.c_int32),
        ("radius", ctypes.c_double), ("delt", ctypes.c_double), ("robert", ctypes.c_double),
        ("h_sigm", _dp),
        ("h_kv", _dp), ("h_kd", _dp), ("h_ktt", _dp), ("h_kq", _dp),
        ("h_vdif_vor", _dp), ("h_vdif_div", _dp), ("h_vdif_t", _dp), ("h_vdif_q", _dp),
        ("tangent_linear", ctypes.c_int32),
    ]


_LIB = None


def load():
    """Load the library once; raise RuntimeError if it is absent."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; "
            "g.build()'` (hipcc --offload-arch=gfx950)")
    # torch first: the library then binds to the HIP runtime torch has
    # already loaded.  Loaded the other way round the process holds two HIP
    # runtimes, and device pointers and streams of one mean nothing to the
    # other.
    import torch  # noqa: F401
    lib = ctypes.CDLL(LIB_PATH)
    P = _vp
    sig = {
        "lbm_plan_create": ([ctypes.POINTER(LbmDesc), ctypes.POINTER(P)], ctypes.c_int),
        "lbm_plan_destroy": ([P], ctypes.c_int),
        "lbm_plan_get": ([P, ctypes.c_char_p, _dp, ctypes.c_int64], ctypes.c_int),
        "lbm_grid_to_wave": ([P, P, P, ctypes.c_int, ctypes.c_int, P], ctypes.c_int),
        "lbm_wave_to_grid": ([P, P, P, ctypes.c_int, ctypes.c_int, P], ctypes.c_int),
        "lbm_uv_to_vordiv": ([P, P, P, P, P, ctypes.c_int, ctypes.c_int, P], ctypes.c_int),
        "lbm_psichi_to_uv": ([P, P, P, P, P, ctypes.c_int, ctypes.c_int, P], ctypes.c_int),
        "lbm_hgrad": ([P, P, P, P, ctypes.c_int, ctypes.c_int, P], ctypes.c_int),
        "lbm_set_basic_state": ([P, P, P, P, P, P, P], ctypes.c_int),
        "lbm_set_forcing": ([P, P, ctypes.c_int, P], ctypes.c_int),
        "lbm_set_state": ([P, P, P, P, P, P, P], ctypes.c_int),
        "lbm_set_lengths": ([P, ctypes.POINTER(ctypes.c_int32), P], ctypes.c_int),
        "lbm_run": ([P, ctypes.c_int, ctypes.c_int, ctypes.c_int, P], ctypes.c_int),
        "lbm_get_state": ([P, P, P, P, P, P, P], ctypes.c_int),
        "lbm_get_spectral": ([P, ctypes.c_int, P, P, P, P, P, P], ctypes.c_int),
        "lbm_steps_taken": ([P, ctypes.POINTER(ctypes.c_int64)], ctypes.c_int),
        "lbm_profile": ([P, ctypes.c_int], ctypes.c_int),
        "lbm_profile_read": ([P, _dp, ctypes.POINTER(ctypes.c_int64), ctypes.c_int], ctypes.c_int),
        "lbm_last_error": ([ctypes.c_char_p, ctypes.c_int64], ctypes.c_int),
        "lbm_set_spectral_state": ([P, P, P, P, P, P, P], ctypes.c_int),
        "lbm_set_response": ([P, _dp, _dp, _dp, ctypes.c_int], ctypes.c_int),
        "lbm_get_response": ([P, _dp], ctypes.c_int),
        "lbm_adjoint_run": ([P, ctypes.c_int, P, P, P, P, P, P, ctypes.c_int, P], ctypes.c_int),
        "lbm_grid_to_wave_adj": ([P, P, P, ctypes.c_int, ctypes.c_int, P], ctypes.c_int),
        "lbm_wave_to_grid_adj": ([P, P, P, ctypes.c_int, ctypes.c_int, P], ctypes.c_int),
        "lbm_uv_to_vordiv_adj": ([P, P, P, P, P, ctypes.c_int, ctypes.c_int, P], ctypes.c_int),
        "lbm_psichi_to_uv_adj": ([P, P, P, P, P, ctypes.c_int, ctypes.c_int, P], ctypes.c_int),
        "lbm_hgrad_adj": ([P, P, P, P, ctypes.c_int, ctypes.c_int, P], ctypes.c_int),
        "lbm_run_to": ([P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, P], ctypes.c_int),
        "lbm_set_history": ([P, ctypes.c_int, ctypes.c_int, ctypes.c_int, P], ctypes.c_int),
        "lbm_get_history": ([P, ctypes.c_int, P, P, P, P, P, P, P, P, P], ctypes.c_int),
        "lbm_get_diagnostics": ([P, P, P, P, P, P], ctypes.c_int),
        "lbm_set_mjo_forcing": ([P, P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_int, ctypes.c_int, P], ctypes.c_int),
        "lbm_run_segment_done": ([P, ctypes.POINTER(ctypes.c_int32)], ctypes.c_int),
        "lbm_spec_resize": ([P, P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, P],
                            ctypes.c_int),
        "lbm_abi_version": ([], ctypes.c_int),
        "lbm_set_steady_forcing": ([P, P, P, P, ctypes.c_int, P], ctypes.c_int),
        "lbm_set_state_keep": ([P, P, P, P, P, P, P], ctypes.c_int),
        "lbm_get_prev_state": ([P, P, P, P, P, P, P], ctypes.c_int),
        "lbm_tendencies": ([P] + [P] * 5 + [P] * 5 + [P], ctypes.c_int),
        "lbm_prepare_dt": ([P, ctypes.c_double], ctypes.c_int),
        "lbm_step": ([P, ctypes.c_double] + [P] * 15 + [P], ctypes.c_int),
        "lbm_robert_asselin": ([ctypes.c_int64, P, P, P, ctypes.c_double, P], ctypes.c_int),
        "lbm_lnps": ([ctypes.c_int64, ctypes.c_int64, P, P, P, ctypes.c_int, P], ctypes.c_int),
        "lbm_grid_stats": ([ctypes.c_int64, P, _dp, P], ctypes.c_int),
        "lbm_grid_stats_noalloc": ([ctypes.c_int64, P, _dp, P], ctypes.c_int),
        "lbm_spec_scale": ([P, P, P, ctypes.c_int, ctypes.c_int, ctypes.c_int, P], ctypes.c_int),
        "lbm_hydrostatic": ([P, P, ctypes.c_int64, ctypes.c_int, ctypes.c_int, P, P, P],
                            ctypes.c_int),
        "lbm_run_steps": ([P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, P],
                          ctypes.c_int),
        "lbm_run_abort": ([P], ctypes.c_int),
        "lbm_autotune": ([P, ctypes.c_int, ctypes.c_char_p, ctypes.c_int64], ctypes.c_int),
        "lbm_set_tuning": ([P, ctypes.POINTER(ctypes.c_int32), ctypes.c_int], ctypes.c_int),
        "lbm_get_tuning": ([P, ctypes.POINTER(ctypes.c_int32), ctypes.c_int], ctypes.c_int),
        "lbm_set_checkpoints": ([P, ctypes.c_int], ctypes.c_int),
    }
    if not hasattr(lib, "lbm_abi_version") or lib.lbm_abi_version() != ABI_VERSION:
        found = lib.lbm_abi_version() if hasattr(lib, "lbm_abi_version") else "none"
        raise RuntimeError(f"{LIB_PATH} has ABI version {found}, these bindings need "
                           f"{ABI_VERSION}: rebuild it (python -c 'import __graft_entry__ as g; "
                           "g.build()')")
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _LIB = lib
    return lib


def last_error():
    lib = load()
    buf = ctypes.create_string_buffer(1024)
    lib.lbm_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(code):
    """Map a status code to the reference's exception types (backend.py:71-77,
    model.py:245-247)."""
    if code >= 0:
        return code
    msg = last_error()
    if code == LBM_ERR_VALUE:
        raise ValueError(msg)
    if code == LBM_ERR_NONFINITE:
        raise FloatingPointError(msg)
    raise RuntimeError(msg)


def ptr(t):
    """Device pointer of a contiguous torch tensor (or None)."""
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


def stream(s=None):
    """``hipStream_t`` of torch's current stream as a ctypes pointer."""
    import torch
    return ctypes.c_void_p((s or torch.cuda.current_stream()).cuda_stream)


def dptr(a):
    """ctypes double* of a contiguous float64 numpy array (or None)."""
    if a is None:
        return None
    return a.ctypes.data_as(_dp)
