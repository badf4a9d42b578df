"""Backend selection with the reference's interface (``pylbm_isca/backend.py``).

The reference maps a device string to an *array module* (numpy or CuPy) and
every hot object stores it as ``self.xp`` (``backend.py:53-77``).  This
package adds the ``"hip"`` device: objects store :data:`hip`, a small array
namespace over PyTorch-ROCm device tensors, and run their hot path through
the gfx950 kernels of ``liblbm_hip.so``.  There is no numpy fallback here;
``"cpu"``/``"numpy"`` belong to the reference's own backend.
"""

import os
import types

import numpy as np

from . import _lib


def _torch():
    import torch
    return torch


def gpu_available():
    """True when a HIP device is visible and the gfx950 library loads."""
    try:
        torch = _torch()
        if not torch.cuda.is_available():
            return False
        _lib.load()
        return True
    except Exception:
        return False


def _asarray(x, dtype=None):
    torch = _torch()
    if isinstance(x, torch.Tensor):
        t = x.to(device="cuda")
    else:
        t = torch.as_tensor(np.asarray(x), device="cuda")
    if dtype is not None:
        t = t.to(dtype)
    return t.contiguous()


def _zeros(shape, dtype=None):
    torch = _torch()
    return torch.zeros(shape, dtype=dtype or torch.float64, device="cuda")


hip = types.SimpleNamespace(name="hip", asarray=_asarray, zeros=_zeros)


def resolve(device):
    """Map a device spec to the array namespace of this backend.

    ``"hip"``/``"gpu"``/``"cuda"``/``"rocm"`` select the MI355X backend;
    ``"auto"``/``None`` honour ``$LBM_DEVICE`` and otherwise require a HIP
    device.  Raises RuntimeError when a GPU is requested but absent and
    ValueError for an unknown device (``backend.py:71-77``).
    """
    if device is None or str(device).lower() == "auto":
        env = os.environ.get("LBM_DEVICE", "").strip().lower()
        if env and env != "auto":
            return resolve(env)
        device = "hip"
    d = str(device).lower()
    if d in ("hip", "gpu", "cuda", "rocm"):
        if not gpu_available():
            raise RuntimeError(
                "device='hip' requested but no HIP device is visible or liblbm_hip.so "
                "failed to load")
        return hip
    if d in ("cpu", "numpy"):
        raise ValueError(
            "lbm_isca_amd implements the 'hip' device only; the numpy backend is the "
            "reference package's own")
    raise ValueError(f"unknown device {device!r}")


def get_array_module(*arrays):
    return hip


def asnumpy(x):
    """Host numpy copy of a device tensor (no-op for numpy input)."""
    torch = _torch()
    if isinstance(x, torch.Tensor):
        return x.detach().cpu().numpy()
    return np.asarray(x)


def to_device(x, xp=None):
    return _asarray(x)


def device_name(xp=None):
    return "hip"


def synchronize(xp=None):
    """Block until queued GPU work finishes (``backend.py:108-114``)."""
    torch = _torch()
    if torch.cuda.is_available():
        torch.cuda.synchronize()
