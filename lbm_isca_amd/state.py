"""Grid-space view of the model state (``pylbm_isca/state.py:26-148``).

On the HIP backend the prognostic state lives in the library's spectral
buffers; a :class:`State` is a snapshot of the grid fields u, v, t, q and
ps' (hPa) synthesised on the device (``state.py:94-99``).
"""

import numpy as np

from .backend import asnumpy, hip


class State:
    __slots__ = ("xp", "u", "v", "t", "q", "ps", "ps_bar")

    def __init__(self, ps_bar, u, v, t, q, ps):
        self.xp = hip
        self.ps_bar, self.u, self.v, self.t, self.q, self.ps = ps_bar, u, v, t, q, ps

    def lnps_pert(self):
        """pi' = ln(ps_bar + ps') - ln(ps_bar) (state.py:77-79)."""
        import torch
        pb = torch.as_tensor(self.ps_bar, device=self.ps.device)
        return torch.log1p(self.ps / pb)

    def max_abs(self):
        return {k: float(getattr(self, k).abs().max()) for k in ("u", "v", "t", "ps", "q")}

    def isfinite(self):
        import torch
        return all(bool(torch.isfinite(getattr(self, k)).all()) for k in ("u", "v", "t", "q", "ps"))

    def to_numpy(self):
        return {k: asnumpy(getattr(self, k)) for k in ("u", "v", "t", "q", "ps")}

    def copy(self):
        return State(np.array(self.ps_bar), *(getattr(self, k).clone()
                                              for k in ("u", "v", "t", "q", "ps")))
