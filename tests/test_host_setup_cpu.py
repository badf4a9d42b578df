"""Host-side FP64 setup of the product (damping coefficients, vertical
diffusion rates) against the oracle's restatement."""

import numpy as np

from lbm_isca_amd.model import Config, _vdif_rate, damping_coefficients
from oracle.integ import Damping, vdif_rate
from oracle.vgrid import VerticalOps


def test_damping_matches_oracle():
    vg = VerticalOps()
    ours = damping_coefficients(Config(), 42, vg.sig)
    ref = Damping(42, vg, ddragv=Config().ddragv, ddragd=Config().ddragd, ddragt=Config().ddragt)
    for a, b in zip(ours, (ref.kv, ref.kd, ref.ktt, ref.kq)):
        np.testing.assert_allclose(a, b, rtol=1e-14, atol=0)


def test_vdif_rate_matches_oracle():
    for v in (None, 1e3, [1e3] * 20, [1e3] * 19 + [0.0]):
        a, b = _vdif_rate(v, 20), vdif_rate(v, 20)
        assert (a is None) == (b is None)
        if a is not None:
            np.testing.assert_allclose(a, b)


def test_bad_drag_profile_raises():
    import pytest
    cfg = Config(ddragv=[1.0] * 19)
    with pytest.raises(ValueError):
        damping_coefficients(cfg, 42, VerticalOps().sig)
