"""lbm_grid_stats_noalloc (no allocation; partial results in a static device
array) against lbm_grid_stats (per-call device counters), bit for bit,
through the C ABI: two calls in a row, the second on a smaller input whose
statistics are smaller than the first's (stale partial results of the first
call would show), then empty and all-non-finite inputs and a bad size."""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def stats(fn, x):
    from lbm_isca_amd import _lib
    out = np.zeros(2)
    _lib.check(fn(x.numel(), _lib.ptr(x) if x.numel() else None, _lib.dptr(out), _lib.stream()))
    return out


def field(n, seed, big=None, bad=0):
    import torch
    rng = np.random.default_rng(seed)
    a = rng.standard_normal(n)
    if big is not None:
        a[rng.integers(n)] = big
    if bad:
        idx = rng.choice(n, bad, replace=False)
        a[idx] = np.array([np.nan, np.inf, -np.inf])[np.arange(bad) % 3]
    return torch.tensor(a, device="cuda"), a


def test_noalloc_stats_match_allocating_entry_point():
    from lbm_isca_amd import _lib
    lib = _lib.load()
    # 3 000 001 doubles: more blocks wanted than the scratch holds (capped
    # grid-stride launch); 1000 doubles: 4 blocks
    cases = [(3_000_001, 1, -7.5e300, 5), (1000, 2, None, 0), (257, 3, 2.0, 1),
             (1, 4, None, 0)]
    for n, seed, big, bad in cases:
        x, a = field(n, seed, big, bad)
        want = stats(lib.lbm_grid_stats, x)
        got = stats(lib.lbm_grid_stats_noalloc, x)
        finite = a[np.isfinite(a)]
        assert want[0] == np.abs(finite).max() and want[1] == bad, (n, want)
        assert got.tobytes() == want.tobytes(), (n, got, want)
        # a second call on the same input: the same answer
        assert stats(lib.lbm_grid_stats_noalloc, x).tobytes() == want.tobytes()
    import torch
    z = torch.zeros(0, dtype=torch.float64, device="cuda")
    assert stats(lib.lbm_grid_stats_noalloc, z).tolist() == [0.0, 0.0]
    nan = torch.full((5000,), float("nan"), dtype=torch.float64, device="cuda")
    assert stats(lib.lbm_grid_stats_noalloc, nan).tolist() == [0.0, 5000.0]
    out = np.zeros(2)
    rc = lib.lbm_grid_stats_noalloc(-1, None, _lib.dptr(out), _lib.stream())
    assert rc == _lib.LBM_ERR_VALUE
    assert "bad arguments" in _lib.last_error()


def test_noalloc_stats_on_a_side_stream():
    """The stream argument is honoured: a non-default stream, the input
    written on it just before the call."""
    import torch
    from lbm_isca_amd import _lib
    lib = _lib.load()
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        x = torch.linspace(-3.0, 2.0, 100_000, dtype=torch.float64, device="cuda")
        out = np.zeros(2)
        _lib.check(lib.lbm_grid_stats_noalloc(x.numel(), _lib.ptr(x), _lib.dptr(out),
                                              ctypes.c_void_p(s.cuda_stream)))
    assert out.tolist() == [3.0, 0.0]
