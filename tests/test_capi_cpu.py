"""The C-ABI library loads and exports every entry point the header declares
(no compute: this runs without a GPU)."""

import ctypes
import os
import re

from lbm_isca_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "lbm_hip.h")).read()
    return sorted(set(re.findall(r"^int\s+(lbm_\w+)\s*\(", text, flags=re.M)))


def test_header_and_binding_agree():
    assert sorted(_lib.EXPORTS) == declared_symbols()


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_last_error_is_callable_without_device():
    assert isinstance(_lib.last_error(), str)


def test_abi_version_of_header_binding_and_library_agree():
    text = open(os.path.join(ROOT, "include", "lbm_hip.h")).read()
    header = int(re.search(r"#define LBM_ABI_VERSION (\d+)", text).group(1))
    assert _lib.ABI_VERSION == header
    assert ctypes.CDLL(_lib.LIB_PATH).lbm_abi_version() == header


def test_plan_create_rejects_a_fourier_slab_beyond_4gib():
    """Validation precedes any device call: one wavenumber of the analysed
    Fourier columns (nlat x (9 nlev + 2) nmem complex) must stay below
    4 GiB, since the analysis addresses it with 32-bit offsets."""
    import numpy as np
    lib = _lib.load()
    nmax, nlev = 42, 20
    sig = np.linspace(1.0, 0.0, nlev + 1)
    damp = np.zeros((nmax + 1) * nlev)
    d = _lib.LbmDesc()
    d.nlon, d.nlat, d.nmax, d.nlev = 128, 64, nmax, nlev
    d.nbasic, d.precision, d.inistp = 1, 0, 2
    d.radius, d.delt, d.robert = 6.371e6, 1200.0, 0.05
    d.h_sigm = _lib.dptr(sig)
    d.h_kv = d.h_kd = d.h_ktt = d.h_kq = _lib.dptr(damp)
    d.nmem = 23100                   # 64 x 182 x 23100 x 16 bytes > 4 GiB
    h = ctypes.c_void_p()
    assert lib.lbm_plan_create(ctypes.byref(d), ctypes.byref(h)) == -1
    assert "exceeds 4 GiB" in _lib.last_error()
    assert not h.value
