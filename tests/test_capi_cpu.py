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
