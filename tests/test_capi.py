"""The C-ABI library loads and exports every symbol include/ declares."""
import ctypes
import re
from pathlib import Path

import pytest

from pycd_amd import _native

HEADER = Path(__file__).resolve().parents[1] / "include" / "pycd_hip.h"


def declared_symbols():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pycd_[a-z0-9_]+)\s*\(", text)))


def test_header_matches_binding():
    assert declared_symbols() == _native.exported_symbols()


def test_library_exports_symbols():
    if not _native.LIB_PATH.exists():
        pytest.fail("libpycd_hip.so not built (python -m pycd_amd.build)")
    lib = ctypes.CDLL(str(_native.LIB_PATH))
    for name in declared_symbols():
        assert hasattr(lib, name), name
    assert lib.pycd_abi_version() == 2


def test_errors_without_context():
    lib = _native.lib()
    # null context is rejected without touching a device
    assert lib.pycd_pair_table(None, None, 0, None, None, None, 0.0, 0.0,
                               0.0, 0.0, 0, None, None) < 0
    assert lib.pycd_msd(None, None, 0, 0, 0, 0, None) < 0
    assert lib.pycd_last_error(None) == b"null context"
