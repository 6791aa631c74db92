"""The C ABI library loads and exports exactly what include/*.h declares (no GPU needed)."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        names |= set(re.findall(r"FL_API\s+[\w\s\*]+?\b(fl_\w+)\s*\(", open(h).read()))
    return names


def test_header_declares_entry_points():
    names = declared()
    assert {"fl_forward_emission_batch", "fl_backward_emission_batch", "fl_mlem_batch", "fl_cal_atten"} <= names


def test_library_exports_every_declared_symbol():
    from fl_correction_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_bindings_match_header():
    from fl_correction_amd import _lib
    assert set(_lib.SIGNATURES) == declared()


def test_version_and_error_without_gpu():
    from fl_correction_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    lib = _lib.load()
    assert lib.fl_version() >> 16 == 1
    # argument checks run on the host before any device call
    rc = lib.fl_forward_emission_batch(None, None, None, None, None, 3, 2, 1, 1, 8, None)
    assert rc == 1 and b"shape" in lib.fl_last_error()
    rc = lib.fl_cal_atten(None, None, None, 1, 1, 4, 4, 1, 1, 1, None)
    assert rc == 1 and b"null" in lib.fl_last_error()
