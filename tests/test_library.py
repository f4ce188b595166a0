"""CPU checks of the C-ABI library: it loads, exports every entry point that
include/sfeng.h declares, and rejects bad arguments with a status code (no
GPU work is started by these calls)."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT


def _declared():
    text = open(os.path.join(ROOT, "include", "sfeng.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sf_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    from segfault_py_amd import engine
    if not os.path.exists(engine.LIB_PATH):
        pytest.fail("libsfeng.so is not built; run python -m segfault_py_amd.build")
    return engine.lib()


def test_exports_every_declared_symbol(lib):
    from segfault_py_amd import engine
    names = _declared()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(engine.exported_symbols()) == names


def test_version_and_error_string(lib):
    assert lib.sf_version() == 100
    assert isinstance(lib.sf_last_error(), bytes)


def test_argument_errors_are_status_codes(lib):
    from segfault_py_amd import engine
    i = np.zeros(4, dtype=np.int32)
    d = np.zeros(4)
    # no functions
    assert lib.sf_eri_components(0, i, i, d, d, d, np.zeros(12, dtype=np.int32), np.zeros(12), 0,
                                 i, d) == 1
    assert b"bad arguments" in lib.sf_last_error()
    # total angular momentum beyond 5 Rys nodes
    shells = np.array([3, 0, 0] * 4, dtype=np.int32)
    q = np.arange(4, dtype=np.int32)
    rc = lib.sf_eri_components(4, np.ones(4, dtype=np.int32), np.arange(4, dtype=np.int32),
                               np.ones(4), np.ones(4), np.ones(4), shells, np.zeros(12), 1, q,
                               np.zeros(1))
    assert rc == 5
    # f shells are outside the J/K engine
    h = ctypes.c_void_p()
    rc = lib.sf_basis_create(1, np.array([3], dtype=np.int32), np.array([1], dtype=np.int32),
                             np.zeros(3), np.ones(1), np.ones(1), 1, np.zeros(1, dtype=np.int32),
                             ctypes.byref(h))
    assert rc == 5 and not h.value
    with pytest.raises(engine.EngineError):
        engine.check(rc)
