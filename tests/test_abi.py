"""The C-ABI library loads and exports every symbol include/csmc.h declares;
the Python binding covers the same set.  No compute calls (no GPU)."""
import ctypes
import os
import re

from conftest import ROOT

import csmc
from csmc import _lib


def _declared():
    src = open(os.path.join(ROOT, "include", "csmc.h")).read()
    return sorted(set(re.findall(r"\b(csmc_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(_lib.LIB_PATH)
    names = _declared()
    assert len(names) > 25
    for n in names:
        assert hasattr(L, n), n


def test_binding_covers_header():
    names = set(_declared()) - {"csmc_last_error"}
    assert names == set(_lib.SIGNATURES), names ^ set(_lib.SIGNATURES)


def test_version_and_error_string():
    assert _lib.lib().csmc_version() == 1
    # a host-side failure sets the thread-local error string
    rc = _lib.lib().csmc_batch_destroy(None, 1)
    assert rc == _lib.EINVAL
    assert b"null context" in _lib.lib().csmc_last_error()


def test_no_device_fails_loudly_or_creates_context():
    # without a GPU the context cannot be created and the library says so;
    # with a GPU it succeeds -- there is no CPU fallback either way
    n = csmc.device_count()
    if n == 0:
        try:
            csmc.Context(0)
        except csmc.CsmcError as e:
            assert "no HIP device" in str(e)
        else:
            raise AssertionError("context created without a device")
