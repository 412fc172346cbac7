"""The C-ABI library loads and exports every symbol include/lmx.h declares;
argument validation runs on the host (no GPU needed)."""
import ctypes
import os
import re

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "lmx.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(lmx_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    syms = declared_symbols()
    for s in ["lmx_dgemm", "lmx_dgetrf", "lmx_dgetrs", "lmx_dgesv", "lmx_dpotrf", "lmx_dpotrs",
              "lmx_dposv", "lmx_dtrsm", "lmx_dgesv_mixed", "lmx_dposv_mixed"]:
        assert s in syms


def test_library_exports_every_declared_symbol():
    import lmx
    lib = ctypes.CDLL(lmx.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert missing == []
    assert lmx.missing_symbols() == []


def test_binding_lists_every_declared_symbol():
    import lmx
    assert sorted(set(lmx.EXPORTED)) == declared_symbols()


def test_version():
    import lmx
    assert lmx.version() >= 100


@pytest.mark.parametrize("call", [
    lambda L: L.lmx_dgemm(b"X", b"N", 2, 2, 2, 1.0, None, 2, None, 2, 0.0, None, 2),
    lambda L: L.lmx_dgemm(b"N", b"N", -1, 2, 2, 1.0, None, 2, None, 2, 0.0, None, 2),
    lambda L: L.lmx_dgemm(b"N", b"N", 4, 4, 4, 1.0, None, 2, None, 4, 0.0, None, 4),
    lambda L: L.lmx_dgetrf(-3, None, 1, None),
    lambda L: L.lmx_dgetrf(4, None, 3, None),
    lambda L: L.lmx_dgesv(4, 2, None, 4, None, None, 1),
    lambda L: L.lmx_dpotrf(5, None, 4),
    lambda L: L.lmx_dtrsm(b"Q", b"N", b"N", 3, 1, None, 3, None, 1),
    lambda L: L.lmx_dtrsm(b"L", b"N", b"Z", 3, 1, None, 3, None, 1),
    lambda L: L.lmx_dgesv_mixed(1, 4, 1, None, 3, None, 1, None, 1, None),
])
def test_bad_arguments_are_rejected_on_the_host(call):
    import lmx
    rc = call(lmx.raw())
    assert rc == lmx.LMX_ERR_ARG
    assert lmx.last_error() != ""


def test_zero_sized_calls_are_no_ops():
    import lmx
    L = lmx.raw()
    assert L.lmx_dgemm(b"N", b"N", 0, 5, 3, 1.0, None, 3, None, 5, 0.0, None, 5) == 0
    assert L.lmx_dgetrf(0, None, 1, None) == 0
    assert L.lmx_dpotrf(0, None, 1) == 0
    assert L.lmx_dtrsm(b"L", b"N", b"N", 0, 3, None, 1, None, 3) == 0
