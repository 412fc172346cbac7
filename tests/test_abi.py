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


@pytest.mark.parametrize("shape,pools,prec,msg", [
    ((0, 0, 1), None, 0, "n must be"),
    ((0, 513, 1), None, 0, "n must be"),
    ((0, 8, 0), None, 0, "nrhs must be"),
    ((0, 8, 9), None, 0, "nrhs must be"),
    ((2, 8, 1), None, 0, "kind"),
    ((0, 8, 1), (63, 8, 8), 0, "A lies outside"),
    ((1, 8, 2), (64, 15, 0), 0, "B lies outside"),
    ((0, 8, 2), (64, 16, 7), 0, "ipiv lies outside"),
    ((0, 8, 1), None, 2, "LMX_PREC_F64 and LMX_PREC_F32"),
])
def test_batch_plan_validation(shape, pools, prec, msg):
    # every check runs before any device call
    import lmx
    probs, na, nb, npv = lmx.batch_layout([shape])
    if pools is not None:
        na, nb, npv = pools
    with pytest.raises(lmx.LmxError) as e:
        lmx.BatchPlan(probs, na, nb, npv, precision=prec)
    assert e.value.code == lmx.LMX_ERR_ARG
    assert msg in str(e.value)


def test_batch_layout_packs_back_to_back():
    import lmx
    probs, na, nb, npv = lmx.batch_layout([(0, 3, 2), (1, 5, 1), (0, 2, 8)])
    assert [p.a_offset for p in probs] == [0, 9, 34]
    assert [p.b_offset for p in probs] == [0, 6, 11]
    assert [p.ipiv_offset for p in probs] == [0, 3, 8]
    assert (na, nb, npv) == (38, 27, 10)
    assert ctypes.sizeof(lmx.BatchProblem) == 40


@pytest.mark.parametrize("kw,msg", [({"slots": 17}, "slots"), ({"small_n": 600}, "small_n"),
                                    ({"budget_bytes": -1}, "budget")])
def test_server_config_validation(kw, msg):
    # checked before any device call
    import lmx
    with pytest.raises(lmx.LmxError) as e:
        lmx.Server(**kw)
    assert e.value.code == lmx.LMX_ERR_ARG
    assert msg in str(e.value)


def test_server_structs_match_the_header():
    import lmx
    assert ctypes.sizeof(lmx.ServerConfig) == 32
    assert ctypes.sizeof(lmx.RequestResult) == 56
    assert ctypes.sizeof(lmx.ServerCounters) == 80
