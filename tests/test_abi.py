"""The C-ABI library loads without a GPU and exports every entry point include/*.h declares."""
import os
import re
import subprocess

from qcmaquis_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header="qcmaquis_amd.h"):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return set(re.findall(r"\b(mq_[a-z0-9_]+)\s*\(", txt))


def exported(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True).stdout
    return set(re.findall(r"\bT (mq_[a-z0-9_]+)", out))


def test_headers_are_known():
    assert sorted(f for f in os.listdir(os.path.join(ROOT, "include")) if f.endswith(".h")) == \
        ["qcmaquis_amd.h", "qcmaquis_amd_h5.h"]


def test_exports_every_declared_symbol():
    names = declared()
    assert len(names) >= 30
    assert names <= exported(abi.LIB_PATH), sorted(names - exported(abi.LIB_PATH))


def test_ctypes_signatures_cover_header():
    assert declared() == {n for n, _, _ in abi.SIGNATURES}


def test_result_file_library_exports_its_header():
    """libqcmaquis_amd_h5.so (host only, HDF5) exports include/qcmaquis_amd_h5.h"""
    import pytest

    from qcmaquis_amd import results
    if not os.path.exists(results._LIB):
        pytest.skip("libqcmaquis_amd_h5.so not built (no HDF5)")
    names = declared("qcmaquis_amd_h5.h")
    assert len(names) == 10 and names <= exported(results._LIB)
    L = results.lib()
    for n in names:
        assert getattr(L, n).argtypes is not None, n


def test_loads_and_reports_version():
    L = abi.lib()
    assert L.mq_version() >= 1
    assert L.mq_last_error(None) == b"null context"
