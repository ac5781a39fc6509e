"""The C-ABI library loads and exports every entry point the header declares
(no GPU calls), and host-side argument checks run without a device."""
import ctypes as C
import os
import re
import subprocess

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "fluxflow_hip.h")
LIB = os.path.join(ROOT, "tsl_array_n_amd", "lib", "libfluxhip.so")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b(flux_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.fail("libfluxhip.so is not built: run __graft_entry__.build()")
    return C.CDLL(LIB)


def test_header_declares_entry_points():
    names = declared()
    assert "flux_init" in names and "flux_pressure3_project" in names
    assert len(names) >= 30


def test_every_declared_symbol_is_exported(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (flux_\w+)", out))
    missing = [n for n in declared() if n not in exported]
    assert not missing, missing
    for n in declared():
        getattr(lib, n)


def test_calls_before_init_fail_cleanly(lib):
    lib.flux_last_error.restype = C.c_char_p
    h = C.c_void_p()
    shape = (C.c_int64 * 1)(4)
    rc = lib.flux_field_create(0, 1, shape, C.byref(h))
    assert rc == 3  # FLUX_ERR_STATE
    assert b"flux_init" in lib.flux_last_error()
