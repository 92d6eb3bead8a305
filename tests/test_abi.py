"""The C-ABI library loads without a GPU and exports every function that
include/cress_hip.h declares; calls that need a device fail cleanly."""
import ctypes as C
import os
import re

import pytest

from cress_amd import host

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "cress_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(cress_[a-z0-9_]+)\s*\(", txt)))


def test_exports_every_declared_symbol():
    L = host.lib()
    names = declared()
    assert len(names) >= 15
    for n in names:
        assert hasattr(L, n), n


def test_error_strings():
    L = host.lib()
    for code in (0, -1, -2, -3, -4, -5, -6):
        assert L.cress_error_string(code)


def test_invalid_config_rejected_before_device_use():
    c = host.make_config(5, 5, 5, 1000.0, 1000.0, 500.0, 5.0, 2.0)   # too small
    ctx = C.c_void_p()
    assert host.lib().cress_ctx_create(C.byref(c), C.byref(ctx)) == -1
    c = host.make_config(20, 20, 20, 1000.0, 1000.0, 500.0, 5.0, 2.0, cphopt=3)
    assert host.lib().cress_ctx_create(C.byref(c), C.byref(ctx)) == -6


def test_nsstp_formula():
    # slvstep.f90:176
    assert host.make_config(20, 20, 20, 1e3, 1e3, 500., 4.0, 1.0).nsstp == 8
    assert host.make_config(20, 20, 20, 1e3, 1e3, 500., 3.0, 1.0).nsstp == 6
    assert host.make_config(20, 20, 20, 1e3, 1e3, 500., 5.0, 2.0).nsstp == 5
