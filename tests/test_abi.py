"""The C-ABI library loads and exports every entry point that
include/bigdft_hip.h declares (no compute call: runs without a GPU)."""
import ctypes
import os
import re

import pytest

import bigdft_hip

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = []
    inc = os.path.join(ROOT, "include")
    for f in os.listdir(inc):
        if f.endswith(".h"):
            src = open(os.path.join(inc, f)).read()
            names += re.findall(r"^\s*(?:int|const char \*)\s*(bdft_\w+)\(", src,
                                re.M)
    return sorted(set(names))


def test_header_matches_binding_list():
    assert declared() == sorted(bigdft_hip.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(bigdft_hip.LIB_PATH)
    for name in declared():
        assert hasattr(lib, name), name


def test_binding_loads():
    lib = bigdft_hip.lib()
    assert lib.bdft_last_error(None) == b"null context"


def test_null_arguments_are_errors():
    lib = bigdft_hip.lib()
    assert lib.bdft_ctx_destroy(None) == 1
    assert lib.bdft_sync(None) == 1
    assert lib.bdft_plan_execute(None) == 1
    assert lib.bdft_lr_destroy(None) == 1
    assert lib.bdft_ctx_create(0, None) == 1
