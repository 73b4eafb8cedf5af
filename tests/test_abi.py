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
    assert lib.bdft_ctx_set_option(None, 1, 0) == 1
    v = ctypes.c_int()
    assert lib.bdft_ctx_get_option(None, 1, ctypes.byref(v)) == 1


def test_option_codes_match_the_header():
    text = open(os.path.join(ROOT, "include", "bigdft_hip.h")).read()
    codes = {m.group(1).lower(): int(m.group(2))
             for m in re.finditer(r"#define BDFT_OPT_(\w+) (\d+)", text)}
    assert codes == bigdft_hip.OPTIONS


@pytest.mark.parametrize("n,nfl,nfu", [((14, 17, 12), (3, 5, 2), (10, 12, 9)),
                                       ((7, 0, 5), (0, 0, 0), (7, 0, 5)),
                                       ((40, 33, 21), (17, 16, 9), (18, 16, 11))])
def test_bounds_layout_matches_oracle(n, nfl, nfu):
    """the library's bounds layout (a host function) against the oracle's"""
    import numpy as np
    import oracle
    off, dims = bigdft_hip.bounds_layout(n, nfl, nfu)
    o_off, o_dims = oracle.bounds_layout(n, nfl, nfu)
    assert np.array_equal(off, o_off) and np.array_equal(dims, o_dims)
    with pytest.raises(bigdft_hip.BigDFTError):
        bigdft_hip.bounds_layout(n, (5, 0, 0), (4, 0, 0))
