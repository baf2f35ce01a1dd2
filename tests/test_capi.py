"""The C-ABI library: it loads, exports every entry point include/ubt.h
declares, and reports errors through status codes and ubt_last_error."""
import ctypes as C
import os
import re

import pytest

from ubolt_amd import capi

HERE = os.path.dirname(os.path.abspath(__file__))
HEADER = os.path.join(HERE, "..", "include", "ubt.h")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^int\s+(ubt_\w+)\s*\(", text, re.M)))


def test_every_declared_symbol_is_exported():
    names = declared()
    assert len(names) >= 30
    L = C.CDLL(capi.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert set(declared()) <= set(capi.SIGNATURES)


def test_status_and_message():
    q = C.c_void_p()
    rc = capi.lib().ubt_quad_create(2, 7, C.byref(q))
    assert rc == 56  # UBT_ERR_SUP
    code, msg = capi.last_error()
    assert code == 56 and "S7" in msg
    rc = capi.lib().ubt_quad_create(2, 4, None)
    assert rc == 85  # UBT_ERR_ARG_NULL
    rc = capi.lib().ubt_quad_create(2, 4, C.byref(q))
    assert rc == 0 and capi.last_error() == (0, "")
    capi.lib().ubt_quad_destroy(q)


def test_unknown_table_name():
    import ubolt_amd as u
    d = u.Disc.structured(u.Quadrature(2, 2), [2, 2], [1, 1])
    with pytest.raises(u.UbtError) as e:
        d.get("no_such_table")
    assert e.value.status == "ERR_ARG_WRONG"


def test_op_create_needs_a_device():
    import torch
    import ubolt_amd as u
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    d = u.Disc.structured(u.Quadrature(2, 2), [2, 2], [1, 1])
    with pytest.raises(u.UbtError) as e:
        u.Operator(d, 1)
    assert e.value.status == "ERR_GPU"
