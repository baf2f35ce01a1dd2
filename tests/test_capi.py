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


def test_struct_layouts_match_header(tmp_path):
    """the ctypes mirrors of the option and result structs have the header's
    size and field offsets (a C program built with gcc reports them)"""
    import ctypes as C
    import subprocess

    from ubolt_amd import capi

    structs = {"ubt_bc": capi.Bc, "ubt_solver_options": capi.SolverOptions,
               "ubt_group_options": capi.GroupOptions, "ubt_group_result": capi.GroupResult}
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "ubt.h"', "int main(void) {"]
    for name, cls in structs.items():
        lines.append(f'  printf("{name} size %zu\\n", sizeof({name}));')
        for field, _ in cls._fields_:
            lines.append(f'  printf("{name} {field} %zu\\n", offsetof({name}, {field}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    subprocess.run(["gcc", "-std=c11", "-I", inc, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    want = {}
    for line in out:
        if line:
            name, key, val = line.split()
            want[(name, key)] = int(val)
    for name, cls in structs.items():
        assert C.sizeof(cls) == want[(name, "size")], name
        for field, _ in cls._fields_:
            assert getattr(cls, field).offset == want[(name, field)], (name, field)
