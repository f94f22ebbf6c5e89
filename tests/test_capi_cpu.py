"""The C-ABI library loads, exports every entry point include/cukd_amd.h
declares, and rejects bad arguments before touching the GPU."""
import ctypes
import os
import re
import subprocess

import pytest

import cudakdtree_amd as ck

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "cukd_amd.h")).read()
    return sorted(set(re.findall(r"\b(cukd_[a-z0-9_]+)\s*\(", txt)))


def test_header_and_binding_agree():
    assert _declared() == sorted(ck.EXPORTED_SYMBOLS)


def test_python_constants_match_header_defines():
    txt = open(os.path.join(ROOT, "include", "cukd_amd.h")).read()
    defs = dict((m.group(1), int(m.group(2)))
                for m in re.finditer(r"#define CUKD_([A-Z0-9_]+) \(?(-?\d+)\)?", txt))
    checked = 0
    for name, val in defs.items():
        if hasattr(ck, name):
            assert getattr(ck, name) == val, name
            checked += 1
    for name in ("QUERY_SORT", "QUERY_AUTOTUNE", "QUERY_ONE_PER_LANE", "STACK_BASED", "CCT",
                 "LIST_HEAP", "BUILDER_ID_INPLACE", "ERR_UNSUPPORTED"):
        assert name in defs and hasattr(ck, name), name
    assert checked >= 20


def test_library_exports_every_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", ck._LIB_PATH],
                         capture_output=True, text=True, check=True).stdout
    names = set(l.split()[-1] for l in out.splitlines() if l.strip())
    for sym in _declared():
        assert sym in names, sym
    lib = ck.lib()
    for sym in _declared():
        assert hasattr(lib, sym)
    assert lib.cukd_version() == 100


def _bad(layout):
    return ck.lib().cukd_build_tree(ctypes.c_void_p(16), 10, ctypes.byref(layout), 0,
                                    None, None, None)


@pytest.mark.parametrize("layout", [
    ck.Layout(0), ck.Layout(9), ck.Layout(3, 2), ck.Layout(3, stride=10),
    ck.Layout(3, stride=8), ck.Layout(3, point_offset=2, stride=16),
    ck.Layout(3, stride=16, dim_offset=12, dim_bytes=3),
    ck.Layout(3, stride=16, dim_offset=8, dim_bytes=4),
])
def test_invalid_layouts_rejected(layout):
    assert _bad(layout) == ck.ERR_INVALID
    assert ck.lib().cukd_last_error()


def test_explicit_dims_need_bounds():
    lay = ck.point_plus_payload(3, explicit_dim=True)
    rc = ck.lib().cukd_build_tree(ctypes.c_void_p(16), 10, ctypes.byref(lay), 0,
                                  None, None, None)
    assert rc == ck.ERR_NO_BOUNDS
    assert b"world bounds" in ck.lib().cukd_last_error()


def test_empty_build_is_noop():
    lay = ck.float_points(3)
    assert ck.lib().cukd_build_tree(None, 0, ctypes.byref(lay), 0, None, None, None) == 0


def test_bad_query_arguments():
    lay = ck.float_points(3)
    L = ck.lib()
    P = ctypes.c_void_p(16)
    # cct without world bounds
    assert L.cukd_fcp(ck.CCT, P, 10, ctypes.byref(lay), None, P, 5, 1.0, 0.0, P, P, 0,
                      None, None) == ck.ERR_INVALID
    # k out of range
    assert L.cukd_knn(0, 0, 65, P, 10, ctypes.byref(lay), None, P, 5, 1.0, P, P, None, 0,
                      None, None) == ck.ERR_INVALID
    # int points have no distance math in the reference
    assert L.cukd_fcp(0, P, 10, ctypes.byref(ck.int_points(3)), None, P, 5, 1.0, 0.0, P, P,
                      0, None, None) == ck.ERR_UNSUPPORTED
    # unknown builder
    assert L.cukd_build_tree(P, 10, ctypes.byref(lay), 7, None, None, None) == ck.ERR_INVALID


def test_temp_bytes():
    lay = ck.float_points(3)
    assert ck.build_temp_bytes(0, lay) == 0
    small = ck.build_temp_bytes(1000, lay)
    big = ck.build_temp_bytes(1_000_000, lay)
    assert 0 < small < big
    # two ping-pong copies of (point, index, key) per element dominate
    assert big >= 2 * 1_000_000 * (12 + 8)


def test_temp_bytes_per_builder():
    """bitonic needs n ints once the tree has levels above the LDS finish;
    in-place needs nothing (builder.h:34-45)."""
    for lay in (ck.float_points(3), ck.point_plus_payload(4, explicit_dim=True)):
        assert ck.build_temp_bytes(1_000_000, lay, ck.BUILDER_BITONIC) == 4_000_000
        assert ck.build_temp_bytes(1_000_000, lay, ck.BUILDER_INPLACE) == 0
        assert ck.build_temp_bytes(100, lay, ck.BUILDER_BITONIC) == 0
        assert ck.build_temp_bytes(1_000_000, lay, ck.BUILDER_THRUST) == \
            ck.build_temp_bytes(1_000_000, lay, ck.BUILDER_DEFAULT)
    with pytest.raises(RuntimeError):
        ck.build_temp_bytes(10, ck.float_points(3), 9)
