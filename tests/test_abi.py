"""C ABI of libqpx.so: the library loads and exports every function include/qpx.h declares, and
argument validation fails cleanly without touching the GPU."""

import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "qpx.h")


def declared_functions():
    text = open(HEADER).read()
    return re.findall(r"QPX_API\s+[\w\s\*]+?\b(qpx_\w+)\s*\(", text)


@pytest.fixture(scope="module")
def lib():
    from qpax_amd import _lib

    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libqpx.so not built")
    return _lib.load()


def test_header_declares_entry_points():
    names = declared_functions()
    for n in ("qpx_solve_qp", "qpx_solve_qp_relaxed", "qpx_diff_qp", "qpx_workspace_size", "qpx_last_error"):
        assert n in names


def test_exports_every_declared_symbol(lib):
    raw = ctypes.CDLL(os.path.join(ROOT, "qpax_amd", "libqpx.so"))
    for n in declared_functions():
        assert hasattr(raw, n), n


def test_binding_covers_header():
    from qpax_amd import _lib

    assert set(declared_functions()) == set(_lib.SIGNATURES)


def test_version_and_workspace(lib):
    assert lib.qpx_version() >= 100
    a = lib.qpx_workspace_size(1, 50, 10, 100)
    b = lib.qpx_workspace_size(4096, 50, 10, 100)
    assert a > 0 and b > a and (b - 4096) == 4096 * (a - 4096)
    assert lib.qpx_workspace_size(-1, 5, 0, 5) == 0


@pytest.mark.parametrize("args,code", [
    ((7, 1, 4, 5, 0, 5), 1),      # bad dtype
    ((0, 9, 4, 5, 0, 5), 1),      # bad mode
    ((0, 1, 4, 0, 0, 5), 1),      # n = 0
    ((0, 1, 4, 5, 0, 0), 1),      # no inequality rows: special-case entry points
])
def test_argument_errors(lib, args, code):
    st = lib.qpx_solve_qp(*args, *([None] * 6), 1e-3, *([None] * 6), None, 1 << 30, None)
    assert st == code
    assert lib.qpx_last_error().decode()


def test_workspace_too_small(lib):
    st = lib.qpx_solve_qp(0, 1, 4, 5, 0, 5, *([None] * 6), 1e-3, *([None] * 6), None, 16, None)
    assert st == 3
    assert b"workspace" in lib.qpx_last_error()
