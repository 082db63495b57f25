"""The C ABI library loads, exports every entry point include/qfe.h declares,
and reports errors through status codes (no compute calls are made here)."""

import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "qfe.h")
LIB = os.path.join(ROOT, "qforte_amd", "libqfe.so")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(qfe_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "qfe_state_create" in names and "qfe_apply_ops" in names and "qfe_program_gradient" in names
    assert len(names) >= 40


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_error_reporting_without_device_work():
    lib = ctypes.CDLL(LIB)
    lib.qfe_last_error.restype = ctypes.c_char_p
    # invalid arguments are rejected before any device call
    assert lib.qfe_state_create(0, ctypes.byref(ctypes.c_void_p())) == 1
    assert b"nqubit" in lib.qfe_last_error()
    assert lib.qfe_set_tuning(20, 0, 0) == 1
    assert b"tile_qubits" in lib.qfe_last_error()
    n = ctypes.c_int(-1)
    assert lib.qfe_device_count(ctypes.byref(n)) == 0
    if n.value == 0:
        # no GPU here: creating a state must fail loudly, never fall back
        h = ctypes.c_void_p()
        rc = lib.qfe_state_create(4, ctypes.byref(h))
        assert rc == 2
        assert b"no GPU" in lib.qfe_last_error()


def test_tile_tuning_default_follows_the_state_size():
    lib = ctypes.CDLL(LIB)
    get = lambda: [v.value for v in vals] if lib.qfe_get_tuning(*[ctypes.byref(v) for v in vals]) == 0 else None  # noqa: E731
    vals = [ctypes.c_int(0) for _ in range(3)]
    saved = get()
    try:
        assert lib.qfe_set_tuning(-1, -1, -1) == 0
        assert get() == [-1, -1, -1]  # the default: shape chosen by state size
        assert lib.qfe_set_tuning(0, 0, 0) == 0 and get() == [-1, -1, -1]
        assert lib.qfe_set_tuning(10, 0, 0) == 0
        assert get() == [10, 4, 24]  # explicit; zeros keep the stored values
        assert lib.qfe_set_tuning(0, 0, 32) == 0 and get() == [10, 4, 32]
        assert lib.qfe_set_tuning(20, 0, 0) == 1  # rejected, state unchanged
        assert get() == [10, 4, 32]
        assert lib.qfe_set_tuning(-1, 0, 0) == 0 and get() == [-1, -1, -1]
    finally:
        lib.qfe_set_tuning(*saved)


def test_python_computer_fails_loudly_without_gpu(qf):
    if qf.device_count() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(RuntimeError, match="no GPU"):
        qf.Computer(3)
