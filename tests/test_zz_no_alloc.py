"""qfe_expval_noalloc (the expectation value that allocates nothing: its tables,
partial sums and results live in a fixed device array) against qfe_expval,
through the C ABI, on two calls in a row (the second smaller), on both the
small-register path and the windowed orbit path, plus the size limit."""

import ctypes
import os

import numpy as np
import pytest

from qforte_amd.engine import workloads as W

pytestmark = pytest.mark.gpu
LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "qforte_amd", "libqfe.so")


class Pauli(ctypes.Structure):
    _fields_ = [("x", ctypes.c_uint64), ("z", ctypes.c_uint64), ("phase", ctypes.c_double * 2),
                ("coeff", ctypes.c_double * 2)]


def mask_terms(terms):
    out = (Pauli * len(terms))()
    for k, (c, word) in enumerate(terms):
        x = z = 0
        ph = 1 + 0j
        for letter, q in word:
            b = 1 << q
            if letter in ("X", "Y"):
                x |= b
            if letter in ("Y", "Z"):
                z |= b
            if letter == "Y":
                ph *= 1j
        out[k].x, out[k].z = x, z
        out[k].phase[0], out[k].phase[1] = ph.real, ph.imag
        out[k].coeff[0], out[k].coeff[1] = complex(c).real, complex(c).imag
    return out


@pytest.fixture(scope="module")
def lib(qf):
    lib = ctypes.CDLL(LIB)
    lib.qfe_last_error.restype = ctypes.c_char_p
    assert lib.qfe_init(0) == 0, lib.qfe_last_error()
    return lib


def make_state(lib, n, seed):
    rng = np.random.default_rng(seed)
    psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    psi /= np.linalg.norm(psi)
    h = ctypes.c_void_p()
    assert lib.qfe_state_create(n, ctypes.byref(h)) == 0, lib.qfe_last_error()
    a = np.ascontiguousarray(psi, dtype=np.complex128)
    assert lib.qfe_state_upload(h, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_uint64(a.size)) == 0
    return h


def both(lib, h, terms):
    arr = mask_terms(terms)
    a = (ctypes.c_double * 2)()
    b = (ctypes.c_double * 2)()
    assert lib.qfe_expval(h, arr, ctypes.c_uint64(len(terms)), a) == 0, lib.qfe_last_error()
    assert lib.qfe_expval_noalloc(h, arr, ctypes.c_uint64(len(terms)), b) == 0, lib.qfe_last_error()
    return complex(a[0], a[1]), complex(b[0], b[1])


@pytest.mark.parametrize("first,second", [((22, 1500), (16, 300)), ((14, 700), (9, 120))])
def test_noalloc_equals_expval_two_calls(lib, first, second):
    # 22 qubits, 1500 terms: 1024 workgroups per pass, so the partial sums of
    # the ~1500 groups take several windows of the fixed array
    for i, (n, nterms) in enumerate((first, second)):
        h = make_state(lib, n, seed=100 + n)
        try:
            terms = W.random_pauli_sum(n, nterms, seed=200 + n, wmax=min(6, n))
            ref, got = both(lib, h, terms)
            print(f"call {i}: n={n} terms={nterms} expval={ref} noalloc={got}", flush=True)
            # the same groups, the same per-group reduction order and the same
            # host sum: bit-identical
            assert got == ref
            assert abs(got - ref) <= 1e-12 * max(1.0, abs(ref))
        finally:
            assert lib.qfe_state_destroy(h) == 0


def test_noalloc_rejects_oversized_operator(lib):
    h = make_state(lib, 4, seed=7)
    try:
        arr = (Pauli * 65537)()
        out = (ctypes.c_double * 2)()
        assert lib.qfe_expval_noalloc(h, arr, ctypes.c_uint64(65537), out) == 1
        assert b"65536" in lib.qfe_last_error()
    finally:
        assert lib.qfe_state_destroy(h) == 0
