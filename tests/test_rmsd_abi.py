"""CPU checks of libtpl_rmsd.so: it loads, exports every function
include/tpl_rmsd.h declares, and validates arguments on the host (no kernel is
launched in these tests)."""
import ctypes

import pytest

from tests import abi_rmsd


@pytest.fixture(scope="module")
def lib():
    return abi_rmsd.load()


def test_exports_every_declared_symbol(lib):
    names = abi_rmsd.declared_functions()
    assert names == ["tpl_rmsd_backward", "tpl_rmsd_forward", "tpl_rmsd_state_bytes",
                     "tpl_rmsd_strerror"], names
    for name in names:
        assert hasattr(lib, name), f"{name} not exported by libtpl_rmsd.so"


def test_strerror_and_state_bytes(lib):
    assert lib.tpl_rmsd_strerror(0) == b"ok"
    for code in (1, 3, 5):
        assert lib.tpl_rmsd_strerror(code) != b"unknown error code"
    assert lib.tpl_rmsd_strerror(42) == b"unknown error code"
    assert lib.tpl_rmsd_state_bytes(0) == 0
    assert lib.tpl_rmsd_state_bytes(1) == 256  # 16 doubles, 256-byte aligned
    assert lib.tpl_rmsd_state_bytes(1000) == (1000 * 128 + 255) // 256 * 256


def test_argument_errors_return_before_launch(lib):
    p = ctypes.c_void_p(0x1000)  # never dereferenced: every call below fails validation
    odd = ctypes.c_void_p(0x1008)  # state not on a 256-byte boundary
    assert lib.tpl_rmsd_forward(7, p, p, p, p, p, None, None, 1, 1, None) == 1
    assert lib.tpl_rmsd_forward(0, p, p, p, p, p, None, None, -1, 1, None) == 1
    assert lib.tpl_rmsd_forward(0, None, p, p, p, p, None, None, 1, 1, None) == 1
    assert lib.tpl_rmsd_forward(0, p, p, None, p, p, None, None, 1, 1, None) == 1
    assert lib.tpl_rmsd_forward(0, p, p, p, p, p, odd, None, 1, 1, None) == 1
    assert lib.tpl_rmsd_forward(0, p, p, p, p, p, None, None, 1, 2 ** 30, None) == 5
    assert lib.tpl_rmsd_backward(0, p, p, p, p, p, None, None, 1, 1, None) == 1
    assert lib.tpl_rmsd_backward(1, p, p, p, p, p, odd, None, 1, 1, None) == 1
    assert lib.tpl_rmsd_backward(3, p, p, p, p, p, p, None, 1, 1, None) == 1
    # empty problems are no-ops
    assert lib.tpl_rmsd_forward(0, None, None, None, None, None, None, None, 0, 5, None) == 0
    assert lib.tpl_rmsd_backward(0, None, None, None, None, None, None, None, 4, 0, None) == 0
