"""CPU checks of the C-ABI library: it loads, exports every function
include/tpl_a2b.h declares, and validates arguments on the host (no kernel is
launched in these tests)."""
import ctypes

import pytest

from tests import abi


@pytest.fixture(scope="module")
def lib():
    return abi.load()


def test_exports_every_declared_symbol(lib):
    names = abi.declared_functions()
    assert len(names) == 11, names
    for name in names:
        assert hasattr(lib, name), f"{name} not exported by libtpl_a2b.so"


def test_strerror(lib):
    assert lib.tpl_a2b_strerror(0) == b"ok"
    for code in range(1, 6):
        assert lib.tpl_a2b_strerror(code) != b"unknown error code"
    assert lib.tpl_a2b_strerror(99) == b"unknown error code"


def test_config_validation_and_sizes(lib):
    c = abi.Config(4, 4, 0, 0)
    assert lib.tpl_a2b_tile_residues(ctypes.byref(c)) == 1024
    for bad in [abi.Config(3, 1, 0, 0), abi.Config(8, 1, 0, 0), abi.Config(1, 3, 0, 0),
                abi.Config(1, 1, -1, 0), abi.Config(1, 1, 0, 1)]:
        assert lib.tpl_a2b_tile_residues(ctypes.byref(bad)) == -1
    def align(x):
        return (x + 255) // 256 * 256

    def state(f32, B, L, ntb):
        # FP32 residual halves [B][9][L rounded up to 8], then (ntb > 1) the
        # backward payload (12 doubles per tile) + flags + tickets + done
        res = align(B * 9 * ((L + 7) // 8 * 8) * 2) if f32 else 0
        return res if ntb <= 1 else align(res + B * ntb * (96 + 4) + 8 * B)

    def sync(B, ntf):  # flags + tickets (zeroed), then 24-double payload per tile
        return 0 if ntf <= 1 else align(4 * (B * ntf + B)) + B * ntf * 24 * 8

    # (4,4): forward tiles of 1024; the backward runs 1 residue per lane, so
    # 256-residue tiles
    assert lib.tpl_a2b_state_bytes(0, 1024, 1000, ctypes.byref(c)) == state(True, 1024, 1000, 4)
    assert lib.tpl_a2b_sync_bytes(0, 1024, 1000, ctypes.byref(c)) == 0
    assert lib.tpl_a2b_state_bytes(0, 64, 16384, ctypes.byref(c)) == state(True, 64, 16384, 64)
    assert lib.tpl_a2b_sync_bytes(0, 64, 16384, ctypes.byref(c)) == sync(64, 16)
    c2 = abi.Config(2, 4, 0, 0)  # forward 512, backward 128 residues
    assert lib.tpl_a2b_state_bytes(0, 512, 500, ctypes.byref(c2)) == state(True, 512, 500, 4)
    assert lib.tpl_a2b_sync_bytes(0, 512, 500, ctypes.byref(c2)) == 0
    assert lib.tpl_a2b_state_bytes(1, 512, 500, ctypes.byref(c2)) == state(False, 512, 500, 4)
    c1 = abi.Config(4, 1, 0, 0)  # 256 both ways: f64 needs no state
    assert lib.tpl_a2b_state_bytes(1, 8, 256, ctypes.byref(c1)) == 0
    assert lib.tpl_a2b_state_bytes(0, 8, 256, ctypes.byref(c1)) == state(True, 8, 256, 1)
    assert lib.tpl_a2b_state_bytes(7, 512, 500, ctypes.byref(c2)) == 0  # unknown dtype
    assert lib.tpl_a2b_plain_frames_bytes(0, 2, 10) == 2 * 30 * 12 * 4
    assert lib.tpl_a2b_plain_frames_bytes(1, 2, 10) == 2 * 30 * 12 * 8


def test_default_config(lib):
    c = abi.Config()
    for L in (1, 200, 500, 1000, 16384):
        assert lib.tpl_a2b_default_config(8, L, ctypes.byref(c)) == 0
        assert lib.tpl_a2b_tile_residues(ctypes.byref(c)) > 0
    assert lib.tpl_a2b_default_config(-1, 10, ctypes.byref(c)) == 1


def test_argument_errors_return_before_launch(lib):
    c = abi.Config(1, 1, 0, 0)
    p = ctypes.c_void_p(0x1000)  # never dereferenced: every call below fails validation
    assert lib.tpl_a2b_forward(7, p, p, p, None, None, None, 1, 1, ctypes.byref(c), None) == 1
    assert lib.tpl_a2b_forward(0, None, p, p, None, None, None, 1, 1, ctypes.byref(c), None) == 1
    assert lib.tpl_a2b_forward(0, p, p, p, None, None, None, -1, 1, ctypes.byref(c), None) == 1
    bad = abi.Config(3, 1, 0, 0)
    assert lib.tpl_a2b_forward(0, p, p, p, None, None, None, 1, 1, ctypes.byref(bad), None) == 2
    # tiles above the register-tile limit: 1024 residues (f32), 512 (f64)
    big = abi.Config(4, 8, 0, 0)
    assert lib.tpl_a2b_forward(0, p, p, p, p, p, None, 1, 1, ctypes.byref(big), None) == 2
    f64big = abi.Config(4, 4, 0, 0)
    assert lib.tpl_a2b_backward(1, p, p, p, p, p, None, 1, 1, ctypes.byref(f64big), None) == 2
    # multi-tile forward without its scratch; backward without its state
    assert lib.tpl_a2b_forward(0, p, p, p, None, None, None, 2, 1000, ctypes.byref(c), None) == 1
    assert lib.tpl_a2b_backward(0, p, p, p, p, None, None, 2, 1000, ctypes.byref(c), None) == 1
    assert lib.tpl_a2b_backward(0, None, p, p, p, None, None, 1, 10, ctypes.byref(c), None) == 1
    assert lib.tpl_a2b_forward_plain(0, p, p, p, None, None, 1, 1, None) == 1
    assert lib.tpl_a2b_backward_plain(5, p, p, p, p, p, None, 1, 1, None) == 1
    # empty problems are no-ops
    assert lib.tpl_a2b_forward(0, None, None, None, None, None, None, 0, 5, ctypes.byref(c), None) == 0
    assert lib.tpl_a2b_backward(0, None, None, None, None, None, None, 3, 0, ctypes.byref(c), None) == 0
    assert lib.tpl_a2b_tune(0, p, p, p, p, p, None, None, 1, 1, ctypes.byref(c), 0, 3,
                            ctypes.byref(c), None, None) == 1
