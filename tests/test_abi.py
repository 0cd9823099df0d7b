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
    assert len(names) == 14, names
    for name in names:
        assert hasattr(lib, name), f"{name} not exported by libtpl_a2b.so"


def test_strerror(lib):
    assert lib.tpl_a2b_strerror(0) == b"ok"
    for code in range(1, 7):
        assert lib.tpl_a2b_strerror(code) != b"unknown error code"
    assert lib.tpl_a2b_strerror(99) == b"unknown error code"


def _align(x):
    return (x + 255) // 256 * 256


def _state(B, L, rb):
    # one FP64 3x4 frame (96 bytes) per backward chunk of 64*rb residues
    return _align(B * -(-L // (64 * rb)) * 96)


def _sync(B, L, rf, rb):
    # per direction: tickets and done counters [8][sbx] (uint32; sbx =
    # ceil(B / 8) rounded up to 64), then each chain's records, nc chunks of
    # 12 64-bit words (forward frames) or 6 (backward sums) rounded up to 32
    # words
    sbx = -(-(-(-B // 8)) // 64) * 64

    def region(nc, pay):
        return _align(_align(2 * 8 * sbx * 4) + B * (-(-nc * pay // 32) * 32) * 8)
    return region(-(-L // (64 * rf)), 12) + region(-(-L // (64 * rb)), 6)


def test_config_validation_and_sizes(lib):
    c = abi.Config(4, 2, 0, 0)
    assert lib.tpl_a2b_chunk_residues(ctypes.byref(c), 0) == 256
    assert lib.tpl_a2b_chunk_residues(ctypes.byref(c), 1) == 128
    for bad in [abi.Config(3, 1, 0, 0), abi.Config(16, 1, 0, 0), abi.Config(1, 8, 0, 0),
                abi.Config(1, 3, 0, 0), abi.Config(1, 1, 1, 0), abi.Config(1, 1, 0, 1)]:
        assert lib.tpl_a2b_chunk_residues(ctypes.byref(bad), 0) == -1
    for dt in (0, 1):
        assert lib.tpl_a2b_state_bytes(dt, 1024, 1000, ctypes.byref(c)) == _state(1024, 1000, 2)
        assert lib.tpl_a2b_sync_bytes(dt, 1024, 1000, ctypes.byref(c)) == _sync(1024, 1000, 4, 2)
    c8 = abi.Config(8, 4, 0, 0)
    assert lib.tpl_a2b_state_bytes(0, 64, 16384, ctypes.byref(c8)) == _state(64, 16384, 4)
    assert lib.tpl_a2b_sync_bytes(0, 64, 16384, ctypes.byref(c8)) == _sync(64, 16384, 8, 4)
    c1 = abi.Config(1, 1, 0, 0)
    assert lib.tpl_a2b_state_bytes(0, 3, 37, ctypes.byref(c1)) == _state(3, 37, 1)
    assert lib.tpl_a2b_state_bytes(7, 512, 500, ctypes.byref(c1)) == 0  # unknown dtype
    assert lib.tpl_a2b_state_bytes(0, 0, 500, ctypes.byref(c1)) == 0
    assert lib.tpl_a2b_plain_frames_bytes(0, 2, 10) == 2 * 30 * 12 * 4
    assert lib.tpl_a2b_plain_frames_bytes(1, 2, 10) == 2 * 30 * 12 * 8


def test_default_config(lib):
    c = abi.Config()
    for L in (1, 64, 200, 500, 1000, 16384):
        assert lib.tpl_a2b_default_config(8, L, ctypes.byref(c)) == 0
        assert lib.tpl_a2b_chunk_residues(ctypes.byref(c), 0) > 0
        assert c.fwd_res_per_lane <= 4 and c.bwd_res_per_lane <= 2  # valid for both dtypes
    assert lib.tpl_a2b_default_config(-1, 10, ctypes.byref(c)) == 1


def test_argument_errors_return_before_launch(lib):
    c = abi.Config(1, 1, 0, 0)
    p = ctypes.c_void_p(0x1000)  # never dereferenced: every call below fails validation
    q = ctypes.c_void_p(0x1001)  # misaligned state / sync
    assert lib.tpl_a2b_forward(7, p, p, p, None, p, None, 1, 1, ctypes.byref(c), None) == 1
    assert lib.tpl_a2b_forward(0, None, p, p, None, p, None, 1, 1, ctypes.byref(c), None) == 1
    assert lib.tpl_a2b_forward(0, p, p, p, None, p, None, -1, 1, ctypes.byref(c), None) == 1
    assert lib.tpl_a2b_forward(0, p, p, p, None, None, None, 1, 1, ctypes.byref(c), None) == 1  # no sync
    assert lib.tpl_a2b_forward(0, p, p, p, q, p, None, 1, 1, ctypes.byref(c), None) == 1
    assert lib.tpl_a2b_forward(0, p, p, p, p, q, None, 1, 1, ctypes.byref(c), None) == 1
    bad = abi.Config(3, 1, 0, 0)
    assert lib.tpl_a2b_forward(0, p, p, p, None, p, None, 1, 1, ctypes.byref(bad), None) == 2
    # register limits of FP64: forward chunks up to 4 residues per lane, backward 2
    f64big = abi.Config(8, 1, 0, 0)
    assert lib.tpl_a2b_forward(1, p, p, p, p, p, None, 1, 1, ctypes.byref(f64big), None) == 2
    f64bwd = abi.Config(1, 4, 0, 0)
    assert lib.tpl_a2b_backward(1, p, p, p, p, p, p, None, 1, 1, ctypes.byref(f64bwd), None) == 2
    # backward without its state or its sync
    assert lib.tpl_a2b_backward(0, p, p, p, p, None, p, None, 2, 1000, ctypes.byref(c), None) == 1
    assert lib.tpl_a2b_backward(0, p, p, p, p, p, None, None, 2, 1000, ctypes.byref(c), None) == 1
    assert lib.tpl_a2b_backward(0, None, p, p, p, p, p, None, 1, 10, ctypes.byref(c), None) == 1
    assert lib.tpl_a2b_forward_plain(0, p, p, p, None, None, 1, 1, None) == 1
    assert lib.tpl_a2b_backward_plain(5, p, p, p, p, p, None, 1, 1, None) == 1
    assert lib.tpl_a2b_sync_clear(None, 256, None) == 1
    assert lib.tpl_a2b_sync_clear(p, 6, None) == 1
    assert lib.tpl_a2b_sync_clear(None, 0, None) == 0
    # too large: 9 * L beyond int32
    assert lib.tpl_a2b_forward(0, p, p, p, None, p, None, 1, 300_000_000, ctypes.byref(c), None) == 5
    # empty problems are no-ops
    assert lib.tpl_a2b_forward(0, None, None, None, None, None, None, 0, 5, ctypes.byref(c), None) == 0
    assert lib.tpl_a2b_backward(0, None, None, None, None, None, None, None, 3, 0, ctypes.byref(c), None) == 0
    assert lib.tpl_a2b_tune(0, p, p, p, p, p, None, None, 1, 1, ctypes.byref(c), 0, 3,
                            ctypes.byref(c), None, None) == 1
