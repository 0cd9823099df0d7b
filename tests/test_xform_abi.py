"""CPU checks of libtpl_xform.so: it loads, exports every function
include/tpl_xform.h declares, and validates arguments on the host."""
import ctypes

from tests import abi_xform


def test_exports_and_argument_errors():
    lib = abi_xform.load()
    names = abi_xform.declared_functions()
    assert names == ["tpl_xform_center", "tpl_xform_center_backward", "tpl_xform_rotate",
                     "tpl_xform_strerror", "tpl_xform_translate", "tpl_xform_translate_backward"], names
    for name in names:
        assert hasattr(lib, name)
    assert lib.tpl_xform_strerror(0) == b"ok"
    p = ctypes.c_void_p(0x1000)  # never dereferenced: every call below fails validation
    assert lib.tpl_xform_translate(9, p, p, p, p, None, 1, 1, None) == 1
    assert lib.tpl_xform_translate(0, p, None, p, p, None, 1, 1, None) == 1
    assert lib.tpl_xform_rotate(0, p, p, None, 0, p, None, 1, 1, None) == 1
    assert lib.tpl_xform_center(0, None, p, p, None, 1, 1, None) == 1
    assert lib.tpl_xform_center_backward(0, p, p, p, None, -1, 1, None) == 1
    assert lib.tpl_xform_translate_backward(0, p, p, p, None, 1, 2 ** 30, None) == 5
    # empty problems are no-ops
    assert lib.tpl_xform_translate(0, None, None, None, None, None, 0, 7, None) == 0
    assert lib.tpl_xform_rotate(1, None, None, None, 1, None, None, 3, 0, None) == 0
