"""CPU checks of libtpl_a2c.so: it loads, exports every function
include/tpl_a2c.h declares, counts atoms as the oracle does, and validates
arguments on the host."""
import ctypes

from oracle import a2c_oracle as O
from tests import abi_a2c


def test_exports_counts_and_argument_errors():
    lib = abi_a2c.load()
    names = abi_a2c.declared_functions()
    assert names == ["tpl_a2c_backward", "tpl_a2c_backward_form", "tpl_a2c_forward", "tpl_a2c_forward_ordered",
                     "tpl_a2c_num_atoms", "tpl_a2c_strerror"], names
    for name in names:
        assert hasattr(lib, name)
    assert lib.tpl_a2c_strerror(0) == b"ok"
    for seq in ("", "G", O.AA20, "WWPPGG", O.AA20 * 7):
        assert lib.tpl_a2c_num_atoms(seq.encode()) == O.num_atoms(seq)
    assert lib.tpl_a2c_num_atoms(b"AXG") == -1
    assert lib.tpl_a2c_num_atoms(None) == -1
    p = ctypes.c_void_p(0x1000)  # never dereferenced: every call below fails validation
    assert lib.tpl_a2c_forward(7, p, 4, p, 4, p, None, None, 10, None, 1, None) == 1
    assert lib.tpl_a2c_forward(0, None, 4, p, 4, p, None, None, 10, None, 1, None) == 1
    assert lib.tpl_a2c_forward(0, p, 4, p, 4, p, p, None, 10, None, 1, None) == 1  # one names pointer
    assert lib.tpl_a2c_forward(0, p, 4, p, 2 ** 29, p, None, None, 10, None, 1, None) == 5
    assert lib.tpl_a2c_backward(1, p, 4, None, p, 10, p, 4, None, 1, None) == 1  # no coordinates
    assert lib.tpl_a2c_backward(0, p, -1, p, p, 10, p, 4, None, 1, None) == 1
    # the kernel forms: unknown form, tiles beyond 2048 residues
    assert lib.tpl_a2c_backward_form(3, None, 0, p, 4, p, p, 10, p, 4, None, 1, None) == 1
    assert lib.tpl_a2c_backward_form(2, p, 0, p, 4, p, p, 10, p, 2049, None, 1, None) == 5
    assert lib.tpl_a2c_backward_form(2, None, 1, p, 4, None, p, 10, p, 4, None, 1, None) == 1
    assert lib.tpl_a2c_forward_ordered(p, 7, p, 4, p, 4, p, None, None, 10, None, 1, None) == 1
    # empty problems are no-ops
    assert lib.tpl_a2c_forward(0, None, 0, None, 0, None, None, None, 0, None, 0, None) == 0
    assert lib.tpl_a2c_backward(1, p, 4, p, p, 10, p, 0, None, 3, None) == 0
    for form in (0, 1, 2):
        assert lib.tpl_a2c_backward_form(form, p, 1, p, 4, p, p, 10, p, 0, None, 3, None) == 0
    assert lib.tpl_a2c_forward_ordered(p, 0, None, 0, None, 0, None, None, None, 0, None, 0, None) == 0
