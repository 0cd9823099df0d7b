"""tpl_a2b_forward_noalloc (the forward that allocates nothing: its look-back
words live in the library's own device array) against tpl_a2b_forward on the
same inputs, through the C ABI: two calls in a row, the second with a smaller
input, both equal to the workspace entry point's output; a problem too large
for the built-in array is rejected."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests import abi
from torchproteinlibrary_amd import synthetic

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _reference_forward(lib, A, N, cfg, stream):
    B, _, L = A.shape
    c = abi.Config(*cfg, 0, 0)
    sy = torch.zeros(lib.tpl_a2b_sync_bytes(0, B, L, ctypes.byref(c)), dtype=torch.uint8, device=DEV)
    X = torch.full((B, 9 * L), float("nan"), device=DEV)
    assert lib.tpl_a2b_forward(0, _p(A), _p(X), _p(N), None, _p(sy), None, B, L, ctypes.byref(c), stream) == 0
    return X


@pytest.mark.parametrize("cfg", [(4, 2), (1, 1)])
def test_noalloc_forward_matches_the_workspace_forward(cfg):
    lib = abi.load()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    c = abi.Config(*cfg, 0, 0)
    cases = [synthetic.workload("mixed1000", batch=256), synthetic.custom([300, 0, 1, 129, 64, 65], 300, seed=2)]
    for k, (a, n, _) in enumerate(cases):  # the second call is smaller
        B, _, L = a.shape
        print(f"noalloc forward: call {k}, B={B} L={L} cfg={cfg}", flush=True)
        A, N = torch.tensor(a, device=DEV), torch.tensor(n, device=DEV)
        X = torch.full((B, 9 * L), float("nan"), device=DEV)
        status = torch.zeros(1, dtype=torch.int32, device=DEV)
        rc = lib.tpl_a2b_forward_noalloc(0, _p(A), _p(X), _p(N), _p(status), B, L, ctypes.byref(c), stream)
        assert rc == 0, lib.tpl_a2b_strerror(rc)
        want = _reference_forward(lib, A, N, cfg, stream)
        torch.cuda.synchronize()
        assert int(status.item()) == 0
        assert torch.equal(X, want), f"call {k}: differs from tpl_a2b_forward"
        x64 = O.forward(a.astype(np.float64), n)
        assert np.abs(X.cpu().numpy() - x64).max() <= 1.2e-7 * max(1.0, np.abs(x64).max()) + 1e-6


def test_noalloc_rejects_problems_beyond_its_scratch():
    lib = abi.load()
    limit = lib.tpl_a2b_noalloc_max_bytes()
    assert limit == 4 << 20
    c = abi.Config(1, 1, 0, 0)
    # 64-residue chunks: 96 bytes of look-back record per chunk slot, 25 MiB here
    B, L = 4096, 64 * 64
    p = ctypes.c_void_p(0x1000)  # never dereferenced: the call fails before any launch
    assert lib.tpl_a2b_forward_noalloc(0, p, p, p, None, B, L, ctypes.byref(c), None) == 6
    assert b"scratch" in lib.tpl_a2b_strerror(6)
