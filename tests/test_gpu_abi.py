"""GPU parity through the C ABI itself (ctypes -> libtpl_a2b.so), without the
torch extension: raw device pointers, torch's current stream."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests import abi
from torchproteinlibrary_amd import synthetic

pytestmark = pytest.mark.gpu


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


@pytest.mark.parametrize("cfg", [(4, 2), (1, 1), (2, 4)])
def test_forward_backward_via_c_abi(cfg):
    lib = abi.load()
    a, n, g = synthetic.custom([300, 17, 0, 256, 1, 299], 300, seed=99)
    dev = "cuda"
    A = torch.tensor(a, device=dev)
    N = torch.tensor(n, device=dev)
    G = torch.tensor(g, device=dev)
    B, _, L = a.shape
    c = abi.Config(*cfg, 0, 0)
    st = torch.empty(max(1, lib.tpl_a2b_state_bytes(0, B, L, ctypes.byref(c))), dtype=torch.uint8, device=dev)
    sy = torch.full((max(1, lib.tpl_a2b_sync_bytes(0, B, L, ctypes.byref(c))),), 7, dtype=torch.uint8,
                    device=dev)
    X = torch.full((B, 9 * L), float("nan"), device=dev)
    GA = torch.full((B, 3, L), float("nan"), device=dev)
    status = torch.zeros(1, dtype=torch.int32, device=dev)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert lib.tpl_a2b_sync_clear(_p(sy), sy.numel(), stream) == 0  # garbage -> zero
    rc = lib.tpl_a2b_forward(0, _p(A), _p(X), _p(N), _p(st), _p(sy), _p(status), B, L,
                             ctypes.byref(c), stream)
    assert rc == 0, lib.tpl_a2b_strerror(rc)
    rc = lib.tpl_a2b_backward(0, _p(GA), _p(G), _p(A), _p(N), _p(st), _p(sy), _p(status), B, L,
                              ctypes.byref(c), stream)
    assert rc == 0, lib.tpl_a2b_strerror(rc)
    torch.cuda.synchronize()
    assert status.item() == 0
    assert int(sy.count_nonzero()) == 0  # the look-back buffer is all zero again
    x = X.cpu().numpy().astype(np.float64)
    ga = GA.cpu().numpy().astype(np.float64)
    x64 = O.forward(a.astype(np.float64), n)
    g64 = O.backward_adjoint(a.astype(np.float64), n, g.astype(np.float64))
    assert np.abs(x - x64).max() <= 1.2e-7 * np.abs(x64).max() + 1e-6
    for b in range(B):
        if n[b] >= 2:
            assert np.abs(ga[b] - g64[b]).max() <= 1e-5 * np.abs(g64[b]).max()
        assert np.all(x[b, 9 * n[b]:] == 0) and np.all(ga[b, :, n[b]:] == 0)


def test_plain_via_c_abi_and_tuner():
    lib = abi.load()
    a, n, g = synthetic.custom([64, 3, 50], 64, seed=3)
    dev = "cuda"
    A, N, G = (torch.tensor(v, device=dev) for v in (a, n, g))
    B, _, L = a.shape
    fr = torch.empty(lib.tpl_a2b_plain_frames_bytes(0, B, L), dtype=torch.uint8, device=dev)
    X = torch.empty(B, 9 * L, device=dev)
    GA = torch.empty(B, 3, L, device=dev)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert lib.tpl_a2b_forward_plain(0, _p(A), _p(X), _p(N), _p(fr), None, B, L, stream) == 0
    assert lib.tpl_a2b_backward_plain(0, _p(GA), _p(G), _p(A), _p(N), _p(fr), None, B, L, stream) == 0
    torch.cuda.synchronize()
    x64 = O.forward(a.astype(np.float64), n)
    assert np.abs(X.cpu().numpy() - x64).max() <= 2e-3
    cands = (abi.Config * 3)(abi.Config(1, 1, 0, 0), abi.Config(4, 1, 0, 0), abi.Config(3, 1, 0, 0))
    best = abi.Config()
    ms = (ctypes.c_float * 3)()
    st = torch.empty(max(lib.tpl_a2b_state_bytes(0, B, L, ctypes.byref(cands[i])) for i in range(2)),
                     dtype=torch.uint8, device=dev)
    sy = torch.zeros(max(1, max(lib.tpl_a2b_sync_bytes(0, B, L, ctypes.byref(cands[i])) for i in range(2))),
                     dtype=torch.uint8, device=dev)
    rc = lib.tpl_a2b_tune(0, _p(A), _p(N), _p(G), _p(X), _p(GA), _p(st), _p(sy), B, L, cands, 3, 3,
                          ctypes.byref(best), ms, stream)
    assert rc == 0
    assert ms[0] > 0 and ms[1] > 0 and ms[2] == -1.0
    assert (best.fwd_res_per_lane, best.bwd_res_per_lane) in [(1, 1), (4, 1)]
