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


_COLD_CALLS = r'''
import ctypes, json, sys
import torch
from tests import abi
from torchproteinlibrary_amd import synthetic
lib = abi.load()
P = lambda t: ctypes.c_void_p(t.data_ptr())
dev = "cuda"
stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
a, n, g = synthetic.workload("long16k")
B, _, L = a.shape
A, N, G = (torch.tensor(v, device=dev) for v in (a, n, g))
X = torch.empty(B, 9 * L, device=dev)
GA = torch.empty(B, 3, L, device=dev)
out, first = [], None
for cfg in [(1, 1), (4, 4), (4, 2), (8, 4), (2, 2), (1, 1)]:
    c = abi.Config(*cfg, 0, 0)
    st = torch.empty(lib.tpl_a2b_state_bytes(0, B, L, ctypes.byref(c)), dtype=torch.uint8, device=dev)
    status = torch.zeros(2, dtype=torch.int32, device=dev)
    sy = torch.zeros(lib.tpl_a2b_sync_bytes(0, B, L, ctypes.byref(c)), dtype=torch.uint8, device=dev)
    assert lib.tpl_a2b_forward(0, P(A), P(X), P(N), P(st), P(sy), P(status), B, L, ctypes.byref(c), stream) == 0
    assert lib.tpl_a2b_backward(0, P(GA), P(G), P(A), P(N), P(st), P(sy),
                                ctypes.c_void_p(status.data_ptr() + 4), B, L, ctypes.byref(c), stream) == 0
    torch.cuda.synchronize()
    same = None
    if cfg == (1, 1):
        if first is None:
            first = (X.clone(), GA.clone())
        else:
            same = bool(torch.equal(X, first[0]) and torch.equal(GA, first[1]))
    out.append([list(cfg), status.cpu().tolist(), same, int(sy.count_nonzero())])
    print(json.dumps(out[-1]), flush=True)
print("RESULT " + json.dumps(out), flush=True)
'''


def test_first_calls_in_a_fresh_process_do_not_time_out():
    """Regression: the first calls of a process (code and look-back buffer
    cold) on long chains with many chunks per chain, where waiting waves used
    to starve the wave they waited for until the look-back gave up
    (TPL_A2B_STATUS_TIMEOUT, typically on the first (1, 1) call).  Run in a
    child process so that nothing is warm; every call must finish with status
    0 and a clean buffer, and the two (1, 1) calls must agree bit for bit."""
    import json
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    print("child process: long16k, configs (1,1) (4,4) (4,2) (8,4) (2,2) (1,1), first calls", flush=True)
    r = subprocess.run([sys.executable, "-c", _COLD_CALLS], cwd=root, capture_output=True, text=True,
                       timeout=110)
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stderr[-2000:]
    res = json.loads([ln for ln in r.stdout.splitlines() if ln.startswith("RESULT ")][-1][7:])
    for cfg, status, same, dirty in res:
        assert status == [0, 0], f"config {cfg}: status forward/backward {status}"
        assert dirty == 0, f"config {cfg}: look-back buffer left {dirty} nonzero bytes"
    assert res[-1][2] is True


def _abi_step(lib, a, n, g, cfg, dtype):
    """Forward and backward through the C ABI (FP64 I/O when dtype is
    torch.float64): coordinates, angle gradients and the status word."""
    dev = "cuda"
    dt = 1 if dtype == torch.float64 else 0
    A = torch.tensor(a, dtype=dtype, device=dev)
    N = torch.tensor(n, device=dev)
    G = torch.tensor(g, dtype=dtype, device=dev)
    B, _, L = a.shape
    c = abi.Config(*cfg, 0, 0)
    st = torch.empty(max(1, lib.tpl_a2b_state_bytes(dt, B, L, ctypes.byref(c))), dtype=torch.uint8, device=dev)
    sy = torch.zeros(max(1, lib.tpl_a2b_sync_bytes(dt, B, L, ctypes.byref(c))), dtype=torch.uint8, device=dev)
    X = torch.full((B, 9 * L), float("nan"), dtype=dtype, device=dev)
    GA = torch.full((B, 3, L), float("nan"), dtype=dtype, device=dev)
    status = torch.zeros(1, dtype=torch.int32, device=dev)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert lib.tpl_a2b_forward(dt, _p(A), _p(X), _p(N), _p(st), _p(sy), _p(status), B, L, ctypes.byref(c),
                               stream) == 0
    assert lib.tpl_a2b_backward(dt, _p(GA), _p(G), _p(A), _p(N), _p(st), _p(sy), _p(status), B, L,
                                ctypes.byref(c), stream) == 0
    torch.cuda.synchronize()
    assert int(sy.count_nonzero()) == 0
    return X.cpu().numpy(), GA.cpu().numpy(), int(status.item())


def test_chain_and_chunk_kernels_agree():
    """The same 64 chains (L=500, config (2,2): at most 4 chunks of 128
    residues per chain, FP64 I/O) alone, where every chain's workgroup is
    resident and the chain kernels run, and at the head of a batch of 1100
    (the rest empty), where the chunk kernels and their look-back run.  Both
    match each other to 1e-12 of the largest value (they compose the chunk
    aggregates in different orders) and the oracle."""
    lib = abi.load()
    lengths = [500, 499, 384, 383, 257, 256, 255, 129, 128, 127, 1, 0] + list(range(13, 500, 9))[:52]
    a, n, g = synthetic.custom(lengths, 500, seed=61)
    a64, g64 = a.astype(np.float64), g.astype(np.float64)
    print(f"chain vs chunk kernels: B={len(lengths)} and 1100, L=500, cfg (2,2), f64", flush=True)
    x1, ga1, s1 = _abi_step(lib, a64, n, g64, (2, 2), torch.float64)
    pad = 1100 - len(lengths)
    x2, ga2, s2 = _abi_step(lib, np.concatenate([a64, np.zeros((pad, 3, 500))]),
                            np.concatenate([n, np.zeros(pad, dtype=n.dtype)]),
                            np.concatenate([g64, np.zeros((pad, 9 * 500))]), (2, 2), torch.float64)
    assert s1 == 0 and s2 == 0
    x2, ga2 = x2[:len(lengths)], ga2[:len(lengths)]
    assert np.abs(x1 - x2).max() <= 1e-12 * np.abs(x1).max()
    assert np.abs(ga1 - ga2).max() <= 1e-12 * np.abs(ga1).max()
    xo = O.forward(a64, n)
    go = O.backward_adjoint(a64, n, g64)
    assert np.abs(x1 - xo).max() <= 1e-9 * np.abs(xo).max()
    for b in range(len(lengths)):
        if n[b] >= 2:
            assert np.abs(ga1[b] - go[b]).max() <= 1e-9 * np.abs(go[b]).max()
        assert np.all(x1[b, 9 * n[b]:] == 0) and np.all(ga1[b, :, n[b]:] == 0)
