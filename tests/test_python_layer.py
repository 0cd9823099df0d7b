"""CPU checks of the Python layer (no kernel runs here)."""
import numpy as np
import pytest
import torch

from torchproteinlibrary_amd import _native, _tuning, synthetic
from torchproteinlibrary_amd.ReducedModel import Angles2Backbone


def test_native_extension_is_the_in_tree_build():
    import os
    from torchproteinlibrary_amd import _build
    assert os.path.samefile(_native.EXT_PATH, _build.ext_path())


def test_cpu_tensors_go_to_the_host_path_float64_only():
    x = Angles2Backbone()(torch.zeros(2, 3, 4, dtype=torch.float64), torch.tensor([4, 2], dtype=torch.int32))
    assert x.shape == (2, 36) and x.dtype == torch.float64
    with pytest.raises(RuntimeError, match="float64"):
        Angles2Backbone()(torch.zeros(2, 3, 4), torch.tensor([4, 2], dtype=torch.int32))


def test_step_graph_needs_gpu_tensors():
    from torchproteinlibrary_amd.ReducedModel.Angles2Backbone import Angles2BackboneStep
    with pytest.raises(RuntimeError, match="GPU tensors"):
        Angles2BackboneStep(torch.zeros(2, 3, 4), torch.tensor([4, 2], dtype=torch.int32))
    with pytest.raises(Exception, match="Incorrect input size"):
        Angles2BackboneStep(torch.zeros(2, 12), torch.tensor([4, 2], dtype=torch.int32))


def test_bad_form():
    with pytest.raises(ValueError):
        Angles2Backbone("fast")


def test_candidates_cover_the_length():
    for L in (1, 30, 64, 65, 200, 700, 1000, 16384):
        for dt, (fmax, bmax) in ((torch.float32, (8, 4)), (torch.float64, (4, 2))):
            c = _tuning.candidates(L, dt)
            assert c and all(f in (1, 2, 4, 8) and b in (1, 2, 4) for f, b in c)
            assert all(f <= fmax and b <= bmax for f, b in c)
            cap = 64
            while cap < L:
                cap *= 2
            assert all(64 * f <= cap and 64 * b <= cap for f, b in c)
    assert (4, 2) in _tuning.candidates(1000, torch.float32)
    assert _tuning.candidates(30, torch.float32) == [(1, 1)]


def test_length_bucket():
    b = _tuning.length_bucket(np.array([30, 40, 100, 300, 1000, 1000, 999, 5]), 1000)
    assert len(b) == 6  # [0,64) [64,128) [128,256) [256,512) [512,1024) [1024,...)
    assert b == (3, 1, 0, 1, 3, 0)
    assert _tuning.length_bucket(np.full(64, 700), 1000) == (0, 0, 0, 0, 8, 0)
    assert _tuning.length_bucket(np.array([], np.int32), 10) == ()


def test_workspace_bytes_binding():
    st, sy = _native.ext.workspace_bytes(False, 512, 200, [4, 1])
    assert st >= 512 * 4 * 96 and sy > 0
    assert _native.ext.workspace_bytes(True, 512, 200, [4, 1]) == (st, sy)  # dtype-independent
    st, sy = _native.ext.workspace_bytes(False, 64, 16384, [4, 2])
    assert st >= 64 * 128 * 96 and sy > 64 * 64 * 24 * 8
    with pytest.raises(RuntimeError):
        _native.ext.workspace_bytes(False, 4, 4, [3, 1])
    with pytest.raises(RuntimeError):
        _native.ext.workspace_bytes(False, 4, 4, [4, 1, 0])


def test_synthetic_workloads_are_deterministic():
    a1, n1, g1 = synthetic.workload("mixed1000")
    a2, n2, g2 = synthetic.workload("mixed1000")
    assert np.array_equal(a1, a2) and np.array_equal(n1, n2) and np.array_equal(g1, g2)
    assert a1.shape == (1024, 3, 1000) and a1.dtype == np.float32
    assert n1.min() >= 30 and n1.max() <= 1000
    assert 240 < n1.mean() < 320  # E[L] ~ 276.6
    for b in range(0, 1024, 97):
        assert np.all(a1[b, :, n1[b]:] == 0) and np.all(g1[b, 9 * n1[b]:] == 0)
    a, n, g = synthetic.workload("long16k")
    assert n.min() >= 4096 and n.max() <= 16384 and a.shape == (64, 3, 16384)
