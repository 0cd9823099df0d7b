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


def test_cpu_tensors_rejected():
    with pytest.raises(RuntimeError, match="GPU only"):
        Angles2Backbone()(torch.zeros(2, 3, 4), torch.tensor([4, 2], dtype=torch.int32))


def test_bad_form():
    with pytest.raises(ValueError):
        Angles2Backbone("fast")


def test_candidates_cover_the_length():
    for L in (1, 30, 64, 200, 700, 1000, 16384):
        for dt, cap in ((torch.float32, 1024), (torch.float64, 512)):
            c = _tuning.candidates(L, dt, cus=256, batch=1024)
            assert c and all(w in (1, 2, 4) and r in (1, 2, 4, 8) for w, r, _ in c)
            assert all(64 * w * r <= cap for w, r, _ in c)
            # persistent grids only where one tile holds a chain
            assert all(g in (0, 512, 768, 1024) for _, _, g in c)
            assert all(g == 0 or 64 * w * r >= L for w, r, g in c)
    assert (4, 4, 768) in _tuning.candidates(1000, torch.float32, cus=256, batch=1024)
    assert all(g == 0 for _, _, g in _tuning.candidates(1000, torch.float32))
    assert all(g == 0 for _, _, g in _tuning.candidates(1000, torch.float32, cus=256, batch=512))


def test_workspace_bytes_binding():
    # float64, one 256-residue tile per chain both ways: nothing to carry
    assert _native.ext.workspace_bytes(True, 512, 200, [4, 1, 0]) == (0, 0)
    # float32: 2 bytes of residual per coordinate even with one tile
    st, sy = _native.ext.workspace_bytes(False, 512, 200, [4, 1, 0])
    assert st >= 512 * 9 * 200 * 2 and sy == 0
    st, sy = _native.ext.workspace_bytes(False, 64, 16384, [4, 4, 0])
    assert st > 64 * 9 * 16384 * 2 + 64 * 32 * 96 and sy > 64 * 16 * 24 * 8
    with pytest.raises(RuntimeError):
        _native.ext.workspace_bytes(False, 4, 4, [3, 1, 0])


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
