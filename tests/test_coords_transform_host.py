"""CPU checks of the host utilities getBBox and getRandomTranslation
(reference CoordsTransform.py:12-33, Math/nUtil.cpp:27-57,88-99)."""
import numpy as np
import torch

from torchproteinlibrary_amd.FullAtomModel import getBBox, getRandomTranslation


def test_bbox_is_the_min_and_max_over_each_structure():
    rng = np.random.default_rng(3)
    B, A = 7, 50
    x = rng.normal(size=(B, 3 * A)) * 20
    n = rng.integers(1, A + 1, size=B).astype(np.int32)
    n[0] = A
    for dtype in (torch.float32, torch.float64):
        a, b = getBBox(torch.tensor(x, dtype=dtype), torch.tensor(n))
        assert a.dtype == dtype and b.dtype == dtype and a.shape == (B, 3)
        xs = torch.tensor(x, dtype=dtype).numpy()
        for i in range(B):
            v = xs[i, : 3 * n[i]].reshape(-1, 3)
            assert np.array_equal(a[i].numpy(), v.min(0))
            assert np.array_equal(b[i].numpy(), v.max(0))


def test_bbox_of_no_atoms_keeps_the_starting_values():
    a, b = getBBox(torch.ones(2, 9), torch.tensor([0, 3], dtype=torch.int32))
    assert torch.all(torch.isinf(a[0]) & (a[0] > 0)) and torch.all(torch.isinf(b[0]) & (b[0] < 0))
    assert torch.equal(a[1], torch.ones(3)) and torch.equal(b[1], torch.ones(3))


def test_bbox_rank_error():
    import pytest
    with pytest.raises(ValueError):
        getBBox(torch.zeros(2, 3, 3), torch.tensor([1, 1], dtype=torch.int32))


def test_random_translation_stays_in_the_volume():
    """Per axis T is uniform in [-d, d], d = max(0, V/2 - (b - a)/2) / 2: a
    centred box moved by T stays inside the cube of side V, and the draws
    reach both ends of the range."""
    gen = torch.Generator().manual_seed(0)
    V = 60.0
    a = torch.tensor([[-5.0, -10.0, -40.0]] * 4000, dtype=torch.float64)
    b = -a
    T = getRandomTranslation(a, b, V, generator=gen)
    d = np.maximum(0.0, V / 2 - (b - a).numpy()[0] / 2) / 2
    assert np.allclose(d, [12.5, 10.0, 0.0])
    t = T.numpy()
    assert np.all(np.abs(t) <= d + 1e-12)
    assert np.all(np.abs(t[:, 2]) == 0.0)  # the box is wider than the volume
    assert np.all(t[:, :2].max(0) > 0.95 * d[:2]) and np.all(t[:, :2].min(0) < -0.95 * d[:2])
    # x and y fit: the moved box stays inside the cube
    assert np.all(np.abs((b.numpy() + t)[:, :2]) <= V / 2) and np.all(np.abs((a.numpy() + t)[:, :2]) <= V / 2)
    assert getRandomTranslation(a.float(), b.float(), V).dtype == torch.float32
