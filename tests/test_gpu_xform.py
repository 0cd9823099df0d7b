"""GPU parity of the coordinate transforms (FullAtomModel.CoordsTransform)
against the FP64 oracle, the reference's own checks, and their composition
with Coords2RMSD."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import rmsd_oracle as R
from oracle import xform_oracle as X
from tests import abi_xform

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def _data(B, A, seed, dtype):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(B, 3 * A)) * 10
    n = rng.integers(1, A + 1, size=B).astype(np.int32)
    n[0] = A
    Rm = np.stack([R.random_rotation(rng) for _ in range(B)])
    T = rng.normal(size=(B, 3)) * 5
    np_dt = np.float32 if dtype == torch.float32 else np.float64
    return [a.astype(np_dt).astype(np.float64) for a in (x, Rm, T)] + [n]


TOL = {torch.float64: 1e-12, torch.float32: 2e-6}


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_parity_with_oracle(dtype):
    from torchproteinlibrary_amd.FullAtomModel.CoordsTransform import (Coords2Center, CoordsRotate,
                                                                          CoordsTranslate)
    x, Rm, T, n = _data(32, 300, 5, dtype)
    print(f"xform parity: B=32 A=300 dtype={dtype}", flush=True)
    t = lambda a: torch.tensor(a, dtype=dtype, device=DEV)  # noqa: E731
    N = torch.tensor(n, device=DEV)
    scale = np.abs(x).max() + np.abs(T).max()
    tol = TOL[dtype] * scale
    # forwards
    X_t = t(x).requires_grad_()
    out_t = CoordsTranslate()(X_t, t(T), N)
    out_r = CoordsRotate()(X_t, t(Rm), N)
    out_c = Coords2Center()(X_t, N)
    assert np.abs(out_t.detach().cpu().numpy() - X.translate(x, T, n)).max() <= tol
    assert np.abs(out_r.detach().cpu().numpy() - X.rotate(x, Rm, n)).max() <= tol
    assert np.abs(out_c.detach().cpu().numpy() - X.center(x, n)).max() <= tol
    # backwards (the three outputs' gradients sum into X_t.grad)
    rng = np.random.default_rng(9)
    g1 = rng.normal(size=x.shape)
    g2 = rng.normal(size=x.shape)
    g3 = rng.normal(size=(32, 3))
    torch.autograd.backward([out_t, out_r, out_c], [t(g1), t(g2), t(g3)])
    want = X.translate_backward(g1, n) + X.rotate(g2, Rm, n, transpose=True) + X.center_backward(g3, n, 300)
    assert np.abs(X_t.grad.cpu().numpy() - want).max() <= TOL[dtype] * 10
    for b in range(32):  # exact zeros past num_atoms
        assert not out_t.detach().cpu().numpy()[b, 3 * int(n[b]):].any()
        assert not X_t.grad.cpu().numpy()[b, 3 * int(n[b]):].any()


def test_reference_checks():
    """The reference's own checks (UnitTests/FullAtomModel/CoordsTransform/test.py):
    the centre against a manual mean (7 places in FP64), translating by minus
    the centre centres the structure, and gradcheck of centre, translate and
    rotate in FP64."""
    from torchproteinlibrary_amd.FullAtomModel.CoordsTransform import (
        Coords2Center, CoordsRotate, CoordsTranslate, getRandomRotation)
    torch.manual_seed(0)
    B, A = 8, 30
    x = torch.randn(B, 3 * A, dtype=torch.float64, device=DEV) * 5
    n = torch.randint(A // 2, A + 1, (B,), dtype=torch.int32, device=DEV)
    c = Coords2Center()(x, n)
    for i in range(B):
        man = x[i, : 3 * int(n[i])].view(-1, 3).mean(0)
        assert torch.allclose(man, c[i], atol=1e-7, rtol=0)
    centred = CoordsTranslate()(x, -c, n)
    assert Coords2Center()(centred, n).abs().max().item() < 1e-7
    xs = torch.randn(2, 3 * 12, dtype=torch.float64, device=DEV, requires_grad=True)
    ns = torch.tensor([12, 7], dtype=torch.int32, device=DEV)
    Rm = getRandomRotation(2).to(DEV)
    Tt = torch.randn(2, 3, dtype=torch.float64, device=DEV)
    assert torch.autograd.gradcheck(lambda v: Coords2Center()(v, ns), (xs,), eps=1e-6, atol=1e-5)
    assert torch.autograd.gradcheck(lambda v: CoordsTranslate()(v, Tt, ns), (xs,), eps=1e-6, atol=1e-5)
    assert torch.autograd.gradcheck(lambda v: CoordsRotate()(v, Rm, ns), (xs,), eps=1e-6, atol=1e-5)


def test_composed_rmsd_matches_fused():
    """The reference's Coords2RMSD Module composes Coords2Center,
    CoordsTranslate and Coords2RMSDFunction (Coords2RMSD.py:71-78); that
    composition with these ops equals the fused layer, value and gradient."""
    from torchproteinlibrary_amd.FullAtomModel.CoordsTransform import Coords2Center, CoordsTranslate
    from torchproteinlibrary_amd.RMSD import Coords2RMSD
    from torchproteinlibrary_amd.RMSD.Coords2RMSD import Coords2RMSDFunction
    x, y, n, g = R.workload(16, 80, seed=23, min_atoms=3)
    N = torch.tensor(n, device=DEV)
    Y = torch.tensor(y, device=DEV)
    X1 = torch.tensor(x, device=DEV, requires_grad=True)
    X2 = torch.tensor(x, device=DEV, requires_grad=True)
    c2c, tr = Coords2Center(), CoordsTranslate()
    ci = tr(X1, -c2c(X1, N), N)
    ct = tr(Y, -c2c(Y, N), N)
    r1, _ = Coords2RMSDFunction.apply(ci, ct, N)
    r2 = Coords2RMSD()(X2, Y, N)
    G = torch.tensor(g, device=DEV)
    r1.backward(G)
    r2.backward(G)
    assert torch.allclose(r1, r2, rtol=1e-10, atol=0)
    assert torch.allclose(X1.grad, X2.grad, rtol=0, atol=1e-10 * X2.grad.abs().max().item())


def test_c_abi_direct():
    lib = abi_xform.load()
    x, Rm, T, n = _data(8, 100, 7, torch.float64)
    Xd, Rd = torch.tensor(x, device=DEV), torch.tensor(Rm, device=DEV)
    Nd = torch.tensor(n, device=DEV)
    out = torch.full_like(Xd, 7.0)
    c = torch.empty(8, 3, dtype=torch.float64, device=DEV)
    torch.cuda.synchronize()
    assert lib.tpl_xform_rotate(1, out.data_ptr(), Xd.data_ptr(), Rd.data_ptr(), 0, Nd.data_ptr(),
                                None, 8, 100, None) == 0
    assert lib.tpl_xform_center(1, c.data_ptr(), Xd.data_ptr(), Nd.data_ptr(), None, 8, 100, None) == 0
    torch.cuda.synchronize()
    assert np.abs(out.cpu().numpy() - X.rotate(x, Rm, n)).max() <= 1e-12 * np.abs(x).max()
    assert np.abs(c.cpu().numpy() - X.center(x, n)).max() <= 1e-12 * np.abs(x).max()


def test_errors():
    from torchproteinlibrary_amd.FullAtomModel.CoordsTransform import Coords2Center, CoordsTranslate
    x = torch.zeros(2, 30, device=DEV)
    n = torch.tensor([10, 10], dtype=torch.int32, device=DEV)
    with pytest.raises(RuntimeError):
        Coords2Center()(x.cpu(), n.cpu())  # GPU only
    with pytest.raises(ValueError):
        Coords2Center()(torch.zeros(2, 10, 3, device=DEV), n)  # rank
    with pytest.raises(ValueError):
        CoordsTranslate()(x, torch.zeros(2, 3, device=DEV), torch.tensor([10, 12], dtype=torch.int32, device=DEV))
    with pytest.raises(Exception, match="forward Nan"):
        Coords2Center()(x, torch.tensor([0, 10], dtype=torch.int32, device=DEV))  # 0/0 as the reference
