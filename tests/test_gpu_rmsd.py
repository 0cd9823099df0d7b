"""GPU parity of Coords2RMSD: the HIP kernels, through the Python layer and
through the C ABI, against the FP64 oracle (oracle/rmsd_oracle.py) on the same
inputs.  The kernels compute in FP64 from the input dtype, so the oracle runs
on the inputs as the GPU sees them (rounded to the dtype)."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import rmsd_oracle as R
from tests import abi_rmsd

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def _module_run(x, y, n, g, dtype):
    from torchproteinlibrary_amd.RMSD import Coords2RMSD
    X = torch.tensor(x, dtype=dtype, device=DEV, requires_grad=True)
    Y = torch.tensor(y, dtype=dtype, device=DEV)
    N = torch.tensor(n, dtype=torch.int32, device=DEV)
    m = Coords2RMSD()
    r = m(X, Y, N)
    r.backward(torch.tensor(g, dtype=dtype, device=DEV))
    torch.cuda.synchronize()
    return r.detach().cpu().numpy(), m.UT.cpu().numpy(), X.grad.cpu().numpy()


def _as_seen(a, dtype):
    return np.asarray(a).astype(np.float32 if dtype == torch.float32 else np.float64).astype(np.float64)


# (rmsd relative, UT absolute, gradient relative to its max) per input dtype:
# FP64 everywhere but the final rounding to the dtype
TOL = {torch.float64: (1e-9, 1e-8, 1e-8), torch.float32: (2e-6, 2e-6, 2e-6)}


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_parity_with_oracle(dtype):
    x, y, n, g = R.workload(48, 300, seed=21, min_atoms=0)
    for b, k in enumerate([0, 1, 2, 3, 300]):  # edge counts; slots past n are 0
        n[b] = k
        x[b, 3 * k:] = 0
        y[b, 3 * k:] = 0
    print(f"rmsd parity: B=48 A=300 dtype={dtype}", flush=True)
    r, ut, ga = _module_run(x, y, n, g, dtype)
    xs, ys, gs = _as_seen(x, dtype), _as_seen(y, dtype), _as_seen(g, dtype)
    r64, ut64 = R.forward(xs, ys, n)
    g64 = R.backward(xs, ys, n, gs)
    tr, tu, tg = TOL[dtype]
    assert r[0] == np.inf and r64[0] == np.inf
    fin = np.isfinite(r64)
    assert np.all(np.abs(r[fin] - r64[fin]) <= tr * r64[fin])
    # With two atoms the rotation about their axis is free (the largest
    # eigenvalue of F is double), so U is compared from three atoms on; the
    # RMSD and the gradient (through U^T y_c) are unique in every case.
    # The double root of two atoms takes the shifted adjugate (rmsd_math.h),
    # whose rotation is good to about 1e-8: its gradient row gets 1e-7.
    many = n >= 3
    assert np.abs(ut[many] - ut64[many]).max() <= tu
    gmax = np.abs(g64).max()
    assert np.abs(ga[many] - g64[many]).max() <= tg * gmax
    assert np.abs(ga[~many] - g64[~many]).max() <= max(tg, 1e-7) * gmax
    for b in range(48):
        assert not ga[b, 3 * int(n[b]):].any(), b  # exact zeros past num_atoms


def test_rigid_copy_has_zero_rmsd():
    """The reference's forward test (UnitTests/RMSD/Coords2RMSD/test.py:30-49):
    a rotated and translated copy has RMSD 0 to 5 places, in FP64 and FP32."""
    rng = np.random.default_rng(4)
    x, _, n, _ = R.workload(16, 64, seed=4, min_atoms=16)
    y = np.zeros_like(x)
    for b in range(16):
        k = int(n[b])
        ys = x[b, : 3 * k].reshape(k, 3) @ R.random_rotation(rng).T + rng.normal(size=3)
        y[b, : 3 * k] = ys.reshape(-1)
    from torchproteinlibrary_amd.RMSD import Coords2RMSD
    for dtype, places in ((torch.float64, 5), (torch.float32, 4)):
        r = Coords2RMSD()(torch.tensor(x, dtype=dtype, device=DEV), torch.tensor(y, dtype=dtype, device=DEV),
                          torch.tensor(n, device=DEV))
        assert np.abs(r.cpu().numpy()).max() < 0.5 * 10 ** (-places), dtype


def test_identical_structures_give_identity_and_zero_gradient():
    """input == target: UT = I, RMSD of the order of sqrt(1e-10 / n) (the
    reference's epsilon; the FP64 rounding of R2 - 2 lambda is of the same
    order and is clamped at 0 from below) and a zero gradient.  The
    quaternion is (1, 0, 0, 0), so three of its four adjugate rows are
    rounding noise: the eigenvector must come from the row of the largest
    |q_i|."""
    from torchproteinlibrary_amd.RMSD import Coords2RMSD
    x, _, n, _ = R.workload(16, 200, seed=5, min_atoms=3)
    for dtype in (torch.float64, torch.float32):
        X = torch.tensor(x, dtype=dtype, device=DEV, requires_grad=True)
        N = torch.tensor(n, device=DEV)
        m = Coords2RMSD()
        r = m(X, X.detach().clone(), N)
        r.backward(torch.ones_like(r))
        torch.cuda.synchronize()
        ut = m.UT.cpu().numpy()
        assert np.abs(ut - np.eye(3)[None]).max() <= 1e-6, dtype
        assert np.all(np.isfinite(r.detach().cpu().numpy())), dtype
        assert np.all(r.detach().cpu().numpy() <= np.sqrt(3e-10 / n)), dtype
        assert np.abs(X.grad.cpu().numpy()).max() <= 1e-3, dtype


@pytest.mark.parametrize("axis,angle", [((0, 0, 1), 0.7), ((1, 0, 0), 0.7), ((0, 1, 0), -1.3),
                                        ((0, 0, 1), np.pi), ((1, 1, 0), np.pi)])
def test_exact_special_rotations(axis, angle):
    """y = Q x for rotations whose quaternions have zero components
    (single-axis turns, half turns): U = Q and RMSD ~ 0, against the oracle
    and the exact rotation."""
    from torchproteinlibrary_amd.RMSD import Coords2RMSD
    ax = np.asarray(axis, np.float64) / np.linalg.norm(axis)
    Q = R.u_from_q(np.concatenate([[np.cos(angle / 2)], np.sin(angle / 2) * ax]))
    x, _, n, g = R.workload(16, 120, seed=6, min_atoms=3)
    y = np.zeros_like(x)
    for b in range(16):
        k = int(n[b])
        y[b, : 3 * k] = (x[b, : 3 * k].reshape(k, 3) @ Q.T).reshape(-1)
    X = torch.tensor(x, device=DEV, requires_grad=True)
    m = Coords2RMSD()
    r = m(X, torch.tensor(y, device=DEV), torch.tensor(n, device=DEV))
    r.backward(torch.tensor(g, device=DEV))
    torch.cuda.synchronize()
    ut = m.UT.cpu().numpy()
    assert np.abs(ut - Q.T[None]).max() <= 1e-8
    assert np.abs(r.detach().cpu().numpy()).max() <= 1e-4
    r64, ut64 = R.forward(x, y, n)
    assert np.abs(ut - ut64).max() <= 1e-8
    g64 = R.backward(x, y, n, g)
    assert np.abs(X.grad.cpu().numpy() - g64).max() <= 1e-6 * max(np.abs(g64).max(), 1.0)


def test_gradcheck_fp64():
    """The reference's backward test (test.py:67-83): torch.autograd.gradcheck
    on the Module, batch 1, up to 16 atoms, eps 1e-6, atol 1e-5, rtol 1e-5."""
    from torchproteinlibrary_amd.RMSD import Coords2RMSD
    torch.manual_seed(0)
    src = torch.randn(1, 48, dtype=torch.double, device=DEV).requires_grad_()
    dst = torch.randn(1, 48, dtype=torch.double, device=DEV)
    num = torch.randint(8, 16, (1,), dtype=torch.int32, device=DEV)
    assert torch.autograd.gradcheck(Coords2RMSD(), (src, dst, num), eps=1e-6, atol=1e-5, rtol=1e-5)


def test_function_returns_rmsd_and_ut_and_centres_itself():
    """Coords2RMSDFunction.apply(centred input, centred target, n) -> (rmsd, UT)
    as in the reference (Coords2RMSD.py:17-38); raw inputs give the same."""
    from torchproteinlibrary_amd.RMSD.Coords2RMSD import Coords2RMSDFunction
    x, y, n, _ = R.workload(8, 40, seed=9, min_atoms=5)
    xc, yc = x.copy(), y.copy()
    for b in range(8):
        k = int(n[b])
        for arr in (xc, yc):
            v = arr[b, : 3 * k].reshape(k, 3)
            arr[b, : 3 * k] = (v - v.mean(0)).reshape(-1)
    t = lambda a: torch.tensor(a, device=DEV)  # noqa: E731
    N = torch.tensor(n, device=DEV)
    r1, u1 = Coords2RMSDFunction.apply(t(xc), t(yc), N)
    r2, u2 = Coords2RMSDFunction.apply(t(x), t(y), N)
    assert torch.allclose(r1, r2, rtol=1e-12, atol=0) and torch.allclose(u1, u2, atol=1e-12)


def test_c_abi_direct():
    """tpl_rmsd_forward / tpl_rmsd_backward called through ctypes on torch's
    buffers match the Python layer bit for bit."""
    lib = abi_rmsd.load()
    x, y, n, g = R.workload(32, 200, seed=13, min_atoms=1)
    X, Y = torch.tensor(x, device=DEV), torch.tensor(y, device=DEV)
    N, G = torch.tensor(n, device=DEV), torch.tensor(g, device=DEV)
    B, A = 32, 200
    rm = torch.empty(B, dtype=torch.float64, device=DEV)
    ut = torch.empty(B, 3, 3, dtype=torch.float64, device=DEV)
    st = torch.empty(lib.tpl_rmsd_state_bytes(B), dtype=torch.uint8, device=DEV)
    gx = torch.full((B, 3 * A), 7.0, dtype=torch.float64, device=DEV)
    status = torch.zeros(1, dtype=torch.int32, device=DEV)
    torch.cuda.synchronize()
    assert lib.tpl_rmsd_forward(1, X.data_ptr(), Y.data_ptr(), N.data_ptr(), rm.data_ptr(), ut.data_ptr(),
                                st.data_ptr(), status.data_ptr(), B, A, None) == 0
    assert lib.tpl_rmsd_backward(1, gx.data_ptr(), G.data_ptr(), X.data_ptr(), Y.data_ptr(), N.data_ptr(),
                                 st.data_ptr(), status.data_ptr(), B, A, None) == 0
    torch.cuda.synchronize()
    assert int(status.item()) == 0
    r, u, ga = _module_run(x, y, n, g, torch.float64)
    assert np.array_equal(rm.cpu().numpy(), r)
    assert np.array_equal(ut.cpu().numpy(), u)
    assert np.array_equal(gx.cpu().numpy(), ga)


def test_errors():
    from torchproteinlibrary_amd.RMSD import Coords2RMSD
    m = Coords2RMSD()
    x = torch.zeros(2, 30, device=DEV)
    n = torch.tensor([10, 10], dtype=torch.int32, device=DEV)
    with pytest.raises(RuntimeError):
        m(x.cpu(), x.cpu(), n.cpu())  # GPU only
    with pytest.raises(ValueError):
        m(torch.zeros(2, 3, 10, device=DEV), torch.zeros(2, 3, 10, device=DEV), n)  # rank
    with pytest.raises(ValueError):
        m(x, x, torch.tensor([10, 11], dtype=torch.int32, device=DEV))  # num_atoms > A
    with pytest.raises(RuntimeError):
        m(x, x.double(), n)  # dtype mismatch
    with pytest.raises(RuntimeError):
        m(x, x, n.long())  # num_atoms must be int32
    x2 = torch.full((1, 9), float("nan"), device=DEV)
    with pytest.raises(Exception, match="forward Nan"):
        m(x2, x2, torch.tensor([3], dtype=torch.int32, device=DEV))


def test_backbone_batch_full_size():
    """Backbone-sized structures: B=1024, up to 3000 atoms (the N/CA/C atoms of
    chains up to 1000 residues), FP32, against the oracle."""
    x, y, n, g = R.workload(1024, 3000, seed=17, min_atoms=90, noise=1.0, dtype=np.float32)
    print("rmsd full size: B=1024 A=3000 f32", flush=True)
    r, ut, ga = _module_run(x, y, n, g, torch.float32)
    xs, ys, gs = _as_seen(x, torch.float32), _as_seen(y, torch.float32), _as_seen(g, torch.float32)
    r64, ut64 = R.forward(xs, ys, n)
    g64 = R.backward(xs, ys, n, gs)
    tr, tu, tg = TOL[torch.float32]
    assert np.all(np.abs(r - r64) <= tr * r64)
    assert np.abs(ut - ut64).max() <= tu
    assert np.abs(ga - g64).max() <= tg * np.abs(g64).max()
