"""GPU parity of FullAtomModel.Angles2Coords (forward and backward) against
the FP64 oracle (oracle/a2c_oracle.py, pinned to the reference's compiled
code on the CPU), the reference's own backward test, and edge cases."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import a2c_oracle as O
from tests import abi_a2c

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def _batch(rng, lengths, stride, dtype):
    seqs = ["".join(rng.choice(list(O.AA20), size=n)) for n in lengths]
    ang = np.zeros((len(seqs), 8, stride))
    for b, n in enumerate(lengths):
        ang[b, 0, :n] = rng.normal(-1.1, 0.5, n)
        ang[b, 1, :n] = rng.normal(2.3, 0.5, n)
        ang[b, 2, :n] = np.pi + rng.normal(0, 0.1, n)
        ang[b, 3:, :n] = rng.uniform(-np.pi, np.pi, (5, n))
    np_dt = np.float32 if dtype == torch.float32 else np.float64
    return seqs, ang.astype(np_dt).astype(np.float64)


def _names(t):
    return [bytes(row).rstrip(b"\0").decode() for row in t.cpu().numpy()]


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_parity_with_oracle(dtype):
    from torchproteinlibrary_amd.FullAtomModel import Angles2Coords
    rng = np.random.default_rng(1)
    lengths = [1, 2, 20, 57, 80, 33, 5, 64, 3, 41, 77, 12]
    print(f"a2c parity: B={len(lengths)} dtype={dtype}", flush=True)
    seqs, ang = _batch(rng, lengths, 83, dtype)
    seqs[2] = O.AA20
    X = torch.tensor(ang, dtype=dtype, device=DEV, requires_grad=True)
    coords, res_names, atom_names, num_atoms = Angles2Coords()(X, seqs)
    want, counts = O.forward(seqs, ang)
    assert num_atoms.cpu().tolist() == counts.tolist()
    assert coords.shape == want.shape and res_names.shape == (len(seqs), counts.max(), 4)
    c = coords.detach().cpu().numpy()
    tol = 1e-10 if dtype == torch.float64 else 3e-5
    assert np.abs(c - want).max() <= tol
    for b, s in enumerate(seqs):
        res3, atoms = O.names_one(s)
        n = counts[b]
        assert _names(atom_names[b, :n]) == atoms and _names(res_names[b, :n]) == res3
        assert not atom_names[b, n:].cpu().numpy().any() and not c[b, 3 * n:].any()
    g = rng.normal(size=c.shape)
    for b, n in enumerate(counts):
        g[b, 3 * n:] = 0.0
    coords.backward(torch.tensor(g, dtype=dtype, device=DEV))
    ga = X.grad.cpu().numpy()
    want_g = O.backward(seqs, ang, g)
    scale = np.abs(want_g).max(axis=(1, 2), keepdims=True)
    rel = np.abs(ga - want_g) / np.maximum(scale, 1e-30)
    assert rel.max() <= (1e-10 if dtype == torch.float64 else 1e-5), rel.max()
    for b, n in enumerate(lengths):  # exact zeros past the sequence and for omega of the last residue
        assert not ga[b, :, n:].any() and ga[b, 2, n - 1] == 0


def test_reference_backward_test():
    """UnitTests/FullAtomModel/Angles2Coords/test.py:83-142 on the GPU: the
    gradient of sum(coords) against forward differences, mean error < 0.01."""
    from torchproteinlibrary_amd.FullAtomModel import Angles2Coords
    seqs = ["ACDEFGHIKLMNPQRSTVWY"]
    x0 = torch.zeros(1, 8, 20, dtype=torch.float64, device=DEV)
    x0[:, 0], x0[:, 1], x0[:, 2], x0[:, 3:] = -1.047, -0.698, np.pi, 110.4 * np.pi / 180.0
    x0.requires_grad_()
    a2c = Angles2Coords()
    y0 = a2c(x0, seqs)[0].sum()
    y0.backward()
    err = []
    with torch.no_grad():
        for i in range(8):
            for j in range(20):
                x1 = x0.detach().clone()
                x1[0, i, j] += 1e-5
                dy = (a2c(x1, seqs)[0].sum().item() - y0.item()) / 1e-5
                err.append(abs(dy - x0.grad[0, i, j].item()))
    assert np.mean(err) < 0.01


def test_gradcheck_fp64():
    from torchproteinlibrary_amd.FullAtomModel.Angles2Coords import (Angles2CoordsFunction,
                                                                     convertStringList)
    seqs = ["MKW", "PG"]
    rng = np.random.default_rng(4)
    _, ang = _batch(rng, [3, 2], 3, torch.float64)
    X = torch.tensor(ang, device=DEV, requires_grad=True)
    S = convertStringList(seqs).to(DEV)
    N = torch.tensor([O.num_atoms(s) for s in seqs], dtype=torch.int32, device=DEV)
    assert torch.autograd.gradcheck(lambda a: Angles2CoordsFunction.apply(a, S, N)[0], (X,), eps=1e-6,
                                    atol=1e-6)


def test_c_abi_direct_matches_layer():
    from torchproteinlibrary_amd.FullAtomModel import Angles2Coords
    from torchproteinlibrary_amd.FullAtomModel.Angles2Coords import convertStringList
    lib = abi_a2c.load()
    rng = np.random.default_rng(2)
    seqs, ang = _batch(rng, [10, 30, 7], 30, torch.float32)
    X = torch.tensor(ang, dtype=torch.float32, device=DEV)
    coords, rn, an, n = Angles2Coords()(X, seqs)
    S = convertStringList(seqs).to(DEV)
    A = coords.shape[1] // 3
    out = torch.full_like(coords, 7.0)
    st = torch.zeros(1, dtype=torch.int32, device=DEV)
    torch.cuda.synchronize()
    assert lib.tpl_a2c_forward(0, S.data_ptr(), S.shape[1], X.data_ptr(), 30, out.data_ptr(), None, None, A,
                               st.data_ptr(), 3, None) == 0
    g = torch.randn_like(coords)
    ga = torch.full_like(X, 7.0)
    assert lib.tpl_a2c_backward(0, S.data_ptr(), S.shape[1], out.data_ptr(), g.data_ptr(), A, ga.data_ptr(), 30,
                                st.data_ptr(), 3, None) == 0
    torch.cuda.synchronize()
    assert int(st.item()) == 0
    assert torch.equal(out, coords)
    X2 = X.clone().requires_grad_()
    Angles2Coords()(X2, seqs)[0].backward(g)
    assert torch.equal(ga, X2.grad)


def test_edge_cases_and_errors():
    from torchproteinlibrary_amd.FullAtomModel import Angles2Coords
    a2c = Angles2Coords()
    X = torch.zeros(3, 8, 5, device=DEV)
    coords, rn, an, n = a2c(X, ["", "G", "WWWWW"])  # an empty sequence gives no atoms
    assert n.tolist() == [0, 4, 70] and coords.shape == (3, 210)
    assert not coords[0].any() and not coords[1, 12:].any()
    with pytest.raises(ValueError):
        a2c(X, ["AXG", "G", "G"])  # unknown letter
    with pytest.raises(ValueError):
        a2c(torch.zeros(1, 8, 3, device=DEV), ["GGGG"])  # longer than the angles
    with pytest.raises(ValueError):
        a2c(torch.zeros(1, 3, 5, device=DEV), ["GG"])  # rank / rows
    with pytest.raises(RuntimeError):
        a2c(torch.zeros(1, 8, 5), ["GG"])  # GPU only


def test_batch_mixed_lengths_full_size():
    """B=256, lengths log-uniform in [30, 1000] (FP32): backbone bond lengths
    of every structure, and oracle parity on a sample of structures."""
    from torchproteinlibrary_amd.FullAtomModel import Angles2Coords
    seqs, ang = O.workload(256, seed=9, dtype=np.float32)
    print(f"a2c full size: B=256 residues={sum(map(len, seqs))}", flush=True)
    X = torch.tensor(ang, dtype=torch.float32, device=DEV, requires_grad=True)
    coords, rn, an, n = Angles2Coords()(X, seqs)
    c = coords.detach().cpu().numpy().astype(np.float64)
    for b in range(0, 256, 7):
        s = seqs[b]
        idx = np.cumsum([0] + [len(O.residue_atoms(a)) for a in s])[:-1]
        X3 = c[b].reshape(-1, 3)
        assert np.abs(np.linalg.norm(X3[idx + 1] - X3[idx], axis=1) - 1.46).max() < 1e-3
    for b in (0, 17, 101):
        want, _, _ = O.forward_one(seqs[b], ang[b])
        assert np.abs(c[b, :want.size] - want).max() < 1e-3
    g = torch.randn_like(coords)
    coords.backward(g)
    assert torch.isfinite(X.grad).all()
    small = [b for b in range(256) if len(seqs[b]) <= 60][:2]
    for b in small:
        want = O.backward_one(seqs[b], ang[b], g[b].cpu().numpy().astype(np.float64))
        got = X.grad[b].cpu().numpy()
        assert np.abs(got - want).max() <= 1e-5 * np.abs(want).max()


def test_large_workgroups_match_small_and_oracle():
    """A few long structures (batch <= 512, stride > 256) run the forward on
    512-thread workgroups; the same structures padded to a batch of 513 with
    empty sequences run it on 256-thread workgroups (and the backward on the
    atom-tile kernel instead of per-thread runs, for batch x stride).  Both
    agree with each other and with the oracle (FP64)."""
    from torchproteinlibrary_amd.FullAtomModel import Angles2Coords
    rng = np.random.default_rng(12)
    lengths = [300, 520, 7, 260]
    print("a2c 512-thread path: lengths", lengths, flush=True)
    seqs, ang = _batch(rng, lengths, 520, torch.float64)
    pad = seqs + [""] * 509
    ang_pad = np.concatenate([ang, np.zeros((509, 8, 520))])
    Xa = torch.tensor(ang, device=DEV, requires_grad=True)
    Xb = torch.tensor(ang_pad, device=DEV, requires_grad=True)
    ca = Angles2Coords()(Xa, seqs)[0]
    cb = Angles2Coords()(Xb, pad)[0]
    assert torch.allclose(ca, cb[:4], rtol=0, atol=1e-10)
    want, counts = O.forward(seqs, ang)
    assert np.abs(ca.detach().cpu().numpy() - want).max() < 1e-9
    g = torch.randn_like(ca)
    ca.backward(g)
    cb.backward(torch.cat([g, torch.zeros(509, g.shape[1], dtype=g.dtype, device=DEV)]))
    ga, gb = Xa.grad, Xb.grad[:4]
    assert (ga - gb).abs().max().item() <= 1e-10 * gb.abs().max().item()
    want_g = O.backward_one(seqs[2], ang[2], g[2].cpu().numpy())
    assert np.abs(ga[2].cpu().numpy() - want_g).max() <= 1e-10 * max(1.0, np.abs(want_g).max())


def test_wave_boundary_lengths_both_workgroup_sizes():
    """Lengths at the 64-thread wave boundaries of the workgroup scans (one
    residue per thread in 512-thread workgroups, one or two in 256-thread
    ones), so that runs end in the last lane of a wave, in the first lane of
    the next, and waves past the structure skip the frame scan.  The batch of
    9 runs the forward on 512 threads; the same structures padded to 513 with
    empty sequences run it on 256 threads.  FP64 against the oracle, both
    directions."""
    from torchproteinlibrary_amd.FullAtomModel import Angles2Coords
    rng = np.random.default_rng(41)
    lengths = [63, 64, 65, 127, 128, 129, 255, 256, 257]
    print("a2c wave-boundary lengths", lengths, flush=True)
    seqs, ang = _batch(rng, lengths, 257, torch.float64)
    nb = len(lengths)
    pad = seqs + [""] * (513 - nb)
    ang_pad = np.concatenate([ang, np.zeros((513 - nb, 8, 257))])
    for ss, aa in ((seqs, ang), (pad, ang_pad)):
        X = torch.tensor(aa, device=DEV, requires_grad=True)
        coords = Angles2Coords()(X, ss)[0]
        g = rng.normal(size=tuple(coords.shape))
        coords.backward(torch.tensor(g, device=DEV))
        c = coords.detach().cpu().numpy()
        for b in range(nb):
            want_c, _, _ = O.forward_one(seqs[b], ang[b])
            assert np.abs(c[b, :want_c.size] - want_c).max() < 1e-9, (len(ss), lengths[b])
            assert not c[b, want_c.size:].any()
            want = O.backward_fast(seqs[b], ang[b], np.asarray(g[b], np.float64))
            rel = np.abs(X.grad[b].cpu().numpy() - want).max() / np.abs(want).max()
            assert rel <= 1e-9, (len(ss), lengths[b], rel)


def test_reference_signature_backward():
    """Angles2Coords_backward with the reference's arguments (angles, not
    coordinates) equals the layer's backward."""
    from torchproteinlibrary_amd._native_xform import ext
    from torchproteinlibrary_amd.FullAtomModel import Angles2Coords
    from torchproteinlibrary_amd.FullAtomModel.Angles2Coords import convertStringList
    rng = np.random.default_rng(6)
    seqs, ang = _batch(rng, [9, 40, 1], 40, torch.float64)
    X = torch.tensor(ang, device=DEV, requires_grad=True)
    coords = Angles2Coords()(X, seqs)[0]
    g = torch.randn_like(coords)
    coords.backward(g)
    ga = torch.empty_like(X)
    ext.Angles2Coords_backward(g, ga, convertStringList(seqs).to(DEV), X.detach())
    assert torch.allclose(ga, X.grad, rtol=0, atol=1e-12)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_long_chain_gradient_full_size(dtype):
    """Two 1000-residue structures (the workloads' longest) against the O(atoms)
    form of the oracle's backward: FP64 to 1e-9 relative, FP32 (whose
    backward works from FP32 coordinates) to 1e-4 relative."""
    from torchproteinlibrary_amd.FullAtomModel import Angles2Coords
    rng = np.random.default_rng(31)
    seqs, ang = _batch(rng, [1000, 1000], 1000, dtype)
    print(f"a2c long chains: 2 x 1000 residues, dtype={dtype}", flush=True)
    X = torch.tensor(ang, dtype=dtype, device=DEV, requires_grad=True)
    coords = Angles2Coords()(X, seqs)[0]
    g = rng.normal(size=tuple(coords.shape))
    coords.backward(torch.tensor(g, dtype=dtype, device=DEV))
    c = coords.detach().cpu().numpy()
    for b in range(2):
        want_c, _, _ = O.forward_one(seqs[b], ang[b])
        assert np.abs(c[b, :want_c.size] - want_c).max() < (1e-9 if dtype == torch.float64 else 1e-3)
        gq = np.asarray(g[b], np.float32 if dtype == torch.float32 else np.float64).astype(np.float64)
        want = O.backward_fast(seqs[b], ang[b], gq)
        rel = np.abs(X.grad[b].cpu().numpy() - want).max() / np.abs(want).max()
        assert rel <= (1e-9 if dtype == torch.float64 else 1e-4), rel


def test_matches_reference_outputs():
    """FP64 forward and backward on the GPU against outputs the reference's own
    compiled CPU code wrote (tests/golden/a2c_reference.npz): coordinates,
    atom names and gradients for the 20-residue sequence and two others."""
    import os
    from torchproteinlibrary_amd.FullAtomModel import Angles2Coords
    z = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "a2c_reference.npz"))
    k = 0
    while f"seq{k}" in z:
        seq = bytes(z[f"seq{k}"]).decode()
        X = torch.tensor(z[f"angles{k}"][None], device=DEV, requires_grad=True)
        coords, _, atom_names, n = Angles2Coords()(X, [seq])
        want = z[f"coords{k}"]
        assert int(n[0]) * 3 == want.size
        assert np.abs(coords[0].detach().cpu().numpy() - want).max() < 1e-10
        assert _names(atom_names[0]) == list(z[f"atom_names{k}"])
        coords.backward(torch.tensor(z[f"grad_coords{k}"][None], device=DEV))
        wg = z[f"grad_angles{k}"]
        assert np.abs(X.grad[0].cpu().numpy() - wg).max() <= 1e-10 * max(1.0, np.abs(wg).max())
        k += 1
    assert k == 3


def _backward_forms(seqs, ang, g, dtype):
    """Gradients by each backward kernel (tpl_a2c_backward_form: 1 = runs of
    residues per thread, 2 = atom tiles) from the layer's coordinates, into
    NaN-filled outputs, with the status word checked."""
    from torchproteinlibrary_amd.FullAtomModel import Angles2Coords
    from torchproteinlibrary_amd.FullAtomModel.Angles2Coords import convertStringList
    lib = abi_a2c.load()
    X = torch.tensor(ang, dtype=dtype, device=DEV)
    coords = Angles2Coords()(X, seqs)[0].contiguous()
    S = convertStringList(seqs).to(DEV)
    A = coords.shape[1] // 3
    G = torch.tensor(g[:, :3 * A], dtype=dtype, device=DEV).contiguous()
    dt = 0 if dtype == torch.float32 else 1
    out = {}
    for form in (1, 2):
        ga = torch.full_like(X, float("nan"))
        st = torch.zeros(1, dtype=torch.int32, device=DEV)
        torch.cuda.synchronize()
        print(f"a2c backward form {form}: B={len(seqs)} stride={ang.shape[2]} dtype={dtype}", flush=True)
        assert lib.tpl_a2c_backward_form(form, None, dt, S.data_ptr(), S.shape[1], coords.data_ptr(), G.data_ptr(),
                                         A, ga.data_ptr(), ang.shape[2], st.data_ptr(), len(seqs), None) == 0
        torch.cuda.synchronize()
        assert int(st.item()) == 0
        out[form] = ga.cpu().numpy().astype(np.float64)
        assert np.isfinite(out[form]).all()
    return out[1], out[2]


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_backward_kernels_match_oracle_small(dtype):
    """Both backward kernels on the small mixed batch (lengths 1-80 and the
    20-residue sequence) against the oracle, with exact zeros past each
    structure."""
    rng = np.random.default_rng(41)
    lengths = [1, 2, 20, 57, 80, 33, 5, 64, 3, 41, 77, 12]
    seqs, ang = _batch(rng, lengths, 83, dtype)
    seqs[2] = O.AA20
    g = rng.normal(size=(len(seqs), 3 * max(O.num_atoms(s) for s in seqs)))
    for b, s in enumerate(seqs):
        g[b, 3 * O.num_atoms(s):] = 0.0
    g = g.astype(np.float32 if dtype == torch.float32 else np.float64).astype(np.float64)
    runs, tiles = _backward_forms(seqs, ang, g, dtype)
    want = O.backward(seqs, ang, g)
    scale = np.abs(want).max(axis=(1, 2), keepdims=True)
    tol = 1e-10 if dtype == torch.float64 else 1e-5
    for got in (runs, tiles):
        assert (np.abs(got - want) / scale).max() <= tol
        for b, n in enumerate(lengths):
            assert not got[b, :, n:].any()
            assert got[b, 2, n - 1] == 0.0


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_backward_kernels_agree_at_tile_limits(dtype):
    """Both backward kernels at the tile kernel's 2048-residue limit: all
    tryptophan (14 atoms a residue, 28 672 atoms, 29 tiles), all glycine (4
    atoms) and a random 1500-residue sequence padded to 2048.  The kernels
    agree to 1e-9 of the largest gradient (both sum in FP64 from the same
    coordinates, in different orders), and the random one matches the
    oracle's O(atoms) backward (FP64 1e-9, FP32 1e-4 relative)."""
    rng = np.random.default_rng(43)
    seqs, ang = _batch(rng, [2048, 2048, 1500], 2048, dtype)
    seqs[0], seqs[1] = "W" * 2048, "G" * 2048
    g = rng.normal(size=(3, 3 * 2048 * 14))
    for b, s in enumerate(seqs):
        g[b, 3 * O.num_atoms(s):] = 0.0
    g = g.astype(np.float32 if dtype == torch.float32 else np.float64).astype(np.float64)
    runs, tiles = _backward_forms(seqs, ang, g, dtype)
    for b in range(3):
        assert np.abs(runs[b] - tiles[b]).max() <= 1e-9 * np.abs(runs[b]).max(), b
    want = O.backward_fast(seqs[2], ang[2], g[2])
    rel = np.abs(tiles[2] - want).max() / np.abs(want).max()
    assert rel <= (1e-9 if dtype == torch.float64 else 1e-4), rel


def test_scheduling_order_changes_nothing():
    """Workgroups taking the structures in the layer's order (most atoms
    first), in a reversed order or in batch order give the same outputs bit
    for bit, through both backward kernels; an order entry outside the batch
    sets the status bit and raises in the layer."""
    from torchproteinlibrary_amd._native_xform import ext
    from torchproteinlibrary_amd.FullAtomModel.Angles2Coords import convertStringList
    from torchproteinlibrary_amd.FullAtomModel.Angles2Coords.Angles2Coords import schedule
    lib = abi_a2c.load()
    rng = np.random.default_rng(47)
    lengths = [int(v) for v in rng.integers(1, 300, size=40)]
    seqs, ang = _batch(rng, lengths, 300, torch.float32)
    natoms = [O.num_atoms(s) for s in seqs]
    A = max(natoms)
    S = convertStringList(seqs).to(DEV)
    X = torch.tensor(ang, dtype=torch.float32, device=DEV)
    G = torch.randn(len(seqs), 3 * A, device=DEV)
    first = schedule(natoms)
    assert first is not None and sorted(first.tolist()) == list(range(len(seqs)))
    assert all(natoms[first[k]] >= natoms[first[k + 1]] for k in range(len(seqs) - 1))
    orders = [None, first.to(DEV), torch.flip(first, [0]).to(DEV)]
    outs = []
    for order in orders:
        c = torch.full((len(seqs), 3 * A), float("nan"), device=DEV)
        rn = torch.full((len(seqs), A, 4), 255, dtype=torch.uint8, device=DEV)
        an = torch.full_like(rn, 255)
        st = torch.zeros(1, dtype=torch.int32, device=DEV)
        ext.Angles2Coords_forward(S, X, c, rn, an, st, order)
        grads = []
        for form in (1, 2):
            ga = torch.full_like(X, float("nan"))
            optr = None if order is None else order.data_ptr()
            torch.cuda.synchronize()
            assert lib.tpl_a2c_backward_form(form, optr, 0, S.data_ptr(), S.shape[1], c.data_ptr(), G.data_ptr(), A,
                                             ga.data_ptr(), 300, st.data_ptr(), len(seqs), None) == 0
            grads.append(ga)
        torch.cuda.synchronize()
        assert int(st.item()) == 0
        outs.append((c, rn, an, *grads))
    for other in outs[1:]:
        for u, v in zip(outs[0], other):
            assert torch.equal(u, v)
    assert torch.isfinite(outs[0][0]).all() and torch.isfinite(outs[0][3]).all()
    bad = first.clone()
    bad[3] = len(seqs)
    st = torch.zeros(1, dtype=torch.int32, device=DEV)
    c = torch.empty(len(seqs), 3 * A, device=DEV)
    ext.Angles2Coords_forward(S, X, c, torch.empty(len(seqs), A, 4, dtype=torch.uint8, device=DEV),
                              torch.empty(len(seqs), A, 4, dtype=torch.uint8, device=DEV), st, bad.to(DEV))
    torch.cuda.synchronize()
    assert int(st.item()) & 8
