"""GPU parity: the HIP path against the FP64 oracle on the same seeded inputs.

Tolerances (stated per test):
* coordinates: the scan composes in FP64 and rounds once to float32, so the
  bound is a fraction of an ulp of the coordinate magnitude:
  |x_gpu - x_64| <= 1.2e-7 * max|x_64| + 1e-6 A  (<= 1e-4 A up to |x| = 800 A);
* gradients: per chain ||g_gpu - g_64||_inf / ||g_64||_inf <= 1e-5;
* exact zeros: padding coordinates, padding gradients, psi[len-1], omega[len-1].
"""
import math

import numpy as np
import pytest
import torch

from oracle import oracle as O
from torchproteinlibrary_amd import _tuning, synthetic
from torchproteinlibrary_amd.ReducedModel import Angles2Backbone

pytestmark = pytest.mark.gpu
DEV = "cuda"


def run(a, n, g, cfg=None, dtype=torch.float32, form="scan"):
    A = torch.tensor(a, device=DEV, dtype=dtype).contiguous().requires_grad_(True)
    N = torch.tensor(np.asarray(n, np.int32), device=DEV)
    if cfg is not None:
        _tuning.set_config(A, cfg)
    x = Angles2Backbone(form)(A, N)
    x.backward(torch.tensor(g, device=DEV, dtype=dtype))
    torch.cuda.synchronize()
    return x.detach().cpu().numpy().astype(np.float64), A.grad.cpu().numpy().astype(np.float64)


def oracle(a, n, g):
    a64, g64 = np.asarray(a, np.float64), np.asarray(g, np.float64)
    return O.forward(a64, n), O.backward_adjoint(a64, n, g64)


def coord_tol(x64):
    return 1.2e-7 * max(1.0, float(np.abs(x64).max())) + 1e-6


def grad_rel_err(got, want, n):
    errs = []
    for b in range(len(n)):
        if n[b] < 2:
            continue
        den = np.abs(want[b]).max()
        errs.append(np.abs(got[b] - want[b]).max() / den if den > 0 else np.abs(got[b]).max())
    return max(errs) if errs else 0.0


def assert_exact_zeros(x, ga, n, L):
    for b in range(len(n)):
        assert np.all(x[b, 9 * n[b]:] == 0.0), f"chain {b}: padding coords not zero"
        assert np.all(ga[b, :, n[b]:] == 0.0), f"chain {b}: padding grads not zero"
        if n[b] > 0:
            assert ga[b, 1, n[b] - 1] == 0.0 and ga[b, 2, n[b] - 1] == 0.0


RAGGED = [300, 0, 1, 2, 64, 65, 127, 128, 257]
# (forward, backward) residues per lane: chunks of 64 to 512 residues, so the
# ragged chains span 1 to 5 chunks (look-back both ways) or fit in one
CONFIGS = [(f, b) for f in (1, 2, 4, 8) for b in (1, 2, 4)]


@pytest.mark.parametrize("cfg", CONFIGS, ids=lambda c: "f%d_b%d" % c)
def test_ragged_batch_all_configs(cfg):
    a, n, g = synthetic.custom(RAGGED, 300, seed=41)
    print(f"ragged batch, config {cfg}", flush=True)
    x, ga = run(a, n, g, cfg)
    x64, g64 = oracle(a, n, g)
    assert np.abs(x - x64).max() <= coord_tol(x64)
    assert grad_rel_err(ga, g64, n) <= 1e-5
    assert_exact_zeros(x, ga, n, 300)


def test_many_chains_every_chunk_count():
    """B=512 chains of 0..1000 residues (every chunk count and the boundaries
    of chunks and lanes), on small and large chunks."""
    rng = np.random.default_rng(7)
    n = rng.integers(0, 1001, size=512)
    n[:10] = [1000, 999, 768, 767, 513, 512, 256, 255, 1, 0]
    a, n, g = synthetic.custom(n, 1000, seed=8)
    x64, g64 = oracle(a, n, g)
    for cfg in [(1, 1), (4, 2), (8, 4), (2, 4)]:
        print(f"B=512 L<=1000, config {cfg}", flush=True)
        x, ga = run(a, n, g, cfg)
        assert np.abs(x - x64).max() <= coord_tol(x64)
        assert grad_rel_err(ga, g64, n) <= 1e-5
        assert_exact_zeros(x, ga, n, 1000)


def test_stride_not_multiple_of_anything():
    a, n, g = synthetic.custom([37, 1, 36, 13, 0, 37, 20, 5, 9], 37, seed=5)
    for cfg in [(1, 1), (2, 1), (8, 2), (4, 4)]:
        x, ga = run(a, n, g, cfg)
        x64, g64 = oracle(a, n, g)
        assert np.abs(x - x64).max() <= coord_tol(x64)
        assert grad_rel_err(ga, g64, n) <= 1e-5
        assert_exact_zeros(x, ga, n, 37)


def test_checkpoints_are_the_frames_and_sync_stays_zero():
    """The forward's checkpoints are the FP64 frames of atoms 3kK (K = the
    backward chunk) as the oracle builds them; after forward and backward the
    look-back buffer is all zero again (it is shared by every later call, of
    any shape and config)."""
    from torchproteinlibrary_amd._native import ext
    a, n, g = synthetic.custom([300, 0, 129, 128, 64, 257, 200, 1], 300, seed=17)
    B, _, L = a.shape
    A, N, G = (torch.tensor(v, device=DEV) for v in (a, n, g))
    fr = O.frames(a.astype(np.float64), n)  # [B, 3L, 3, 4]: rows of the 3x4 affine part
    for cfg in [(1, 1), (4, 2), (2, 4)]:
        st_b, sy_b = ext.workspace_bytes(False, B, L, list(cfg))
        st = torch.zeros(st_b, dtype=torch.uint8, device=DEV)
        sy = torch.zeros(sy_b, dtype=torch.uint8, device=DEV)
        X = torch.empty(B, 9 * L, device=DEV)
        GA = torch.empty(B, 3, L, device=DEV)
        for _ in range(2):
            ext.Angles2BackboneGPU_forward(A, X, N, st, sy, None, list(cfg))
            ext.Angles2BackboneGPU_backward(GA, G, A, N, st, sy, None, list(cfg))
            torch.cuda.synchronize()
            assert int(sy.count_nonzero()) == 0, cfg
        K = 64 * cfg[1]
        nk = -(-L // K)
        ck = st[: B * nk * 96].view(torch.float64).reshape(B, nk, 12).cpu().numpy()
        for b in range(B):
            for k in range(1, nk):
                if k * K >= n[b]:
                    continue
                want = fr[b, 3 * k * K]  # rows: R[i][0..2] | t[i]
                got = ck[b, k].reshape(4, 3).T  # stored by columns: R e_x, R e_y, R e_z, t
                assert np.abs(got - want).max() <= 1e-9 * max(1.0, np.abs(want).max()), (cfg, b, k)


def test_mixed1000_full_size():
    """B=1024, lengths log-uniform in [30,1000]: tuned launch shape."""
    a, n, g = synthetic.workload("mixed1000")
    x, ga = run(a, n, g)
    x64, g64 = oracle(a, n, g)
    err = np.abs(x - x64).max()
    print(f"mixed1000: max coord err {err:.3e} A (max |x| {np.abs(x64).max():.1f}), "
          f"grad rel err {grad_rel_err(ga, g64, n):.3e}")
    assert err <= 1e-4
    assert grad_rel_err(ga, g64, n) <= 1e-5
    assert_exact_zeros(x, ga, n, 1000)


def test_uniform200_reference_cpu_config():
    """B=32, L=200 (the reference CPU benchmark's shape family,
    Benchmark/reducedmodel.py:15-23): float32 on the GPU and float64 I/O,
    both against the FP64 oracle."""
    a, n, g = synthetic.workload("uniform200")
    x, ga = run(a, n, g)
    x64, g64 = oracle(a, n, g)
    err = np.abs(x - x64).max()
    print(f"uniform200: max coord err {err:.3e} A, grad rel err {grad_rel_err(ga, g64, n):.3e}", flush=True)
    assert err <= coord_tol(x64)
    assert grad_rel_err(ga, g64, n) <= 1e-5
    assert_exact_zeros(x, ga, n, 200)
    x, ga = run(a.astype(np.float64), n, g.astype(np.float64), dtype=torch.float64)
    assert np.abs(x - x64).max() <= 1e-10
    assert grad_rel_err(ga, g64, n) <= 1e-10


def test_backbone_matches_angles2coords_polygly():
    """The reference's forward cross-check (UnitTests/ReducedModel/Angles2Backbone/test.py:32-70):
    poly-glycine, B=16, L_max=32, lengths random_(16, 32), angles randn;
    Angles2Backbone's N/CA/C against the backbone atoms Angles2Coords places
    (both on the GPU here); mean per-atom error below 5e-6 A (places=5)."""
    from torchproteinlibrary_amd.FullAtomModel import Angles2Coords
    torch.manual_seed(3)
    length, batch = 32, 16
    ang = torch.randn(batch, 3, length, dtype=torch.float, device=DEV)
    num_aa = torch.zeros(batch, dtype=torch.int, device=DEV).random_(length // 2, length)
    xb = Angles2Backbone()(ang, num_aa).cpu().reshape(batch, 3 * length, 3).numpy()
    seqs = ["G" * int(k) for k in num_aa.cpu()]
    fa = torch.zeros(batch, 8, length, dtype=torch.float, device=DEV)
    fa[:, 0:3, :] = ang
    coords, _, atom_names, num_atoms = Angles2Coords()(fa, seqs)
    amax = int(num_atoms.max())
    xf = coords.detach().cpu().reshape(batch, amax, 3).numpy()
    names = atom_names.cpu().numpy()
    err, cnt = 0.0, 0
    for j in range(batch):
        k = 0
        for i in range(int(num_atoms[j])):
            nm = bytes(names[j, i]).rstrip(b"\0")
            if nm in (b"N", b"CA", b"C"):
                err += np.linalg.norm(xb[j, k] - xf[j, i])
                k += 1
        assert k == 3 * int(num_aa[j])
        cnt += k
    print(f"poly-Gly Angles2Backbone vs Angles2Coords: mean error {err / cnt:.3e} A", flush=True)
    assert err / cnt < 5e-6


def test_uniform700_with_finite_differences():
    """B=256, L=700: oracle parity plus central FD (h=1e-3, FP64 oracle) on 64 angles."""
    a, n, g = synthetic.workload("uniform700")
    x, ga = run(a, n, g)
    x64, g64 = oracle(a, n, g)
    assert np.abs(x - x64).max() <= 1e-4
    assert grad_rel_err(ga, g64, n) <= 1e-5
    rng = np.random.default_rng(64)
    a64, g64in = a.astype(np.float64), g.astype(np.float64)
    h = 1e-3
    worst = 0.0
    for _ in range(64):
        b = int(rng.integers(0, 256))
        k = int(rng.integers(0, 3))
        j = int(rng.integers(0, n[b] - 1 if k else n[b]))
        ap, am = a64[b:b + 1].copy(), a64[b:b + 1].copy()
        ap[0, k, j] += h
        am[0, k, j] -= h
        fd = (O.forward(ap, n[b:b + 1])[0] - O.forward(am, n[b:b + 1])[0]) @ g64in[b] / (2 * h)
        scale = np.abs(g64[b]).max()
        worst = max(worst, abs(ga[b, k, j] - fd) / scale)
    print(f"uniform700 FD: max |g - fd| / ||g||_inf = {worst:.3e}")
    assert worst <= 1e-4


def test_long_chains_lookback():
    """B=64 chains of 4096-16384 residues: 16 to 256 chunks per chain, so the
    look-back walks several 64-chunk windows with (1, 1); report coordinate
    error growth with L."""
    a, n, g = synthetic.workload("long16k")
    for cfg in [(4, 2), (8, 4), (1, 1)]:
        x, ga = run(a, n, g, cfg)
        x64, g64 = oracle(a, n, g)
        per_chain = np.abs(x - x64).max(axis=1)
        order = np.argsort(n)
        print("long16k", cfg, "coord err vs L:",
              [(int(n[i]), float(per_chain[i])) for i in order[:: max(1, len(order) // 6)]])
        assert np.all(per_chain <= 1.2e-7 * np.abs(x64).max(axis=1) + 1e-6)
        assert grad_rel_err(ga, g64, n) <= 1e-5
        assert_exact_zeros(x, ga, n, 16384)


def test_float64_io():
    a, n, g = synthetic.custom([50, 3, 0, 49, 17], 50, seed=8)
    a = a.astype(np.float64) + 0.01  # non-f32-representable values
    for i, k in enumerate(n):
        a[i, :, k:] = 0
    for cfg in [(1, 1), (4, 2)]:
        x, ga = run(a, n, g.astype(np.float64), cfg, dtype=torch.float64)
        x64, g64 = oracle(a, n, g)
        assert np.abs(x - x64).max() <= 1e-10
        assert grad_rel_err(ga, g64, n) <= 1e-10
        assert_exact_zeros(x, ga, n, 50)


def test_plain_form():
    """The baseline kernels (FP32 serial forward, O(L^2) per-pair backward)."""
    a, n, g = synthetic.custom([200, 120, 1, 0, 77, 2, 199, 64], 200, seed=12)
    x, ga = run(a, n, g, form="plain")
    x64, g64 = oracle(a, n, g)
    assert np.abs(x - x64).max() <= 5e-3  # FP32 serial recurrence
    assert grad_rel_err(ga, g64, n) <= 1e-3
    assert_exact_zeros(x, ga, n, 200)
    x, ga = run(a, n, g.astype(np.float64), dtype=torch.float64, form="plain")
    assert np.abs(x - x64).max() <= 1e-10
    assert grad_rel_err(ga, g64, n) <= 1e-9


def test_reference_gradcheck_fp32():
    """UnitTests/ReducedModel/Angles2Backbone/test.py:72-82: FP32 gradcheck,
    B=8, L=8, lengths in [4,8), eps=1e-3, atol=1e-2, rtol=0.01."""
    torch.manual_seed(0)
    A = torch.randn(8, 3, 8, dtype=torch.float, device=DEV).requires_grad_()
    N = torch.zeros(8, dtype=torch.int, device=DEV).random_(4, 8)
    assert torch.autograd.gradcheck(Angles2Backbone(), (A, N), eps=1e-3, atol=1e-2, rtol=0.01)


def test_gradcheck_fp64_tight():
    torch.manual_seed(1)
    A = torch.randn(4, 3, 7, dtype=torch.float64, device=DEV).requires_grad_()
    N = torch.tensor([7, 5, 1, 3], dtype=torch.int, device=DEV)
    for cfg in [(1, 1), (4, 2)]:
        _tuning.set_config(A, cfg)
        assert torch.autograd.gradcheck(Angles2Backbone(), (A, N), eps=1e-6, atol=1e-8, rtol=1e-7)


def test_errors():
    m = Angles2Backbone()
    a, n, _ = synthetic.custom([5, 4], 5, seed=1)
    A = torch.tensor(a, device=DEV)
    N = torch.tensor(n, device=DEV)
    bad = A.clone()
    bad[0, 1, 2] = float("nan")
    with pytest.raises(Exception, match="forward Nan"):
        m(bad, N)
    with pytest.raises(ValueError):
        m(A, torch.tensor([6, 1], dtype=torch.int32, device=DEV))
    with pytest.raises(RuntimeError):
        m(A.transpose(1, 2).contiguous().transpose(1, 2), N)
    with pytest.raises(RuntimeError):
        m(A, N.long())
    with pytest.raises(Exception, match="Incorrect input size"):
        m(A[0], N)
    with pytest.raises(RuntimeError):
        m(A, N[:1])


@pytest.mark.skip(reason="torch.cuda.graph capture of the fwd+bwd step ended in a host "
                         "segfault inside CUDAGraph.capture_end on the MI355X host; the cause "
                         "is not shown yet, so the capture path stays switched off")
def test_graph_capture_replays_match_eager():
    a, n, g = synthetic.workload("fit500", batch=64)
    A = torch.tensor(a, device=DEV).requires_grad_(True)
    N = torch.tensor(n, device=DEV)
    G = torch.tensor(g, device=DEV)
    m = Angles2Backbone()
    x_eager = m(A, N)
    x_eager.backward(G)
    want_x, want_g = x_eager.detach().clone(), A.grad.detach().clone()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        for _ in range(2):
            A.grad = None
            m(A, N).backward(G)
    torch.cuda.current_stream().wait_stream(s)
    graph = torch.cuda.CUDAGraph()
    A.grad = None
    with torch.cuda.graph(graph):
        x_static = m(A, N)
        x_static.backward(G)
    for rep in range(4):
        junk = torch.randn(1000, device=DEV).sum()  # eager work between replays
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(x_static, want_x), f"replay {rep}: coords differ"
        assert torch.equal(A.grad, want_g), f"replay {rep}: grads differ"
        del junk


@pytest.mark.parametrize("cfg", [(4, 2), (1, 1)])
def test_hip_graph_step_replays_match_eager(cfg):
    """A forward + backward per buffer set captured into one HIP graph on a
    private stream (the extension's bench_graph) and replayed: every replay
    writes exactly the eager outputs into both sets, also with eager work in
    between, and with chains of several chunks ((1, 1): up to 8 per chain,
    look-back words reused by every replay)."""
    from torchproteinlibrary_amd._native import ext
    a, n, g = synthetic.workload("fit500", batch=64)
    B, _, L = a.shape
    A, N, G = (torch.tensor(v, device=DEV) for v in (a, n, g))
    st, sy = ext.workspace_bytes(False, B, L, list(cfg))

    def buffers():
        return [A, N, G, torch.zeros(B, 9 * L, device=DEV), torch.zeros(B, 3, L, device=DEV),
                torch.zeros(max(st, 1), dtype=torch.uint8, device=DEV),
                torch.zeros(max(sy, 1), dtype=torch.uint8, device=DEV)]
    e = buffers()
    ext.Angles2BackboneGPU_forward(A, e[3], N, e[5], e[6], None, list(cfg))
    ext.Angles2BackboneGPU_backward(e[4], G, A, N, e[5], e[6], None, list(cfg))
    gs, gs2 = buffers(), buffers()  # two buffer sets: one graph holds a step for each
    for rep in range(3):
        junk = torch.randn(1000, device=DEV).sum()  # eager work between replays
        for b_ in (gs, gs2):
            b_[3].zero_()
            b_[4].zero_()
        _, _, done = ext.bench_graph([gs, gs2], list(cfg), 4, 2)
        torch.cuda.synchronize()
        assert done == 4
        for k, b_ in enumerate((gs, gs2)):
            assert torch.equal(b_[3], e[3]), f"replay round {rep}, set {k}: coords differ"
            assert torch.equal(b_[4], e[4]), f"replay round {rep}, set {k}: grads differ"
            assert int(b_[6].count_nonzero()) == 0
        del junk
    x64, g64 = oracle(a, n, g)
    assert np.abs(gs[3].cpu().numpy() - x64).max() <= coord_tol(x64)


@pytest.mark.parametrize("cfg", [(4, 2), (1, 1)])
def test_step_graph_replays_match_eager(cfg):
    """Angles2BackboneStep: one forward+backward captured once into a HIP
    graph through the library's Python API and replayed.  The loop updates
    the angles and the incoming gradient in place between replays, with other
    eager work in between; every replay must write exactly what the layer
    computes eagerly for the same inputs and config, and leave the look-back
    words zero."""
    from torchproteinlibrary_amd.ReducedModel.Angles2Backbone import Angles2BackboneStep
    a, n, g = synthetic.workload("fit500", batch=64)
    A = torch.tensor(a, device=DEV)
    N = torch.tensor(np.asarray(n, np.int32), device=DEV)
    G = torch.tensor(g, device=DEV)
    step = Angles2BackboneStep(A, N, config=cfg)
    for rep in range(4):
        junk = torch.randn(1000, device=DEV).sum()  # eager work between replays
        with torch.no_grad():
            A.add_(0.01)
        step.grad_coords.copy_(G * (1.0 + rep))
        step.replay()
        Ae = A.clone().requires_grad_(True)
        _tuning.set_config(Ae, cfg)
        x = Angles2Backbone()(Ae, N)
        x.backward(G * (1.0 + rep))
        torch.cuda.synchronize()
        assert torch.equal(step.coords, x.detach()), f"replay {rep}: coords differ"
        assert torch.equal(step.grad_angles, Ae.grad), f"replay {rep}: grads differ"
        assert int(step._sync.count_nonzero()) == 0, f"replay {rep}: look-back words left set"
        del junk
    x64, g64 = oracle(A.cpu().numpy(), n, (G * 4.0).cpu().numpy())
    assert np.abs(step.coords.cpu().numpy().astype(np.float64) - x64).max() <= coord_tol(x64)
    assert grad_rel_err(step.grad_angles.cpu().numpy().astype(np.float64), g64, n) <= 1e-5


def test_tuner_picks_a_candidate():
    a, n, g = synthetic.workload("fit500", batch=128)
    A = torch.tensor(a, device=DEV)
    N = torch.tensor(n, device=DEV)
    cfg = _tuning.tune_now(A, N)
    assert cfg in [tuple(c) for c in _tuning.candidates(500, torch.float32)]
    t = _tuning.timings(A, N)
    assert t and all(v > 0 for v in t.values())
    print("fit500 B=128 tuner timings (ms):", t)


def test_fit_loop_tracks_fp64():
    """The fit500 usage pattern (docs/Examples/fitBackbone.py:44-63): Adam on the
    angles toward FP64 target coordinates.  30 steps in float32 and in float64
    from the same start; the loss falls and float32 tracks float64."""
    a, n, _ = synthetic.workload("fit500", batch=64)
    rng = np.random.default_rng(7)
    start = a.astype(np.float64) + rng.normal(0.0, 0.3, a.shape)
    start *= (np.arange(a.shape[2])[None, None, :] < n[:, None, None])
    N = torch.tensor(n, device=DEV)
    m = Angles2Backbone()
    with torch.no_grad():
        target = m(torch.tensor(a, dtype=torch.float64, device=DEV), N)
    losses = {}
    for dt in (torch.float32, torch.float64):
        ang = torch.tensor(start, dtype=dt, device=DEV).requires_grad_(True)
        opt = torch.optim.Adam([ang], lr=1e-2)
        hist = []
        for _ in range(30):
            opt.zero_grad()
            loss = ((m(ang, N) - target.to(dt)) ** 2).sum()
            loss.backward()
            opt.step()
            hist.append(float(loss))
        losses[dt] = hist
    l32, l64 = losses[torch.float32], losses[torch.float64]
    print(f"fit loop: loss {l64[0]:.4e} -> {l64[-1]:.4e} (f64), {l32[-1]:.4e} (f32)")
    assert l64[-1] < 0.9 * l64[0]
    assert abs(l32[-1] - l64[-1]) / l64[-1] < 1e-3


def test_fit500_full_loop_f32_tracks_f64():
    """fit500 at full size (docs/Examples/fitBackbone.py:44-63): B=512,
    lengths uniform in [50,500], targets from the FP64 forward of the seeded
    angles, start = target + N(0, 0.3) rad, loss = masked sum of squared
    coordinate differences, 1000 Adam steps (lr 1e-2) in float32 and in
    float64 from the same start.  Both losses fall by more than 100x; the
    final losses of the two precisions agree within 15%.  The loop is not
    convex, so rounding differences of 1e-7 relative per step grow over 1000
    steps into different (equally good) trajectories: the bound is on the
    outcome, not on the path."""
    a, n, _ = synthetic.workload("fit500")
    rng = np.random.default_rng(55)
    start = a.astype(np.float64) + rng.normal(0.0, 0.3, a.shape)
    start *= (np.arange(a.shape[2])[None, None, :] < n[:, None, None])
    N = torch.tensor(n, device=DEV)
    m = Angles2Backbone()
    with torch.no_grad():
        target = m(torch.tensor(a, dtype=torch.float64, device=DEV), N)
    out = {}
    for dt in (torch.float32, torch.float64):
        ang = torch.tensor(start, dtype=dt, device=DEV).requires_grad_(True)
        opt = torch.optim.Adam([ang], lr=1e-2)
        tgt = target.to(dt)
        first = None
        for i in range(1000):
            opt.zero_grad(set_to_none=True)
            loss = ((m(ang, N) - tgt) ** 2).sum()
            loss.backward()
            opt.step()
            if i == 0:
                first = float(loss)
        with torch.no_grad():
            out[dt] = (first, float(((m(ang, N) - tgt) ** 2).sum()))
    (f32_0, f32_1), (f64_0, f64_1) = out[torch.float32], out[torch.float64]
    rel = abs(f32_1 - f64_1) / f64_1
    print(f"fit500 B=512 x 1000 Adam steps: loss {f64_0:.4e} -> {f64_1:.4e} (f64), "
          f"{f32_0:.4e} -> {f32_1:.4e} (f32); final rel diff {rel:.3e}", flush=True)
    assert abs(f32_0 - f64_0) / f64_0 < 1e-5
    assert f64_1 < 1e-2 * f64_0 and f32_1 < 1e-2 * f32_0
    assert rel < 0.15
