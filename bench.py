"""Angles2Backbone fwd+bwd benchmark on one MI355X.

A step is one forward (angles -> coords) plus one backward (grad coords ->
grad angles) of ``Angles2Backbone`` over one batch of the mixed1000 workload:
B=1024 chains, lengths log-uniform in [30,1000], padded to 1000, float32
angles/gradients, inputs resident in HBM.  Steps are launched back to back
from a native loop (the extension's ``bench_steps``: forward launch, backward
launch, next buffer set) and timed with HIP events on the launch stream.  The
loop rotates over enough buffer sets to exceed the 256 MB Infinity Cache, so
every step streams from HBM.

Prints one JSON line.  ``value`` is residues/s (sum of chain lengths per
second).  The line also carries:
* ``roofline``: the slower of the two kernels, timed alone with HIP events on
  the stream it runs on; its compulsory bytes per launch (``alg_bytes``: the
  dense-tensor API's reads and writes, nothing else) / average duration
  against the MI355X 8 TB/s HBM peak, and ``traffic``, the HBM bytes per launch
  that two ``rocprofv3 --pmc`` runs of this script's probe mode count in the
  same run (FETCH_SIZE, then WRITE_SIZE, each scaled by what it counts for a
  256 MiB device copy in the same run);
* ``graph``: the same step replayed from a HIP graph that holds one forward +
  backward per buffer set;
* ``plain``: the same step with the plain kernels (serial forward, O(L^2)
  backward) and the speedup of the scan path over it;
* ``cpu_baseline``: the reference's own CPU code (built from its sources into
  oracle/_ref) on its CPU configuration, B=32 chains of 200 residues, FP64,
  one thread, one warm-up then the median of 5; with the host's CPU model;
* ``cpu_path``: this repository's host path (libtpl_a2b_cpu.so) on the same
  configuration, all host threads;
* ``fit500``: the optimisation loop of docs/Examples/fitBackbone.py at B=512
  (lengths 50-500), 1000 Adam steps in float32 and in float64: steps/s and
  the final losses;
* ``rmsd``: the Coords2RMSD loss (forward + backward) on the workload's
  backbone atoms against a moved, noisy copy, with its own roofline;
* ``xform``: the coordinate transforms (translate, rotate, centre and its
  backward) on the same atoms, per-kernel times and bandwidth;
* ``a2c``: the full-atom builder Angles2Coords (forward + backward) on the
  reference's own benchmark shape (32 random sequences of 350 residues) and
  on a batch of 1024 random sequences with the workload's lengths, with the
  reference's CPU code timed beside it;
* ``pdb``: the PDB reader (host code) on synthetic PDB files, atoms/s, with
  the Python restatement of the reference's loader timed beside it.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "fwd+bwd residues/s, mixed-L B=1024; batched adjoint vs plain; % HBM peak"
HBM_PEAK_GBS = 8000.0
MALL_BYTES = 256 << 20


def alg_bytes(n, L, io_bytes=4):
    """Compulsory HBM bytes of one forward and one backward launch: what the
    layer's dense-tensor API must read and write, nothing else.

    forward:  read angles (3 per residue), write coords (9 per slot, padding
              included), read lengths;
    backward: read grad coords (9 per residue), re-read angles (3 per residue,
              the minimal forward->backward state), write grad angles (3 per
              slot, padding included), read lengths.
    The checkpoint frames the forward saves (96 bytes per backward chunk) are
    not credited."""
    import numpy as np
    B = len(n)
    S = int(np.sum(n))
    fwd = io_bytes * (3 * S + 9 * B * L) + 4 * B
    bwd = io_bytes * (9 * S + 3 * S + 3 * B * L) + 4 * B
    return fwd, bwd


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=2000)
    p.add_argument("--warmup", type=int, default=200)
    p.add_argument("--workload", default="mixed1000")
    p.add_argument("--sets", type=int, default=0, help="buffer sets (0: enough to exceed the MALL)")
    p.add_argument("--plain-steps", type=int, default=20)
    p.add_argument("--kernel-reps", type=int, default=200)
    p.add_argument("--no-plain", action="store_true")
    p.add_argument("--no-rmsd", action="store_true", help="skip the Coords2RMSD line")
    p.add_argument("--no-a2c", action="store_true", help="skip the Angles2Coords line")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-seconds", type=float, default=15.0,
                   help="time budget of the Angles2Coords CPU baseline")
    p.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 counter runs")
    p.add_argument("--no-fit", action="store_true", help="skip the fit500 line")
    p.add_argument("--probe", action="store_true",
                   help="(internal) launch each kernel --steps times and 5 device copies, for rocprofv3")
    p.add_argument("--config", default="", help="pin config 'fwd_res_per_lane,bwd_res_per_lane'")
    p.add_argument("--fit", action="store_true",
                   help="run the fit500 Adam optimisation loop instead (own JSON line)")
    p.add_argument("--fit-steps", type=int, default=1000)
    return p.parse_args()


def time_kernels(sets, cfg, reps):
    """Duration of the forward and of the backward launch, HIP events on
    torch's current stream (the stream the extension launches on):
    avg_ms = one event pair around `reps` back-to-back launches of that kernel
    alone (rotating buffer sets), divided by reps; isolated_median_ms = the
    median over launches each bracketed by its own event pair, which adds the
    event round trip to every launch (an empty kernel measures about 6 us
    that way against 3 us back to back on MI355X)."""
    import torch
    from torchproteinlibrary_amd._native import ext
    cur = torch.cuda.current_stream()
    res = {}
    for name in ("forward", "backward"):
        def launch(i):
            A, N, G, X, GA, st, sy = sets[i % len(sets)]
            if name == "forward":
                ext.Angles2BackboneGPU_forward(A, X, N, st, sy, None, list(cfg))
            else:
                ext.Angles2BackboneGPU_backward(GA, G, A, N, st, sy, None, list(cfg))
        for i in range(10):
            launch(i)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(cur)
        for i in range(reps):
            launch(i)
        e1.record(cur)
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
               for _ in range(reps)]
        for i in range(reps):
            evs[i][0].record(cur)
            launch(i)
            evs[i][1].record(cur)
        torch.cuda.synchronize()
        ms = sorted(a.elapsed_time(b) for a, b in evs)
        res[name] = {"avg_ms": e0.elapsed_time(e1) / reps, "isolated_median_ms": ms[len(ms) // 2]}
    return res


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or "unknown"


def cpu_baseline():
    """The reference's CPU path on its CPU configuration (uniform200: B=32,
    L=200, FP64): its own kernels compiled from its sources (oracle/_ref) if
    present, else this repository's C restatement with the same
    derivative-tensor behaviour.  One thread; buffers zero-filled inside the
    timed call as the reference's Function does; one warm-up, then the median
    of 5 forward+backward calls."""
    import numpy as np
    from oracle import oracle as O
    from torchproteinlibrary_amd import synthetic
    a, n, g = synthetic.workload("uniform200")
    a, g = a.astype(np.float64), g.astype(np.float64)
    kind = "reference" if O.ref_available() else "port"

    def once():
        if kind == "reference":
            O.ref_forward_backward(a, n, g)
        else:
            O.forward(a, n)
            O.backward_dense(a, n, g)
    once()
    ts = []
    for _ in range(5):
        t0 = time.perf_counter()
        once()
        ts.append(time.perf_counter() - t0)
    t = sorted(ts)[2]
    return {
        "value": int(n.sum()) / t,
        "unit": "residues/s",
        "cores": 1,
        "kind": kind,
        "sample": f"uniform200 (B=32, L=200, 6400 residues), FP64 forward+backward with the "
                  f"O(L^2) derivative tensor zero-filled and reduced as the reference does; "
                  f"median of 5 after a warm-up: {t:.3f} s per call",
        "cpu_model": cpu_model(),
        "nproc": os.cpu_count(),
    }


def cpu_path_line():
    """This repository's host path (libtpl_a2b_cpu.so through the layer, CPU
    float64 tensors) on uniform200, forward+backward, all host threads (at
    most 16): median of 20 after a warm-up."""
    import numpy as np
    import torch
    from torchproteinlibrary_amd import synthetic
    from torchproteinlibrary_amd.ReducedModel import Angles2Backbone
    a, n, g = synthetic.workload("uniform200")
    threads = min(16, os.cpu_count() or 1)
    old = torch.get_num_threads()
    torch.set_num_threads(threads)
    try:
        A = torch.tensor(a, dtype=torch.float64, requires_grad=True)
        N = torch.tensor(n)
        G = torch.tensor(g, dtype=torch.float64)
        m = Angles2Backbone()
        m(A, N).backward(G)
        ts = []
        for _ in range(20):
            t0 = time.perf_counter()
            m(A, N).backward(G)
            ts.append(time.perf_counter() - t0)
    finally:
        torch.set_num_threads(old)
    t = sorted(ts)[len(ts) // 2]
    return {"value": int(n.sum()) / t, "unit": "residues/s", "threads": threads,
            "ms_per_step": 1e3 * t,
            "sample": "uniform200 (B=32, L=200), FP64, forward + O(L) adjoint backward"}


def probe(args):
    """Counter-probe mode (run under rocprofv3 by traffic()): each kernel alone
    --steps times over rotating buffer sets, then 5 device copies of 256 MiB
    for the counters' calibration."""
    import numpy as np
    import torch
    from torchproteinlibrary_amd import synthetic
    from torchproteinlibrary_amd._native import ext
    dev = torch.device("cuda:0")
    a, n, g = synthetic.workload(args.workload)
    B, _, L = a.shape
    cfg = [int(v) for v in args.config.split(",")]
    st_b, sy_b = ext.workspace_bytes(False, B, L, cfg)
    sets = []
    for _ in range(6):
        sets.append((torch.tensor(a, device=dev), torch.tensor(n, device=dev), torch.tensor(g, device=dev),
                     torch.empty(B, 9 * L, device=dev), torch.empty(B, 3, L, device=dev),
                     torch.empty(st_b, dtype=torch.uint8, device=dev)))
    sy = torch.zeros(sy_b, dtype=torch.uint8, device=dev)
    for A, N, G, X, GA, st in sets:
        ext.Angles2BackboneGPU_forward(A, X, N, st, sy, None, cfg)
    for i in range(args.steps):
        A, N, G, X, GA, st = sets[i % 6]
        ext.Angles2BackboneGPU_forward(A, X, N, st, sy, None, cfg)
    for i in range(args.steps):
        A, N, G, X, GA, st = sets[i % 6]
        ext.Angles2BackboneGPU_backward(GA, G, A, N, st, sy, None, cfg)
    # Calibration, the last 5 dispatches: an elementwise kernel that reads
    # 256 MiB and writes 256 MiB (float32, 64 Mi elements).
    src = torch.ones(64 << 20, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(5):
        torch.add(src, 1.0, out=dst)
    torch.cuda.synchronize()
    print("probe done", flush=True)


def _kernel_side(name):
    """"forward" or "backward" for the Angles2Backbone kernels of either form
    (chunk kernels, or one workgroup per chain), else None."""
    if "a2b_fwd_chunk" in name or "a2b_fwd_chain" in name:
        return "forward"
    if "a2b_bwd_chunk" in name or "a2b_bwd_chain" in name:
        return "backward"
    return None


def _pmc_medians(outdir, counter, last=5):
    """({kernel name: median counter value over its dispatches}, median over
    the last `last` dispatches) from a rocprofv3 --output-format csv run."""
    import csv
    import glob
    import statistics
    rows = []
    for f in glob.glob(os.path.join(outdir, "**", "*counter_collection.csv"), recursive=True):
        with open(f, newline="") as fh:
            for row in csv.DictReader(fh):
                if row.get("Counter_Name") == counter:
                    rows.append((int(row.get("Dispatch_Id", 0)), row.get("Kernel_Name", ""),
                                 float(row["Counter_Value"])))
    rows.sort()
    vals = {}
    for _, k, v in rows:
        vals.setdefault(k, []).append(v)
    tail = [v for _, _, v in rows[-last:]]
    return ({k: statistics.median(v) for k, v in vals.items()},
            statistics.median(tail) if len(tail) == last else None)


def _trace_avg_ms(outdir):
    """{"forward": ms, "backward": ms}: mean end - start of the Angles2Backbone
    kernels' dispatches in a rocprofv3 --kernel-trace csv run."""
    import csv
    import glob
    durs = {"forward": [], "backward": []}
    for f in glob.glob(os.path.join(outdir, "**", "*kernel_trace.csv"), recursive=True):
        with open(f, newline="") as fh:
            for row in csv.DictReader(fh):
                name = row.get("Kernel_Name", "")
                k = _kernel_side(name)
                if k:
                    durs[k].append((int(row["End_Timestamp"]) - int(row["Start_Timestamp"])) * 1e-6)
    if not (durs["forward"] and durs["backward"]):
        return None
    return {k: sum(v) / len(v) for k, v in durs.items()}


def traffic(args, cfg):
    """HBM bytes per launch of the forward and backward kernels: two
    rocprofv3 --pmc runs of the probe mode (FETCH_SIZE, then WRITE_SIZE, in
    KiB per dispatch), each scaled by what it counts for the probe's last
    dispatches, an elementwise kernel that reads and writes 256 MiB.  None if
    rocprofv3 is not there or a run fails."""
    import shutil
    import subprocess
    import tempfile
    prof = shutil.which("rocprofv3")
    if prof is None:
        return None
    out = {}
    with tempfile.TemporaryDirectory() as d:
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            od = os.path.join(d, counter)
            cmd = [prof, "--pmc", counter, "--output-format", "csv", "-d", od, "-o", "p", "--",
                   sys.executable, os.path.abspath(__file__), "--probe", "--steps", "24",
                   "--workload", args.workload, "--config", ",".join(str(c) for c in cfg)]
            print(f"bench: counter run {counter}", file=sys.stderr, flush=True)
            try:
                r = subprocess.run(cmd, capture_output=True, text=True, timeout=120,
                                   env=dict(os.environ, TMPDIR=d))
            except subprocess.TimeoutExpired:
                return None
            if "probe done" not in r.stdout:
                return None
            med, calib = _pmc_medians(od, counter)
            fwd = [v for k, v in med.items() if _kernel_side(k) == "forward"]
            bwd = [v for k, v in med.items() if _kernel_side(k) == "backward"]
            if not (fwd and bwd and calib):
                return None
            scale = 262144.0 / calib  # the calibration kernel reads and writes 256 MiB
            out[counter] = {"forward_kib": fwd[0], "backward_kib": bwd[0], "copy_kib": calib, "scale": scale}
        # Kernel durations of the same probe run, from the kernel trace.
        od = os.path.join(d, "trace")
        cmd = [prof, "--kernel-trace", "--output-format", "csv", "-d", od, "-o", "p", "--",
               sys.executable, os.path.abspath(__file__), "--probe", "--steps", "24",
               "--workload", args.workload, "--config", ",".join(str(c) for c in cfg)]
        print("bench: kernel trace run", file=sys.stderr, flush=True)
        durations = None
        try:
            r = subprocess.run(cmd, capture_output=True, text=True, timeout=120,
                               env=dict(os.environ, TMPDIR=d))
            if "probe done" in r.stdout:
                durations = _trace_avg_ms(od)
        except subprocess.TimeoutExpired:
            pass
    res = {"fetch_scale": out["FETCH_SIZE"]["scale"], "write_scale": out["WRITE_SIZE"]["scale"]}
    if durations:
        res["rocprof_avg_ms"] = durations
    for k in ("forward", "backward"):
        res[k] = 1024.0 * (out["FETCH_SIZE"][k + "_kib"] * out["FETCH_SIZE"]["scale"] +
                           out["WRITE_SIZE"][k + "_kib"] * out["WRITE_SIZE"]["scale"])
        res[k + "_fetch_kib"] = out["FETCH_SIZE"][k + "_kib"]
        res[k + "_write_kib"] = out["WRITE_SIZE"][k + "_kib"]
    return res


def fit_loop(steps: int):
    """The fit500 optimisation loop: B=512 chains, lengths uniform in [50,500].
    Targets are the FP64 coordinates of the seeded angles.  The start angles are
    target + N(0, 0.3) rad on valid slots.  Loss = sum of squared coordinate
    differences (padding is zero in both, so the sum is masked).  Adam,
    lr 1e-2.  The same loop runs in float32 and in float64 from the same
    start; the line reports end-to-end steps/s of each and the final losses.
    Eager steps, with the per-call status read switched off."""
    import numpy as np
    import torch

    from torchproteinlibrary_amd import synthetic
    from torchproteinlibrary_amd.ReducedModel import Angles2Backbone

    os.environ["TPL_A2B_CHECKS"] = "0"
    dev = torch.device("cuda:0")
    a, n, _ = synthetic.workload("fit500")
    rng = np.random.default_rng(55)
    start = a.astype(np.float64) + rng.normal(0.0, 0.3, a.shape)
    start *= (np.arange(a.shape[2])[None, None, :] < n[:, None, None])
    N = torch.tensor(n, device=dev)
    m = Angles2Backbone()
    with torch.no_grad():
        target64 = m(torch.tensor(a, dtype=torch.float64, device=dev), N)
    res = {}
    for name, dt in (("f32", torch.float32), ("f64", torch.float64)):
        target = target64.to(dt)
        ang = torch.tensor(start, dtype=dt, device=dev).requires_grad_(True)
        opt = torch.optim.Adam([ang], lr=1e-2)
        loss0 = None

        def step():
            opt.zero_grad(set_to_none=True)
            loss = ((m(ang, N) - target) ** 2).sum()
            loss.backward()
            opt.step()
            return loss

        for _ in range(3):  # warm-up (tuning, allocator)
            step()
        with torch.no_grad():
            ang.copy_(torch.tensor(start, dtype=dt, device=dev))
        opt = torch.optim.Adam([ang], lr=1e-2)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(steps):
            loss = step()
            if i == 0:
                loss0 = loss.detach()
        torch.cuda.synchronize()
        dt_s = time.perf_counter() - t0
        with torch.no_grad():
            final = float(((m(ang, N) - target) ** 2).sum())
        res[name] = {"steps_per_s": steps / dt_s, "loss_start": float(loss0), "loss_final": final}
    return res


def rmsd_bench(ksets, n, L, steps, warmup, reps, workload):
    """Coords2RMSD fwd+bwd on each buffer set's forward coordinates (the
    backbone atoms N, CA, C: num_atoms = 3 * length, stride 3 * L) against a
    rotated, shifted copy with 0.3 A noise.  Buffers rotate like the main
    loop's; times are HIP events around native launch loops."""
    import torch
    from torchproteinlibrary_amd._native_rmsd import ext as rext

    dev = ksets[0][3].device
    B, A = len(n), 3 * L
    natoms = torch.tensor(3 * n, dtype=torch.int32, device=dev)
    atoms = int(3 * n.sum())
    gen = torch.Generator(device=dev).manual_seed(5)
    valid = (torch.arange(A, device=dev)[None, :] < natoms[:, None]).unsqueeze(-1)
    rsets = []
    for k in ksets:
        x = k[3]
        q = torch.randn(4, generator=gen, device=dev, dtype=torch.float64)
        q = q / q.norm()
        w, i, j, kk = q.tolist()
        Q = torch.tensor([[w * w + i * i - j * j - kk * kk, 2 * (i * j - w * kk), 2 * (i * kk + w * j)],
                          [2 * (i * j + w * kk), w * w - i * i + j * j - kk * kk, 2 * (j * kk - w * i)],
                          [2 * (i * kk - w * j), 2 * (j * kk + w * i), w * w - i * i - j * j + kk * kk]],
                         dtype=x.dtype, device=dev)
        y = x.view(B, A, 3) @ Q.T + 5.0
        y = y + 0.3 * torch.randn(y.shape, generator=gen, device=dev, dtype=x.dtype)
        y = torch.where(valid, y, torch.zeros_like(y)).reshape(B, 3 * A).contiguous()
        rsets.append([x, y, natoms, torch.empty(B, dtype=x.dtype, device=dev),
                      torch.empty(B, 3, 3, dtype=x.dtype, device=dev),
                      torch.empty(max(rext.state_bytes(B), 1), dtype=torch.uint8, device=dev),
                      torch.randn(B, generator=gen, device=dev, dtype=x.dtype),
                      torch.empty(B, 3 * A, dtype=x.dtype, device=dev)])
    torch.cuda.synchronize()
    step_ms, f_ms, b_ms = rext.bench(rsets, steps, warmup, reps)
    # Compulsory bytes (FP32): both inputs' valid atoms in each direction; the
    # backward writes every gradient slot (zeros past num_atoms); per structure
    # num_atoms, rmsd, UT and the 16-double state.
    fwd_b = 24 * atoms + B * (4 + 4 + 36 + 128)
    bwd_b = 24 * atoms + 12 * B * A + B * (4 + 4 + 128)
    kern = {"forward": {"avg_ms": f_ms, "alg_bytes": fwd_b, "GBs": fwd_b / (f_ms * 1e-3) / 1e9},
            "backward": {"avg_ms": b_ms, "alg_bytes": bwd_b, "GBs": bwd_b / (b_ms * 1e-3) / 1e9}}
    dom = max(kern, key=lambda k: kern[k]["avg_ms"])
    ms_step = step_ms / steps
    traffic = None
    tf = os.path.join(ROOT, "profiles", f"traffic_rmsd_{workload}.json")
    if os.path.exists(tf):
        with open(tf) as f:
            traffic = json.load(f).get(dom)
    return {
        "workload": f"Coords2RMSD fwd+bwd, backbone atoms of the workload: B={B}, "
                    f"atoms_stride={A}, num_atoms=3*length, f32",
        "atoms": atoms,
        "value": atoms / (ms_step * 1e-3),
        "unit": "atoms/s",
        "ms_per_step": ms_step,
        "steps": steps,
        "buffer_sets": len(rsets),
        "kernels": kern,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": kern[dom]["GBs"],
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": kern[dom]["GBs"] / HBM_PEAK_GBS, "traffic": traffic,
                     "alg_bytes_per_launch": kern[dom]["alg_bytes"]},
    }


def xform_bench(ksets, n, L, reps):
    """The coordinate transforms on each buffer set's forward coordinates as
    backbone structures (num_atoms = 3 * length, stride 3 * L), FP32: average
    launch time of translate, rotate, centre and centre-backward (HIP events
    around native launch loops, buffers rotating)."""
    import torch
    from torchproteinlibrary_amd._native_xform import ext as xext

    dev = ksets[0][3].device
    B, A = len(n), 3 * L
    natoms = torch.tensor(3 * n, dtype=torch.int32, device=dev)
    atoms = int(3 * n.sum())
    from torchproteinlibrary_amd.FullAtomModel.CoordsTransform import getRandomRotation
    Rm = getRandomRotation(B).to(device=dev, dtype=torch.float32)
    T = torch.randn(B, 3, device=dev)
    xsets = [[k[3], torch.empty(B, 3 * A, device=dev), T, Rm, torch.empty(B, 3, device=dev), natoms]
             for k in ksets]
    torch.cuda.synchronize()
    tr, ro, ce, cb = xext.bench(xsets, reps)
    # compulsory bytes: read the valid atoms, write every slot (zeros past
    # num_atoms); per structure num_atoms and T/R/centre
    rw = 12 * atoms + 12 * B * A
    byt = {"translate": rw + B * 16, "rotate": rw + B * 40, "center": 12 * atoms + B * 16,
           "center_backward": 12 * B * A + B * 16}
    ms = {"translate": tr, "rotate": ro, "center": ce, "center_backward": cb}
    kern = {k: {"avg_ms": ms[k], "alg_bytes": byt[k], "GBs": byt[k] / (ms[k] * 1e-3) / 1e9} for k in ms}
    dom = max(kern, key=lambda k: kern[k]["avg_ms"])  # the longest kernel
    return {"workload": f"CoordsTransform on the workload's backbone atoms: B={B}, atoms_stride={A}, "
                        "num_atoms=3*length, f32",
            "atoms": atoms, "kernels": kern,
            "roofline": {"bound": "hbm", "kernel": dom, "achieved": kern[dom]["GBs"], "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": kern[dom]["GBs"] / HBM_PEAK_GBS}}


AA20 = "ACDEFGHIKLMNPQRSTVWY"


def a2c_cpu_baseline(seqs, ang, seconds):
    """The reference's own full-atom code (FP64, one thread, compiled from its
    sources into oracle/_ref) on structures of the given set, forward and
    backward per structure, until the time budget is used."""
    import numpy as np
    from oracle import a2c_ref
    if not a2c_ref.available():
        return None
    done, t_total, k = 0, 0.0, 0
    while k < len(seqs) and t_total < seconds:
        a1 = ang[k].astype(np.float64)
        t0 = time.perf_counter()
        c = a2c_ref.forward_one(seqs[k], a1)
        a2c_ref.backward_one(seqs[k], a1, np.ones(c.size))
        t_total += time.perf_counter() - t0
        done += len(seqs[k])
        k += 1
    return {"value": done / t_total, "unit": "residues/s", "cores": 1, "kind": "reference",
            "sample": f"{k} structures of the ref350 set, forward + backward each"}


def a2c_bench(lengths_mixed, steps, warmup, cpu_seconds, with_cpu):
    """Angles2Coords fwd+bwd on the GPU: HIP events around a native loop of
    forward and backward launches over rotating buffer sets (names not
    written).  Random sequences (uniform over the 20 amino acids) and angles
    phi ~ N(-1.1, 0.3), psi ~ N(2.3, 0.4), omega = pi + N(0, 0.05), side-chain
    angles ~ U(-pi, pi).  The structures go to the GPU in the order the layer
    uses (``schedule``: the most atoms first when sizes differ)."""
    import numpy as np
    import torch
    from torchproteinlibrary_amd._native_xform import ext as xext
    from torchproteinlibrary_amd.FullAtomModel.Angles2Coords import convertStringList
    from torchproteinlibrary_amd.FullAtomModel.Angles2Coords.Angles2Coords import schedule

    rng = np.random.default_rng(7)
    dev = torch.device("cuda:0")

    def make(lengths, stride):
        seqs = ["".join(rng.choice(list(AA20), size=int(n))) for n in lengths]
        ang = np.zeros((len(seqs), 8, stride), np.float32)
        for b, n in enumerate(lengths):
            ang[b, 0, :n] = rng.normal(-1.1, 0.3, n)
            ang[b, 1, :n] = rng.normal(2.3, 0.4, n)
            ang[b, 2, :n] = np.pi + rng.normal(0, 0.05, n)
            ang[b, 3:, :n] = rng.uniform(-np.pi, np.pi, (5, n))
        return seqs, ang

    out = {}
    cases = [("ref350_f32", [350] * 32, 350, torch.float32, 8),
             ("ref350_f64", [350] * 32, 350, torch.float64, 8),
             ("mixed1000_f32", list(lengths_mixed), int(max(lengths_mixed)), torch.float32, 3)]
    ref_seqs = ref_ang = None
    for name, lengths, stride, dtype, nsets in cases:
        seqs, ang = make(lengths, stride)
        if name == "ref350_f32":
            ref_seqs, ref_ang = seqs, ang
        natoms = [xext.getSeqNumAtoms(q) for q in seqs]
        A = max(natoms)
        S = convertStringList(seqs).to(dev)
        order = schedule(natoms)
        sets = []
        for _ in range(nsets):
            X = torch.tensor(ang, dtype=dtype, device=dev)
            sets.append([S, X, torch.empty(len(seqs), 3 * A, dtype=dtype, device=dev),
                         torch.randn(len(seqs), 3 * A, dtype=dtype, device=dev), torch.empty_like(X)])
            if order is not None:
                sets[-1].append(order.to(dev))
        torch.cuda.synchronize()
        st, f, b = xext.a2c_bench(sets, steps, warmup)
        residues, atoms = int(sum(lengths)), int(sum(natoms))
        io = 4 if dtype == torch.float32 else 8
        # compulsory bytes: sequences and the 8 angle rows of every slot read,
        # coordinates written for every slot (zeros past each structure); the
        # backward reads the gradient of the valid atoms and writes all angles
        fwd_b = len(seqs) * (S.shape[1] + 8 * stride * io + 3 * A * io)
        bwd_b = len(seqs) * (S.shape[1] + 16 * stride * io) + 3 * atoms * io
        out[name] = {"B": len(seqs), "residues": residues, "atoms": atoms, "ms_per_step": st,
                     "fwd_ms": f, "bwd_ms": b, "residues_per_s": residues / (st * 1e-3),
                     "atoms_per_s": atoms / (st * 1e-3),
                     "hbm_frac_fwd": fwd_b / (f * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "hbm_frac_bwd": bwd_b / (b * 1e-3) / 1e9 / HBM_PEAK_GBS}
    if with_cpu:
        cb = a2c_cpu_baseline(ref_seqs, ref_ang, cpu_seconds)
        if cb is not None:
            out["cpu_baseline"] = cb
            out["ref350_f64"]["speedup_over_cpu"] = out["ref350_f64"]["residues_per_s"] / cb["value"]
    return out


def pdb_bench(files, atoms_per_file, with_cpu, cpu_seconds):
    """PDB2CoordsUnordered on synthetic files (ATOM records of random
    backbone-like coordinates, 10 % hydrogens, TER/END lines), written to a
    temporary directory first; the files are read once untimed to warm the
    page cache."""
    import tempfile

    import numpy as np
    from torchproteinlibrary_amd.FullAtomModel import PDB2CoordsUnordered

    rng = np.random.default_rng(5)
    res3 = ["GLY", "ALA", "SER", "LEU", "LYS", "TRP"]
    names = ["N", "CA", "C", "O", "CB", "H"]
    with tempfile.TemporaryDirectory() as d:
        paths, heavy = [], 0
        for f in range(files):
            xyz = rng.normal(scale=30.0, size=(atoms_per_file, 3))
            lines = []
            for j in range(atoms_per_file):
                nm = names[j % 6]
                heavy += nm != "H"
                lines.append("ATOM  %5d  %-3s %3s %c%4d    %8.3f%8.3f%8.3f  1.00 20.00\n" % (
                    j % 100000, nm, res3[(j // 6) % 6], "AB"[j // 50000 % 2], (j // 6) % 10000,
                    xyz[j, 0], xyz[j, 1], xyz[j, 2]))
            lines.append("TER\nEND\n")
            p = os.path.join(d, f"s{f}.pdb")
            with open(p, "w") as fh:
                fh.write("".join(lines))
            paths.append(p)
        p2c = PDB2CoordsUnordered()
        p2c(paths)
        t0 = time.perf_counter()
        out = p2c(paths)
        dt = time.perf_counter() - t0
        assert int(out[5].sum()) == heavy
        res = {"files": files, "records": files * atoms_per_file, "heavy_atoms": heavy, "seconds": dt,
               "atoms_per_s": heavy / dt, "threads": __import__("torch").get_num_threads()}
        if with_cpu:
            from oracle import pdb_oracle
            done, t_total, k = 0, 0.0, 0
            while k < len(paths) and t_total < cpu_seconds:
                t0 = time.perf_counter()
                done += len(pdb_oracle.read_file(paths[k])[4])
                t_total += time.perf_counter() - t0
                k += 1
            res["cpu_baseline"] = {"value": done / t_total, "unit": "atoms/s", "cores": 1, "kind": "port",
                                   "sample": f"{k} of the files, the Python restatement of the "
                                             "reference's loader (its C++ loader does not build here)"}
    return res


def main():
    args = parse()
    if args.probe:
        probe(args)
        return
    if args.fit:
        out = {"metric": "fit500 optimisation loop: Adam steps/s (fwd+bwd+update), B=512",
               "unit": "steps/s", "steps": args.fit_steps, "n_gpus": 1, "data": "synthetic",
               "config": {"workload": "fit500", "batch": 512, "max_length": 500, "lr": 1e-2,
                          "optimizer": "Adam"}}
        out["runs"] = fit_loop(args.fit_steps)
        out["value"] = out["runs"]["f32"]["steps_per_s"]
        out["final_loss_rel_diff_f32_vs_f64"] = (
            abs(out["runs"]["f32"]["loss_final"] - out["runs"]["f64"]["loss_final"])
            / max(out["runs"]["f64"]["loss_final"], 1e-30))
        print(json.dumps(out), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (got --gpus {args.gpus}); "
              "chains are independent, so batches shard across GPUs without exchange, "
              "but no multi-GPU path is built or measured.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from torchproteinlibrary_amd import _tuning, synthetic
    from torchproteinlibrary_amd._native import ext
    from torchproteinlibrary_amd.ReducedModel import Angles2Backbone

    os.environ.setdefault("TPL_A2B_CHECKS", "0")
    dev = torch.device("cuda:0")
    a, n, g = synthetic.workload(args.workload)
    B, _, L = a.shape
    residues = int(n.sum())
    fwd_b, bwd_b = alg_bytes(n, L)
    set_bytes = a.nbytes + g.nbytes + 4 * B + B * 9 * L * 4 + B * 3 * L * 4
    sets_n = args.sets or max(2, -(-2 * MALL_BYTES // set_bytes))

    A0 = torch.tensor(a, device=dev)
    N0 = torch.tensor(n, device=dev)
    if args.config:
        _tuning.set_config(A0, tuple(int(v) for v in args.config.split(",")))
    cfg = _tuning.config_for(A0, N0)  # tunes on the real inputs (eager)

    # Buffer sets: autograd tensors (gsets) and raw launch buffers (ksets).
    gsets, ksets = [], []
    st_b, sy_b = ext.workspace_bytes(False, B, L, list(cfg))
    for _ in range(sets_n):
        A = A0.clone().requires_grad_(True)
        N = N0.clone()
        G = torch.tensor(g, device=dev)
        gsets.append((A, N, G))
        ksets.append((A.detach(), N, G, torch.empty(B, 9 * L, device=dev),
                      torch.empty(B, 3, L, device=dev),
                      torch.empty(max(st_b, 1), dtype=torch.uint8, device=dev),
                      torch.zeros(max(sy_b, 1), dtype=torch.uint8, device=dev)))

    m = Angles2Backbone()
    for A, N, G in gsets[:1]:  # one eager autograd step through the public API
        m(A, N).backward(G)
    torch.cuda.synchronize()
    ev_ms, wall_ms = ext.bench_steps([list(k) for k in ksets], list(cfg), args.steps,
                                     args.warmup, False)
    t_ev = ev_ms / 1e3
    value = residues * args.steps / t_ev
    ms_step = ev_ms / args.steps
    t_wall = wall_ms / 1e3

    kern = time_kernels(ksets, cfg, args.kernel_reps)
    kern["forward"]["alg_bytes"], kern["backward"]["alg_bytes"] = fwd_b, bwd_b
    for k in ("forward", "backward"):
        kern[k]["alg_GBs"] = kern[k]["alg_bytes"] / (kern[k]["avg_ms"] * 1e-3) / 1e9
    # The roofline object describes the kernel furthest below its roofline
    # (the lower algorithmic GB/s).  Forward and backward take about the same
    # time per launch here (16-17 us at mixed1000), so "the longer kernel"
    # flips between them with run-to-run noise; this choice is stable and the
    # conservative one.  Both kernels' figures stay under "kernels".
    dom = min(("forward", "backward"), key=lambda k: kern[k]["alg_GBs"])
    dom_bytes = fwd_b if dom == "forward" else bwd_b
    achieved = dom_bytes / (kern[dom]["avg_ms"] * 1e-3) / 1e9

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "residues/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": args.workload,
            "batch": B,
            "max_length": L,
            "residues": residues,
            "io_dtype": "f32",
            "res_per_lane": {"forward": cfg[0], "backward": cfg[1]},
            "launch": "native loop, back-to-back fwd+bwd launches",
            "buffer_sets": sets_n,
            "step_alg_bytes": fwd_b + bwd_b,
            "step_alg_GBs": (fwd_b + bwd_b) / (ms_step * 1e-3) / 1e9,
            "wall_s": t_wall,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "alg_bytes_per_launch": dom_bytes,
            "avg_ms": kern[dom]["avg_ms"],
            "kernels": kern,
        },
    }

    # The same step replayed from a HIP graph holding one fwd+bwd per buffer set.
    gms, _, gsteps = ext.bench_graph([list(k) for k in ksets], list(cfg), args.steps, args.warmup)
    out["graph"] = {"value": residues * gsteps / (gms / 1e3), "ms_per_step": gms / gsteps,
                    "steps": gsteps,
                    "launch": f"hipGraph replay, one graph of {len(ksets)} fwd+bwd steps (one per buffer set)"}

    if not args.no_rmsd:
        out["rmsd"] = rmsd_bench(ksets, n, L, min(args.steps, 1000), min(args.warmup, 100),
                                 args.kernel_reps, args.workload)

    if not args.no_rmsd:
        out["xform"] = xform_bench(ksets, n, L, args.kernel_reps)

    if not args.no_a2c:
        out["a2c"] = a2c_bench(n, min(args.steps, 200), min(args.warmup, 20), args.cpu_seconds / 3,
                               not args.no_cpu_baseline)

    if not args.no_a2c:
        out["pdb"] = pdb_bench(32, 20000, not args.no_cpu_baseline, 3.0)

    if not args.no_plain:
        fr_b = ext.plain_frames_bytes(False, B, L)
        psets = [[A, N, G, X, GA, torch.empty(fr_b, dtype=torch.uint8, device=dev), st]
                 for (A, N, G, X, GA, _, st) in ksets[:2]]
        pms, _ = ext.bench_steps(psets, list(cfg), args.plain_steps, 2, True)
        pt = pms / 1e3
        pv = residues * args.plain_steps / pt
        out["plain"] = {"value": pv, "ms_per_step": 1e3 * pt / args.plain_steps,
                        "speedup_scan_over_plain": value / pv}

    if not args.no_fit:
        fit = fit_loop(1000)
        out["fit500"] = {
            "workload": "fit500: B=512, lengths uniform 50-500, 1000 Adam steps (lr 1e-2) of "
                        "forward + loss + backward + update, eager, per-call checks off",
            "steps_per_s_f32": fit["f32"]["steps_per_s"], "steps_per_s_f64": fit["f64"]["steps_per_s"],
            "loss_start": fit["f64"]["loss_start"],
            "loss_final_f32": fit["f32"]["loss_final"], "loss_final_f64": fit["f64"]["loss_final"],
            "final_loss_rel_diff": abs(fit["f32"]["loss_final"] - fit["f64"]["loss_final"])
            / max(fit["f64"]["loss_final"], 1e-30)}

    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline()
        out["cpu_path"] = cpu_path_line()
        out["cpu_path"]["speedup_over_reference_cpu"] = out["cpu_path"]["value"] / out["cpu_baseline"]["value"]

    # Counted HBM bytes of the two kernels: rocprofv3 runs of the probe mode,
    # after every other GPU measurement of this process.
    if not args.no_traffic:
        torch.cuda.synchronize()
        tr = traffic(args, cfg)
        if tr is not None:
            out["roofline"]["traffic"] = tr[dom]
            out["roofline"]["traffic_over_alg"] = tr[dom] / dom_bytes
            out["roofline"]["counters"] = tr
            if "rocprof_avg_ms" in tr:
                out["roofline"]["rocprof_avg_ms"] = tr["rocprof_avg_ms"][dom]

    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
