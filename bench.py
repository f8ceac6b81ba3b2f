"""Benchmark: 3D cone-beam distance-driven projection + backprojection on one MI355X.

Workload (one "step"): A then B of the cone-flat geometry with a 256x256x64 volume at
1 mm, a flat 512x64 detector at 1 mm pitch, SOD 500 mm, SDD 1000 mm and 720 views over
2 pi, on a synthetic phantom of 16 seeded random ellipsoids (fp32), inputs resident in
HBM.  value = voxel-views/s = Nx*Ny*Nz*V / (t_A + t_B).

The JSON line also carries:
  roofline      the dominant gather kernel's algorithmic FLOP rate, F = 4 * nnz per
                launch (nnz = exact (voxel, view, cell) pairs with positive overlap,
                tests/golden/nnz.json), over its HIP-event-timed average launch, against
                the FP32 vector peak 157.3 TFLOP/s;
  cpu_baseline  the reference's own CPU kernels (oracle/_ref, std::thread over all host
                cores) timed on a bounded sample of the same workload (a subset of views,
                scaled linearly in views) -- a reported baseline, not the target;
  adjoint       the A + A^T pair (autograd use) in the same unit.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "3D cone DD proj+backproj voxel-views/s, batched+adjoint vs plain; % roofline"
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0
WORKLOAD = "cone_flat_256x256x64_512x64_720v"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-views", type=int, default=72, help="views in the CPU baseline sample")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run 3 device-to-device copies of 512 MiB (calibrates FETCH_SIZE/WRITE_SIZE counters)")
    ap.add_argument("--summarize", nargs="+", metavar="DB",
                    help="only print a per-kernel summary of rocprofv3 SQLite result files")
    ap.add_argument("--verbose", action="store_true", help="print autotuner timings")
    ap.add_argument("--no-batched", action="store_true", help="skip the 16-instance batched section")
    ap.add_argument("--no-configs", action="store_true", help="skip the 2D fan-flat and cone-arc sections")
    ap.add_argument("--cone-arc-gathers", action="store_true",
                    help="only run A, B and A^T of the cone-arc 512^2x256 workload twice (for counter passes)")
    return ap.parse_args()


def batched_instances(seed=1003):
    """16 uneven cone-flat instances: volumes from {128^2x32, 192^2x48, 256^2x64, 384^2x96}
    at 1 mm, flat detectors of 2W columns x D rows at 1 mm pitch, SOD 500 / SDD 1000 mm,
    V uniform in 180..984 over 2 pi with a random start angle per instance."""
    rng = np.random.default_rng(seed)
    shapes = [(128, 32), (192, 48), (256, 64), (384, 96)]
    out = []
    for i in range(16):
        W, D = shapes[int(rng.integers(0, 4))]
        V = int(rng.integers(180, 985))
        start = np.float32(rng.uniform(0, 2 * np.pi))
        out.append(dict(W=W, D=D, Nu=2 * W, Nv=D, V=V, start=start, seed=seed * 100 + i))
    return out


def bench_batched(dev, steps, verbose=False):
    """A + B over the 16 uneven instances: the plain per-instance loop (each instance's own
    autotuned tiles) vs one batched launch per pass vs the batched passes replayed from a
    captured HIP graph.  Synthetic phantoms: seeded random ellipsoids."""
    import torch

    import xrecon_amd
    from xrecon_amd import _backend as be
    from xrecon_amd import autotune
    from xrecon_amd.batched import BatchedProjector3d
    from xrecon_amd.phantom import random_ellipsoids

    A, B = be.OP_PROJECTION, be.OP_BACKPROJECTION
    inst = batched_instances()
    projs, angs, vols = [], [], []
    for it in inst:
        W, D = it["W"], it["D"]
        projs.append(xrecon_amd.Projector3d((W, W, D), (it["Nu"], it["Nv"]), 1.0, 1.0, 500.0, 1000.0))
        angs.append(it["start"] + torch.arange(it["V"], dtype=torch.float32) * np.float32(2 * np.pi / it["V"]))
        ph = random_ellipsoids((W, W, D), (1.0, 1.0, 1.0), count=16, semi_axes=(5.0, 80.0), seed=it["seed"])
        vols.append(torch.from_numpy(ph).to(dev).view(1, 1, D, W, W))
    stream = torch.cuda.current_stream(dev)
    sid = stream.cuda_stream
    bp = BatchedProjector3d(projs, angs, device=dev)
    vv = bp.voxel_views()

    # plain loop: per-instance plans, workspaces and tuned tiles
    plans = bp.plans
    sinos = [torch.empty(bp.sino_shape(i), device=dev) for i in range(16)]
    recs = [torch.empty(bp.volume_shape(i), device=dev) for i in range(16)]
    wsa = [torch.empty(max(p.workspace_bytes(A, 1), 1), dtype=torch.uint8, device=dev) for p in plans]
    wsb = [torch.empty(max(p.workspace_bytes(B, 1), 1), dtype=torch.uint8, device=dev) for p in plans]
    ta, tb = [], []
    for i, p in enumerate(plans):
        p.execute_into(A, vols[i], sinos[i], wsa[i], sid, 0, 0, 0)
        ta.append(autotune.tune(p, A, vols[i], sinos[i], wsa[i]))
        tb.append(autotune.tune(p, B, sinos[i], recs[i], wsb[i]))

    def plain():
        for i, p in enumerate(plans):
            p.execute_into(A, vols[i], sinos[i], wsa[i], sid, ta[i][0], ta[i][1], 0)
            p.execute_into(B, sinos[i], recs[i], wsb[i], sid, tb[i][0], tb[i][1], 0)

    bta = bp.tune(A, vols, verbose=verbose)
    bsino = bp.run(A, vols)
    btb = bp.tune(B, bsino, verbose=verbose)
    brec = bp.empty_outputs(B)

    def batched():
        bp.run(A, vols, bsino)
        bp.run(B, bsino, brec)

    def timeit(fn, n):
        fn()
        torch.cuda.synchronize()
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record(stream)
        for _ in range(n):
            fn()
        e.record(stream)
        torch.cuda.synchronize()
        return s.elapsed_time(e) / n

    t_plain = timeit(plain, steps)
    t_batch = timeit(batched, steps)
    bp.capture([A, B], vols)
    t_graph = timeit(bp.replay, steps)
    # the batched path must reproduce the plain loop with the same tiles, bitwise
    for i, p in enumerate(plans):
        p.execute_into(A, vols[i], sinos[i], wsa[i], sid, bta[0], bta[1], 0)
        p.execute_into(B, sinos[i], recs[i], wsb[i], sid, btb[0], btb[1], 0)
    bp.run(A, vols, bsino)
    bp.run(B, bsino, brec)
    torch.cuda.synchronize()
    same = all(torch.equal(bsino[i], sinos[i]) and torch.equal(brec[i], recs[i]) for i in range(16))
    return {"workload": "16 uneven cone-flat instances (volumes 128^2x32..384^2x96, detectors 2W x D, "
                        "V in 180..984, random start angles), A+B",
            "voxel_views_per_operator": vv,
            "plain_loop": {"ms": t_plain, "value": vv / (t_plain * 1e-3)},
            "batched": {"ms": t_batch, "value": vv / (t_batch * 1e-3), "tiles": [bta[0], btb[1]]},
            "batched_graph": {"ms": t_graph, "value": vv / (t_graph * 1e-3)},
            "speedup_vs_plain": t_plain / t_graph,
            "batched_equals_plain_bitwise": bool(same),
            "unit": "voxel-views/s"}


def summarize_rocpd(paths):
    """Per-kernel summary of rocprofv3 SQLite outputs (kernel-trace and/or --pmc runs):
    dispatch count, mean/min/max duration, and the mean of each collected counter."""
    import sqlite3
    from collections import defaultdict

    for path in paths:
        con = sqlite3.connect(path)
        print(f"# {path}")
        rows = con.execute("select name, duration from kernels").fetchall()
        dur = defaultdict(list)
        for name, d in rows:
            dur[name].append(d)
        pmc = defaultdict(lambda: defaultdict(list))
        try:
            for name, cname, val in con.execute("select name, counter_name, counter_value from pmc_events"):
                pmc[name][cname].append(val)
        except sqlite3.OperationalError:
            pass
        print(f"{'kernel':70s} {'calls':>5s} {'mean_us':>10s} {'min_us':>10s} {'max_us':>10s}  counters(mean)")
        for name in sorted(dur, key=lambda n: -sum(dur[n])):
            d = dur[name]
            ctr = " ".join(f"{c}={sum(v) / len(v):.1f}" for c, v in sorted(pmc[name].items()))
            short = name.replace("(anonymous namespace)::", "").split("(")[0][-70:]
            print(f"{short:70s} {len(d):5d} {sum(d) / len(d) / 1e3:10.2f} {min(d) / 1e3:10.2f} {max(d) / 1e3:10.2f}  {ctr}")


def _time_gpu(fn, reps, stream):
    """Average ms of fn() over `reps` calls, HIP events on `stream` (after one warm-up)."""
    import torch

    fn()
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record(stream)
    for _ in range(reps):
        fn()
    e.record(stream)
    torch.cuda.synchronize()
    return s.elapsed_time(e) / reps


def bench_cfg1(dev, reps, with_cpu):
    """2D fan-flat: 256^2 image at 1 mm, 256 detectors at 2 mm (1 mm virtual), SOD 500 mm,
    SDD 1000 mm, 720 views, 10 seeded random ellipses; projection A then FBP (cosine
    weight, ramp filter by rFFT, weighted backprojection).  Unit: pixel-views/s."""
    import torch

    import xrecon_amd
    from xrecon_amd.phantom import random_ellipses

    W = H = 256
    Nu, V = 256, 720
    proj = xrecon_amd.Projector2d(W, Nu, 1.0, 2.0, 500.0, 1000.0, scan_type="flat").to(dev)
    angles = (torch.arange(V, dtype=torch.float32) * np.float32(2 * np.pi / V)).to(dev)
    img = random_ellipses((W, H), (1.0, 1.0), count=10, semi_axes=(5.0, 60.0), seed=1001)
    x = torch.from_numpy(img).to(dev).view(1, 1, H, W)
    stream = torch.cuda.current_stream(dev)
    sino = proj.projection(x, angles)
    t_a = _time_gpu(lambda: proj.projection(x, angles), reps, stream)
    t_f = _time_gpu(lambda: proj.filtered_backprojection(sino, angles), reps, stream)
    pv = W * H * V
    out = {"workload": "2D fan-flat A + FBP: 256^2 @1mm, 256 detectors @2mm (1mm virtual), SOD 500, SDD 1000, "
                       "720 views, 10 random ellipses (seed 1001)",
           "value": pv / ((t_a + t_f) * 1e-3), "unit": "pixel-views/s", "ms_A": t_a, "ms_FBP": t_f,
           "path": "Projector2d API (plan cache, per-call allocation)"}
    if with_cpu:
        import oracle

        if oracle.ref_available():
            o = [float(v) for v in proj._host_options]
            a = angles.cpu().numpy()
            oracle.ref_op2d("A", img[None], a[:4], o)
            t0 = time.perf_counter()
            y = oracle.ref_op2d("A", img[None], a, o)
            t1 = time.perf_counter()
            oracle.ref_op2d("BW", y, a, o)
            t2 = time.perf_counter()
            out["cpu_baseline"] = {"value": pv / (t2 - t0), "unit": "pixel-views/s", "cores": os.cpu_count(),
                                   "kind": "reference", "sample": f"all 720 views: A {t1 - t0:.3f}s + weighted B "
                                   f"{t2 - t1:.3f}s (filtering not included)"}
    return out


CFG5_WORKLOAD = "cone_arc_512x512x256_736x128_1152v"


def cfg5_setup(dev):
    """Config 5: cone-arc 512^2 x 256 at 0.7 mm, arc detector 736 x 128 (1.2 mm / 1.0 mm),
    SOD 595 mm, SDD 1085.6 mm (du = 1.105e-3 rad), 1152 views over 2 pi, batch 4: four
    phantoms of 32 seeded random ellipsoids (intensities 0.0-0.05 per mm, seeds 1005..1008)."""
    import torch

    import xrecon_amd
    from xrecon_amd.phantom import random_ellipsoids

    W = H = 512
    D, Nu, Nv, V, NB = 256, 736, 128, 1152, 4
    proj = xrecon_amd.Projector3d((W, H, D), (Nu, Nv), 0.7, (1.2, 1.0), 595.0, 1085.6, scan_type="arc")
    angles = torch.arange(V, dtype=torch.float32) * np.float32(2 * np.pi / V)
    x = torch.stack([random_ellipsoids((W, H, D), (0.7, 0.7, 0.7), count=32, semi_axes=(5.0, 120.0),
                                       intensity=(0.0, 0.05), seed=1005 + b, device=dev) for b in range(NB)])
    return proj, angles, x.view(NB, 1, D, H, W)


def bench_cfg5(dev, reps):
    """Config 5 operators: A + B and the adjoint pair A + A^T at batch 4 with the default
    tiles (no autotuning at this size); the dominant gather's algorithmic FLOP rate from
    the counted nnz (an estimate from 36 evenly spaced views, scaled to 1152)."""
    import torch

    from xrecon_amd import _backend as be

    proj, angles, x = cfg5_setup(dev)
    NB, _, D, H, W = x.shape
    Nu, Nv = proj.num_det
    V = angles.numel()
    plan = proj.plan(angles, dev)
    sid = torch.cuda.current_stream(dev).cuda_stream
    stream = torch.cuda.current_stream(dev)
    A, B, AT = be.OP_PROJECTION, be.OP_BACKPROJECTION, be.OP_PROJECTION_T
    sino = torch.empty((NB, 1, V, Nv, Nu), device=dev)
    rec = torch.empty_like(x)
    ws = torch.empty(max(plan.workspace_bytes(A, NB), plan.workspace_bytes(B, NB), plan.workspace_bytes(AT, NB)),
                     dtype=torch.uint8, device=dev)
    t_a = _time_gpu(lambda: plan.execute_into(A, x, sino, ws, sid, 0, 0, 0), reps, stream)
    t_b = _time_gpu(lambda: plan.execute_into(B, sino, rec, ws, sid, 0, 0, 0), reps, stream)
    t_at = _time_gpu(lambda: plan.execute_into(AT, sino, rec, ws, sid, 0, 0, 0), reps, stream)
    vv = W * H * D * V * NB
    out = {"workload": "cone-arc 512^2x256 @0.7mm, 736x128 detector (1.2/1.0 mm), SOD 595, SDD 1085.6, "
                       "1152 views, B=4, 32 random ellipsoids each (intensities 0-0.05/mm, seeds 1005-1008)",
           "voxel_views_per_operator": vv, "value": vv / ((t_a + t_b) * 1e-3), "unit": "voxel-views/s",
           "ms_A": t_a, "ms_B": t_b, "ms_AT": t_at, "adjoint_value": vv / ((t_a + t_at) * 1e-3),
           "tiles": "defaults (8 rows, LDS-staged cell gather; 32 slices in halves of 16)"}
    try:
        nnz = nnz_of(CFG5_WORKLOAD) * NB
    except (KeyError, OSError):
        nnz = None
    if nnz:
        t_dom, name = max((t_a, "cell_gather(A)"), (t_b, "voxel_gather(B)"))
        achieved = 4.0 * nnz / (t_dom * 1e-3) / 1e12
        out["roofline"] = {"bound": "valu", "kernel": name + ", whole pass (prefix included)",
                           "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                           "frac": achieved / FP32_PEAK_TFLOPS, "frac_nonpacked": achieved / (FP32_PEAK_TFLOPS / 2),
                           "nnz_per_operator": nnz,
                           "nnz_note": "counted by the FP64 oracle on 36 evenly spaced views, scaled to 1152, x4"}
    return out


def bench_cfg4(dev, proj, angles_dev, x_true, steps=50):
    """Config 4: differentiable reconstruction on the cone-flat 256^2 x 64, 720-view
    geometry.  y = -ln(max(Poisson(I0 exp(-A x*)), 1) / I0) at I0 = 1e5 (the clamp at one
    count is the recorded choice); then `steps` Adam steps (lr 1e-2) on
    0.5 ||A x - y||^2 + 1e-3 TV(x), x fp32 with requires_grad, gradients through the
    hand-written adjoint (A^T) kernels.  Timed on the host clock around synchronisations."""
    import torch

    from xrecon_amd.recon import reconstruct, simulate_transmission

    y = simulate_transmission(proj, x_true, angles_dev, I0=1e5, seed=1004)
    losses = []
    reconstruct(proj, y, angles_dev, steps=2, lr=1e-2, tv_weight=1e-3)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rec = reconstruct(proj, y, angles_dev, steps=steps, lr=1e-2, tv_weight=1e-3,
                      callback=lambda k, loss: losses.append(loss))
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    losses = [float(v) for v in losses]
    W, H, D = proj.image_size
    vv = W * H * D * angles_dev.numel()
    return {"workload": "cone-flat 256^2x64 / 720 views, Poisson I0=1e5 + log (counts clamped at 1), "
                        "Adam lr 1e-2 on 0.5||Ax-y||^2 + 1e-3 TV(x)",
            "adam_steps": steps, "ms_per_step": t * 1e3 / steps, "value": vv / (t / steps),
            "unit": "voxel-views/s (per operator, A + A^T + TV + Adam per step)",
            "loss_first": losses[0], "loss_last": losses[-1],
            "loss_decreasing": bool(losses[-1] < losses[0]),
            "loss_monotone_steps": int(sum(b <= a for a, b in zip(losses, losses[1:]))),
            "finite": bool(torch.isfinite(rec).all().item())}


def bench_cfg5_unrolled(dev, iterations=10):
    """Config 5 unrolled reconstruction: `iterations` gradient steps
    x <- x - a_k A^T (A x - y) at batch 4 through autograd (a_k learnable, initialised
    to 1/L with L from a power iteration on A^T A), loss ||x_K - x*||^2, forward and
    backward; with per-iteration checkpointing and, for comparison, without.  Reports time
    and peak device memory of each."""
    import torch

    from xrecon_amd.recon import UnrolledGradientDescent

    proj, angles, x_true = cfg5_setup(dev)
    angles = angles.to(dev)
    with torch.no_grad():
        y = proj.projection(x_true, angles)
        v = x_true[:1].clone()
        for _ in range(6):
            w = proj.projection_t(proj.projection(v, angles), angles)
            lip = (w.norm() / v.norm()).item()
            v = w / w.norm()
        del v, w
    out = {"workload": "cone-arc 512^2x256 / 1152 views, batch 4, %d unrolled gradient iterations" % iterations,
           "step_init": 1.0 / lip}
    for ckpt in (True, False):
        net = UnrolledGradientDescent(proj, angles, iterations=iterations, step=1.0 / lip,
                                      use_checkpoint=ckpt).to(dev)
        torch.cuda.synchronize()
        torch.cuda.reset_peak_memory_stats(dev)
        base = torch.cuda.memory_allocated(dev)
        t0 = time.perf_counter()
        rec = net(y)
        loss = ((rec - x_true) ** 2).sum()
        loss.backward()
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        key = "checkpointed" if ckpt else "stored_activations"
        out[key] = {"seconds_forward_backward": t,
                    "peak_memory_GiB": (torch.cuda.max_memory_allocated(dev) - base) / 2**30,
                    "loss": float(loss), "loss_initial": float((x_true ** 2).sum()),
                    "grad_finite": bool(torch.isfinite(net.steps.grad).all().item())}
        del net, rec, loss
        torch.cuda.empty_cache()
    return out


# Sources that decide the single-plan gather kernels' code and tables: a traffic figure
# measured on other sources is reported as not current.
KERNEL_SOURCES = ("xrecon_amd/csrc/dd_body.h", "xrecon_amd/csrc/dd_kernels.hip", "xrecon_amd/csrc/xr_internal.h",
                  "xrecon_amd/csrc/plan.cpp")


def kernel_sources_digest():
    """Digest of the code of KERNEL_SOURCES, comments and layout left out (a comment edit
    does not make a measurement stale)."""
    import hashlib
    import re

    h = hashlib.sha256()
    for rel in KERNEL_SOURCES:
        with open(os.path.join(ROOT, rel)) as f:
            text = re.sub(r"/\*.*?\*/", " ", f.read(), flags=re.S)
        text = re.sub(r"//[^\n]*", " ", text)
        h.update(" ".join(text.split()).encode())
    return h.hexdigest()[:16]


def profile_digest(path):
    """The kernel-sources digest a profile summary was taken with ("# kernel sources: D")."""
    with open(path) as f:
        for line in f:
            if line.startswith("# kernel sources:"):
                return line.split(":", 1)[1].strip()
    return None


def traffic_from_profiles(kernel):
    """HBM bytes per launch of `kernel` (a symbol, or a list of symbols whose launches
    together make one pass: their bytes add) from the committed rocprofv3 summaries: a
    line "traffic <kernel> fetch_bytes=F write_bytes=W total_bytes=T source=<pmc runs>"
    in profiles/*.txt, written from FETCH_SIZE / WRITE_SIZE passes scaled by the
    copy-kernel calibration.  None when a summary is missing for any of them."""
    if isinstance(kernel, (list, tuple)):
        total, srcs = 0.0, []
        for k in kernel:
            t, src = traffic_from_profiles(k)
            if t is None:
                return None, None
            total += t
            if src not in srcs:
                srcs.append(src)
        return total, ", ".join(srcs)
    pdir = os.path.join(ROOT, "profiles")
    if not os.path.isdir(pdir):
        return None, None
    for name in sorted(os.listdir(pdir)):
        if not name.endswith(".txt"):
            continue
        with open(os.path.join(pdir, name)) as f:
            for line in f:
                if not line.startswith("traffic "):
                    continue
                head, _, rest = line[len("traffic "):].partition(" fetch_bytes=")
                if head.strip() != kernel:
                    continue
                fields = dict(kv.split("=", 1) for kv in ("fetch_bytes=" + rest).split() if "=" in kv)
                return float(fields["total_bytes"]), f"profiles/{name}"
    return None, None


def pipes_from_profiles(kernels):
    """VALU issue, TA busy and L2 hit rate of each kernel from the committed pipe-counter
    summary (lines "pipe <kernel> valu_issue=V ta_busy=T l2_hit=H"), with whether that
    summary was taken on the current kernel sources; None when one is missing."""
    pdir = os.path.join(ROOT, "profiles")
    found, current = {}, True
    if not os.path.isdir(pdir):
        return None
    for name in sorted(os.listdir(pdir)):
        if not name.endswith(".txt"):
            continue
        path = os.path.join(pdir, name)
        with open(path) as f:
            for line in f:
                if not line.startswith("pipe "):
                    continue
                head, _, rest = line[len("pipe "):].partition(" valu_issue=")
                if head.strip() in kernels:
                    found[head.strip()] = {k: float(v) for k, v in
                                           (kv.split("=", 1) for kv in ("valu_issue=" + rest).split())}
                    current = current and profile_digest(path) == kernel_sources_digest()
    if any(k not in found for k in kernels):
        return None
    return {"kernels": found, "current": current}


def calibrate_copy():
    import torch

    n = 512 * 1024 * 1024 // 4
    src = torch.ones(n, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read_per_launch": n * 4, "copy_bytes_written_per_launch": n * 4, "launches": 3}))


def workload():
    W, H, D, Nu, Nv, V = 256, 256, 64, 512, 64, 720
    return dict(W=W, H=H, D=D, Nu=Nu, Nv=Nv, V=V, pix=1.0, det=1.0, sod=500.0, sdd=1000.0)


def nnz_of(name):
    path = os.path.join(ROOT, "tests", "golden", "nnz.json")
    with open(path) as f:
        return int(json.load(f)[name]["nnz"])


def cpu_baseline(proj_opts, phantom, angles, nviews):
    """Reference CPU kernels (oracle/_ref) on `nviews` evenly spaced views, A + B."""
    import oracle

    if not oracle.ref_available():
        return None
    V = angles.shape[0]
    idx = np.linspace(0, V, nviews, endpoint=False).astype(np.int64)
    sub = np.ascontiguousarray(angles[idx])
    x = phantom[None]
    oracle.ref_op3d("A", x, sub[:2], proj_opts)  # warm-up (thread start-up, page faults)
    t0 = time.perf_counter()
    y = oracle.ref_op3d("A", x, sub, proj_opts)
    t1 = time.perf_counter()
    oracle.ref_op3d("B", y, sub, proj_opts)
    t2 = time.perf_counter()
    vv = phantom.size * nviews
    return {"value": vv / (t2 - t0), "unit": "voxel-views/s", "cores": os.cpu_count(), "kind": "reference",
            "sample": f"A+B of the same volume over {nviews} of {V} views (evenly spaced), "
                      f"scaled linearly in views; t_A={t1 - t0:.2f}s t_B={t2 - t1:.2f}s",
            "threads": "std::thread x hardware_concurrency (reference cpu3d_common.h:17-22)"}


def main():
    args = parse()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.summarize:
        summarize_rocpd(args.summarize)
        return
    if args.cone_arc_gathers:
        import torch

        print(json.dumps(bench_cfg5(torch.device("cuda:0"), 1)))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build runs on a single MI355X and has no multi-GPU path "
              f"(--gpus {args.gpus} requested)", file=sys.stderr)
        sys.exit(2)
    import torch

    import xrecon_amd
    from xrecon_amd import _backend as be
    from xrecon_amd.phantom import random_ellipsoids

    wl = workload()
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    proj = xrecon_amd.Projector3d((wl["W"], wl["H"], wl["D"]), (wl["Nu"], wl["Nv"]), wl["pix"], wl["det"],
                                  wl["sod"], wl["sdd"], scan_type="flat")
    angles = (torch.arange(wl["V"], dtype=torch.float32) * np.float32(2 * np.pi / wl["V"]))
    phantom = random_ellipsoids((wl["W"], wl["H"], wl["D"]), (wl["pix"],) * 3, count=16, semi_axes=(5.0, 80.0),
                                intensity=(0.1, 1.0), seed=1002)
    x = torch.from_numpy(phantom).to(dev).view(1, 1, wl["D"], wl["H"], wl["W"])
    plan = proj.plan(angles, dev)
    stream = torch.cuda.current_stream(dev)
    sid = stream.cuda_stream
    proj = proj.to(dev)
    angles_dev = angles.to(dev)

    A, B, AT = be.OP_PROJECTION, be.OP_BACKPROJECTION, be.OP_PROJECTION_T
    sino = torch.empty((1, 1, wl["V"], wl["Nv"], wl["Nu"]), device=dev)
    recon = torch.empty_like(x)
    grad = torch.empty_like(x)
    ws_a = torch.empty(plan.workspace_bytes(A, 1), dtype=torch.uint8, device=dev)
    ws_b = torch.empty(plan.workspace_bytes(B, 1), dtype=torch.uint8, device=dev)

    from xrecon_amd import autotune

    # Tile shapes chosen by the autotuner (untimed; the sinogram is produced first so B is
    # tuned on real data).
    plan.execute_into(A, x, sino, ws_a, sid, 0, 0, 0)
    tune_a = autotune.tune(plan, A, x, sino, ws_a, verbose=args.verbose)
    tune_b = autotune.tune(plan, B, sino, recon, ws_b, verbose=args.verbose)

    def step(ev=None):
        # A: prefix pass, then the cell-gather kernel (events bracket the gather only)
        plan.execute_into(A, x, sino, ws_a, sid, tune_a[0], tune_a[1], 1)
        if ev: ev[0].record(stream)
        plan.execute_into(A, x, sino, ws_a, sid, tune_a[0], tune_a[1], 2)
        if ev: ev[1].record(stream)
        plan.execute_into(B, sino, recon, ws_b, sid, tune_b[0], tune_b[1], 1)
        if ev: ev[2].record(stream)
        plan.execute_into(B, sino, recon, ws_b, sid, tune_b[0], tune_b[1], 2)
        if ev: ev[3].record(stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    K = args.steps
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in range(K)]
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record(stream)
    for k in range(K):
        step(evs[k])
    end.record(stream)
    torch.cuda.synchronize()
    total_ms = start.elapsed_time(end)
    ms_step = total_ms / K
    t_cell = sum(e[0].elapsed_time(e[1]) for e in evs) / K
    t_vox = sum(e[2].elapsed_time(e[3]) for e in evs) / K

    # adjoint pair A + A^T (the autograd pair), same unit, untimed-kernel breakdown
    ws_at = torch.empty(plan.workspace_bytes(AT, 1), dtype=torch.uint8, device=dev)
    tune_at = autotune.tune(plan, AT, sino, grad, ws_at, verbose=args.verbose)
    for _ in range(2):
        plan.execute_into(AT, sino, grad, ws_at, sid, tune_at[0], tune_at[1], 0)
    torch.cuda.synchronize()
    s2, e2 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s2.record(stream)
    for _ in range(K):
        plan.execute_into(A, x, sino, ws_a, sid, tune_a[0], tune_a[1], 0)
        plan.execute_into(AT, sino, grad, ws_at, sid, tune_at[0], tune_at[1], 0)
    e2.record(stream)
    torch.cuda.synchronize()
    ms_adj = s2.elapsed_time(e2) / K

    vox_views = wl["W"] * wl["H"] * wl["D"] * wl["V"]
    value = vox_views / (ms_step * 1e-3)
    nnz = nnz_of(WORKLOAD)
    flops = 4.0 * nnz
    if t_cell >= t_vox:
        # one launch per span class of the column groups (two-pixel at the tuned rows,
        # three-pixel and general at most 8 rows)
        ncls = plan.cell_classes()
        r_a = tune_a[0] & 0xFF
        rows = (r_a, min(r_a, 8), min(r_a, 8))
        # LDS-staged classes 2 and 3 at 4 rows and more: as tuned (this table is below the
        # 128 MiB the library stages by default)
        staged = bool(tune_a[0] & autotune.CELL_LDS)
        blocks = [-(-n // 4) * -(-wl["Nv"] // r) for r, n in zip(rows, ncls)]
        if -(-sum(ncls) // 4) * -(-wl["Nv"] // r_a) < 4096:  # small grids: one launch for all classes
            symbol = [f"k_cell<0, 1, {r_a}, -1, false, false>"]
        else:  # a class launch stages only with 4096 workgroups or more
            symbol = [f"k_cell<0, 1, {r}, {span}, false, "
                      f"{'true' if staged and span and r >= 4 and nb >= 4096 else 'false'}>"
                      for r, span, n, nb in zip(rows, (2, 3, 0), ncls, blocks) if n > 0]
        dominant, t_dom = "cell_gather(A)", t_cell
    else:
        kc = tune_b[1]
        dominant, t_dom = "voxel_gather(B)", t_vox
        symbol = f"k_voxel<0, 1, {kc & 0xff}, {'true' if kc & 0x100 else 'false'}>"
    achieved = flops / (t_dom * 1e-3) / 1e12
    traffic, traffic_src = traffic_from_profiles(symbol)
    # the summary's kernels are these only if it was taken on the same kernel sources
    traffic_current = None
    if traffic_src:
        traffic_current = all(profile_digest(os.path.join(ROOT, src.strip())) == kernel_sources_digest()
                              for src in traffic_src.split(","))

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(list(proj._host_options), phantom, angles.numpy(), args.cpu_views)

    # FDK through the public API on the same sinogram (weights, rFFT ramp filter, FDK-weighted
    # backprojection, scaling)
    fdk_ms = _time_gpu(lambda: proj.filtered_backprojection(sino, angles_dev), max(3, K // 2), stream)
    batched = None if args.no_batched else bench_batched(dev, max(2, K // 2), args.verbose)
    configs = None
    if not args.no_configs:
        configs = {"cfg1_fan_flat_2d": bench_cfg1(dev, max(3, K), not args.no_cpu)}
        configs["cfg4_adam_reconstruction"] = bench_cfg4(dev, proj, angles_dev, x)
        torch.cuda.empty_cache()
        configs["cfg5_cone_arc"] = bench_cfg5(dev, 2)
        torch.cuda.empty_cache()
        configs["cfg5_unrolled"] = bench_cfg5_unrolled(dev)
        torch.cuda.empty_cache()
    compulsory = 4.0 * (vox_views / wl["V"] + wl["V"] * wl["Nv"] * wl["Nu"])

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "voxel-views/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: 16 seeded random ellipsoids (semi-axes 5-80 mm, intensities 0.1-1.0, seed 1002)",
        "config": {"workload": "cone-flat A+B: volume 256x256x64 @1mm, flat detector 512x64 @1mm, "
                               "SOD 500 mm, SDD 1000 mm, 720 views over 2pi, B*C=1",
                   "voxel_views_per_operator": vox_views, "nnz_per_operator": nnz},
        "roofline": {"bound": "valu", "kernel": dominant, "symbol": symbol, "achieved": achieved,
                     "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS,
                     "peak_nonpacked": FP32_PEAK_TFLOPS / 2, "frac_nonpacked": achieved / (FP32_PEAK_TFLOPS / 2),
                     "traffic": traffic, "traffic_source": traffic_src, "traffic_current": traffic_current,
                     "pipes": pipes_from_profiles(symbol if isinstance(symbol, list) else [symbol]),
                     "hbm_frac": None if traffic is None else traffic / (t_dom * 1e-3) / 8e12,
                     "compulsory_bytes": compulsory,
                     "compulsory_hbm_frac": compulsory / (t_dom * 1e-3) / 8e12,
                     "note": "FP32 vector-ALU roof (packed FP32 peak); no dense contraction, so no MFMA; "
                             "F = 4*nnz per launch; traffic = HBM bytes per launch (rocprofv3 FETCH_SIZE x2.0 "
                             "+ WRITE_SIZE x1.0, copy-kernel calibrated)"},
        "tiles": {"A_tile_rows": tune_a[0] & 0xFF, "A_cell_staged": bool(tune_a[0] & autotune.CELL_LDS),
                  "B_voxel_chunk": tune_b[1]},
        "kernels_ms": {"cell_gather_A": t_cell, "voxel_gather_B": t_vox,
                       "prefix_passes": ms_step - t_cell - t_vox},
        "adjoint": {"ms_per_pair": ms_adj, "value": vox_views / (ms_adj * 1e-3), "unit": "voxel-views/s",
                    "pair": "A + A^T"},
        "fdk": {"ms": fdk_ms, "value": vox_views / (fdk_ms * 1e-3), "unit": "voxel-views/s",
                "path": "Projector3d.filtered_backprojection (API, includes the filter stage)"},
        "cpu_baseline": cpu,
        "batched": batched,
        "configs": configs,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
