#!/usr/bin/env python3
"""NFFT throughput on one MI355X: NU points/s of forward + adjoint.

Workload (the headline configuration): 3-D NFFT, N = 128^3, M = 2^21 nodes
uniform in [-1/2, 1/2)^3, ComplexF32 data ~ N(0,1), Kaiser-Bessel m = 4,
sigma = 2 (oversampled grid 256^3), LINEAR window table; FP32 spreading,
interpolation and FFT with FP64 node indexing and deconvolution factors.

One step = one forward mul! plus one adjoint mul! of the FP32 plan on inputs
already resident in HBM (plan creation excluded).  value = M * steps / time.
The same workload on the FP64 plan is timed for the FP32/FP64 ratio, errors
are reported against the FP64 plan (full rel-l2) and against the FP64 direct
NDFT on sampled outputs, per-stage device times come from HIP events recorded
on the plan's stream inside a second timed pass, and the CPU baseline is the
FP64 C++/OpenMP restatement of the reference plan (oracle/) on the host cores.

Other workloads (--workload): "1d" -- N = 2^24, M = 2^24 nodes
Beta(1/2, 1/2) - 1/2 clustered at the edges, m = 6, sigma = 1.25, TENSOR;
"2d-spiral" -- N = 256^2, M = 65,536 nodes on an 8-interleave Archimedean
spiral, m = 4, sigma = 2, POLYNOMIAL, where one step also computes the node
gradient of ||A(k) f - y||^2.

"batch" -- 64 2-D plans with N drawn from {64^2, 128^2, 256^2, 512^2} and M
log-uniform in [1e4, 1e6] (uniform nodes), plus 32 images sharing one 256^2
golden-angle radial trajectory (402 spokes x 512 samples), m = 4, sigma = 2,
LINEAR, run as one autotuned batch (persistent launches, HIP graphs) and, for
comparison, plan by plan.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload 3d|1d|2d-spiral|batch]
"""
import argparse
import ctypes
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [REPO, os.path.join(REPO, "nfft.jl_amd")]

import numpy as np  # noqa: E402

HBM_PEAK = 8.0e12  # B/s, MI355X
F32_TOL = 1e-5  # rel-l2 bar of every FP32 result (against the FP64 plan and the FP64 NDFT)

# HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE
# (scaled by a 256 MiB elementwise copy in the same run: factors 2.0 / 1.0);
# profiles/nfft3d_128cubed_M2p21_hbm_traffic.md holds the passes and the
# command.  Keyed by (workload, stage) for the default FP32 plan of that
# workload, together with the digest of the sources the 3-D gridding kernels
# are built from (source_digest(TRAFFIC_SOURCES)) when they were measured:
# when those sources change, the figure is reported as stale (traffic null)
# until it is measured again.
TRAFFIC_SOURCES = ("k_tile3.hip", "kernels.hpp", "launch.hpp", "fft_line.hpp")
MEASURED_TRAFFIC = {("3d", "spreading"): (1086917875, "53e2f819487db85e"),
                    ("3d", "interpolation"): (644968774, "53e2f819487db85e")}

# Stage order of nfftm_profile_read.  "fold": the per-item padded tiles summed
# into the grid; on 3-D power-of-two grids with one work item per tile it is
# fused with the first adjoint FFT pass (lines along dimension 3).
STAGES = ("deconvolve", "fft_forward", "interpolation", "fold", "spreading", "fft_backward",
          "deconvolve_transpose", "node_gradient")

# Issue cost of the LDS instructions of the 3-D gridding taps, measured live
# (nfftm_lds_issue_cycles: per-CU shader cycles of one wavefront instruction).
LDS_OPS = {"ds_read_b64": 0, "ds_read_b128": 1, "ds_add_u64": 2, "ds_add_f64": 3}


def lds_roofline(L, stage, ms, D, W, M, fp32):
    """Secondary (LDS-issue) bound of the 3-D gridding kernels: the time the
    taps' LDS instructions take at the measured per-CU issue cost, against
    the kernel's measured time.  Spreading (k_spread3): a wavefront takes one
    node, its lanes the W^2 footprint columns (ceil(W^2 / 64) instructions),
    each adding W cells along z -- packed fixed point, one ds_add_u64 per cell
    (FP32), or two ds_add_f64 (FP64).  Interpolation (k_interp3): 16 lanes per
    node, each reading ceil(W^2 / 16) columns of W cells (ds_read_b64), 4
    nodes per wavefront."""
    if ms is None or D != 3 or stage not in ("spreading", "interpolation"):
        return None
    cyc = (ctypes.c_double * 4)()
    hz = ctypes.c_double()
    ncu = ctypes.c_int()
    if L.nfftm_lds_issue_cycles(cyc, ctypes.byref(hz), ctypes.byref(ncu)) != 0:
        return None
    if stage == "interpolation":
        op, per_node = "ds_read_b64", -(-W * W // 16) * W / 4.0
    elif fp32:
        op, per_node = "ds_add_u64", -(-W * W // 64) * W
    else:
        op, per_node = "ds_add_f64", 2 * -(-W * W // 64) * W
    c = cyc[LDS_OPS[op]]
    floor_s = M * per_node * c / (ncu.value * hz.value)
    return {"bound": "lds", "kernel": stage, "instruction": op, "instructions_per_node": per_node,
            "cycles_per_instruction": c, "clock_hz": hz.value, "cus": ncu.value,
            "floor_ms": floor_s * 1e3, "avg_ms": ms, "frac": floor_s * 1e3 / ms,
            "issue_cycles_measured": {k: cyc[i] for k, i in LDS_OPS.items()}}


def source_digest(names=None):
    """sha256 (16 hex digits) of the HIP sources the kernels are built from
    (all of them, or the files in names)."""
    import hashlib
    h = hashlib.sha256()
    d = os.path.join(REPO, "nfft.jl_amd", "csrc")
    for name in sorted(os.listdir(d)):
        if (name in names) if names else name.endswith((".hip", ".hpp", ".inc")):
            with open(os.path.join(d, name), "rb") as fh:
                h.update(name.encode() + b"\0" + fh.read())
    return h.hexdigest()[:16]


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def stage_bytes(D, P, Q, M, sc, sk=8):
    """Algorithmic HBM bytes per launch of each stage.  The fold is credited
    with the grid it produces (Q s_c); in 3-D it is fused with the first
    adjoint FFT pass (lines along dimension 3), whose algorithmic bytes are
    one grid read and the half of the grid the next pass reads: 1.5 Q s_c."""
    return {
        "deconvolve": P * sc + Q * sc,
        "fft_forward": 2 * Q * sc,
        "interpolation": M * (D * sk + sc) + Q * sc,
        "fold": (1.5 if D == 3 else 1.0) * Q * sc,
        "spreading": M * (D * sk + sc) + Q * sc,
        "fft_backward": 2 * Q * sc,
        "deconvolve_transpose": 2 * P * sc,
        "node_gradient": M * (D * sk + sc + D * 8) + Q * sc,
    }


def spiral_nodes(interleaves=8, samples=8192, turns=16):
    t = np.arange(samples) / samples
    ks = []
    for q in range(interleaves):
        th = 2 * np.pi * turns * t + 2 * np.pi * q / interleaves
        ks.append(np.stack([0.5 * t * np.cos(th), 0.5 * t * np.sin(th)]))
    return np.concatenate(ks, axis=1)


def workload(name):
    """(description, N, M, m, sigma, precompute name, rng, nodes)"""
    if name == "3d":
        rng = np.random.default_rng([123, 2])
        N, M = (128, 128, 128), 2 ** 21
        k = rng.random((3, M)) - 0.5
        return ("3D NFFT N=128^3 M=2^21 uniform nodes, ComplexF32, Kaiser-Bessel m=4 sigma=2 "
                "(grid 256^3), LINEAR table, forward+adjoint", N, M, 4, 2.0, "LINEAR", rng, k)
    if name == "1d":
        rng = np.random.default_rng([123, 5])
        N, M = (2 ** 24,), 2 ** 24
        k = (rng.beta(0.5, 0.5, size=M) - 0.5).reshape(1, M)
        return ("1D NFFT N=2^24 M=2^24 nodes Beta(1/2,1/2)-1/2 (edge-clustered), ComplexF32, "
                "Kaiser-Bessel m=6 sigma=1.25 (grid 20971520), TENSOR, forward+adjoint",
                N, M, 6, 1.25, "TENSOR", rng, k)
    if name == "2d-spiral":
        rng = np.random.default_rng([123, 4])
        k = spiral_nodes()
        return ("2D NFFT N=256^2 M=65536 nodes on an 8-interleave Archimedean spiral, ComplexF32, "
                "Kaiser-Bessel m=4 sigma=2, POLYNOMIAL, forward+adjoint+node gradient",
                (256, 256), k.shape[1], 4, 2.0, "POLYNOMIAL", rng, k)
    raise SystemExit(f"unknown workload {name}")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def radial_nodes(spokes, samples):
    phi = (1 + 5 ** 0.5) / 2
    th = np.arange(spokes) * np.pi / phi
    r = -0.5 + np.arange(samples) / samples
    return np.stack([np.outer(np.cos(th), r).ravel(), np.outer(np.sin(th), r).ravel()])


def run_batch(args, torch, nfftm, dev):
    """Batched uneven plans plus shared-trajectory images (see the module
    docstring).  value = total NU points of all slots per second of batch
    forward + adjoint."""
    rng = np.random.default_rng([123, 3])
    specs = []
    for _ in range(64):
        n = int(rng.choice([64, 128, 256, 512]))
        M = int(round(10 ** rng.uniform(4, 6)))
        specs.append(((n, n), rng.random((2, M)) - 0.5, 1))
    specs.append(((256, 256), radial_nodes(402, 512), 32))

    def build(dtype):
        plans = [nfftm.plan_nfft(k, N, m=4, sigma=2.0, precompute=nfftm.LINEAR, dtype=dtype)
                 for N, k, _ in specs]
        return plans

    def data(plans, b, cd, seed):
        g = torch.Generator(device=dev)
        g.manual_seed(seed)
        # drawn in FP64 and rounded, so both precisions see the same data
        fs = [torch.randn(plans[i].N, dtype=torch.complex128, device=dev, generator=g).to(cd)
              for i, _ in b.slots]
        fh = [torch.randn((plans[i].J,), dtype=torch.complex128, device=dev, generator=g).to(cd)
              for i, _ in b.slots]
        return fs, fh

    def timed(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps

    res = {}
    for name, dtype, cd in (("f32", torch.float32, torch.complex64), ("f64", torch.float64, torch.complex128)):
        log(f"bench: batch {name}: plans")
        plans = build(dtype)
        t0 = time.perf_counter()
        b = nfftm.BatchPlan(plans, [n for _, _, n in specs], autotune=True)
        t_setup = time.perf_counter() - t0
        fs, fh = data(plans, b, cd, 5)
        ys = [torch.empty((plans[i].J,), dtype=cd, device=dev) for i, _ in b.slots]
        xs = [torch.empty(plans[i].N, dtype=cd, device=dev) for i, _ in b.slots]

        def step_batch():
            b.forward(fs, ys)
            b.adjoint(fh, xs)

        def step_single():
            for s, (i, _) in enumerate(b.slots):
                plans[i].forward(fs[s], ys[s])
                plans[i].adjoint(fh[s], xs[s])

        log(f"bench: batch {name}: timing")
        tb = timed(step_batch, args.steps, args.warmup)
        yb = [y.clone() for y in ys]
        xb = [x.clone() for x in xs]
        ts = timed(step_single, max(2, args.steps // 4), 1)
        err = max(max(float(torch.linalg.vector_norm(a - c) / torch.linalg.vector_norm(c))
                      for a, c in zip(yb, ys)),
                  max(float(torch.linalg.vector_norm(a - c) / torch.linalg.vector_norm(c))
                      for a, c in zip(xb, xs)))
        shapes = sorted({(plans[i].Ntilde, plans[i].tile_shape()) for i, _ in b.slots})
        res[name] = {"t_batch": tb, "t_single": ts, "setup_s": t_setup, "graph": b.graph,
                     "batch_vs_single_relerr": err, "tiles": [str(x) for x in shapes],
                     "y": [y.cpu() for y in (yb[0], yb[-1])], "x": [x.cpu() for x in (xb[0], xb[-1])]}
        del b, plans, fs, fh, ys, xs
        torch.cuda.empty_cache()
    Mtot = sum(k.shape[1] * n for _, k, n in specs)
    # algorithmic bytes of one forward + adjoint of every slot (FP32, s_c = 8)
    sc = 8
    bytes_step = 0
    for N, k, n in specs:
        P, Q, M = N[0] * N[1], 4 * N[0] * N[1], k.shape[1]
        bytes_step += n * ((P * sc + 4 * Q * sc + M * (2 * 8 + sc)) + (2 * P * sc + 3 * Q * sc + M * (2 * 8 + sc)))
    achieved = bytes_step / res["f32"]["t_batch"]
    cpu = None
    if not args.no_cpu:
        # a bounded sample of the same workload: the shared radial trajectory
        # (one of the 32 coils) and the first 8 uneven plans, FP64 blocked
        # LINEAR oracle plans on the host cores, forward + adjoint
        import oracle as O
        threads = O.lib().orc_num_threads()
        log(f"bench: CPU baseline (oracle restatement, {threads} threads)")
        sample = [specs[-1]] + specs[:8]
        t_cpu, m_cpu = 0.0, 0
        for N, k, _ in sample:
            op = O.OraclePlan(k, N, 4, 2.0, O.LINEAR, True, workers=threads)
            rng_c = np.random.default_rng(11)
            fc = rng_c.standard_normal(N) + 1j * rng_c.standard_normal(N)
            fhc = rng_c.standard_normal(k.shape[1]) + 1j * rng_c.standard_normal(k.shape[1])
            op.forward(fc)
            op.adjoint(fhc)
            best = float("inf")
            for _ in range(3):
                t0 = time.perf_counter()
                op.forward(fc)
                op.adjoint(fhc)
                best = min(best, time.perf_counter() - t0)
            t_cpu += best
            m_cpu += k.shape[1]
        cpu = {"value": m_cpu / t_cpu, "unit": "NU points/s", "cores": threads, "kind": "port",
               "cpu_model": cpu_model(),
               "sample": f"{len(sample)} of the batch's plans (the radial trajectory once and the first 8 "
                         f"uneven plans, {m_cpu} nodes), FP64 blocked LINEAR restatement, one plan at a time, "
                         "min of 3 forward+adjoint runs each, FFT by scipy.fft"}
    e32 = max(float(np.linalg.norm(a.numpy() - c.numpy()) / np.linalg.norm(c.numpy()))
              for a, c in zip(res["f32"]["y"] + res["f32"]["x"], res["f64"]["y"] + res["f64"]["x"]))
    out = {
        "metric": "NU points/s fwd+adj, batch of 64 uneven 2-D plans + 32 radial images (FP32)",
        "value": Mtot / res["f32"]["t_batch"],
        "unit": "NU points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["f32"]["t_batch"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "64 2-D plans N in {64,128,256,512}^2, M log-uniform [1e4,1e6], uniform "
                               "nodes; + 32 images on a 256^2 radial trajectory (402x512, golden angle); "
                               "m=4 sigma=2 LINEAR, autotuned, one batch, forward+adjoint",
                   "slots": 96, "total_nodes": Mtot},
        "single_plans_ms_per_step": res["f32"]["t_single"] * 1e3,
        "batch_speedup_vs_single": res["f32"]["t_single"] / res["f32"]["t_batch"],
        "fp64": {"value": Mtot / res["f64"]["t_batch"], "ms_per_step": res["f64"]["t_batch"] * 1e3},
        "f32_over_f64": res["f64"]["t_batch"] / res["f32"]["t_batch"],
        "errors": {"f32_vs_f64_sampled_slots": e32,
                   "batch_vs_single_f32": res["f32"]["batch_vs_single_relerr"],
                   "batch_vs_single_f64": res["f64"]["batch_vs_single_relerr"]},
        "graph": res["f32"]["graph"],
        "setup_s": {"f32": res["f32"]["setup_s"], "f64": res["f64"]["setup_s"]},
        "tiles_f32": res["f32"]["tiles"],
        "roofline": {"bound": "hbm", "kernel": "whole batch step (every stage of all 96 slots)",
                     "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": None, "bytes_per_step": bytes_step},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def run_tools(args, torch, nfftm, dev):
    """NFFTTools on the plan, 256^2 image on the golden-angle radial trajectory
    (402 spokes x 512 samples): sampling density compensation (sdc, 20
    iterations) and A'A products through the Toeplitz kernel (2x oversampled
    FFTs) against the same products as forward + adjoint NFFT.  value = A'A
    products per second through the Toeplitz kernel (FP32)."""
    import nfftm.tools as T
    k = radial_nodes(402, 512)
    N = (256, 256)
    M = k.shape[1]

    def timed(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps

    res = {}
    for name, dtype, cd in (("f32", torch.float32, torch.complex64), ("f64", torch.float64, torch.complex128)):
        log(f"bench: tools {name}")
        p = nfftm.plan_nfft(k, N, m=4, sigma=2.0, precompute=nfftm.LINEAR, dtype=dtype)
        w = torch.ones(M, dtype=dtype, device=dev)

        def run_sdc():
            w.fill_(1)
            T.sdc_(p, 20, w)

        t_sdc = timed(run_sdc, max(2, args.steps // 4), 1)
        lam = T.calculateToeplitzKernel(N, k.astype(np.float32 if dtype == torch.float32 else np.float64))
        t = T.Toeplitz(N, cd)
        g = torch.Generator(device=dev)
        g.manual_seed(9)
        x = torch.randn(N, dtype=torch.complex128, device=dev, generator=g).to(cd)
        y = x.clone()
        ya = torch.empty(M, dtype=cd, device=dev)
        xa = torch.empty(N, dtype=cd, device=dev)

        def toep():
            y.copy_(x)
            t.apply(lam, y)

        def nfft_pair():
            p.forward(x, ya)
            p.adjoint(ya, xa)

        t_toep = timed(toep, args.steps, args.warmup)
        t_pair = timed(nfft_pair, args.steps, args.warmup)
        toep()
        nfft_pair()
        torch.cuda.synchronize()
        err = float(torch.linalg.vector_norm(y - xa) / torch.linalg.vector_norm(xa))
        res[name] = {"t_sdc": t_sdc, "t_toep": t_toep, "t_pair": t_pair, "toep_vs_nfft": err,
                     "w": w.double().cpu().numpy()}
        del p, lam, t
        torch.cuda.empty_cache()
    wf, wd = res["f32"]["w"], res["f64"]["w"]
    Q = 4 * N[0] * N[1]
    P = N[0] * N[1]
    # pad (P read, Q write), FFT (Q read + write), product (2Q read, Q write),
    # inverse FFT (Q read + write), crop (P read + write): 8 B per point
    bytes_prod = (P + Q + 2 * Q + 3 * Q + 2 * Q + 2 * P) * 8
    cpu = None
    if not args.no_cpu:
        import oracle as O
        orc = O.OraclePlan(k, N, m=4, sigma=2.0, precompute=O.LINEAR)
        t0 = time.perf_counter()
        O.sdc(orc, iters=2, scale=False)
        t_it = (time.perf_counter() - t0) / 2
        cpu = {"value": 1.0 / (t_it * 20), "unit": "sdc runs/s (20 iterations)", "cores": 1, "kind": "port",
               "sample": "2 sdc iterations of the FP64 C restatement (convolve_transpose! + convolve!) on "
                         "the same trajectory, scaled to 20 iterations"}
    out = {
        "metric": "A'A products/s via the Toeplitz kernel, 256^2 radial 402x512 (FP32)",
        "value": 1.0 / res["f32"]["t_toep"],
        "unit": "products/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["f32"]["t_toep"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "NFFTTools: 256^2 image, golden-angle radial trajectory 402 spokes x 512 "
                               "samples (M=205,824), m=4 sigma=2 LINEAR; sdc 20 iterations; Toeplitz "
                               "A'A (2x oversampled FFTs) vs forward+adjoint NFFT",
                   "N": list(N), "M": M},
        "nfft_pair_ms": res["f32"]["t_pair"] * 1e3,
        "toeplitz_speedup_vs_nfft_pair": res["f32"]["t_pair"] / res["f32"]["t_toep"],
        "sdc_ms": {"f32": res["f32"]["t_sdc"] * 1e3, "f64": res["f64"]["t_sdc"] * 1e3},
        "fp64": {"value": 1.0 / res["f64"]["t_toep"], "ms_per_step": res["f64"]["t_toep"] * 1e3},
        "errors": {"toeplitz_vs_nfft_pair_f32": res["f32"]["toep_vs_nfft"],
                   "toeplitz_vs_nfft_pair_f64": res["f64"]["toep_vs_nfft"],
                   "sdc_f32_vs_f64": float(np.linalg.norm(wf - wd) / np.linalg.norm(wd))},
        "roofline": {"bound": "hbm", "kernel": "whole Toeplitz product (pad, FFT, product, inverse FFT, crop)",
                     "achieved": bytes_prod / res["f32"]["t_toep"] / 1e9, "peak": 8000.0, "unit": "GB/s",
                     "frac": bytes_prod / res["f32"]["t_toep"] / 1e9 / 8000.0, "traffic": None},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--ndft-fwd", type=int, default=1024, help="sampled nodes for the NDFT check")
    ap.add_argument("--ndft-adj", type=int, default=64, help="sampled grid points for the NDFT check")
    ap.add_argument("--workload", default="3d", choices=("3d", "1d", "2d-spiral", "batch", "tools"))
    ap.add_argument("--tune", action="store_true",
                    help="autotune the tile shape of both plans before timing (plan set-up)")
    ap.add_argument("--copy-calib", action="store_true",
                    help="also run a plain 256 MiB device copy (reference kernel for HBM counters)")
    ap.add_argument("--sparse-baseline", action="store_true",
                    help="also time the reference GPU backend's algorithm on the same workload: "
                         "precompute FULL as a device sparse matrix (FP32)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (got --gpus {args.gpus}); the multi-GPU "
              "sharding of the NFFT is a design note, not a measured configuration.",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import nfftm

    torch.cuda.init()
    dev = torch.device("cuda", 0)
    L = nfftm._L.load()
    if args.workload == "batch":
        return run_batch(args, torch, nfftm, dev)
    if args.workload == "tools":
        return run_tools(args, torch, nfftm, dev)

    desc, N, M, m, sigma, pre_name, rng, k = workload(args.workload)
    pre = getattr(nfftm, pre_name)
    grad = args.workload == "2d-spiral"
    D = len(N)
    P = int(np.prod(N))
    f_h = (rng.standard_normal(N) + 1j * rng.standard_normal(N)).astype(np.complex64)
    fh_h = (rng.standard_normal(M) + 1j * rng.standard_normal(M)).astype(np.complex64)

    log("bench: creating FP32 and FP64 plans")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    # the small 2-D workload is launch-bound: its transforms replay HIP graphs
    graphs = args.workload == "2d-spiral"
    p32 = nfftm.plan_nfft(k, N, m=m, sigma=sigma, precompute=pre, dtype=torch.float32, graphs=graphs)
    torch.cuda.synchronize()
    pre32 = time.perf_counter() - t0
    # nodes! on the same nodes (device binning, records): the per-update cost
    # of a trajectory that changes between transforms
    t_nodes = float("inf")
    for _ in range(3):
        t0 = time.perf_counter()
        p32.nodes_(k)
        torch.cuda.synchronize()
        t_nodes = min(t_nodes, time.perf_counter() - t0)
    if args.tune:
        log(f"bench: tuned FP32 tiles {p32.tune()}")
    Q = int(np.prod(p32.Ntilde))
    f32 = torch.as_tensor(f_h, device=dev)
    fh32 = torch.as_tensor(fh_h, device=dev)
    y32 = torch.empty(M, dtype=torch.complex64, device=dev)
    x32 = torch.empty(N, dtype=torch.complex64, device=dev)
    kb32 = torch.empty((M, D), dtype=torch.float64, device=dev)

    # the plan's grid: after a forward it holds FFT(deconvolve(f)), from which
    # the node gradient needs only its gather (an rrule's pullback reusing its
    # forward pass)
    g32 = ctypes.c_void_p()
    L.nfftm_grid(p32._handle, ctypes.byref(g32))

    def step32():
        p32.forward(f32, y32)
        if grad:  # node cotangent of ||A(k) f - y||^2 with y := fHat (same buffers)
            L.nfftm_grad_nodes_grid(p32._handle, g32, ctypes.c_void_p(fh32.data_ptr()),
                                    ctypes.c_void_p(kb32.data_ptr()))
        p32.adjoint(fh32, x32)

    for _ in range(args.warmup):
        step32()
    torch.cuda.synchronize()
    log(f"bench: timing {args.steps} FP32 steps")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step32()
    torch.cuda.synchronize()
    t32 = (time.perf_counter() - t0) / args.steps

    # per-stage device time: HIP events on the plan's stream, same workload
    L.nfftm_profile(p32._handle, 1)
    for _ in range(args.steps):
        step32()
    L.nfftm_profile(p32._handle, 0)
    ms = (ctypes.c_double * 8)()
    cnt = (ctypes.c_int64 * 8)()
    L.nfftm_profile_read(p32._handle, ms, cnt)
    # mean device time per launch of each stage (the FFT also runs inside the
    # node gradient, so its count is per launch, not per step)
    stage_ms = {s: (ms[i] / cnt[i] if cnt[i] else None) for i, s in enumerate(STAGES)}

    # FP64 plan on the same workload
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    p64 = nfftm.plan_nfft(k, N, m=m, sigma=sigma, precompute=pre, dtype=torch.float64, graphs=graphs)
    torch.cuda.synchronize()
    pre64 = time.perf_counter() - t0
    if args.tune:
        log(f"bench: tuned FP64 tiles {p64.tune()}")
    f64 = f32.to(torch.complex128)
    fh64 = fh32.to(torch.complex128)
    y64 = torch.empty(M, dtype=torch.complex128, device=dev)
    x64 = torch.empty(N, dtype=torch.complex128, device=dev)
    kb64 = torch.empty((M, D), dtype=torch.float64, device=dev)

    g64 = ctypes.c_void_p()
    L.nfftm_grid(p64._handle, ctypes.byref(g64))

    def step64():
        p64.forward(f64, y64)
        if grad:
            L.nfftm_grad_nodes_grid(p64._handle, g64, ctypes.c_void_p(fh64.data_ptr()),
                                    ctypes.c_void_p(kb64.data_ptr()))
        p64.adjoint(fh64, x64)

    for _ in range(max(1, args.warmup)):
        step64()
    torch.cuda.synchronize()
    n64 = max(3, args.steps // 2)
    t0 = time.perf_counter()
    for _ in range(n64):
        step64()
    torch.cuda.synchronize()
    t64 = (time.perf_counter() - t0) / n64

    # the reference GPU backend's algorithm (FULL weights as a device sparse
    # matrix, gather / atomic-scatter products) on the same workload
    sparse = None
    if args.sparse_baseline:
        sparse = {}
        for name, rdt, cdt, fin, fhin in (("f32", torch.float32, torch.complex64, f32, fh32),
                                          ("f64", torch.float64, torch.complex128, f64, fh64)):
            log(f"bench: sparse-matrix baseline ({name}, precompute FULL)")
            t_build = time.perf_counter()
            psp = nfftm.plan_nfft(k, N, m=m, sigma=sigma, precompute=nfftm.FULL, dtype=rdt,
                                  sparse_matrix=True)
            torch.cuda.synchronize()
            t_build = time.perf_counter() - t_build
            ysp = torch.empty(M, dtype=cdt, device=dev)
            xsp = torch.empty(N, dtype=cdt, device=dev)

            def stepsp():
                psp.forward(fin, ysp)
                psp.adjoint(fhin, xsp)

            stepsp()
            torch.cuda.synchronize()
            nsp = max(3, args.steps // 4)
            t0 = time.perf_counter()
            for _ in range(nsp):
                stepsp()
            torch.cuda.synchronize()
            tsp = (time.perf_counter() - t0) / nsp
            step64()
            torch.cuda.synchronize()
            sparse[name] = {
                "value": M / tsp, "ms_per_step": tsp * 1e3, "steps": nsp, "plan_s": t_build,
                "matrix_bytes": M * (2 * m) ** D * (4 + (8 if name == "f64" else 4)),
                "fwd_vs_f64": float(torch.linalg.norm(ysp.to(torch.complex128) - y64) / torch.linalg.norm(y64)),
                "adj_vs_f64": float(torch.linalg.norm(xsp.to(torch.complex128) - x64) / torch.linalg.norm(x64)),
                "default_f32_over_this": tsp / t32}
            del psp, ysp, xsp
            torch.cuda.empty_cache()

    # accuracy: FP32 vs FP64 plan (full), and vs the FP64 NDFT on samples
    step32()
    step64()
    torch.cuda.synchronize()
    y32n, x32n = y32.cpu().numpy(), x32.cpu().numpy()
    y64n, x64n = y64.cpu().numpy(), x64.cpu().numpy()
    err = {
        "fwd_f32_vs_f64": float(np.linalg.norm(y32n - y64n) / np.linalg.norm(y64n)),
        "adj_f32_vs_f64": float(np.linalg.norm(x32n - x64n) / np.linalg.norm(x64n)),
    }
    if grad:
        a, b = kb32.cpu().numpy(), kb64.cpu().numpy()
        err["grad_f32_vs_f64"] = float(np.linalg.norm(a - b) / np.linalg.norm(b))
    import oracle

    # the direct sums cost P (resp. M) terms per sample: keep them to seconds
    nfwd = min(args.ndft_fwd, max(16, int(2e9 // P)))
    nadj = min(args.ndft_adj, max(16, int(2e9 // M)))
    sel = np.sort(rng.choice(M, size=nfwd, replace=False))
    ref = oracle.ndft(k, f_h.astype(np.complex128), sel)
    err["fwd_f32_vs_ndft"] = float(np.linalg.norm(y32n[sel] - ref) / np.linalg.norm(ref))
    err["fwd_f64_vs_ndft"] = float(np.linalg.norm(y64n[sel] - ref) / np.linalg.norm(ref))
    gsel = np.sort(rng.choice(P, size=nadj, replace=False))
    ref = oracle.ndft_adjoint(k, N, fh_h.astype(np.complex128), gsel)
    err["adj_f32_vs_ndft"] = float(np.linalg.norm(x32n.ravel()[gsel] - ref) / np.linalg.norm(ref))
    err["adj_f64_vs_ndft"] = float(np.linalg.norm(x64n.ravel()[gsel] - ref) / np.linalg.norm(ref))
    err["samples"] = {"fwd_nodes": nfwd, "adj_grid_points": nadj}
    accuracy_ok = all(v <= F32_TOL for key, v in err.items() if "f32" in key)

    # roofline of the dominant hand-written kernel
    sc = 8
    sb = stage_bytes(D, P, Q, M, sc)
    mine = ("deconvolve", "interpolation", "spreading", "deconvolve_transpose", "fold",
            "node_gradient")
    dom = max((s for s in mine if stage_ms[s]), key=lambda s: stage_ms[s])
    achieved = sb[dom] / (stage_ms[dom] * 1e-3)
    digest = source_digest()
    tdigest = source_digest(TRAFFIC_SOURCES)
    meas = None if args.tune else MEASURED_TRAFFIC.get((args.workload, dom))
    traffic = meas[0] if meas is not None and meas[1] == tdigest else None
    roofline = {"bound": "hbm", "kernel": dom, "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": traffic,
                "bytes_per_launch": sb[dom], "avg_ms": stage_ms[dom],
                "traffic_over_algorithmic": None if traffic is None else traffic / sb[dom],
                "traffic_source_digest": None if meas is None else meas[1], "gridding_source_digest": tdigest,
                "source_digest": digest}
    B_total = (P * sc + 4 * Q * sc + M * (D * 8 + sc)) + (2 * P * sc + 3 * Q * sc + M * (D * 8 + sc))

    cpu = None
    if not args.no_cpu:
        threads = oracle.lib().orc_num_threads()
        log(f"bench: CPU baseline (oracle restatement, {threads} threads)")
        op = oracle.OraclePlan(k, N, m, sigma, getattr(oracle, pre_name), True, workers=threads)
        fc = f_h.astype(np.complex128)
        fhc = fh_h.astype(np.complex128)
        op.forward(fc)
        op.adjoint(fhc)
        best = float("inf")
        reps = 0
        tstart = time.perf_counter()
        while reps < 5 and (reps < 2 or time.perf_counter() - tstart < 20):
            t0 = time.perf_counter()
            op.forward(fc)
            op.adjoint(fhc)
            best = min(best, time.perf_counter() - t0)
            reps += 1
        cpu = {"value": M / best, "unit": "NU points/s", "cores": threads, "kind": "port",
               "cpu_model": cpu_model(),
               "sample": f"full workload ({desc}; FP64 blocked {pre_name} plan), min of {reps} "
                         "forward+adjoint runs, FFT by scipy.fft"}

    metric = {"3d": "NU points/s fwd+adj 3D 128³ M=2²¹: FP32 vs FP64 GPU at rel-l2≤1e-5; HBM %",
              "1d": "NU points/s fwd+adj 1D N=M=2²⁴ edge-clustered, m=6 σ=1.25 TENSOR (FP32)",
              "2d-spiral": "NU points/s fwd+adj+node gradient 2D 256² spiral M=65536 (FP32)"}[args.workload]
    out = {
        "metric": metric,
        "value": M / t32,
        "unit": "NU points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t32 * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": desc, "N": list(N), "M": M, "m": m, "sigma": sigma,
                   "precompute": pre_name, "hip_graphs": graphs},
        "hbm_frac_total": B_total / t32 / HBM_PEAK,
        "fp64": {"value": M / t64, "ms_per_step": t64 * 1e3},
        "f32_over_f64": t64 / t32,
        "errors": err,
        "accuracy_ok": accuracy_ok,
        "plan_s": {"f32": pre32, "f64": pre64, "nodes_f32": t_nodes},
        "stage_ms": stage_ms,
        "tiles": {"f32": list(p32.tile_shape()), "f64": list(p64.tile_shape())},
        "roofline": roofline,
        "roofline_lds": lds_roofline(L, dom, stage_ms[dom], D, 2 * m, M, True),
        "roofline_lds_interpolation": lds_roofline(L, "interpolation", stage_ms["interpolation"], D, 2 * m, M,
                                                   True),
        "cpu_baseline": cpu,
    }
    if sparse is not None:
        out["sparse_matrix_baseline"] = sparse
    print(json.dumps(out))
    if not accuracy_ok:
        log(f"bench: FP32 error above {F32_TOL:g}: {err}")
        sys.exit(3)
    if args.copy_calib:
        # a device-to-device copy of exactly 256 MiB read + 256 MiB written:
        # the counters it produces calibrate FETCH_SIZE / WRITE_SIZE per byte
        src = torch.ones(2 ** 26, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.neg(src, out=dst)  # one elementwise kernel: reads 256 MiB, writes 256 MiB
        torch.cuda.synchronize()


if __name__ == "__main__":
    main()
