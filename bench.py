#!/usr/bin/env python3
"""Benchmark: Rayleigh–Sommerfeld far field, forward + adjoint, on one MI355X.

Workload (the headline shape): 303 x 303 = 91,809 pillars at 0.35 um, lambda 0.633,
f = 100 um hyperbolic phase with +-0.05 rad noise and |t| ~ U[0.7, 1.0], a 161 x 161
focal grid over +-4 lambda/NA at z = f.  One step = the forward sum (complex field)
plus the adjoint sum (gradient of -encircled-energy with respect to every pillar's
complex transmission), FP32-compensated, inputs resident in HBM.

Prints one JSON line.  Throughput = pillar-pixel pairs / (forward + adjoint time),
in G pairs/s.  Roofline: 21 + 21 = 42 FLOP per pair (dx, dy, R^2, sqrt, 1/R, k R,
sin, cos, complex multiply, 1/R scale, accumulate; same count transposed) against
the MI355X FP32 vector peak of 157.3 TFLOP/s, using the two pair-sum kernels'
durations measured with HIP events on their stream.  That peak is one wave64 FMA per
2 cycles per SIMD; without packed FP32 (switched off in this build) an FMA whose two
multiplicands sit in different VGPR banks issues at 2.2 cycles per SIMD
(celeris_amd/_lib/issue_probe), so the peak is within reach of unpacked code.

`--kernel rsi` runs the same step with the Rayleigh-Sommerfeld I kernel
exp(ikR)/R (z/R)(1/R - ik), which the reference does not have: 28 FLOP per pair each
way (the obliquity z/R, the complex factor (1/R - ik) and a complex instead of a real
scale), 56 per pair; no CPU baseline is reported for it.
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import math
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FLOP_PER_PAIR = 42.0
FLOP_PER_PAIR_RSI = 56.0
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--grid", type=int, default=161)
    ap.add_argument("--mode", default="fp32c", choices=["fp32c", "fp64", "fp32"])
    ap.add_argument("--kernel", default="ref", choices=["ref", "rsi"],
                    help="ref: the reference's exp(ikR)/R (default); rsi: Rayleigh-Sommerfeld I "
                         "exp(ikR)/R (z/R)(1/R - ik), no CPU baseline")
    ap.add_argument("--no-compare", action="store_true", help="skip FP64/FP32 GPU comparison legs")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="approximate CPU time budget of the cpu_baseline sample")
    ap.add_argument("--no-tune", action="store_true", help="skip the launch-split autotuner")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="also run a plain device copy of BYTES (reference for HBM counters)")
    ap.add_argument("--lens", default="headline", choices=["headline", "config1"],
                    help="headline: 91,809 pillars (303 x 303 at 0.35 um, lambda 0.633, f 100) "
                         "on a --grid focal grid; config1: the small case, 4,096 pillars (64 x 64 "
                         "at 0.4 um, lambda 0.532, f 50) on 81 x 81 over +-5 um")
    ap.add_argument("--psfbench", action="store_true",
                    help="run the reference's psfbench instead (cli/bench.cpp:96-170)")
    ap.add_argument("--diameter", type=float, default=120.0, help="psfbench lens diameter (um)")
    ap.add_argument("--focal", type=float, default=50.0, help="psfbench focal length (um)")
    return ap.parse_args()


def cpu_model() -> str:
    try:
        out = subprocess.run(["lscpu"], stdout=subprocess.PIPE, text=True).stdout
        for ln in out.splitlines():
            if ln.startswith("Model name"):
                return ln.split(":", 1)[1].strip()
    except Exception:
        pass
    return platform.processor()


def _every(stride: int, what: str) -> str:
    if stride == 1:
        return f"every {what}"
    suffix = "th" if 10 <= stride % 100 <= 20 else {1: "st", 2: "nd", 3: "rd"}.get(stride % 10, "th")
    return f"every {stride}{suffix} {what}"


def cpu_baseline(lens, u_np, budget_s: float) -> dict:
    """The oracle (FP64 restatement of focal.cpp:205-238 + its transposed adjoint) on
    the host cores: rows striped over threads for the forward, pillars for the
    adjoint, each on a bounded deterministic sample scaled by its pair count."""
    import oracle

    threads = oracle.num_cpus()
    planes = lens.planes
    n = planes[0].nx
    # Calibrate on one row / a sparse pillar sample, then size the sample to the budget.
    t0 = time.perf_counter()
    p_cal = oracle.forward_pairs(lens.px, lens.py, lens.t, planes, threads, row_stride=n)
    t_cal = time.perf_counter() - t0
    # One row runs on one thread; the per-pair rate with all threads is ~threads x.
    per_pair_1t = t_cal / max(p_cal, 1)
    total_pairs = lens.npil * n * n
    est_full = per_pair_1t * total_pairs / threads
    frac = min(1.0, 0.5 * budget_s / max(est_full, 1e-9))
    row_stride = max(1, int(round(1.0 / frac)))
    t0 = time.perf_counter()
    p_fwd = oracle.forward_pairs(lens.px, lens.py, lens.t, planes, threads, row_stride=row_stride)
    t_fwd = time.perf_counter() - t0
    pil_stride = row_stride
    t0 = time.perf_counter()
    _, p_adj = oracle.adjoint(lens.px, lens.py, u_np, planes, 1, threads, pillar_stride=pil_stride,
                              return_pairs=True)
    t_adj = time.perf_counter() - t0
    tau = t_fwd / p_fwd + t_adj / p_adj  # seconds per pair, forward + adjoint
    # One thread, for reference: a few rows forward and a pillar sample adjoint.
    t0 = time.perf_counter()
    p1 = oracle.forward_pairs(lens.px, lens.py, lens.t, planes, 1, row_stride=max(1, n // 4))
    t1 = time.perf_counter() - t0
    t0 = time.perf_counter()
    _, p1a = oracle.adjoint(lens.px, lens.py, u_np, planes, 1, 1, pillar_stride=2000,
                            return_pairs=True)
    t1a = time.perf_counter() - t0
    return {
        "value": 1.0 / tau / 1e9,
        "unit": "G pairs/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"oracle FP64 forward on {_every(row_stride, 'focal row')} ({p_fwd:.3e} pairs, "
                   f"{t_fwd:.2f} s) + adjoint on {_every(pil_stride, 'pillar')} ({p_adj:.3e} "
                   f"pairs, {t_adj:.2f} s), {threads} threads, scaled per pair"),
        "cpu_model": cpu_model(),
        "fwd_s_per_pair": t_fwd / p_fwd,
        "adj_s_per_pair": t_adj / p_adj,
        "single_thread_G_pairs_s": 1.0 / (t1 / p1 + t1a / p1a) / 1e9,
    }


def reference_cpu_forward(lens, budget_s: float):
    """The reference's own CPU forward, celeris::propagate_pillars (focal.cpp:205-238,
    rows over std::thread::hardware_concurrency() workers; built from its sources into
    oracle/_ref by oracle/ref.mk with the reference's Release flags), on every k-th pillar
    of the workload over the full focal grid, k sized from a calibration run so the timed
    run takes about budget_s.  Returns (pairs, seconds, k) or None where that build is
    absent."""
    import oracle

    if not oracle.ref_available():
        return None
    plane = lens.planes[0]
    n = plane.nx
    W = plane.sx * (n - 1) / 2.0
    cx, cy = plane.x0 + W, plane.y0 + W
    lam = lens.wavelengths[0]

    def run(k):
        px, py, t = lens.px[::k], lens.py[::k], lens.t.reshape(-1)[: lens.npil][::k]
        t0 = time.perf_counter()
        oracle.ref_propagate_pillars(px, py, t, cx, cy, plane.z0, lam, n, W)
        return px.size * n * n, time.perf_counter() - t0

    p_cal, t_cal = run(64)
    k = max(1, int(round(64 * t_cal / max(budget_s, 1e-3))))
    pairs, secs = run(k)
    return pairs, secs, k


def profiled_traffic(launches):
    """HBM bytes of one forward + one adjoint pair-kernel launch, from the rocprofv3
    FETCH_SIZE / WRITE_SIZE summary in profiles/ (scaled by its copy-kernel calibration)
    for exactly these kernels and work-group counts; (None, None) if not profiled."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            prof = json.load(f)
    except (OSError, ValueError):
        return None, None
    total = 0.0
    for name, groups in launches.values():
        entry = prof.get("kernels", {}).get(f"{name}|{groups}")
        if entry is None:
            return None, None
        total += entry["read_bytes"] + entry["write_bytes"]
    return total, os.path.relpath(path, ROOT)


def psfbench_cpu_baseline(L, max_pairs: float = 6e9) -> tuple:
    """psfbench's CPU leg, measured for real: the oracle's FP64 restatement of
    compute_psf's CPU loop (focal.cpp:205-238, rows over every host thread).  Maps above
    max_pairs pairs are timed on every k-th row and the time scaled by the row count.
    Returns (ms for the whole map, threads, |E|^2 with unsampled rows left 0, k)."""
    import oracle
    from celeris_amd import Plane

    threads = oracle.num_cpus()
    plane = Plane.psf(0.0, 0.0, L.focal_um, 2 * math.pi / L.wavelength_um, L.grid, L.half_window_um)
    stride = max(1, int(math.ceil(L.px.size * L.grid * L.grid / max_pairs)))
    t0 = time.perf_counter()
    _, inten = oracle.forward(L.px, L.py, L.t, [plane], threads=threads, intensity=True,
                              row_stride=stride)
    ms = (time.perf_counter() - t0) * 1e3
    rows = len(range(0, L.grid, stride))
    return ms * L.grid / rows, threads, inten.reshape(L.grid, L.grid), stride


def run_psfbench(args) -> None:
    """The reference's psfbench (cli/bench.cpp:96-170) on this build: same lens
    construction, grid and window; the CPU leg timed on the CPU and compared against
    (the reference timed its GPU path there and compared GPU with GPU)."""
    from celeris_amd import psfbench as pb

    print("PSF propagation benchmark (building lens...)")
    L = pb.psfbench_lens(args.diameter, args.focal, args.grid)
    print(f"  {L.lens.n_cells} x {L.lens.n_cells} cells, {L.px.size} pillars in aperture, "
          f"{L.grid}x{L.grid} focal grid")
    pairs = L.px.size * L.grid * L.grid
    gpu = pb.run_gpu(L)
    cpu_ms, threads, cpu_I, stride = psfbench_cpu_baseline(L)
    rows = slice(0, L.grid, stride)  # agreement over the rows the CPU leg computed
    print("")
    sample = "" if stride == 1 else f", {_every(stride, 'row')}, scaled"
    print(f"  CPU (double, parallel) : {cpu_ms:8.1f} ms   ({threads} threads, "
          f"{pairs / cpu_ms / 1e6:.2f} G pairs/s{sample})")
    rec = {"workload": "psfbench", "pillars": int(L.px.size), "grid": L.grid,
           "pairs": int(pairs), "diameter_um": L.diameter_um, "focal_um": L.focal_um,
           "wavelength_um": L.wavelength_um, "half_window_um": L.half_window_um,
           "library": "stand-in (no RCWA): 16 fills, phase ramp over 2 pi, |t| 0.86-0.95",
           "cpu_baseline": {"ms": cpu_ms, "threads": threads, "kind": "port",
                            "row_stride": stride, "cpu_model": cpu_model()},
           "gpu": {}}
    for mode, label in (("fp32", "float kernel"), ("fp32c", "FP32-comp"), ("fp64", "double")):
        g = gpu[mode]
        diff = pb.max_normalized_diff(cpu_I[rows], g["intensity"][rows])
        print(f"  GPU ({label:<12}) : {g['call_ms']:8.1f} ms   ({cpu_ms / g['call_ms']:.1f}x faster)"
              f"  kernel {g['kernel_ms']:.3f} ms = {pairs / g['kernel_ms'] / 1e6:.0f} G pairs/s")
        print(f"  agreement ({mode:<5})      : max|d normalized PSF| = {diff:.2e}  ok=true")
        rec["gpu"][mode] = {"call_ms": g["call_ms"], "kernel_ms": g["kernel_ms"],
                            "kernel_G_pairs_s": pairs / g["kernel_ms"] / 1e6,
                            "max_normalized_diff_vs_cpu": diff}
    print(json.dumps(rec))


def main():
    args = parse()
    if args.psfbench and args.gpus == 1:
        run_psfbench(args)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "the forward shards over pixels and the adjoint over pillars with no exchange "
              "beyond one gradient all-reduce, but no multi-GPU figure is measured here.",
              file=sys.stderr)
        sys.exit(2)
    import torch

    from celeris_amd import KERNEL_RSI, Context, configs
    from celeris_amd.autograd import encircled_energy

    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)
    dev = torch.device("cuda:0")
    lens = configs.config2(n=args.grid) if args.lens == "headline" else configs.config1()
    plane = lens.planes[0]
    flop_per_pair = FLOP_PER_PAIR
    if args.kernel == "rsi":
        plane = dataclasses.replace(plane, kernel=KERNEL_RSI)
        flop_per_pair = FLOP_PER_PAIR_RSI
    n_grid = plane.nx
    npix = plane.npix
    pairs = lens.npil * npix
    px = torch.from_numpy(lens.px).to(dev)
    py = torch.from_numpy(lens.py).to(dev)
    t = torch.from_numpy(lens.t).to(dev)
    ctx = Context(0)
    r0 = 1.22 * lens.wavelengths[0] / lens.na

    # Cotangent of -EE at the initial design: u = dL/dRe E + i dL/dIm E = 2 w E.
    E0, _ = ctx.forward(px, py, t, [plane], mode="fp64")
    Ev = E0.clone().requires_grad_(True)
    inten = Ev.real ** 2 + Ev.imag ** 2
    loss = -encircled_energy(inten, plane, r0)
    loss.backward()
    u = Ev.grad.detach().contiguous()

    tuning = None
    if not args.no_tune:
        from celeris_amd.tune import autotune

        tuning = autotune(ctx, px, py, t, [plane], u, mode=args.mode, use_cache=False)
        print(f"[bench] tuned: forward {tuning.get('fwd_shape')} x {tuning['fwd']} work-group target, "
              f"adjoint {tuning.get('adj_shape')} x {tuning.get('adj')} (0: built-in size)",
              file=sys.stderr, flush=True)

    def step(mode):
        f, _ = ctx.forward(px, py, t, [plane], mode=mode, field=True)
        g = ctx.adjoint(px, py, u, [plane], nsets=1, mode=mode)
        return f, g

    def time_steps(mode, steps, warmup):
        for _ in range(warmup):
            step(mode)
        torch.cuda.synchronize()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(steps):
            step(mode)
        b.record()
        torch.cuda.synchronize()
        return a.elapsed_time(b) / steps

    def kernel_ms(mode, reps):
        ctx.set_timing(True)
        fw, ad = [], []
        for _ in range(reps):
            ctx.forward(px, py, t, [plane], mode=mode, field=True)
            fw.append(ctx.last_kernel_ms(False))
            ctx.adjoint(px, py, u, [plane], nsets=1, mode=mode)
            ad.append(ctx.last_kernel_ms(True))
        ctx.set_timing(False)
        return float(np.median(fw)), float(np.median(ad))

    ms = time_steps(args.mode, args.steps, args.warmup)
    k_fwd, k_adj = kernel_ms(args.mode, max(10, min(args.steps, 20)))
    launches = {"fwd": ctx.last_launch(False), "adj": ctx.last_launch(True)}
    traffic, traffic_src = profiled_traffic(launches)
    value = pairs / (ms * 1e-3) / 1e9
    achieved = pairs * flop_per_pair / ((k_fwd + k_adj) * 1e-3) / 1e12
    print(f"[bench] {args.mode}: {ms:.3f} ms/step, kernels fwd {k_fwd:.3f} ms adj {k_adj:.3f} ms, "
          f"{value:.1f} G pairs/s", file=sys.stderr, flush=True)

    extra = {}
    f32c, g32c = step(args.mode)
    torch.cuda.synchronize()
    if not args.no_compare:
        # Accuracy against the FP64 GPU kernel (itself pinned to the oracle by the tests),
        # and the plain FP64 / FP32 GPU kernels' throughput on the same step.
        f64, g64 = step("fp64")
        torch.cuda.synchronize()
        rel = lambda a, b: float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))
        extra["rel_l2_field_vs_fp64"] = rel(f32c, f64)
        extra["rel_l2_intensity_vs_fp64"] = rel(f32c.abs() ** 2, f64.abs() ** 2)
        extra["max_abs_field_vs_fp64_over_peak"] = float((f32c - f64).abs().max() / f64.abs().max())
        extra["rel_l2_grad_vs_fp64"] = rel(g32c, g64)
        ms64 = time_steps("fp64", 2, 1)
        ms32 = time_steps("fp32", 3, 1)
        extra["fp64_plain_G_pairs_s"] = pairs / (ms64 * 1e-3) / 1e9
        extra["fp32_plain_G_pairs_s"] = pairs / (ms32 * 1e-3) / 1e9
        extra["speedup_vs_fp64_plain"] = ms64 / ms
        print(f"[bench] compare: relL2 field {extra['rel_l2_field_vs_fp64']:.3e} grad "
              f"{extra['rel_l2_grad_vs_fp64']:.3e}; fp64 {ms64:.2f} ms, fp32 {ms32:.2f} ms",
              file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu and args.kernel == "ref":
        port = cpu_baseline(lens, u.cpu().numpy(), args.cpu_seconds)
        cpu = port
        ref = reference_cpu_forward(lens, 0.5 * args.cpu_seconds)
        if ref is not None:
            # The reference's forward per pair, with the adjoint leg (which the reference
            # does not have) timed on the oracle's FP64 restatement.
            p_ref, t_ref, k_ref = ref
            tau = t_ref / p_ref + port["adj_s_per_pair"]
            cpu = {
                "value": 1.0 / tau / 1e9,
                "unit": "G pairs/s",
                "cores": port["cores"],
                "kind": "reference",
                "sample": (f"reference propagate_pillars (focal.cpp:205-238, oracle/_ref build, "
                           f"{port['cores']} hardware threads) on {_every(k_ref, 'pillar')} x "
                           f"{n_grid}^2 ({p_ref:.3e} pairs, {t_ref:.2f} s) for the forward; the "
                           f"adjoint, absent from the reference, from the oracle's FP64 "
                           f"restatement ({port['adj_s_per_pair'] * 1e9:.3f} ns per pair); "
                           f"scaled per pair"),
                "cpu_model": port["cpu_model"],
                "fwd_s_per_pair": t_ref / p_ref,
                "adj_s_per_pair": port["adj_s_per_pair"],
                "port": {k: port[k] for k in ("value", "kind", "sample", "fwd_s_per_pair",
                                               "adj_s_per_pair", "single_thread_G_pairs_s")},
            }
        print(f"[bench] cpu baseline {cpu['value']:.3f} G pairs/s ({cpu['kind']}) on {cpu['cores']} "
              f"threads; port {port['value']:.3f}", file=sys.stderr, flush=True)

    if args.calibrate_copy > 0:
        n = args.calibrate_copy // 4
        src = torch.ones(n, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()

    out = {
        "metric": "G pillar·pixel pairs/s fwd+adjoint, FP32-comp vs FP64 at ≤1e-5 rel L2",
        "value": value,
        "unit": "G pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.mode != "fp64" else "f64",
        "data": "synthetic",
        "config": {
            "workload": (f"RS far field fwd+adjoint, {lens.npil} pillars "
                         + ("(303x303 @0.35um, lambda 0.633, f 100)" if args.lens == "headline"
                            else "(64x64 @0.4um, lambda 0.532, f 50)")
                         + f" x {plane.nx}x{plane.ny} focal grid, mode {args.mode}"
                         + (", RS-I kernel" if args.kernel == "rsi" else "")),
            "pillars": lens.npil,
            "grid": plane.nx,
            "pairs_per_step": pairs,
            "mode": args.mode,
            "kernel": args.kernel,
            "parallelism": "single GPU",
            "tuning": tuning and {"fwd_target_blocks": tuning["fwd"],
                                  "adj_target_blocks": tuning.get("adj"),
                                  "fwd_block": tuning.get("fwd_shape"),
                                  "adj_block": tuning.get("adj_shape")},
        },
        "roofline": {
            "bound": "valu",
            "achieved": achieved,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            "traffic": traffic,
            "traffic_source": traffic_src,
            "kernels": {launches["fwd"][0]: k_fwd, launches["adj"][0]: k_adj},
            "workgroups": {launches["fwd"][0]: launches["fwd"][1],
                           launches["adj"][0]: launches["adj"][1]},
            "flop_per_pair": flop_per_pair,
        },
        "cpu_baseline": cpu,
        **extra,
    }
    print(json.dumps(out))
    ctx.close()


if __name__ == "__main__":
    main()
