"""Autotuner for the FP32C block shape and the launch size of the forward and adjoint sums.

Two knobs decide how a call maps onto the 256 CUs:

* the block shape (FP32C only): pixels per forward lane (4 x 4, 4 x 8 or 8 x 4) and per
  adjoint block (4 or 8 each way).  Larger blocks spread each pillar's FP64 distance
  work over more pairs but hold more accumulators (the 4 x 8 forward runs 3 waves per
  SIMD instead of 4); a plane whose pair-error bound does not admit a shape falls back to
  4 x 4, then to pair-by-pair evaluation, so every candidate is exact to the same bound;
* the work-group count: the pieces a call's work is cut into (FP32C: pieces of falling
  cost; FP32 / FP64: pillar chunks or pixel-row runs).  Few pieces write few partial
  sums; more pieces let the CUs that finish early take more work.

Each candidate's whole call (staging, pair kernels, reduction) is timed with stream
events over back-to-back calls.  Among the candidates within TIE of the fastest, the one
with the fewest work-groups wins (fewer partial slots to write and reduce for the same
time); it is installed on the context (crs_set_block_shape, crs_set_tuning) and
remembered in a JSON cache keyed by the shape of the problem.
"""
from __future__ import annotations

import json
import os
from typing import Dict, Sequence

from .rs import Context, Plane

TARGETS = (0, 2048, 4096, 8192, 16384)  # 0: the library's built-in size
TIE = 0.005  # candidates this close to the fastest count as equally fast
FWD_SHAPES = ((4, 4), (4, 8), (8, 4))
ADJ_SHAPES = ((4, 4), (4, 8), (8, 4), (8, 8))
CACHE = os.environ.get("CELERIS_RS_TUNE_CACHE",
                       os.path.join(os.path.expanduser("~"), ".cache", "celeris_rs_tune.json"))


def shape_key(mode: str, npil: int, nsets: int, planes: Sequence[Plane], arch: str) -> str:
    pix = sorted((p.nx, p.ny) for p in planes)
    return f"{arch}|{mode}|{npil}|{nsets}|{len(planes)}|{pix[0]}|{pix[-1]}|{sum(a * b for a, b in pix)}"


def _load_cache() -> Dict[str, dict]:
    try:
        with open(CACHE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store_cache(cache: Dict[str, dict]) -> None:
    try:
        os.makedirs(os.path.dirname(CACHE), exist_ok=True)
        tmp = CACHE + ".tmp"
        with open(tmp, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
        os.replace(tmp, CACHE)
    except OSError:
        pass


def _install(ctx: Context, mode: str, best: dict) -> None:
    if mode == "fp32c":
        ctx.set_block_shape(fwd=tuple(best["fwd_shape"]), adj=tuple(best.get("adj_shape", (0, 0))))
    ctx.set_tuning(mode, best["fwd"], best.get("adj", 0))


def pick(table: Dict[str, dict], tie: float = TIE) -> dict:
    """The candidate with the fewest work-groups among those within `tie` of the fastest
    (then the fastest of those)."""
    fastest = min(r["ms"] for r in table.values())
    near = [r for r in table.values() if r["ms"] <= fastest * (1.0 + tie)]
    return min(near, key=lambda r: (r.get("groups", 0), r["ms"]))


def autotune(ctx: Context, px, py, t, planes: Sequence[Plane], u=None, mode: str = "fp32c",
             targets: Sequence[int] = TARGETS, reps: int = 10, use_cache: bool = True) -> dict:
    """Time the forward (and, if a cotangent u is given, the adjoint) of this problem for
    each candidate (block shape x work-group target; targets only for FP32 / FP64),
    install the fastest on ctx and return {"fwd": target, "fwd_shape": (bx, by),
    "adj": target, "adj_shape": (bx, by), "fwd_ms": {...}, "adj_ms": {...}}; the timing
    tables are keyed "BXxBY/target" and name the kernel that ran."""
    import torch

    npil = px.numel()
    nsets = t.numel() // max(npil, 1)
    arch = torch.cuda.get_device_properties(px.device).gcnArchName.split(":")[0]
    key = shape_key(mode, npil, nsets, planes, arch)
    cache = _load_cache() if use_cache else {}
    if use_cache and key in cache:
        best = cache[key]
        _install(ctx, mode, best)
        return best

    def timed(fn):
        fn()  # plan
        torch.cuda.synchronize()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            fn()
        b.record()
        b.synchronize()
        return a.elapsed_time(b) / reps

    fwd_shapes = FWD_SHAPES if mode == "fp32c" else ((0, 0),)
    adj_shapes = ADJ_SHAPES if mode == "fp32c" else ((0, 0),)
    # The clocks ramp up over the first few milliseconds of work; without a warm-up the
    # first candidate reads several percent slow.
    ctx.set_tuning(mode, 0, 0)
    for _ in range(20):
        ctx.forward(px, py, t, planes, mode=mode)
    torch.cuda.synchronize()
    fwd_ms, adj_ms = {}, {}
    for shape in fwd_shapes:
        for target in targets:
            if mode == "fp32c":
                ctx.set_block_shape(fwd=shape, adj=(0, 0))
            ctx.set_tuning(mode, target, 0)
            ms = timed(lambda: ctx.forward(px, py, t, planes, mode=mode))
            name, groups = ctx.last_launch(False)
            fwd_ms[f"{shape[0]}x{shape[1]}/{target}"] = {"ms": ms, "kernel": name, "groups": groups,
                                                          "shape": shape, "target": target}
    if u is not None:
        for shape in adj_shapes:
            for target in targets:
                if mode == "fp32c":
                    ctx.set_block_shape(fwd=(0, 0), adj=shape)
                ctx.set_tuning(mode, 0, target)
                ms = timed(lambda: ctx.adjoint(px, py, u, planes, nsets=nsets, mode=mode))
                name, groups = ctx.last_launch(True)
                adj_ms[f"{shape[0]}x{shape[1]}/{target}"] = {"ms": ms, "kernel": name, "groups": groups,
                                                              "shape": shape, "target": target}
    bf = pick(fwd_ms)
    best = {"fwd": bf["target"], "fwd_shape": list(bf["shape"]), "fwd_ms": fwd_ms}
    if adj_ms:
        ba = pick(adj_ms)
        best.update({"adj": ba["target"], "adj_shape": list(ba["shape"]), "adj_ms": adj_ms})
    _install(ctx, mode, best)
    if use_cache:
        cache = _load_cache()
        cache[key] = best
        _store_cache(cache)
    return best
