"""Autotuner for the launch split of the forward and adjoint sums.

The forward splits its pillar sum into chunks and the adjoint its pixel sum into quad
runs so that a launch has enough work-groups for the 256 CUs; too few leaves CUs
idle in the last wave, too many adds partial sums (more traffic for the reduction
kernel) and per-block setup.  The best count depends on the shape (pillars, pixels
per plane, planes, mode), so it is measured: each candidate's whole call (staging,
pair kernel, reduction) is timed with stream events over back-to-back calls, and the
fastest is installed on the context (crs_set_tuning) and remembered in a JSON cache
keyed by the shape.
"""
from __future__ import annotations

import json
import os
from typing import Dict, Optional, Sequence

from .rs import Context, Plane

CANDIDATES = (1024, 2048, 3072, 4096, 6144, 8192, 12288, 16384, 24576)
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


def autotune(ctx: Context, px, py, t, planes: Sequence[Plane], u=None, mode: str = "fp32c",
             candidates: Sequence[int] = CANDIDATES, reps: int = 10,
             use_cache: bool = True) -> dict:
    """Time the forward (and, if a cotangent u is given, the adjoint) of this shape for
    each candidate work-group target, install the fastest on ctx, and return
    {"fwd": target, "adj": target, "fwd_ms": {...}, "adj_ms": {...}}."""
    import torch

    npil = px.numel()
    nsets = t.numel() // max(npil, 1)
    arch = torch.cuda.get_device_properties(px.device).gcnArchName.split(":")[0]
    key = shape_key(mode, npil, nsets, planes, arch)
    cache = _load_cache() if use_cache else {}
    if use_cache and key in cache:
        best = cache[key]
        ctx.set_tuning(mode, best["fwd"], best.get("adj", 0))
        return best
    fwd_ms, adj_ms = {}, {}

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

    # The clocks ramp up over the first few milliseconds of work; without a warm-up the
    # first candidate reads several percent slow.
    ctx.set_tuning(mode, candidates[0], candidates[0])
    for _ in range(20):
        ctx.forward(px, py, t, planes, mode=mode)
    torch.cuda.synchronize()
    for cand in candidates:
        ctx.set_tuning(mode, cand, cand)
        fwd_ms[cand] = timed(lambda: ctx.forward(px, py, t, planes, mode=mode))
        if u is not None:
            adj_ms[cand] = timed(lambda: ctx.adjoint(px, py, u, planes, nsets=nsets, mode=mode))
    best = {"fwd": min(fwd_ms, key=fwd_ms.get), "fwd_ms": fwd_ms}
    if adj_ms:
        best["adj"] = min(adj_ms, key=adj_ms.get)
        best["adj_ms"] = adj_ms
    ctx.set_tuning(mode, best["fwd"], best.get("adj", 0))
    if use_cache:
        cache = _load_cache()
        cache[key] = {k: v for k, v in best.items()}
        _store_cache(cache)
    return best
