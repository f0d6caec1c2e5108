"""Rayleigh–Sommerfeld far-field propagation on MI355X.

Two layers, mirroring include/celeris_rs.h:

* Drop-in functions with the reference's semantics (``celeris::cuda`` in
  src/celeris/cuda/propagate.hpp:16-38): :func:`available`, :func:`device_name`,
  :func:`propagate_psf`, :func:`propagate_zx`.  Host numpy arrays in, ``|E|^2``
  out; ``None`` where the reference returns ``false``.
* :class:`Context`: the batched device API (many planes / wavelengths /
  transmission sets per launch, complex field, adjoint, arithmetic modes) on
  torch tensors that live on the GPU.

Units are microns; ``k = 2*pi/lambda`` in rad/um.
"""
from __future__ import annotations

import ctypes
import math
import threading
from dataclasses import dataclass, replace
from typing import Iterable, Optional, Sequence

import numpy as np

from . import _lib
from ._lib import MODE_FP32, MODE_FP32C, MODE_FP64, MODES, crs_plane


# Plane kernels (crs_kernel): the reference's exp(ikR)/R, and Rayleigh-Sommerfeld I,
# exp(ikR)/R (z/R)(1/R - ik) (in FP32C in 4 x 4 blocks where its series bound admits
# them, else pair by pair at FP64 distances).
KERNEL_REF = 0
KERNEL_RSI = 1


def _mode(mode) -> int:
    if isinstance(mode, str):
        try:
            return MODES[mode]
        except KeyError:
            raise ValueError(f"unknown mode {mode!r}; expected one of {sorted(MODES)}") from None
    return int(mode)


@dataclass(frozen=True)
class Plane:
    """A focal plane: pixel (i, j) at fx = x0 + i*sx, fy = y0 + j*sy, z = z0 + j*sz."""

    k: float
    x0: float
    sx: float
    y0: float
    sy: float
    z0: float
    sz: float
    nx: int
    ny: int
    tset: int = 0
    offset: int = -1  # -1: placed after the previous plane by pack_planes
    kernel: int = 0   # KERNEL_REF: exp(ikR)/R (the reference's); KERNEL_RSI: RS-I, opt-in

    @property
    def npix(self) -> int:
        return self.nx * self.ny

    @staticmethod
    def _step(span: float, n: int) -> float:
        # The reference's step = 2W/(n-1) (propagate.cu:143, focal.cpp:223) in IEEE
        # arithmetic: n == 1 gives inf (NaN coordinates downstream), as there.
        return float(np.float64(span) / np.float64(n - 1)) if n == 1 else span / (n - 1)

    @classmethod
    def psf(cls, cx: float, cy: float, z: float, k: float, n: int, half_window: float,
            tset: int = 0, offset: int = -1) -> "Plane":
        """The n x n window of propagate_psf / propagate_pillars (focal.cpp:223-227)."""
        with np.errstate(divide="ignore", invalid="ignore"):
            step = cls._step(2.0 * half_window, n)
        return cls(k, cx - half_window, step, cy - half_window, step, z, 0.0, n, n, tset, offset)

    @classmethod
    def zx(cls, xmin: float, xmax: float, nx: int, zmin: float, zmax: float, nz: int, k: float,
           tset: int = 0, offset: int = -1) -> "Plane":
        """The y = 0 caustic slice of propagate_zx (propagate.cu:205-206, :101-102)."""
        return cls(k, xmin, (xmax - xmin) / (nx - 1), 0.0, 0.0, zmin, (zmax - zmin) / (nz - 1),
                   nx, nz, tset, offset)

    @classmethod
    def axis(cls, zmin: float, zmax: float, nz: int, k: float, tset: int = 0,
             offset: int = -1) -> "Plane":
        """On-axis z scan I(0, 0, z) (throughfocus.cpp:68-75, chromatic.cpp:30-37)."""
        sz = (zmax - zmin) / (nz - 1) if nz > 1 else 0.0
        return cls(k, 0.0, 0.0, 0.0, 0.0, zmin, sz, 1, nz, tset, offset)


def pack_planes(planes: Sequence[Plane]):
    """ctypes array of crs_plane with offsets resolved; returns (array, total_pixels)."""
    arr = (crs_plane * max(1, len(planes)))()
    cursor = 0
    total = 0
    for idx, p in enumerate(planes):
        off = cursor if p.offset < 0 else p.offset
        arr[idx] = crs_plane(p.k, p.x0, p.sx, p.y0, p.sy, p.z0, p.sz, p.nx, p.ny, p.tset, p.kernel, off)
        cursor = off + p.npix
        total = max(total, off + p.npix)
    return arr, total


def resolve_offsets(planes: Sequence[Plane]) -> list:
    out, cursor = [], 0
    for p in planes:
        off = cursor if p.offset < 0 else p.offset
        out.append(replace(p, offset=off))
        cursor = off + p.npix
    return out


# --------------------------------------------------------------------------------------
# Drop-in layer.

def available() -> bool:
    """celeris::cuda::available() (propagate.hpp:20)."""
    return bool(_lib.load().crs_available())


def device_name() -> str:
    """celeris::cuda::device_name() (propagate.hpp:21)."""
    return _lib.load().crs_device_name().decode()


def set_default_mode(mode) -> str:
    """Arithmetic used by propagate_psf / propagate_zx; returns the previous mode name."""
    prev = _lib.load().crs_set_default_mode(_mode(mode))
    if prev < 0:
        raise ValueError(f"unknown mode {mode!r}")
    return {v: k for k, v in MODES.items()}[prev]


def _host_inputs(px, py, t):
    px = np.ascontiguousarray(px, dtype=np.float64)
    py = np.ascontiguousarray(py, dtype=np.float64)
    t = np.ascontiguousarray(t, dtype=np.complex128)
    if not (px.ndim == py.ndim == t.ndim == 1 and px.shape == py.shape == t.shape):
        raise ValueError("px, py, t must be 1-D arrays of one length")
    dp = ctypes.POINTER(ctypes.c_double)
    return px, py, t, px.ctypes.data_as(dp), py.ctypes.data_as(dp), t.ctypes.data_as(dp)


def propagate_psf(px, py, t, cx: float, cy: float, z: float, k: float, n: int,
                  half_window: float) -> Optional[np.ndarray]:
    """celeris::cuda::propagate_psf (propagate.hpp:27-30): |E|^2 as an (n, n) array
    (row j = y), or None where the reference returns false."""
    px, py, t, ppx, ppy, ppt = _host_inputs(px, py, t)
    out = np.empty((max(n, 0), max(n, 0)), dtype=np.float64)
    ok = _lib.load().crs_propagate_psf(ppx, ppy, ppt, int(px.size), float(cx), float(cy), float(z),
                                      float(k), int(n), float(half_window),
                                      out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return out if ok else None


def propagate_zx(px, py, t, xmin: float, xmax: float, nx: int, zmin: float, zmax: float,
                 nz: int, k: float) -> Optional[np.ndarray]:
    """celeris::cuda::propagate_zx (propagate.hpp:35-38): |E|^2 as an (nz, nx) array
    in the plane y = 0, or None where the reference returns false."""
    px, py, t, ppx, ppy, ppt = _host_inputs(px, py, t)
    out = np.empty((max(nz, 0), max(nx, 0)), dtype=np.float64)
    ok = _lib.load().crs_propagate_zx(ppx, ppy, ppt, int(px.size), float(xmin), float(xmax),
                                     int(nx), float(zmin), float(zmax), int(nz), float(k),
                                     out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return out if ok else None


# --------------------------------------------------------------------------------------
# Batched device layer.

def _torch():
    import torch  # deferred: the drop-in layer needs no torch

    return torch


class Context:
    """A crs_context on one GPU; tensors passed in must live on that GPU.

    Work is enqueued on torch's current stream of the device unless ``stream`` (a
    torch.cuda.Stream) is given; calls return without synchronizing, like any torch op.
    A C context serves one host thread at a time; this wrapper takes a lock around every
    library call, so one Context may be shared by several threads (the autograd engine's
    backward thread and the caller's, say).
    """

    def __init__(self, device: int = 0):
        lib = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(lib.crs_context_create(int(device), ctypes.byref(h)), "crs_context_create")
        self._lock = threading.RLock()
        self._h = h
        self.device = int(device)

    def close(self) -> None:
        lock = getattr(self, "_lock", None)
        if lock is None:
            return
        with lock:
            if getattr(self, "_h", None):
                _lib.load().crs_context_destroy(self._h)
                self._h = None

    def _call(self, what: str, fn, *args) -> None:
        """fn(handle, *args) under the context's lock; raises on a failure status."""
        with self._lock:
            if not getattr(self, "_h", None):
                raise RuntimeError(f"{what}: the context is closed")
            _lib.check(fn(self._h, *args), what)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _stream(self, stream):
        torch = _torch()
        s = stream if stream is not None else torch.cuda.current_stream(self.device)
        return ctypes.c_void_p(s.cuda_stream)

    def _check_dev(self, name, x, dtype, min_numel: int = 0):
        """x must be a contiguous `dtype` tensor on this context's GPU with at least
        `min_numel` elements (the kernels read and write through its raw pointer)."""
        torch = _torch()
        if not isinstance(x, torch.Tensor) or x.device.type != "cuda" or x.device.index != self.device:
            raise ValueError(f"{name} must be a CUDA tensor on device {self.device}")
        if x.dtype != dtype:
            raise ValueError(f"{name} must be {dtype}, got {x.dtype}")
        if not x.is_contiguous():
            raise ValueError(f"{name} must be contiguous")
        if x.numel() < min_numel:
            raise ValueError(f"{name} has {x.numel()} elements, {min_numel} needed")

    def set_tuning(self, mode, fwd_target_blocks: int = 0, adj_target_blocks: int = 0) -> None:
        self._call("crs_set_tuning", _lib.load().crs_set_tuning, _mode(mode),
                   int(fwd_target_blocks), int(adj_target_blocks))

    def set_block_shape(self, fwd=(0, 0), adj=(0, 0)) -> None:
        """FP32C pixels per forward lane and per adjoint block, (columns, rows); (0, 0)
        restores the built-in shape (forward 4 x 4, adjoint 8 x 8)."""
        self._call("crs_set_block_shape", _lib.load().crs_set_block_shape, int(fwd[0]),
                   int(fwd[1]), int(adj[0]), int(adj[1]))

    def set_timing(self, enable: bool = True) -> None:
        self._call("crs_set_timing", _lib.load().crs_set_timing, int(bool(enable)))

    def last_kernel_ms(self, adjoint: bool = False) -> float:
        """Duration of the last pair-sum kernel (HIP events on its stream); waits for it."""
        ms = ctypes.c_float()
        self._call("crs_last_kernel_ms", _lib.load().crs_last_kernel_ms, int(bool(adjoint)),
                   ctypes.byref(ms))
        return float(ms.value)

    def last_launch(self, adjoint: bool = False):
        """(kernel name as a profiler lists it, work-groups) of the last pair-sum launch."""
        name, groups = ctypes.c_char_p(), ctypes.c_int()
        self._call("crs_last_launch", _lib.load().crs_last_launch, int(bool(adjoint)),
                   ctypes.byref(name), ctypes.byref(groups))
        return name.value.decode(), int(groups.value)

    def forward(self, px, py, t, planes: Sequence[Plane], mode="fp32c", field: bool = True,
                intensity: bool = False, stream=None, out_field=None, out_intensity=None):
        """Complex field (complex128, total pixels) and/or |E|^2 (float64)."""
        torch = _torch()
        self._check_dev("px", px, torch.float64)
        self._check_dev("py", py, torch.float64)
        self._check_dev("t", t, torch.complex128)
        npil = px.numel()
        if py.numel() != npil or t.numel() % max(npil, 1):
            raise ValueError("px, py, t sizes disagree")
        nsets = t.numel() // max(npil, 1)
        arr, total = pack_planes(planes)
        dev = px.device
        f = out_field
        if f is not None:
            self._check_dev("out_field", f, torch.complex128, total)
        elif field:
            f = torch.empty(total, dtype=torch.complex128, device=dev)
        it = out_intensity
        if it is not None:
            self._check_dev("out_intensity", it, torch.float64, total)
        elif intensity:
            it = torch.empty(total, dtype=torch.float64, device=dev)
        self._call("crs_forward", _lib.load().crs_forward, self._stream(stream), _mode(mode),
                   ctypes.c_void_p(px.data_ptr()), ctypes.c_void_p(py.data_ptr()),
                   ctypes.c_void_p(t.data_ptr()), npil, nsets, arr, len(planes),
                   ctypes.c_void_p(f.data_ptr() if f is not None else None),
                   ctypes.c_void_p(it.data_ptr() if it is not None else None))
        return f, it

    def adjoint(self, px, py, u, planes: Sequence[Plane], nsets: int = 1, mode="fp32c",
                stream=None, out=None):
        """g[s, q] = sum_p conj(G_pq) u_p over the planes of set s (complex128 [nsets, npil])."""
        torch = _torch()
        self._check_dev("px", px, torch.float64)
        self._check_dev("py", py, torch.float64)
        self._check_dev("u", u, torch.complex128)
        npil = px.numel()
        if py.numel() != npil:
            raise ValueError("px, py sizes disagree")
        arr, total = pack_planes(planes)
        if u.numel() < total:
            raise ValueError(f"u has {u.numel()} pixels, planes need {total}")
        if out is not None:
            self._check_dev("out", out, torch.complex128, int(nsets) * npil)
        g = out if out is not None else torch.empty((nsets, npil), dtype=torch.complex128,
                                                    device=px.device)
        self._call("crs_adjoint", _lib.load().crs_adjoint, self._stream(stream), _mode(mode),
                   ctypes.c_void_p(px.data_ptr()), ctypes.c_void_p(py.data_ptr()), npil,
                   int(nsets), arr, len(planes), ctypes.c_void_p(u.data_ptr()),
                   ctypes.c_void_p(g.data_ptr()))
        return g


class Graph:
    """An instantiated HIP graph of one forward or adjoint (crs_graph_*)."""

    def __init__(self, handle, keep):
        self._h = handle
        self._keep = keep  # tensors bound by pointer: keep them alive

    def launch(self, stream=None) -> None:
        torch = _torch()
        s = stream if stream is not None else torch.cuda.current_stream()
        _lib.check(_lib.load().crs_graph_launch(self._h, ctypes.c_void_p(s.cuda_stream)),
                   "crs_graph_launch")

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib.load().crs_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def graph_forward(ctx: "Context", px, py, t, planes: Sequence[Plane], field, intensity=None,
                  mode="fp32c") -> Graph:
    """Capture ctx.forward(px, py, t, planes) writing into the given output tensors."""
    torch = _torch()
    arr, total = pack_planes(planes)
    npil = px.numel()
    ctx._check_dev("px", px, torch.float64)
    ctx._check_dev("py", py, torch.float64, npil)
    ctx._check_dev("t", t, torch.complex128, npil)
    if t.numel() % max(npil, 1):
        raise ValueError("t must hold whole transmission sets of px.numel() pillars")
    if field is None and intensity is None:
        raise ValueError("graph_forward needs field and/or intensity")
    if field is not None:
        ctx._check_dev("field", field, torch.complex128, total)
    if intensity is not None:
        ctx._check_dev("intensity", intensity, torch.float64, total)
    nsets = t.numel() // max(npil, 1)
    h = ctypes.c_void_p()
    ctx._call("crs_graph_forward", _lib.load().crs_graph_forward, _mode(mode),
              ctypes.c_void_p(px.data_ptr()), ctypes.c_void_p(py.data_ptr()),
              ctypes.c_void_p(t.data_ptr()), px.numel(), nsets, arr, len(planes),
              ctypes.c_void_p(field.data_ptr() if field is not None else None),
              ctypes.c_void_p(intensity.data_ptr() if intensity is not None else None),
              ctypes.byref(h))
    return Graph(h, (px, py, t, field, intensity))


def graph_adjoint(ctx: "Context", px, py, u, planes: Sequence[Plane], g, nsets: int = 1,
                  mode="fp32c") -> Graph:
    """Capture ctx.adjoint(px, py, u, planes) writing into g (complex128 [nsets, npil])."""
    torch = _torch()
    arr, total = pack_planes(planes)
    npil = px.numel()
    ctx._check_dev("px", px, torch.float64)
    ctx._check_dev("py", py, torch.float64, npil)
    ctx._check_dev("u", u, torch.complex128, total)
    ctx._check_dev("g", g, torch.complex128, int(nsets) * npil)
    h = ctypes.c_void_p()
    ctx._call("crs_graph_adjoint", _lib.load().crs_graph_adjoint, _mode(mode),
              ctypes.c_void_p(px.data_ptr()), ctypes.c_void_p(py.data_ptr()), px.numel(),
              int(nsets), arr, len(planes), ctypes.c_void_p(u.data_ptr()),
              ctypes.c_void_p(g.data_ptr()), ctypes.byref(h))
    return Graph(h, (px, py, u, g))


def kernel_name(mode, adjoint: bool = False) -> str:
    return _lib.load().crs_kernel_name(_mode(mode), int(bool(adjoint))).decode()
