"""The C++ drop-in, executed: tests/cpp/dropin_driver.cpp is compiled (by build()) against
include/celeris/cuda/propagate.hpp — the reference's propagate.hpp:16-38 names and
signatures — and linked to libceleris_rs.so, as the reference's analysis code would be.
It calls celeris::cuda::propagate_psf / propagate_zx on the inputs of the reference's
own propagate_pillars outputs (tests/golden/propagate_pillars.npz) and, from 8 host
threads at once, on 8 different lenses (the reference calls these concurrently from
std::async trials, tolerance.cpp:24-35, and GUI threads).  The default mode is FP32C;
|E|^2 is checked at 2e-5 of its peak."""
import math
import os
import struct
import subprocess

import numpy as np
import pytest

import oracle
from celeris_amd import Plane
from celeris_amd.build import DRIVER

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _driver():
    if not os.path.exists(DRIVER):
        raise RuntimeError(f"{DRIVER} is not built (run __graft_entry__.build())")
    return DRIVER


def _psf_job(px, py, t, cx, cy, z, k, n, W) -> bytes:
    t = np.asarray(t, np.complex128)
    return (struct.pack("<ii5d", px.size, n, cx, cy, z, k, W)
            + np.ascontiguousarray(px, "<f8").tobytes() + np.ascontiguousarray(py, "<f8").tobytes()
            + np.ascontiguousarray(t).view("<f8").tobytes())


def _run(tmp_path, cmd, payload, sizes):
    fin, fout = tmp_path / f"{cmd}.in", tmp_path / f"{cmd}.out"
    fin.write_bytes(payload)
    print(f"[dropin_driver] {cmd} ({len(payload)} bytes in)", flush=True)
    subprocess.run([_driver(), cmd, str(fin), str(fout)], check=True, timeout=120)
    raw = fout.read_bytes()
    res, pos = [], 0
    for m in sizes:
        ok = struct.unpack_from("<i", raw, pos)[0]
        pos += 4
        res.append((ok, np.frombuffer(raw, "<f8", m, pos).copy()))
        pos += 8 * m
    assert pos == len(raw)
    return res


def _golden():
    d = np.load(os.path.join(GOLDEN, "propagate_pillars.npz"))
    names = sorted({k.split("__")[0] for k in d.files})
    return [(nm, d[f"{nm}__px"], d[f"{nm}__py"], d[f"{nm}__t"], d[f"{nm}__args"], d[f"{nm}__I"])
            for nm in names]


def test_cpp_driver_info_runs():
    """The compiled caller links and runs (available() is 0 without a device)."""
    out = subprocess.run([_driver(), "info"], check=True, stdout=subprocess.PIPE, text=True,
                         timeout=60).stdout
    assert out.startswith("available ")


@pytest.mark.gpu
def test_cpp_dropin_info_on_device(torch_cuda):
    out = subprocess.run([_driver(), "info"], check=True, stdout=subprocess.PIPE, text=True,
                         timeout=60).stdout
    assert "available 1" in out and "device " in out


@pytest.mark.gpu
@pytest.mark.parametrize("case", _golden(), ids=lambda c: c[0])
def test_cpp_propagate_psf_matches_reference_output(torch_cuda, tmp_path, case):
    """celeris::cuda::propagate_psf through the compiled C++ caller against the
    reference's own propagate_pillars |E|^2 (focal.cpp:205-238, built from its sources)."""
    name, px, py, t, args, I_ref = case
    cx, cy, z, lam, n, W = args
    n = int(n)
    (ok, I), = _run(tmp_path, "psf", _psf_job(px, py, t, cx, cy, z, 2 * math.pi / lam, n, W), [n * n])
    assert ok == 1
    err = np.max(np.abs(I - I_ref.ravel())) / np.max(np.abs(I_ref))
    print(f"{name}: C++ drop-in vs reference |E|^2 {err:.2e}")
    assert err <= 2e-5


@pytest.mark.gpu
def test_cpp_propagate_zx_and_rejections(torch_cuda, tmp_path):
    """propagate_zx against the oracle; the reference's false returns leave out untouched
    (propagate.cu:139, :201): npil = 0, n = 0, nx = 1."""
    rng = np.random.default_rng(8)
    px, py = rng.uniform(-4, 4, 150), rng.uniform(-4, 4, 150)
    t = rng.uniform(0.5, 1, 150) * np.exp(1j * rng.uniform(-3, 3, 150))
    k = 2 * math.pi / 0.6
    payload = (struct.pack("<iii5d", 150, 31, 21, -3.0, 3.0, 10.0, 30.0, k)
               + px.tobytes() + py.tobytes() + t.view("<f8").tobytes())
    (ok, I), = _run(tmp_path, "zx", payload, [31 * 21])
    ref = np.abs(oracle.forward(px, py, t[None], [Plane.zx(-3.0, 3.0, 31, 10.0, 30.0, 21, k)])) ** 2
    assert ok == 1 and np.max(np.abs(I - ref)) / np.max(ref) <= 2e-5
    (ok, I), = _run(tmp_path, "psf", _psf_job(px[:0], py[:0], t[:0], 0, 0, 20, k, 5, 2.0), [25])
    assert ok == 0 and np.all(I == -1.0)
    (ok, I), = _run(tmp_path, "psf", _psf_job(px, py, t, 0, 0, 20, k, 0, 2.0), [0])
    assert ok == 0
    payload = (struct.pack("<iii5d", 150, 1, 21, -3.0, 3.0, 10.0, 30.0, k)
               + px.tobytes() + py.tobytes() + t.view("<f8").tobytes())
    (ok, I), = _run(tmp_path, "zx", payload, [21])
    assert ok == 0 and np.all(I == -1.0)


@pytest.mark.gpu
def test_cpp_dropin_concurrent_threads(torch_cuda, tmp_path):
    """8 host threads call propagate_psf at once, each on its own lens (pillar count,
    positions, transmissions, window, distance and wavelength differ), 3 times each;
    every result against the oracle."""
    jobs, refs, sizes = [], [], []
    for i in range(8):
        rng = np.random.default_rng(100 + i)
        npil = 200 + 150 * i
        px, py = rng.uniform(-6, 6, npil), rng.uniform(-6, 6, npil)
        t = rng.uniform(0.4, 1, npil) * np.exp(1j * rng.uniform(-3, 3, npil))
        lam, z, n, W = 0.45 + 0.03 * i, 15.0 + 4 * i, 17 + 6 * i, 1.5 + 0.3 * i
        cx, cy = 0.2 * i - 0.7, -0.1 * i
        k = 2 * math.pi / lam
        jobs.append(_psf_job(px, py, t, cx, cy, z, k, n, W))
        step = 2 * W / (n - 1)
        plane = Plane(k, cx - W, step, cy - W, step, z, 0.0, n, n)
        refs.append(np.abs(oracle.forward(px, py, t[None], [plane])) ** 2)
        sizes.append(n * n)
    res = _run(tmp_path, "threads", struct.pack("<ii", 8, 3) + b"".join(jobs), sizes)
    for i, ((ok, I), ref) in enumerate(zip(res, refs)):
        err = np.max(np.abs(I - ref)) / np.max(ref)
        print(f"thread {i}: ok {ok}, vs oracle {err:.2e}")
        assert ok == 1 and err <= 2e-5
