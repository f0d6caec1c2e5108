"""psfbench's GPU legs (celeris_amd.psfbench.run_gpu) on the reference's lens at a
reduced grid, against the oracle's FP64 map: FP64 to 1e-12, FP32C to 1e-5, the
reference kernel's FP32 arithmetic to 1e-3 (max |d normalized PSF|, the reference's own
agreement figure, bench.cpp:156-169)."""
import math

import numpy as np
import pytest

import oracle
from celeris_amd import Plane
from celeris_amd import psfbench as pb

pytestmark = pytest.mark.gpu


def test_psfbench_gpu_legs_agree_with_oracle(torch_cuda):
    L = pb.psfbench_lens(grid=41)
    print(f"[psfbench] {L.px.size} pillars x {L.grid}^2", flush=True)
    gpu = pb.run_gpu(L, reps=2)
    plane = Plane.psf(0.0, 0.0, L.focal_um, 2 * math.pi / L.wavelength_um, L.grid, L.half_window_um)
    _, ref = oracle.forward(L.px, L.py, L.t, [plane], intensity=True)
    ref = ref.reshape(L.grid, L.grid)
    for mode, tol in (("fp64", 1e-12), ("fp32c", 1e-5), ("fp32", 1e-3)):
        g = gpu[mode]
        d = pb.max_normalized_diff(ref, g["intensity"])
        print(f"[psfbench] {mode}: call {g['call_ms']:.2f} ms kernel {g['kernel_ms']:.3f} ms "
              f"max|d normalized| {d:.2e}", flush=True)
        assert d <= tol
        assert g["kernel_ms"] > 0 and g["call_ms"] >= g["kernel_ms"]
