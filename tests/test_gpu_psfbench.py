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


@pytest.mark.parametrize("grid,stride", [(161, 10), (321, 40), (1024, 128)])
def test_psfbench_full_grids_vs_oracle_rows(torch_cuda, grid, stride):
    """psfbench's own grids (161^2 - 1024^2, bench.cpp:103): the drop-in call's |E|^2 in
    FP32C and FP64 against the oracle on every stride-th row (the CPU leg's sampling)."""
    from celeris_amd import rs

    L = pb.psfbench_lens(grid=grid)
    k = 2 * math.pi / L.wavelength_um
    print(f"[psfbench] {L.px.size} pillars x {grid}^2, oracle every {stride}th row", flush=True)
    plane = Plane.psf(0.0, 0.0, L.focal_um, k, grid, L.half_window_um)
    _, ref = oracle.forward(L.px, L.py, L.t, [plane], intensity=True, row_stride=stride)
    ref = ref.reshape(grid, grid)[::stride]
    prev = rs.set_default_mode("fp32c")
    try:
        for mode, tol in (("fp64", 1e-12), ("fp32c", 1e-5)):
            rs.set_default_mode(mode)
            got = rs.propagate_psf(L.px, L.py, L.t, 0.0, 0.0, L.focal_um, k, grid, L.half_window_um)
            assert got is not None
            # normalised by each map's own peak over the sampled rows, as psfbench compares
            d = pb.max_normalized_diff(ref, got.reshape(grid, grid)[::stride])
            print(f"[psfbench] {grid}^2 {mode}: max|d normalized| {d:.2e}", flush=True)
            assert d <= tol
    finally:
        rs.set_default_mode(prev)
