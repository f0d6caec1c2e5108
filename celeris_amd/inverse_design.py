"""Inverse design on the differentiable far field: pillar phases that maximise the
encircled energy of a focal window.

The loop the autograd bridge exists for.  Each step runs the forward sum for the
current transmissions t = a exp(i phi), the loss -EE (the reference's encircled-energy
ratio, focal.cpp:126-137, with a chosen radius), and its gradient with respect to every
pillar phase through the hand-written adjoint kernel; torch's Adam takes the step.  The
amplitudes stay fixed (a real design would map phase back to a pillar fill through a
unit-cell library: ``celeris_amd.design.UnitCellLibrary.lookup_weighted``).
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Callable, List, Optional

import numpy as np

from .autograd import encircled_energy, rs_field
from .rs import Plane


@dataclass
class PhaseDesign:
    phase: np.ndarray          # optimised pillar phases (rad)
    encircled_energy: List[float]  # EE before each step, and after the last


def optimize_phase(px, py, amplitude, phase0, plane: Plane, radius: float, steps: int = 50,
                   lr: float = 0.05, mode: str = "fp32c", device: str = "cuda:0",
                   callback: Optional[Callable[[int, float], None]] = None) -> PhaseDesign:
    """Adam on the pillar phases to maximise the fraction of the plane's power within
    ``radius`` of its centre (cx = cy = 0).  Returns the phases and the EE history."""
    import torch

    dev = torch.device(device)
    tpx = torch.as_tensor(np.ascontiguousarray(px, np.float64), device=dev)
    tpy = torch.as_tensor(np.ascontiguousarray(py, np.float64), device=dev)
    amp = torch.as_tensor(np.ascontiguousarray(amplitude, np.float64), device=dev)
    phi = torch.as_tensor(np.array(phase0, np.float64), device=dev).clone().requires_grad_(True)
    opt = torch.optim.Adam([phi], lr=lr)
    history: List[float] = []

    def ee_of(p):
        t = torch.polar(amp, p)
        E = rs_field(t, tpx, tpy, [plane], mode=mode)
        inten = E.real ** 2 + E.imag ** 2
        return encircled_energy(inten, plane, radius)

    for it in range(steps):
        opt.zero_grad()
        ee = ee_of(phi)
        (-ee).backward()
        history.append(float(ee.detach()))
        if callback is not None:
            callback(it, history[-1])
        opt.step()
    with torch.no_grad():
        history.append(float(ee_of(phi)))
    return PhaseDesign(phi.detach().cpu().numpy(), history)
