import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); runs through the C ABI")


def rel_l2(a, b) -> float:
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def max_rel(a, b) -> float:
    """max |a - b| / max |b| (the max-abs error scaled by the field's peak)."""
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.fixture(scope="session")
def torch_cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.fixture(scope="session")
def rsctx(torch_cuda):
    from celeris_amd import Context

    ctx = Context(0)
    yield ctx
    ctx.close()


def achromatic_golden():
    """(name, kind, px, py, t_bands, lam, center_index, center_lambda, f, reference out,
    library, design) for every case of tests/golden/achromatic.npz (written by the reference's
    verify_achromatic_focus and verify_pb_achromatic_focus)."""
    from celeris_amd import design as dz
    d = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                             "achromatic.npz"))
    out = []
    for name in sorted({k.split("__")[0] for k in d.files}):
        g = lambda k: d[f"{name}__{k}"]  # noqa: E731
        lam, ci, lc, f = g("lam"), int(g("ci")), float(g("lc")), float(g("f"))
        lib = dz.DispersiveLibrary(lam, lc, ci, [dz.DispersiveAtom(a, ph) for a, ph in
                                                 zip(g("amp"), g("phase"))])
        des = dz.AchromaticDesign(int(g("n_cells")), float(g("period")), g("atom"), f,
                                  float(g("D")))
        px, py, t = dz.achromatic_band_transmissions(lib, des)
        out.append((name, "achro", px, py, t, lam, ci, lc, f, g("achro"), lib, des))
        plib = dz.DispersivePbLibrary(lam, lc, ci, [dz.DispersivePbAtom(c) for c in g("cross")])
        pdes = dz.PbAchromaticDesign(int(g("n_cells")), float(g("period")), g("pb_atom"),
                                     g("rotation"), f, float(g("D")), int(g("hand")))
        px, py, t = dz.pb_achromatic_band_transmissions(plib, pdes)
        out.append((name, "pbachro", px, py, t, lam, ci, lc, f, g("pbachro"), plib, pdes))
    return out


def dispersive_golden():
    """(lens, build_lib_at, f, lam, D, lmin, lmax, n, reference result, calls) from
    tests/golden/chromatic_dispersive.npz (written by the reference's
    analyze_chromatic_dispersive with a table-backed library builder); calls records the
    wavelengths build_lib_at was asked for."""
    from celeris_amd.design import MetalensDesign, UnitCellLibrary
    d = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                             "chromatic_dispersive.npz"))
    lens = MetalensDesign(int(d["n_cells"]), float(d["period"]), d["fill_map"])
    calls = []

    def build_lib_at(wl):
        calls.append(wl)
        i = int(np.argmin(np.abs(d["lib_wavelengths"] - wl)))
        assert abs(d["lib_wavelengths"][i] - wl) <= 1e-12 * wl
        return UnitCellLibrary(float(d["period"]), d["lib_fill"], d["lib_phases"][i], d["lib_amps"][i])

    lmin, lmax, n = d["args"]
    return (lens, build_lib_at, float(d["f"]), float(d["lam"]), float(d["D"]), float(lmin),
            float(lmax), int(n), d["result"].reshape(-1, 3), calls)
