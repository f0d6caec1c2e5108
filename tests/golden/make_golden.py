#!/usr/bin/env python3
"""Writes tests/golden/*.npz: inputs and outputs of the REFERENCE's own CPU code.

Run where the reference sources are present (they are compiled by oracle/ref.mk into
oracle/_ref/ref_driver, from src/celeris/analysis/{focal,throughfocus,chromatic,mtf,field}.cpp
and src/celeris/design/{metalens,phase_profile,achromatic,pb_achromatic}.cpp).  The fixtures are data only:
seeded inputs and the reference's outputs, used by tests/test_oracle.py to pin the
FP64 restatement in oracle/rs_oracle.c where the reference build is absent.

    python tests/golden/make_golden.py            # every fixture
    python tests/golden/make_golden.py achromatic # only achromatic.npz (or: tolerance, design)
"""
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from celeris_amd import configs  # noqa: E402


def pillars_cases():
    """(name, px, py, t, cx, cy, z, lam, n, W) cases for celeris::propagate_pillars."""
    cases = []
    px, py = configs.square_grid(12, 0.4)
    a = 0.8 + 0.2 * configs.uniform(configs.SEED, configs.STREAM_AMP, px.size)
    t = a * np.exp(1j * configs.hyperbolic_phase(px, py, 0.532, 20.0))
    cases.append(("lens12_center", px, py, t, 0.0, 0.0, 20.0, 0.532, 9, 2.0))
    cases.append(("lens12_offaxis", px, py, t, 1.3, -0.7, 18.5, 0.532, 6, 1.5))
    px2, py2 = configs.square_grid(7, 0.35)
    u = configs.uniform(configs.SEED, configs.STREAM_U, 2 * px2.size)
    t2 = (0.5 + 0.5 * u[: px2.size]) * np.exp(2j * np.pi * u[px2.size:])
    cases.append(("random49_two_pixel", px2, py2, t2, 0.1, 0.2, 7.0, 0.633, 2, 0.9))
    cases.append(("random49_wide", px2, py2, t2, 0.0, 0.0, 3.0, 0.45, 11, 6.0))
    return cases


def lens_case():
    """A small metalens described the reference's way: a fill map over a square cell
    lattice plus a 16-entry unit-cell library (phase ramp, |t| in [0.75, 0.95]);
    pillars inside a circle of diameter D (focal.cpp:45-57)."""
    n_cells, period, lam, f, D = 21, 0.4, 0.532, 15.0, 8.0
    nlib = 16
    lib_fill = np.linspace(0.1, 0.85, nlib)
    lib_phase = np.linspace(-math.pi, math.pi, nlib, endpoint=False) + 0.5 * math.pi / nlib
    lib_amp = 0.75 + 0.2 * configs.uniform(configs.SEED, 11, nlib)
    c = (n_cells - 1) / 2.0
    iy, ix = np.meshgrid(np.arange(n_cells), np.arange(n_cells), indexing="ij")
    x = (ix - c) * period
    y = (iy - c) * period
    k = 2 * math.pi / lam
    phi = -k * (np.sqrt(x * x + y * y + f * f) - f)
    wrapped = np.angle(np.exp(1j * phi))
    # nearest library phase (circular), then that entry's fill
    d = np.angle(np.exp(1j * (wrapped[..., None] - lib_phase[None, None, :])))
    entry = np.argmin(np.abs(d), axis=-1)
    fill_map = lib_fill[entry]
    return dict(n_cells=n_cells, period=period, lam=lam, f=f, D=D, lib_fill=lib_fill,
                lib_phase=lib_phase, lib_amp=lib_amp, fill_map=fill_map)


def dispersive_libraries(L, wavelengths):
    """Per-wavelength unit-cell libraries on the design library's fill grid, standing in
    for build_lib_at(lambda)'s RCWA sweeps: each entry's phase scales with lambda0 / lambda
    plus a fill-dependent waveguide term, its amplitude drifts with lambda (deterministic
    formulas, no RCWA)."""
    lam0 = L["lam"]
    out = []
    for wl in wavelengths:
        d = (wl - lam0) / lam0
        phase = L["lib_phase"] * (lam0 / wl) + 0.8 * d * np.sin(2 * math.pi * L["lib_fill"])
        amp = np.clip(L["lib_amp"] * (1.0 - 0.6 * d * (L["lib_fill"] - 0.5)), 0.05, 1.0)
        out.append((wl, L["lib_fill"].copy(), phase, amp))
    return out


def write_chromatic_dispersive():
    """analyze_chromatic_dispersive (chromatic.cpp:60-127) on the lens case, with the
    table-backed library builder of dispersive_libraries."""
    L = lens_case()
    lmin, lmax, n = 0.45, 0.65, 6
    wls = [L["lam"]] + [lmin + (lmax - lmin) * i / (n - 1) for i in range(n)]
    libs = dispersive_libraries(L, wls)
    blk = oracle.lens_block(L["n_cells"], L["period"], L["fill_map"], L["lib_fill"], L["lib_phase"],
                            L["lib_amp"])
    vals = [blk, [L["f"], L["lam"], L["D"], lmin, lmax, n, len(libs)]]
    for wl, fill, phase, amp in libs:
        vals += [[wl], fill, phase, amp]
    out = {k: np.asarray(v) for k, v in L.items()}
    out["args"] = np.array([lmin, lmax, n])
    out["lib_wavelengths"] = np.array([w for w, *_ in libs])
    out["lib_phases"] = np.array([p for _, _, p, _ in libs])
    out["lib_amps"] = np.array([a for *_, a in libs])
    out["result"] = oracle.ref_run("chromatic_disp", np.concatenate(vals))
    np.savez_compressed(os.path.join(HERE, "chromatic_dispersive.npz"), **out)


def mtf_case():
    """A larger lens for analyze_mtf alone: 81 x 81 cells at 0.35 um with seeded random
    fills (a poor lens, so the MTF curve is far from the ideal one) inside a D = 27 um
    aperture (the mask cuts the corner cells), f = 30, lambda 0.6."""
    n_cells, period, lam, f, D = 81, 0.35, 0.6, 30.0, 27.0
    L = lens_case()
    fills = L["lib_fill"][np.floor(configs.uniform(configs.SEED, 12, n_cells * n_cells) *
                                    len(L["lib_fill"])).astype(int)]
    return dict(n_cells=n_cells, period=period, lam=lam, f=f, D=D, lib_fill=L["lib_fill"],
                lib_phase=L["lib_phase"], lib_amp=L["lib_amp"],
                fill_map=fills.reshape(n_cells, n_cells))


def achromatic_cases():
    """Dispersive libraries and designs for verify_achromatic_focus (achromatic.cpp:288)
    and verify_pb_achromatic_focus (pb_achromatic.cpp:255).  Atoms carry a seeded phase
    offset, a dispersion slope and |t| per band; each site takes the atom whose centre-band
    phase is nearest the hyperbolic target (PB: a seeded atom, and the rotation that puts
    the centre-band cross phase on the target), so the lens focuses near f at the centre
    wavelength and elsewhere at the other bands.  Sites outside the aperture get seeded
    atoms too; the checks must skip them."""
    rng = np.random.default_rng(20261203)
    cases = []
    for name, n_cells, period, D, f, lam, ci, natoms, hand in (
            ("band7", 45, 0.35, 15.0, 20.0, np.linspace(0.46, 0.64, 7), 3, 32, -1),
            ("band1", 17, 0.4, 6.0, 9.0, np.array([0.532]), 0, 8, 1)):
        nb = lam.size
        lc = float(lam[ci])
        phase0 = rng.uniform(-math.pi, math.pi, natoms)
        tau = rng.uniform(0.0, 1.5, natoms)
        phase = phase0[:, None] + 2 * math.pi * tau[:, None] * (lc / lam[None, :] - 1.0)
        amp = rng.uniform(0.6, 1.0, (natoms, nb))
        c = (n_cells - 1) / 2.0
        iy, ix = np.meshgrid(np.arange(n_cells), np.arange(n_cells), indexing="ij")
        x, y = (ix - c) * period, (iy - c) * period
        target = -2 * math.pi / lc * (np.sqrt(x * x + y * y + f * f) - f)
        d = np.angle(np.exp(1j * (target[..., None] - phase0[None, None, :])))
        atom = np.argmin(np.abs(d), axis=-1).ravel()
        outside = np.sqrt(x * x + y * y).ravel() > D / 2
        atom[outside] = rng.integers(0, natoms, int(outside.sum()))
        cross = amp * np.exp(1j * phase)
        pb_atom = rng.integers(0, natoms, n_cells * n_cells)
        geo = target.ravel() - np.angle(cross[pb_atom, ci])
        rotation = -geo / (2.0 * hand)
        cases.append(dict(name=name, n_cells=n_cells, period=period, D=D, f=f, lam=lam, ci=ci,
                          lc=lc, hand=hand, amp=amp, phase=phase, atom=atom, cross=cross,
                          pb_atom=pb_atom, rotation=rotation))
    return cases


def write_achromatic():
    out = {}
    for C in achromatic_cases():
        n, nb = C["n_cells"], C["lam"].size
        head = [n, C["period"], C["D"], C["f"]]
        band = [nb, *C["lam"], C["ci"], C["lc"], C["amp"].shape[0]]
        achro_in = np.concatenate([head, band, np.concatenate([C["amp"], C["phase"]], 1).ravel(),
                                   C["atom"]])
        pb_in = np.concatenate([head, [C["hand"]], band,
                                np.stack([C["cross"].real, C["cross"].imag], -1).ravel(),
                                C["pb_atom"], C["rotation"]])
        for k in ("n_cells", "period", "D", "f", "lam", "ci", "lc", "hand", "amp", "phase", "atom",
                  "cross", "pb_atom", "rotation"):
            out[f"{C['name']}__{k}"] = np.asarray(C[k])
        out[f"{C['name']}__achro"] = oracle.ref_run("achro", achro_in).reshape(-1, 3)
        out[f"{C['name']}__pbachro"] = oracle.ref_run("pbachro", pb_in).reshape(-1, 3)
    np.savez_compressed(os.path.join(HERE, "achromatic.npz"), **out)


def write_tolerance():
    """analyze_tolerance (tolerance.cpp:12-48) on the golden lens: the CLI's sweep
    (sigmas 0, 5, 10, 20 nm, 12 trials, seed 12345; cmd_design.cpp:478-480), and a
    single-trial sweep (trials = 0 means one) with another seed."""
    L = lens_case()
    blk = oracle.lens_block(L["n_cells"], L["period"], L["fill_map"], L["lib_fill"], L["lib_phase"],
                            L["lib_amp"])
    base = np.concatenate([blk, [L["f"], L["lam"], L["D"]]])
    out = {k: np.asarray(v) for k, v in L.items()}
    for name, sig, trials, seed in (("cli", [0.0, 5.0, 10.0, 20.0], 12, 12345),
                                    ("single", [3.0, 40.0], 0, 4000000000)):
        out[f"{name}__args"] = np.array([trials, seed], np.float64)
        out[f"{name}__sigmas"] = np.array(sig)
        out[f"{name}__out"] = oracle.ref_run("tolerance", np.concatenate(
            [base, [len(sig)], sig, [trials, seed]])).reshape(-1, 4)
    np.savez_compressed(os.path.join(HERE, "tolerance.npz"), **out)


def write_design():
    """design_metalens (metalens.cpp:230-273) on psfbench's lens (D 120, f 50, lambda
    0.532, p 0.35 with the stand-in 16-sample library of celeris_amd.psfbench) and on a
    seeded random library with amplitude weights 0 and 1; fill maps stored as library
    indices."""
    from celeris_amd.psfbench import psfbench_library
    rng = np.random.default_rng(8)
    lib = psfbench_library()
    rand_phase = rng.uniform(-math.pi, math.pi, 12)
    cases = [("psfbench", lib.fill, lib.phase, lib.amplitude, 0.532, 0.35, 50.0, 120.0, 0.25),
             ("rand_w0", np.linspace(0.1, 0.9, 12), rand_phase, rng.uniform(0.5, 1.0, 12), 0.633,
              0.4, 20.0, 16.3, 0.0),
             ("rand_w1", np.linspace(0.1, 0.9, 12), rand_phase, rng.uniform(0.5, 1.0, 12), 0.633,
              0.4, 20.0, 16.3, 1.0)]
    out = {}
    for name, fill, phase, amp, lam, per, f, D, w in cases:
        r = oracle.ref_run("design", np.concatenate([[len(fill)], fill, phase, amp,
                                                     [lam, per, f, D, w]]))
        n = int(r[0])
        fm = r[3:]
        idx = np.array([int(np.flatnonzero(fill == v)[0]) for v in fm], np.uint8)
        out[f"{name}__lib"] = np.stack([fill, phase, amp])
        out[f"{name}__args"] = np.array([lam, per, f, D, w])
        out[f"{name}__n_rms_amp"] = r[:3]
        out[f"{name}__fill_index"] = idx.reshape(n, n)
    np.savez_compressed(os.path.join(HERE, "design_metalens.npz"), **out)


def main():
    if not oracle.ref_available():
        sys.exit("oracle/_ref/ref_driver is not built (make -C oracle ref)")
    only = {"achromatic": write_achromatic, "tolerance": write_tolerance, "design": write_design,
            "dispersive": write_chromatic_dispersive}
    if sys.argv[1:] and sys.argv[1] in only:
        only[sys.argv[1]]()
        return
    write_achromatic()
    write_tolerance()
    write_design()
    write_chromatic_dispersive()
    out = {}
    for name, px, py, t, cx, cy, z, lam, n, W in pillars_cases():
        I = oracle.ref_propagate_pillars(px, py, t, cx, cy, z, lam, n, W)
        out[f"{name}__px"] = px
        out[f"{name}__py"] = py
        out[f"{name}__t"] = t
        out[f"{name}__args"] = np.array([cx, cy, z, lam, n, W])
        out[f"{name}__I"] = I
    np.savez_compressed(os.path.join(HERE, "propagate_pillars.npz"), **out)

    L = lens_case()
    blk = oracle.lens_block(L["n_cells"], L["period"], L["fill_map"], L["lib_fill"], L["lib_phase"],
                            L["lib_amp"])
    base = np.concatenate([blk, [L["f"], L["lam"], L["D"]]])
    lens_out = {k: np.asarray(v) for k, v in L.items()}
    lens_out["focus"] = oracle.ref_run("focus", base)
    lens_out["psf_n"] = np.array([33, 3.0])
    lens_out["psf"] = oracle.ref_run("psf", np.concatenate([base, [33, 3.0]]))
    lens_out["throughfocus"] = oracle.ref_run("throughfocus", base)
    lens_out["chromatic_args"] = np.array([0.45, 0.65, 5])
    lens_out["chromatic"] = oracle.ref_run("chromatic", np.concatenate([base, [0.45, 0.65, 5]]))
    lens_out["mtf"] = oracle.ref_run("mtf", base)
    lens_out["fieldgrid_args"] = np.array([8.0, 1, 41])
    lens_out["fieldgrid"] = oracle.ref_run("fieldgrid", np.concatenate([base, [8.0, 1, 41]]))
    on_axis_peak = float(np.max(lens_out["psf"]))
    lens_out["psffield_args"] = np.array([5.0, 33, 3.0, on_axis_peak])
    lens_out["psffield"] = oracle.ref_run("psffield", np.concatenate([base, [5.0, 33, 3.0, on_axis_peak]]))
    np.savez_compressed(os.path.join(HERE, "lens_analyses.npz"), **lens_out)

    fov = {k: np.asarray(v) for k, v in L.items()}
    fov["fov_angles"] = np.array([0.0, 3.0, 8.0, 15.0])
    fov["fov"] = oracle.ref_run("fov", np.concatenate([base, [fov["fov_angles"].size],
                                                       fov["fov_angles"]]))
    fov["widefov_args"] = np.array([4.0, 10.0])  # stop diameter, stop distance
    fov["widefov_angles"] = np.array([0.0, 5.0, 12.0, 20.0])
    fov["widefov"] = oracle.ref_run("widefov", np.concatenate(
        [base, fov["widefov_args"], [fov["widefov_angles"].size], fov["widefov_angles"]]))
    np.savez_compressed(os.path.join(HERE, "field_of_view.npz"), **fov)

    # PGM images (io/image.cpp) of the golden PSF, a caustic-like map, signed and
    # all-zero maps, with the CLI's gammas.
    rng = np.random.default_rng(5)
    caustic = np.asarray(oracle.ref_run("throughfocus", base))
    Nz = int(caustic[0])
    cnx, cnz = int(caustic[3 + 2 * Nz]), int(caustic[4 + 2 * Nz])
    maps = [(33, 33, lens_out["psf"], 2.2), (cnx, cnz, caustic[5 + 2 * Nz:5 + 2 * Nz + cnx * cnz], 1.6),
            (7, 5, rng.normal(size=35), 1.0), (4, 3, np.zeros(12), 2.2),
            (16, 9, rng.uniform(0, 1, 144) ** 3, 0.5)]
    pgm = {}
    for i, (nx, ny, v, g) in enumerate(maps):
        pgm[f"case{i}__args"] = np.array([nx, ny, g])
        pgm[f"case{i}__values"] = np.asarray(v, np.float64)
        pgm[f"case{i}__pgm"] = np.frombuffer(oracle.ref_pgm(nx, ny, v, g), dtype=np.uint8)
    np.savez_compressed(os.path.join(HERE, "pgm.npz"), **pgm)

    M = mtf_case()
    blk = oracle.lens_block(M["n_cells"], M["period"], M["fill_map"], M["lib_fill"], M["lib_phase"],
                            M["lib_amp"])
    mtf_out = {k: np.asarray(v) for k, v in M.items()}
    mtf_out["mtf"] = oracle.ref_run("mtf", np.concatenate([blk, [M["f"], M["lam"], M["D"]]]))
    np.savez_compressed(os.path.join(HERE, "mtf_large.npz"), **mtf_out)
    print("wrote", sorted(os.listdir(HERE)))


if __name__ == "__main__":
    main()
