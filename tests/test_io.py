"""PGM output (celeris_amd.io) against the reference's write_pgm bytes
(tests/golden/pgm.npz, written by oracle/_ref/ref_driver from io/image.cpp)."""
import os

import numpy as np
import pytest

from celeris_amd import io

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "pgm.npz")


def cases():
    d = np.load(GOLDEN)
    names = sorted({k.split("__")[0] for k in d.files})
    return [(nm, d[f"{nm}__args"], d[f"{nm}__values"], d[f"{nm}__pgm"].tobytes()) for nm in names]


@pytest.mark.parametrize("case", cases(), ids=lambda c: c[0])
def test_pgm_bytes_match_reference(case, tmp_path):
    _, args, values, want = case
    nx, ny, gamma = int(args[0]), int(args[1]), float(args[2])
    assert io.pgm_bytes(nx, ny, values, gamma) == want
    path = tmp_path / "x.pgm"
    assert io.write_pgm(str(path), nx, ny, values, gamma)
    assert path.read_bytes() == want


def test_pgm_rejects_what_the_reference_rejects(tmp_path):
    assert io.pgm_bytes(0, 3, [], 1.0) is None
    assert io.pgm_bytes(2, 2, [1.0, 2.0, 3.0], 1.0) is None
    assert not io.write_pgm(str(tmp_path / "no" / "dir.pgm"), 1, 1, [1.0])
    # half-way values round away from zero, as std::lround
    assert io.pgm_bytes(2, 1, [0.5 / 255.0 * 2, 2.0], 1.0)[-2:] == bytes([1, 255])
