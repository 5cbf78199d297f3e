"""Frame files out/mNNNNN.dat (bsam_amd/frames.py) against the reference's
own writer (WriteQ, lib/bsaminputoutput.f90:51-168).

The reference program (oracle/_ref/bsam) writes frames of the shipped
Cahn-Hilliard deck; reading each and writing it back must give the same
bytes, and so must the hierarchy rebuilt from the frame as a Layout (parents,
mbounds and xlower re-derived as MakeNewGrid derives them). The number
formats are checked against flang's own output for values the frames do not
reach (three-digit exponents, signs, zeros).
"""

import os
import subprocess

import numpy as np
import pytest

import bsam_program as bp
from bsam_amd import frames


def test_reference_frames_round_trip(tmp_path):
    if not os.path.exists(bp.REF_PROGRAM):
        pytest.skip("reference program not built (make -C oracle run)")
    ref, _, _ = bp.run(bp.REF_PROGRAM, bp.shipped_deck(2), str(tmp_path / "run"))
    assert sorted(ref) == [0, 1, 2]
    for n, data in ref.items():
        path = str(tmp_path / "run" / "out" / f"m{n:05d}.dat")
        time, finest, patches = frames.read_frame(path)
        assert finest == 2 and patches[0].level == 0
        frames.write_frame(str(tmp_path / "w.dat"), time, finest, patches)
        assert (tmp_path / "w.dat").read_bytes() == data, n
        lay, ids = frames.layout_from_frame(patches, (2, 2, 10, 10, 10, 10), minlevel=-5)
        assert lay.finestlevel == finest
        assert sum(len(lay.level_ids(l)) for l in range(0, finest + 1)) == len(patches)
        rebuilt = frames.frame_from_arrays(lay, {ids[i]: p.q for i, p in enumerate(patches)},
                                           patches[0].xupper)
        frames.write_frame(str(tmp_path / "w2.dat"), time, finest, rebuilt)
        assert (tmp_path / "w2.dat").read_bytes() == data, n


FLANG = "/opt/rocm/lib/llvm/bin/flang"
VALUES = [0.0, 1.0, -1.0, 1.001152265268, -8.612898522527e-04, 1.5e-99, 2.5e-100, -3.75e+101,
          6.02214076e+23, 1.0e+200, 9.999999999999e-01, 9.9999999999999e-01, 123456.789]


def test_number_formats_match_flang(tmp_path):
    if not os.path.exists(FLANG):
        pytest.skip("flang not available")
    src = tmp_path / "fmt.f90"

    def lit(v):  # a double-precision Fortran literal of the same value
        m, _, e = f"{v:.17e}".partition("e")
        return f"{m}d{int(e)}"

    lines = [f"  write(*,'(es20.12)') {lit(v)}" for v in VALUES]
    lines += [f"  write(*,'(f25.12)') {lit(v)}" for v in VALUES if abs(v) < 1e12]
    src.write_text("program fmt\n" + "\n".join(lines) + "\nend program fmt\n")
    exe = tmp_path / "fmt"
    subprocess.run([FLANG, str(src), "-o", str(exe)], check=True, capture_output=True)
    got = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    want = [frames._es20(v) for v in VALUES] + [frames._f25(v) for v in VALUES if abs(v) < 1e12]
    assert got.splitlines() == want
    # the library's formatter, which writes the frames' values
    lib_rows = (frames.format_rows(np.array(VALUES)[:, None], 0).splitlines() +
                frames.format_rows(np.array([v for v in VALUES if abs(v) < 1e12])[:, None],
                                   1).splitlines())
    assert lib_rows == want


def test_write_read_3d(tmp_path):
    """3D frames hold f25.12 values; the reader takes either layout."""
    rng = np.random.default_rng(3)
    q = rng.uniform(-1, 1, (6, 4, 5, 2))
    p = frames.FramePatch(0, 3, 2, (0.5, 0.5, 0.5), (0.0, 0.0, 0.0), (2.0, 1.0, 1.5),
                          (4, 2, 3), ((1, 4), (1, 2), (1, 3)), q)
    path = str(tmp_path / "m3.dat")
    frames.write_frame(path, 0.25, 0, [p])
    t, fin, ps = frames.read_frame(path)
    assert t == 0.25 and fin == 0 and ps[0].mx == (4, 2, 3)
    np.testing.assert_allclose(ps[0].q, q, rtol=0, atol=5e-13)
