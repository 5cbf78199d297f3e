"""The uniform-mesh output (WriteUniformMeshQ, lib/bsaminputoutput.f90:438-541)
built on the GPU against the reference's own routine.

The engine builds the uniform grids in a context of one domain-wide patch
per level (bsam_amd.driver.uniform_layout): bsam_overlay_level copies the
hierarchy's patches, ghosts included, into it level by level, with
bsam_init_fields' mass-corrected prolongation between levels. The reference
side is WriteUniformMeshQ itself, called by the reference harness
(oracle/_ref/harness, op 'uniform') on the same hierarchy and arrays. The
arrays are random in every cell, ghosts included, so a ghost cell that lies
over a neighbour's interior differs from it: the file can only match if the
later patch of the level list wins there, as in the reference's copies.
"""

import os

import pytest

import refrun
from bsam_amd import frames
from bsam_amd.driver import uniform_layout, uniform_mesh
from bsam_amd.engine import Engine
from helpers import (layout_amr2d, layout_amr2d_tiles, layout_amr3d, layout_periodic2d,
                     layout_single, random_fields)

pytestmark = pytest.mark.gpu

PLAN_FREE = {"fuse_ghost": 0, "tail_cells": 0, "graph": 0}
KW = dict(dt=1e-3, eps=0.03, device=0, tune=PLAN_FREE)

LAYOUTS = {
    "single": layout_single,
    "amr2d": layout_amr2d,
    "amr2d_tiles": layout_amr2d_tiles,
    "amr2d_periodic": layout_periodic2d,
    "amr3d": layout_amr3d,
}


def reference_file(lay, fields, tmp):
    refrun.run(lay, fields, [("uniform", 0, 0)], keep=tmp)
    with open(os.path.join(tmp, "out", "u00000.dat"), "rb") as fh:
        return fh.read()


def engine_file(u, eng, ulay, top, path):
    lay = eng.layout
    nd = lay.ndims
    q = uniform_mesh(u, eng, top)
    # the header as WriteUniformMeshQ prints it for the harness's root
    # (dx the same in every dimension, xupper = xlower + dx * mx)
    root = lay.patches[0]
    mxuf = ulay.patches[top].mx[:nd]
    dx = root.dx
    for _ in range(lay.finestlevel):
        dx = dx / 2
    xupper = tuple(root.xlower[d] + root.dx * float(root.mx[d]) for d in range(nd))
    fp = frames.FramePatch(0, nd, 2, (dx,) * nd, tuple(root.xlower[:nd]), xupper,
                           tuple(mxuf), tuple((1, m) for m in mxuf), q)
    frames.write_frame(path, 0.0, 0, [fp])
    with open(path, "rb") as fh:
        return fh.read()


@pytest.mark.parametrize("name", sorted(LAYOUTS))
def test_uniform_mesh_matches_reference(tmp_path, name):
    """Two hierarchies' arrays in a row through the same uniform context
    (the second call overwrites what the first left), each equal to the
    reference's file byte for byte."""
    if not refrun.available():
        pytest.skip("reference harness not built (make -C oracle ref)")
    lay = LAYOUTS[name]()
    ulay, top = uniform_layout(lay)
    u = Engine(ulay, **KW)
    eng = Engine(lay, **KW)
    for k, seed in enumerate((71, 72)):
        fields = random_fields(lay, seed=seed)
        want = reference_file(lay, fields, str(tmp_path / f"ref{k}"))
        for n, p in enumerate(eng.patches):
            eng.upload(n, "q", fields[p.id]["q"])
        got = engine_file(u, eng, ulay, top, str(tmp_path / f"u{k}.dat"))
        assert got == want, (seed, first_difference(got, want))
    _, finest, ps = frames.read_frame(str(tmp_path / "u1.dat"))
    assert finest == 0 and len(ps) == 1
    assert ps[0].mx == tuple(m * 2 ** lay.finestlevel for m in lay.patches[0].mx[:lay.ndims])
    u.close()
    eng.close()


def first_difference(a: bytes, b: bytes) -> str:
    la, lb = a.decode().splitlines(), b.decode().splitlines()
    for n, (x, y) in enumerate(zip(la, lb)):
        if x != y:
            return f"line {n + 1}: gpu {x!r} reference {y!r}"
    return f"lengths {len(la)} vs {len(lb)} lines"
