"""Frame files ``out/mNNNNN.dat``: writer and reader of the reference's text
format, and conversion to and from a :class:`Layout` plus patch arrays.

The format is the one ``WriteQ``/``OutputQ``/``WriteQ2D``/``WriteQ3D`` write
(lib/bsaminputoutput.f90:51-168):

* the time (``f25.12``) and the finest level (``i3``);
* per patch, level by level from the root, in level-list order (youngest
  first): a line holding one blank (list-directed ``' '``), then
  ``level ndims 2 nrvars`` (``I3,3(1x,I3)``), ``dx``, ``xlower``, ``xupper``
  (``f25.12`` separated by one blank), ``mx`` and ``mglobal`` (``i8``
  separated by one blank), then q over the whole patch including its ghost
  layer, one cell a line with i fastest: ``es20.12`` values separated by two
  blanks in 2D, ``f25.12`` in 3D.

The writer reproduces those bytes from the same numbers. The reader takes
the values by whitespace, so it reads both the 2D (``es20.12``) and the 3D
(``f25.12``) layout; it does not inherit the reference reader's mismatch
(``ReadQ2D`` reads 2D frames with ``f25.12``, :402-405, which is not what
``WriteQ2D`` writes).
"""

from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass
from typing import Dict, List, Sequence, Tuple

import numpy as np

from .layout import INTERNAL_BC, Layout


@dataclass
class FramePatch:
    level: int
    ndims: int
    nrvars: int
    dx: Tuple[float, ...]
    xlower: Tuple[float, ...]
    xupper: Tuple[float, ...]
    mx: Tuple[int, ...]
    mglobal: Tuple[Tuple[int, int], ...]
    q: np.ndarray  # (mx+2 per dimension, nrvars), Fortran order


def _f25(x: float) -> str:
    return f"{x:25.12f}"


def _es20(x: float) -> str:
    """Fortran ES20.12: d.dddddddddddd with a two-digit exponent E+nn; for
    exponents of three digits the letter E is dropped (+nnn)."""
    if x == 0.0 or not math.isfinite(x):
        return f"{x:20.12E}"
    s = f"{x:.12E}"
    mant, exp = s.split("E")
    e = int(exp)
    if abs(e) >= 100:
        body = f"{mant}{'+' if e > 0 else '-'}{abs(e):03d}"
    else:
        body = f"{mant}E{'+' if e >= 0 else '-'}{abs(e):02d}"
    return f"{body:>20s}"


def format_rows(vals: np.ndarray, kind: int) -> str:
    """Rows of values (a 2D array, one row a line) in the frame files'
    number format: es20.12 (kind 0) or f25.12 (kind 1), two blanks apart;
    formatted by the library (bsam_format_rows)."""
    from .engine import lib
    a = np.asarray(vals, dtype=np.float64)
    if a.ndim != 2:
        raise ValueError("rows of values expected")
    nrows, ncols = a.shape
    L = lib()
    w = 20 if kind == 0 else 25
    cap = nrows * (ncols * (w + 2) - 2 + 1) + 1
    buf = C.create_string_buffer(cap)
    n = L.bsam_format_rows(a.ctypes.data_as(C.POINTER(C.c_double)), nrows, ncols,
                           a.strides[0] // 8, a.strides[1] // 8, kind, buf, cap)
    if n < 0:
        raise RuntimeError("bsam_format_rows failed")
    return buf.raw[:n].decode()


def write_frame(path: str, time: float, finestlevel: int,
                patches: Sequence[FramePatch]) -> None:
    """Writes a frame as WriteQ does; ``patches`` in frame order."""
    out: List[str] = [_f25(time), f"{finestlevel:3d}"]
    for p in patches:
        nd = p.ndims
        out.append("  ")
        out.append(f"{p.level:3d} {nd:3d} {2:3d} {p.nrvars:3d}")
        for vec in (p.dx, p.xlower, p.xupper):
            out.append(" ".join(_f25(v) for v in vec[:nd]))
        out.append(" ".join(f"{v:8d}" for v in p.mx[:nd]))
        out.append(" ".join(f"{v:8d}" for d in range(nd) for v in p.mglobal[d]))
        q = np.asarray(p.q, dtype=np.float64)
        flat = q.reshape(-1, p.nrvars, order="F")
        out.append(format_rows(flat, 0 if nd == 2 else 1).rstrip("\n"))
    with open(path, "w") as fh:
        fh.write("\n".join(out) + "\n")


def read_frame(path: str):
    """Returns (time, finestlevel, [FramePatch, ...]) in frame order."""
    with open(path) as fh:
        lines = fh.read().split("\n")
    time = float(lines[0])
    finest = int(lines[1])
    pos = 2
    patches: List[FramePatch] = []
    while pos < len(lines) and lines[pos].strip() == "" and pos + 1 < len(lines) \
            and lines[pos + 1].strip():
        level, nd, _r, nv = (int(v) for v in lines[pos + 1].split())
        dx = tuple(float(v) for v in lines[pos + 2].split())
        xl = tuple(float(v) for v in lines[pos + 3].split())
        xu = tuple(float(v) for v in lines[pos + 4].split())
        mx = tuple(int(v) for v in lines[pos + 5].split())
        mg = [int(v) for v in lines[pos + 6].split()]
        pos += 7
        ncell = 1
        for m in mx:
            ncell *= m + 2
        vals = np.array(" ".join(lines[pos:pos + ncell]).split(), dtype=np.float64)
        pos += ncell
        q = vals.reshape((ncell, nv)).reshape(tuple(m + 2 for m in mx) + (nv,), order="F")
        patches.append(FramePatch(level, nd, nv, dx, xl, xu, mx,
                                  tuple((mg[2 * d], mg[2 * d + 1]) for d in range(nd)), q))
    return time, finest, patches


def layout_from_frame(patches: Sequence[FramePatch], mthbc: Sequence[int],
                      minlevel: int = 1) -> Tuple[Layout, Dict[int, int]]:
    """The hierarchy a frame describes (levels >= 0) as a :class:`Layout`,
    with the root's boundary codes ``mthbc`` (frames do not hold them), and
    the map frame index -> layout patch id. The parent of each refined patch
    is the patch of the level below that covers it; the level-list order of
    the frame is kept (the layout creates each level's patches eldest first,
    so that they traverse youngest first as in the frame)."""
    root = patches[0]
    nd = root.ndims
    lay = Layout(nd, root.mx, root.dx[0], xlower=root.xlower, mthbc=mthbc,
                 minlevel=minlevel, nrvars=root.nrvars)
    ids = {0: 0}
    by_level: Dict[int, List[int]] = {}
    for n, p in enumerate(patches):
        by_level.setdefault(p.level, []).append(n)
    for level in sorted(by_level):
        if level == 0:
            continue
        for n in reversed(by_level[level]):
            p = patches[n]
            # coarse global index range of the patch on the level below
            cg = [((p.mglobal[d][0] + 1) // 2, p.mglobal[d][1] // 2) for d in range(nd)]
            parent = None
            for m in by_level[level - 1]:
                q = patches[m]
                if all(q.mglobal[d][0] <= cg[d][0] and cg[d][1] <= q.mglobal[d][1]
                       for d in range(nd)):
                    parent = m
                    break
            if parent is None:
                raise ValueError(f"frame patch {n} has no covering parent")
            pm = patches[parent]
            mb = [(cg[d][0] - pm.mglobal[d][0] + 1, cg[d][1] - pm.mglobal[d][0] + 1)
                  for d in range(nd)]
            ids[n] = lay.add_child(ids[parent], mb)
    return lay, ids


def frame_from_arrays(layout: Layout, q_by_id: Dict[int, np.ndarray],
                      xupper_root: Sequence[float]) -> List[FramePatch]:
    """Frame patches (levels >= 0, frame order) of a layout with its q arrays
    (``Layout.field_shape`` 'q', indexed by patch id). xupper follows
    MakeNewGrid (lib/bsamroutines.f90:1670-1673): mbounds(2) * parent dx +
    parent xlower; the root's is given."""
    nd = layout.ndims
    out: List[FramePatch] = []
    xupper: Dict[int, Tuple[float, ...]] = {0: tuple(float(v) for v in xupper_root[:nd])}
    for level in range(0, layout.finestlevel + 1):
        for pid in layout.level_ids(level):
            p = layout.patches[pid]
            if pid not in xupper:
                par = layout.patches[p.parent_id]
                xupper[pid] = tuple(float(p.mbounds[d][1]) * par.dx + par.xlower[d]
                                    for d in range(nd))
            q = np.asarray(q_by_id[pid])
            q = q.reshape(tuple(p.mx[d] + 2 for d in range(nd)) + (layout.nrvars,), order="F")
            out.append(FramePatch(p.level, nd, layout.nrvars, (p.dx,) * nd, p.xlower[:nd],
                                  xupper[pid], p.mx[:nd],
                                  tuple(p.mglobal[d] for d in range(nd)), q))
    return out


def frame_from_engine(engine, xupper_root: Sequence[float]) -> List[FramePatch]:
    """Frame patches of an engine's current q (downloaded from the device)."""
    q = {p.id: engine.download(n, "q") for n, p in enumerate(engine.patches) if p.level >= 0}
    return frame_from_arrays(engine.layout, q, xupper_root)


__all__ = ["FramePatch", "write_frame", "read_frame", "layout_from_frame",
           "frame_from_arrays", "frame_from_engine", "INTERNAL_BC"]
