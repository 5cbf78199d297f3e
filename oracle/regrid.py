"""Oracle (test infrastructure): plain-Python restatement of the regrid's
integer core for one level, used only to check the engine's host clustering
(bsam_regrid_level) on the CPU.

Follows lib/bsamroutines.f90:
* ErrFlag's BufferAndList (:1159-1205): a flagged cell whose buffer box fits
  in its patch flags that box; otherwise its global coordinates join the
  tag list (both decided on the flags before buffering);
* InflateEdgeTags / BufferTaggedCells (:1207-1280): every listed tag, newest
  first, flags its buffer box in every patch of the level, and again in its
  periodic images (GetPeriodicTagOffset, boundary.f90:1046-1201, including
  the x3/mxmax(level,2) comparison of offset 9);
* RefineGrid / NewSubGrids / GetSignatures / GridFlagRatio (:1334-1643):
  Berger-Rigoutsos per patch in level-list order.

Flags are numpy int arrays (mx1, mx2, mx3) indexed [i-1, j-1, k-1].
"""

from __future__ import annotations

from typing import Dict, List, Sequence, Tuple

import numpy as np

MAXSUBGRIDS = 1024
MAXSPLITPASSES = 15


def periodic_offsets(root_mthbc: Sequence[int], ndims: int) -> List[int]:
    """PeriodicSetup's offset list (boundary.f90:893-963)."""
    px = root_mthbc[0] == 2
    py = root_mthbc[2] == 2
    pz = ndims == 3 and root_mthbc[4] == 2
    if px and py and pz:
        return list(range(1, 14))
    if px and py:
        return [1, 2, 4, 5]
    if px and pz:
        return [1, 3, 8, 9]
    if px:
        return [1]
    if py and pz:
        return [2, 3, 6, 7]
    if py:
        return [2]
    if pz:
        return [3]
    return []


def periodic_tag_offset(coord, level, off, ndims, ib, root_mx):
    """GetPeriodicTagOffset: (shifted coordinate, True) or (coordinate, False)."""
    c = list(coord)
    mg = [[1, 1], [1, 1], [1, 1]]
    mm = [1, 1, 1]
    for d in range(ndims):
        mg[d] = [c[d] - ib, c[d] + ib]
        mm[d] = root_mx[d] * 2 ** level
    up = [mg[d][1] > mm[d] for d in range(3)]
    dn = [mg[d][0] < 1 for d in range(3)]
    # (dims, signs of the first test) per offset: +1 upper edge, -1 lower edge
    table = {1: ((0,), (1,)), 2: ((1,), (1,)), 3: ((2,), (1,)),
             4: ((0, 1), (1, 1)), 5: ((0, 1), (1, -1)), 6: ((1, 2), (1, 1)),
             7: ((1, 2), (1, -1)), 8: ((0, 2), (1, 1)), 9: ((0, 2), (1, -1)),
             10: ((0, 1, 2), (1, 1, 1)), 11: ((0, 1, 2), (1, 1, -1)),
             12: ((0, 1, 2), (-1, 1, 1)), 13: ((0, 1, 2), (1, -1, 1))}
    dims, signs = table[off]

    def test(sg):
        ok = True
        for d, s in zip(dims, sg):
            ok = ok and (up[d] if s > 0 else dn[d])
        return ok

    first = test(signs)
    if off == 9:
        # second test of offset 9 compares x3's upper edge with mxmax(level, 2)
        second = dn[0] and mg[2][1] > mm[1]
    else:
        second = test([-s for s in signs])
    if first:
        for d, s in zip(dims, signs):
            c[d] += -mm[d] if s > 0 else mm[d]
        return c, True
    if second:
        for d, s in zip(dims, signs):
            c[d] += mm[d] if s > 0 else -mm[d]
        return c, True
    return c, False


def buffer_and_inflate(ndims, level, ib, patches, flags, root_mx, root_mthbc):
    """BufferAndList on every patch, then InflateEdgeTags. `patches`: list of
    (mx, mglobal) in level-list order; `flags`: list of arrays, modified."""
    tags = []
    if ib > 0:
        for (mx, mg), fl in zip(patches, flags):
            tmp = fl.copy()
            for k in range(1, mx[2] + 1):
                for j in range(1, mx[1] + 1):
                    for i in range(1, mx[0] + 1):
                        if tmp[i - 1, j - 1, k - 1] != 1:
                            continue
                        idx = (i, j, k)
                        lo = [1, 1, 1]
                        hi = [1, 1, 1]
                        for d in range(ndims):
                            lo[d], hi[d] = idx[d] - ib, idx[d] + ib
                        if any(lo[d] < 1 for d in range(3)) or any(hi[d] > mx[d] for d in range(3)):
                            c = [1, 1, 1]
                            for d in range(ndims):
                                c[d] = idx[d] + mg[d][0] - 1
                            tags.append(c)
                        else:
                            fl[lo[0] - 1:hi[0], lo[1] - 1:hi[1], lo[2] - 1:hi[2]] = 1

    def buffer_tag(c):
        for (mx, mg), fl in zip(patches, flags):
            lo = [1, 1, 1]
            hi = [1, 1, 1]
            for d in range(ndims):
                lo[d] = max(c[d] - ib, mg[d][0])
                hi[d] = min(c[d] + ib, mg[d][1])
            if any(hi[d] - lo[d] < 0 for d in range(3)):
                continue
            loc = [(lo[d] - mg[d][0] + 1, hi[d] - mg[d][0] + 1) if d < ndims else (1, 1)
                   for d in range(3)]
            fl[loc[0][0] - 1:loc[0][1], loc[1][0] - 1:loc[1][1], loc[2][0] - 1:loc[2][1]] = 1

    offs = periodic_offsets(root_mthbc, ndims)
    for c in reversed(tags):
        buffer_tag(c)
        for off in offs:
            cc, ok = periodic_tag_offset(c, level, off, ndims, ib, root_mx)
            if ok:
                buffer_tag(cc)


def _ratio(fl, ndims, b):
    total = 1
    for d in range(ndims):
        total *= b[d][1] - b[d][0] + 1
    flagged = int(fl[b[0][0] - 1:b[0][1], b[1][0] - 1:b[1][1], b[2][0] - 1:b[2][1]].sum())
    return flagged / total


def _signature(fl, ndims, b):
    box = fl[b[0][0] - 1:b[0][1], b[1][0] - 1:b[1][1], b[2][0] - 1:b[2][1]]
    sig = []
    for n in range(ndims):
        axes = tuple(a for a in range(3) if a != n)
        s = box.sum(axis=axes)
        sig.append({b[n][0] + t: int(s[t]) for t in range(len(s))})
    return sig


def new_subgrids(fl, mx, ndims, mgp, desfill):
    """NewSubGrids on one patch: list of boxes [[lo, hi] x 3]."""
    mb = [[1, mx[d] if d < ndims else 1] for d in range(3)]
    if _ratio(fl, ndims, mb) < 1.0e-10:
        return []
    sub = [[list(r) for r in mb]]
    cansplit = [True]
    igrid = 0
    while len(sub) < MAXSUBGRIDS and igrid < len(sub):
        npass = 0
        while cansplit[igrid] and npass < MAXSPLITPASSES:
            npass += 1
            sig = _signature(fl, ndims, sub[igrid])
            for n in range(ndims):
                lo, hi = sub[igrid][n]
                i1, i2 = lo, hi
                while sig[n][i1] == 0 and i1 < hi and i2 - i1 + 1 > mgp:
                    i1 += 1
                while sig[n][i2] == 0 and i2 > lo and i2 - i1 + 1 > mgp:
                    i2 -= 1
                sub[igrid][n] = [i1, i2]
            if not _ratio(fl, ndims, sub[igrid]) < desfill:
                cansplit[igrid] = False
                continue
            sig = _signature(fl, ndims, sub[igrid])
            isplit = [0, 0, 0]
            have = False
            for n in range(ndims):
                i1, i2 = sub[igrid][n]
                for i in range(i1 + mgp, i2 - mgp + 1):
                    if sig[n][i] == 0 and min(i - i1 + 1, i2 - i) >= mgp:
                        isplit[n] = i
                        have = True
                        break
                if have:
                    break
            if not have:
                dd = []
                for n in range(ndims):
                    i1, i2 = sub[igrid][n]
                    dd.append({i: sig[n][i - 1] - 2 * sig[n][i] + sig[n][i + 1]
                               for i in range(i1 + 1, i2)})
                inflect, dist0 = 0, 0
                for n in range(ndims):
                    i1, i2 = sub[igrid][n]
                    for i in range(i1 + max(2, mgp), i2 - max(1, mgp - 1) + 1):
                        delta = abs(dd[n][i] - dd[n][i - 1])
                        if delta > inflect:
                            inflect = delta
                            isplit = [0, 0, 0]
                            isplit[n] = i - 1
                            have = True
                            dist0 = min(i - i1, i2 - i + 1)
                        elif delta == inflect and inflect > 0:
                            dist1 = min(i - i1, i2 - i + 1)
                            if dist1 > dist0:
                                isplit = [0, 0, 0]
                                isplit[n] = i - 1
                                have = True
                                dist0 = dist1
            if not have:
                cansplit[igrid] = False
                continue
            for n in range(ndims):
                s = isplit[n]
                if s > 0 and min(sub[igrid][n][1] - s, s - sub[igrid][n][0] + 1) >= mgp:
                    nb = [list(r) for r in sub[igrid]]
                    nb[n][0] = s + 1
                    sub[igrid][n][1] = s
                    sub.append(nb)
                    cansplit.append(True)
                    break
        igrid += 1
    for b in sub:
        if min(b[d][1] - b[d][0] + 1 for d in range(ndims)) < mgp:
            raise ValueError("grid smaller than minimumgridpoints")
    return sub


def regrid_level(ndims, level, ib, mgp, desfill, patches, flags, root_mx, root_mthbc):
    """New subgrids of a level in creation order: [(parent index, mbounds)],
    mbounds [[lo, hi] x 3]. `flags` are buffered in place."""
    buffer_and_inflate(ndims, level, ib, patches, flags, root_mx, root_mthbc)
    out = []
    for k, ((mx, _), fl) in enumerate(zip(patches, flags)):
        for b in new_subgrids(fl, mx, ndims, mgp, desfill):
            out.append((k, [tuple(r) for r in b]))
    return out


def qinit2d(mx, h, xlower) -> np.ndarray:
    """QInit2D of the Cahn-Hilliard module (problems/simple_cahn-hilliard/
    problem.f90:58-101): q(0:mx+1, 0:my+1, 1) is 1 inside a set of
    rectangles (and below y = 0.5), 0 elsewhere; q(:, :, 2) = 0."""
    q = np.zeros((mx[0] + 2, mx[1] + 2, 1, 2), order="F")
    rects = [(0.8, 1.7, 0.5, 5.0), (2.8, 3.7, 0.5, 5.0), (4.3, 5.2, 1.5, 4.0),
             (3.5, 4.5, 1.5, 2.0), (5.8, 6.7, 0.5, 5.0), (4.0, 7.5, 4.3, 5.0),
             (9.8, 10.7, 0.5, 6.0), (9.0, 11.5, 5.3, 6.0)]
    for i in range(mx[0] + 2):
        x = xlower[0] + (i - 0.5) * h
        for j in range(mx[1] + 2):
            y = xlower[1] + (j - 0.5) * h
            inside = any(x < x1 and x > x0 and y < y1 and y > y0 for x0, x1, y0, y1 in rects)
            if inside or y < 0.5:
                q[i, j, 0, 0] = 1.0
    return q
