"""Host hooks of the Cahn-Hilliard problem module that the time-step driver
needs besides the solve: the initial condition (QInit2D/QInit3D,
problems/simple_cahn-hilliard/problem.f90:58-101, :365-397, called through
QInit, lib/bsamroutines.f90:540-566).

The initial condition is evaluated once per initial grid, on the host, into
the patch's q array (the reference's layout, ``Layout.field_shape(p, 'q')``),
before it is uploaded.
"""

from __future__ import annotations

import math
from typing import Sequence

import numpy as np

# QInit2D's rectangles (x0, x1, y0, y1): q1 = 1 where x0 < x < x1 and
# y0 < y < y1, in the order of its if/elseif chain, or where y < 0.5
_RECTS_2D = ((0.8, 1.7, 0.5, 5.0), (2.8, 3.7, 0.5, 5.0), (4.3, 5.2, 1.5, 4.0),
             (3.5, 4.5, 1.5, 2.0), (5.8, 6.7, 0.5, 5.0), (4.0, 7.5, 4.3, 5.0),
             (9.8, 10.7, 0.5, 6.0), (9.0, 11.5, 5.3, 6.0))


def qinit2d(mx: Sequence[int], h: float, xlower: Sequence[float], nrvars: int = 2) -> np.ndarray:
    """q(0:mx+1, 0:my+1, 1, 1:nrvars) of QInit2D: cell centres
    x = xlower + (i - 0.5) h over the ghost-extended patch."""
    ex, ey = mx[0] + 2, mx[1] + 2
    x = xlower[0] + (np.arange(ex, dtype=np.float64) - 0.5) * h
    y = xlower[1] + (np.arange(ey, dtype=np.float64) - 0.5) * h
    X, Y = np.meshgrid(x, y, indexing="ij")
    one = Y < 0.5
    for x0, x1, y0, y1 in _RECTS_2D:
        one |= (X < x1) & (X > x0) & (Y < y1) & (Y > y0)
    q = np.zeros((ex, ey, 1, nrvars), order="F")
    q[:, :, 0, 0] = np.where(one, 1.0, 0.0)
    return q


def qinit3d(mx: Sequence[int], h: float, xlower: Sequence[float], eps: float,
            nrvars: int = 2) -> np.ndarray:
    """q(0:mx+1, 0:my+1, 0:mz+1, 1:nrvars) as QInit leaves it in 3D.

    QInit passes QInit3D the section q(1:mx, 1:my, 1:mz+1, :) while QInit3D
    declares its dummy q(0:, 0:, 0:, 1:) and loops i, j, k over 0..mx+1: the
    dummy's (i, j, k, 1) is the array element one step further along each
    dimension of the ghost-extended array (0-based offset (i+1) + px (j+1) +
    px py (k+1)), past the section's extent in every dimension, so the
    values land at those offsets of the whole array, running into the next
    row, plane and variable. The section itself (1..mx, 1..my, 1..mz+1, all
    variables) is zeroed first (QInit3D's q = 0)."""
    ex, ey, ez = mx[0] + 2, mx[1] + 2, mx[2] + 2
    q = np.zeros((ex, ey, ez, nrvars), order="F")
    flat = q.reshape(-1, order="F")
    s2 = 2.0 * math.sqrt(2.0) * eps
    for i in range(ex):
        x = (i - 0.5) * h + xlower[0]
        for j in range(ey):
            y = (j - 0.5) * h + xlower[1]
            for k in range(ez):
                z = (k - 0.5) * h + xlower[2]
                # x**2 is x*x in the reference's arithmetic
                a, b, c = x - 1.6, y - 1.6, z - 1.6
                r1 = math.sqrt(a * a + b * b + c * c)
                b, c = y - 1.2, z - 1.2
                r2 = math.sqrt(a * a + b * b + c * c)
                tmp = 1.0 - 0.5 * (1.0 - math.tanh((r1 - 1.0) / s2))
                tmp = tmp * (1.0 - 0.5 * (1.0 - math.tanh((r2 - 0.5) / s2)))
                off = (i + 1) + ex * (j + 1) + ex * ey * (k + 1)
                if off < flat.size:
                    flat[off] = 1.0 - tmp
    return q
