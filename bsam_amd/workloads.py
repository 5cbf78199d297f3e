"""Seeded synthetic inputs of the benchmark configurations.

Initial conditions use a counter-based SplitMix64 generator keyed by
(seed, level, global cell index), so every patch of every level draws the
same values however the hierarchy is cut into patches. Hierarchies are built
on the host the way a regrid would: tag cells, cover the tags with aligned
blocks, create child patches. These benchmark hierarchies stay fixed over
their steps; the reference's own regrid (Berger-Rigoutsos every step) is
bsam_regrid_level with bsam_amd/driver.py.
"""

from __future__ import annotations

import math
from typing import Callable, Dict, List, Tuple

import numpy as np

from .layout import Layout, Patch

M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x: np.ndarray) -> np.ndarray:
    """SplitMix64 finaliser of counters x (uint64, wrapping arithmetic)."""
    with np.errstate(over="ignore"):
        z = x.astype(np.uint64) + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform(seed: int, level: int, gidx: np.ndarray, lo: float, hi: float) -> np.ndarray:
    """U(lo, hi) for global cell indices gidx of a level."""
    with np.errstate(over="ignore"):
        key = (np.uint64(seed) * np.uint64(0xD1B54A32D192ED03)
               + np.uint64(level + 64) * np.uint64(0x8CB92BA72F3D8DD7))
        z = splitmix64(gidx.astype(np.uint64) ^ key)
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return lo + (hi - lo) * u


def global_index(lay: Layout, p: Patch) -> np.ndarray:
    """Global linear cell index over the patch interior (Fortran order)."""
    nd = lay.ndims
    root = lay.patches[0]
    nglob = [root.mx[d] * (2 ** p.level) if p.level >= 0 else root.mx[d] // 2 ** (-p.level)
             for d in range(nd)]
    ax = [np.arange(p.mglobal[d][0], p.mglobal[d][1] + 1, dtype=np.int64) - 1
          for d in range(nd)]
    grids = np.meshgrid(*ax, indexing="ij")
    g = grids[0].copy()
    mul = 1
    for d in range(1, nd):
        mul *= nglob[d - 1]
        g += grids[d] * mul
    return g


def centers(p: Patch, nd: int, ghost: int = 0) -> List[np.ndarray]:
    ax = [p.xlower[d] + (np.arange(1 - ghost, p.mx[d] + 1 + ghost) - 0.5) * p.dx
          for d in range(nd)]
    return np.meshgrid(*ax, indexing="ij")


def q_from_phi(lay: Layout, p: Patch, phi: np.ndarray) -> np.ndarray:
    """A ghost-extended q array with interior phi and mu = 0."""
    q = np.zeros(lay.field_shape(p, "q"), order="F")
    if lay.ndims == 2:
        q[1:-1, 1:-1, 0, 0] = phi
    else:
        q[1:-1, 1:-1, 1:-1, 0] = phi
    return q


def noise_q(lay: Layout, seed: int, mean: float, amp: float) -> Dict[int, np.ndarray]:
    out = {}
    for p in lay.ordered():
        if p.level < 0:
            out[p.id] = np.zeros(lay.field_shape(p, "q"), order="F")
            continue
        phi = uniform(seed, p.level, global_index(lay, p), mean - amp, mean + amp)
        out[p.id] = q_from_phi(lay, p, phi)
    return out


def analytic_q(lay: Layout, fn: Callable[..., np.ndarray]) -> Dict[int, np.ndarray]:
    out = {}
    for p in lay.ordered():
        if p.level < 0:
            out[p.id] = np.zeros(lay.field_shape(p, "q"), order="F")
            continue
        out[p.id] = q_from_phi(lay, p, fn(*centers(p, lay.ndims)))
    return out


# ---- regrid by tagging ---------------------------------------------------------
def _blocks(tags: np.ndarray, block: int, maxw: int) -> List[Tuple[Tuple[int, int], ...]]:
    """Aligned blocks of `block` cells holding a tag, merged along dim 1 into
    runs of at most `maxw` cells. Returns 1-based inclusive bounds."""
    nd = tags.ndim
    nb = [math.ceil(tags.shape[d] / block) for d in range(nd)]
    marked = np.zeros(nb, bool)
    for idx in np.ndindex(*nb):
        sl = tuple(slice(idx[d] * block, min((idx[d] + 1) * block, tags.shape[d]))
                   for d in range(nd))
        marked[idx] = tags[sl].any()
    out = []
    per = max(1, maxw // block)
    for rest in np.ndindex(*nb[1:]):
        b = 0
        while b < nb[0]:
            if not marked[(b,) + rest]:
                b += 1
                continue
            e = b
            while e + 1 < nb[0] and marked[(e + 1,) + rest] and e + 1 - b < per:
                e += 1
            lo = [b * block + 1] + [r * block + 1 for r in rest]
            hi = [min((e + 1) * block, tags.shape[0])] + \
                 [min((r + 1) * block, tags.shape[d + 1]) for d, r in enumerate(rest)]
            out.append(tuple((lo[d], hi[d]) for d in range(nd)))
            b = e + 1
    return out


def _dilate(t: np.ndarray, n: int) -> np.ndarray:
    for _ in range(n):
        u = t.copy()
        for d in range(t.ndim):
            u |= np.roll(t, 1, d) | np.roll(t, -1, d)
        t = u
    return t


def refine_by_gradient(lay: Layout, fn: Callable[..., np.ndarray], nlevels: int,
                       frac: float = 0.5, buffer: int = 1, block: int = 8,
                       maxw: int = 32) -> None:
    """Add `nlevels` refined levels where |grad phi| > frac * max over the
    level (central differences of the analytic phi), tags dilated by
    `buffer` cells and covered by aligned blocks of block..maxw coarse cells
    (2*block .. 2*maxw fine cells per side)."""
    nd = lay.ndims
    for lev in range(0, nlevels):
        parents = [lay.patches[i] for i in lay.level_ids(lev)]
        grads = []
        for p in parents:
            X = centers(p, nd, ghost=1)
            phi = fn(*X)
            g2 = 0.0
            for d in range(nd):
                sl_p = [slice(1, -1)] * nd
                sl_m = [slice(1, -1)] * nd
                sl_p[d] = slice(2, None)
                sl_m[d] = slice(0, -2)
                g2 = g2 + ((phi[tuple(sl_p)] - phi[tuple(sl_m)]) / (2 * p.dx)) ** 2
            grads.append(np.sqrt(g2))
        gmax = max(float(g.max()) for g in grads)
        for p, g in zip(parents, grads):
            tags = _dilate(g > frac * gmax, buffer)
            for mb in _blocks(tags, block, maxw):
                lay.add_child(p.id, mb)


# ---- the benchmark configurations --------------------------------------------
def circles_phi(seed: int, eps: float, n: int = 5, radius: float = 0.25):
    rng = np.random.default_rng(seed)
    cen = rng.uniform(0.2, 0.8, size=(n, 2))

    def phi(X, Y):
        r = np.full(X.shape, np.inf)
        for cx, cy in cen:
            r = np.minimum(r, np.hypot(X - cx, Y - cy))
        return np.tanh((radius - r) / (math.sqrt(2.0) * eps))
    return phi


C2_PARAMS = dict(eps=0.03, dt=1e-3, well="plusminus", nsmoothingpasses=2, qerrortol=1e-8)


def config_c2(root: int = 256, seed: int = 2):
    """2D AMR, plain FP64: root 256^2 on [0,1]^2 (Neumann) plus 2 levels
    refined where |grad phi| > 0.5 max, blocks of 16..64 fine cells per side;
    phi0 = tanh((0.25 - r)/(sqrt(2) eps)), r the distance to 5 seeded centres."""
    eps = C2_PARAMS["eps"]
    lay = Layout(2, (root, root), 1.0 / root)
    phi = circles_phi(seed, eps)
    refine_by_gradient(lay, phi, nlevels=2)
    return lay, analytic_q(lay, phi)


def config_c1(root: int = 128, seed: int = 1):
    """2D single level 128^2, phi0 = U(-0.05, 0.05)."""
    lay = Layout(2, (root, root), 1.0 / root)
    return lay, noise_q(lay, seed, 0.0, 0.05)


def nested_random_layout(root: int, nlevels: int, counts, seed: int,
                         smin: int = 8, smax: int = 128, margin: int = 1,
                         tries_per_patch: int = 4000) -> Layout:
    """A properly nested synthetic 2D hierarchy: on each level, `counts[l]`
    patches whose side lengths (fine cells, even) are drawn independently
    per dimension log-uniform in [smin, smax], whatever the parents' sizes.

    The level's patches are placed largest first. Each goes to a random free
    position, at least `margin` coarse cells from the parent's edges and not
    overlapping a sibling, in a parent drawn in proportion to its remaining
    free area among the parents large enough to hold it. A patch that finds
    no room in `tries_per_patch` attempts is left out (none is on the
    benchmark configuration; `layout.dropped` counts them per level)."""
    rng = np.random.default_rng(seed)
    lay = Layout(2, (root, root), 1.0 / root)
    occ = {0: np.zeros((root, root), bool)}
    lay.dropped = []
    for lev in range(1, nlevels + 1):
        parents = [lay.patches[i] for i in lay.level_ids(lev - 1)]
        sides = np.exp(rng.uniform(np.log(smin), np.log(smax), size=(counts[lev - 1], 2)))
        cs = np.maximum(smin // 2, np.round(sides / 2.0).astype(int))   # coarse cells
        order = np.argsort(-(cs[:, 0] * cs[:, 1]), kind="stable")
        free = np.array([float(p.mx[0] * p.mx[1]) for p in parents])
        dropped = 0
        for n in order:
            cx, cy = int(cs[n, 0]), int(cs[n, 1])
            cand = [i for i, p in enumerate(parents)
                    if p.mx[0] - 2 * margin >= cx and p.mx[1] - 2 * margin >= cy]
            placed = False
            if cand:
                w = free[cand] / free[cand].sum()
                for _ in range(tries_per_patch):
                    pi = cand[rng.choice(len(cand), p=w)]
                    par = parents[pi]
                    lo0 = margin + 1 + int(rng.integers(0, par.mx[0] - 2 * margin - cx + 1))
                    lo1 = margin + 1 + int(rng.integers(0, par.mx[1] - 2 * margin - cy + 1))
                    sl = (slice(lo0 - 1, lo0 - 1 + cx), slice(lo1 - 1, lo1 - 1 + cy))
                    o = occ[par.id]
                    if o[sl].any():
                        continue
                    o[sl] = True
                    free[pi] -= cx * cy
                    cid = lay.add_child(par.id, ((lo0, lo0 + cx - 1), (lo1, lo1 + cy - 1)))
                    occ[cid] = np.zeros((2 * cx, 2 * cy), bool)
                    placed = True
                    break
            dropped += not placed
        lay.dropped.append(dropped)
    return lay


C3_PARAMS = dict(eps=0.03, dt=1e-3, well="plusminus", nsmoothingpasses=2, qerrortol=1e-8)


def config_c3(root: int = 512, seed: int = 3, counts=(300, 640, 1060)):
    """2D root 512^2 plus 3 refined levels of 2000 uneven patches (sides
    8..128 fine cells, log-uniform per dimension on every level: about
    3.9 M refined cells), phi0 = 0.1 + U(-0.05, 0.05)."""
    lay = nested_random_layout(root, 3, counts, seed)
    return lay, noise_q(lay, seed, 0.1, 0.05)


C4_PARAMS = dict(eps=0.02, dt=5e-4, well="plusminus", nsmoothingpasses=2, qerrortol=1e-8)


def nested_block_layout_3d(root: int, nlevels: int, fill, seed: int, sides=(16, 32),
                           margin: int = 1, max_tries: int = 100000) -> Layout:
    """A properly nested synthetic 3D hierarchy: on each level, cubic patches
    of side (fine cells) drawn from `sides`, placed at random inside a parent
    (at least `margin` coarse cells from its faces, not overlapping siblings)
    until they cover `fill[l]` of the parent level's volume."""
    rng = np.random.default_rng(seed)
    lay = Layout(3, (root,) * 3, 1.0 / root)
    occ = {0: np.zeros((root,) * 3, bool)}
    for lev in range(1, nlevels + 1):
        parents = [lay.patches[i] for i in lay.level_ids(lev - 1)]
        vol = np.array([np.prod(p.mx) for p in parents], float)
        target = fill[lev - 1] * vol.sum()
        covered = 0.0
        tries = 0
        while covered < target and tries < max_tries:
            tries += 1
            par = parents[rng.choice(len(parents), p=vol / vol.sum())]
            cs = int(rng.choice(sides)) // 2          # coarse cells per side
            room = [par.mx[d] - 2 * margin - cs for d in range(3)]
            if min(room) < 0:
                continue
            lo = [margin + 1 + int(rng.integers(0, room[d] + 1)) for d in range(3)]
            sl = tuple(slice(lo[d] - 1, lo[d] - 1 + cs) for d in range(3))
            o = occ[par.id]
            if o[sl].any():
                continue
            o[sl] = True
            cid = lay.add_child(par.id, tuple((lo[d], lo[d] + cs - 1) for d in range(3)))
            occ[cid] = np.zeros((2 * cs,) * 3, bool)
            covered += cs ** 3
    return lay


def config_c4(root: int = 128, seed: int = 4, fill=(0.25, 0.25)):
    """3D root 128^3 plus 2 refined levels of 16^3 and 32^3 blocks (each
    level covering about a quarter of its parent level), phi0 = U(-0.05, 0.05)."""
    lay = nested_block_layout_3d(root, 2, fill, seed)
    return lay, noise_q(lay, seed, 0.0, 0.05)

