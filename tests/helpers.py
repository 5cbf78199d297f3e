"""Shared layouts and inputs for the tests."""

import numpy as np

from bsam_amd.layout import Layout


def layout_single(n=32, ndims=2, mthbc=(1, 1, 1, 1, 1, 1)):
    return Layout(ndims, (n,) * ndims, 1.0 / n, mthbc=mthbc)


def layout_amr2d(mthbc=(1, 1, 1, 1)):
    """16^2 root, two level-1 patches (one touching the left boundary edge
    region), level-2 patches including siblings sharing faces."""
    L = Layout(2, (16, 16), 1.0 / 16, mthbc=mthbc)
    c1 = L.add_child(0, ((3, 10), (5, 12)))
    L.add_child(0, ((11, 14), (5, 8)))
    L.add_child(c1, ((2, 5), (3, 8)))
    L.add_child(c1, ((6, 9), (3, 6)))
    c5 = L.add_child(0, ((1, 2), (13, 16)))
    L.add_child(c5, ((1, 2), (1, 4)))
    return L


def layout_amr2d_tiles(mthbc=(1, 1, 1, 1)):
    """Root 32^2 with a 2x2 block of touching level-1 patches covering a
    corner of the domain, and level-2 patches crossing sibling boundaries."""
    L = Layout(2, (32, 32), 1.0 / 32, mthbc=mthbc)
    a = L.add_child(0, ((1, 8), (1, 8)))
    b = L.add_child(0, ((9, 16), (1, 8)))
    c = L.add_child(0, ((1, 16), (9, 12)))
    L.add_child(a, ((5, 8), (5, 8)))
    L.add_child(b, ((1, 4), (5, 8)))
    L.add_child(c, ((3, 10), (1, 4)))
    return L


def random_fields(layout, seed=0, names=("q", "f", "qc")):
    rng = np.random.default_rng(seed)
    out = {}
    for p in layout.ordered():
        out[p.id] = {n: np.asfortranarray(rng.uniform(-1, 1, layout.field_shape(p, n)))
                     for n in names}
    return out


def noise_q(layout, seed=0, mean=0.5, amp=0.05):
    rng = np.random.default_rng(seed)
    out = {}
    for p in layout.ordered():
        q = np.zeros(layout.field_shape(p, "q"), order="F")
        q[..., 0] = mean + rng.uniform(-amp, amp, q.shape[:3])
        out[p.id] = q
    return out


def layout_amr3d(mthbc=(1, 1, 1, 1, 1, 1)):
    """8^3 root, two touching level-1 blocks, level-2 blocks (one touching a
    physical face of its parent)."""
    L = Layout(3, (8, 8, 8), 1.0 / 8, mthbc=mthbc)
    c1 = L.add_child(0, ((2, 5), (3, 6), (1, 4)))
    L.add_child(0, ((6, 8), (3, 6), (1, 4)))
    L.add_child(c1, ((2, 5), (3, 6), (2, 5)))
    L.add_child(c1, ((6, 8), (1, 4), (3, 6)))
    return L


def layout_periodic2d(mthbc=(2, 2, 1, 1)):
    """Children touching periodic faces, with periodic images of each other."""
    L = Layout(2, (16, 16), 1.0 / 16, mthbc=mthbc)
    a = L.add_child(0, ((1, 4), (3, 10)))
    b = L.add_child(0, ((13, 16), (5, 12)))
    L.add_child(0, ((6, 10), (13, 16)))
    L.add_child(0, ((6, 9), (1, 2)))
    L.add_child(a, ((1, 4), (1, 8)))
    L.add_child(b, ((5, 8), (3, 10)))
    return L


def layout_periodic3d(mthbc=(2, 2, 2, 2, 2, 2)):
    L = Layout(3, (8, 8, 8), 1.0 / 8, mthbc=mthbc)
    a = L.add_child(0, ((1, 3), (2, 5), (1, 4)))
    b = L.add_child(0, ((6, 8), (3, 6), (5, 8)))
    L.add_child(a, ((1, 3), (1, 4), (1, 4)))
    L.add_child(b, ((4, 6), (3, 6), (5, 8)))
    return L
