"""Layouts and inputs of the golden fixtures (shared by make_golden.py and
the tests that read the fixtures back)."""

from __future__ import annotations

import math

import numpy as np

from bsam_amd.layout import Layout


def _amr2d():
    L = Layout(2, (16, 16), 1.0 / 16)
    c1 = L.add_child(0, ((3, 10), (5, 12)))
    L.add_child(0, ((11, 14), (5, 8)))
    L.add_child(c1, ((2, 5), (3, 8)))
    L.add_child(c1, ((6, 9), (3, 6)))
    c5 = L.add_child(0, ((1, 2), (13, 16)))
    L.add_child(c5, ((1, 2), (1, 4)))
    return L


def _amr2d_userbc():
    L = Layout(2, (16, 16), 1.0 / 16, mthbc=(10, 10, 10, 10))
    a = L.add_child(0, ((1, 8), (1, 8)))
    b = L.add_child(0, ((9, 16), (1, 8)))
    L.add_child(0, ((1, 16), (9, 16)))
    L.add_child(a, ((1, 8), (1, 8)))
    L.add_child(b, ((1, 8), (3, 8)))
    return L


def _rect2d():
    # non-square root (the shipped example is 128 x 64) and a non-unit domain
    L = Layout(2, (24, 12), 0.5, xlower=(1.0, -2.0))
    c = L.add_child(0, ((4, 15), (2, 9)))
    L.add_child(c, ((3, 6), (3, 12)))
    L.add_child(c, ((7, 12), (1, 4)))
    return L


def _amr3d(mthbc=(1, 1, 1, 1, 1, 1)):
    L = Layout(3, (8, 8, 8), 1.0 / 8, mthbc=mthbc)
    c1 = L.add_child(0, ((2, 5), (3, 6), (1, 4)))
    L.add_child(0, ((6, 8), (3, 6), (1, 4)))
    L.add_child(c1, ((2, 5), (3, 6), (2, 5)))
    L.add_child(c1, ((6, 8), (1, 4), (3, 6)))
    return L


def _periodic2d(mthbc=(2, 2, 2, 2)):
    # children on the periodic faces, images of one another across them
    L = Layout(2, (16, 16), 1.0 / 16, mthbc=mthbc)
    a = L.add_child(0, ((1, 4), (3, 10)))
    b = L.add_child(0, ((13, 16), (5, 12)))
    L.add_child(0, ((6, 10), (13, 16)))
    L.add_child(0, ((6, 9), (1, 2)))
    L.add_child(a, ((1, 4), (1, 8)))
    L.add_child(b, ((5, 8), (3, 10)))
    return L


def _periodic3d(mthbc=(2, 2, 2, 2, 2, 2)):
    L = Layout(3, (8, 8, 8), 1.0 / 8, mthbc=mthbc)
    a = L.add_child(0, ((1, 3), (2, 5), (1, 4)))
    b = L.add_child(0, ((6, 8), (3, 6), (5, 8)))
    L.add_child(a, ((1, 3), (1, 4), (1, 4)))
    L.add_child(b, ((4, 6), (3, 6), (5, 8)))
    return L


LAYOUTS = {
    "single32": lambda: Layout(2, (32, 32), 1.0 / 32),
    "c1_128": lambda: Layout(2, (128, 128), 1.0 / 128),
    "amr2d": _amr2d,
    "amr2d_userbc": _amr2d_userbc,
    "rect2d": _rect2d,
    "amr3d": _amr3d,
    "amr3d_userbc": lambda: _amr3d((1, 10, 1, 10, 10, 1)),
    "single16_3d": lambda: Layout(3, (16, 16, 16), 1.0 / 16),
    "periodic2d": _periodic2d,
    "periodic2d_x_userbc": lambda: _periodic2d((2, 2, 10, 10)),
    "periodic3d": _periodic3d,
    "periodic3d_xz": lambda: _periodic3d((2, 2, 1, 1, 2, 2)),
}

OPS = [("setghost", 1, 0), ("setghost", 2, 1), ("setghost", 2, 2), ("relax", 2, 1),
       ("relax", 1, 2), ("restrict", 2), ("setghost", 1, 0), ("coarseghost", 2),
       ("coarseload", 2), ("correct", 2), ("filldown", 2), ("filldown", 1)]

# case -> (layout, seed, op sequence): one fixture per case with the
# inputs and the fields after the whole sequence
USERFLAG_OPS = [("setghost", 0, 0), ("userflags", 0, 900), ("setghost", 1, 0),
                ("userflags", 1, 1200), ("setghost", 2, 0), ("userflags", 2, 1500),
                ("stats", 0, 0)]

OPS_CASES = {
    "amr2d_seq": ("amr2d", 1, OPS),
    "amr2d_vcycle": ("amr2d", 2, [("vcycle", 2)]),
    "userbc_seq": ("amr2d_userbc", 3, OPS),
    "rect2d_vcycle": ("rect2d", 4, [("vcycle", 2)]),
    "amr3d_seq": ("amr3d", 7, OPS),
    "amr3d_userbc_vcycle": ("amr3d_userbc", 8, [("vcycle", 2)]),
    "periodic2d_seq": ("periodic2d", 9, OPS),
    "periodic2d_x_userbc_vcycle": ("periodic2d_x_userbc", 10, [("vcycle", 2)]),
    "periodic3d_seq": ("periodic3d", 11, OPS),
    "periodic3d_xz_vcycle": ("periodic3d_xz", 12, [("vcycle", 2)]),
    # EstimateLevelErrors' truncation error on the two refined levels
    "amr2d_truncerr": ("amr2d", 13, [("truncerr", 2), ("truncerr", 1), ("truncerr", 0)]),
    "userbc_truncerr": ("amr2d_userbc", 14, [("truncerr", 2), ("truncerr", 1), ("truncerr", 0)]),
    "amr3d_truncerr": ("amr3d", 15, [("truncerr", 2), ("truncerr", 1), ("truncerr", 0)]),
    "periodic2d_truncerr": ("periodic2d", 16, [("truncerr", 2), ("truncerr", 1), ("truncerr", 0)]),
    "periodic3d_truncerr": ("periodic3d", 17, [("truncerr", 1), ("truncerr", 0)]),
    # ErrFlag's user tagging (qtolerance = N * 1e-3) after SetGhost(level, 0),
    # as AMR runs it, then the after-step statistics (mass, 2D energy)
    "amr2d_userflags": ("amr2d", 18, USERFLAG_OPS),
    "userbc_userflags": ("amr2d_userbc", 19, USERFLAG_OPS),
    "amr3d_userflags": ("amr3d", 20, USERFLAG_OPS),
    "periodic2d_userflags": ("periodic2d", 21, USERFLAG_OPS),
    "periodic3d_userflags": ("periodic3d", 22, USERFLAG_OPS),
    # quasi-periodic offsets (QPO below) on the periodic layouts
    "periodic2d_qpo_seq": ("periodic2d", 23, OPS),
    "periodic2d_x_userbc_qpo_vcycle": ("periodic2d_x_userbc", 24, [("vcycle", 2)]),
    "periodic3d_qpo_seq": ("periodic3d", 25, OPS),
    "periodic3d_xz_qpo_vcycle": ("periodic3d_xz", 26, [("vcycle", 2)]),
}

# qpo(1:nrvars) of the cases that set it (0 for the others): the values
# copied across the periodic x boundary gain +qpo(n) on the high side and
# -qpo(n) on the low side
QPO = {
    "periodic2d_qpo_seq": (0.375, -1.25),
    "periodic2d_x_userbc_qpo_vcycle": (0.125, 0.5),
    "periodic3d_qpo_seq": (-0.625, 0.25),
    "periodic3d_xz_qpo_vcycle": (0.25, -0.75),
    "periodic2d_qpo": (0.02, 0.0),
    "periodic3d_qpo": (-0.015, 0.001),
}


def noise_init(mean, amp, seed):
    def init(lay):
        rng = np.random.default_rng(seed)
        out = {}
        for p in lay.ordered():
            q = np.zeros(lay.field_shape(p, "q"), order="F")
            q[..., 0] = mean + rng.uniform(-amp, amp, q.shape[:3])
            out[p.id] = {"q": q, "f": np.zeros(lay.field_shape(p, "f"), order="F"),
                         "qc": np.zeros(lay.field_shape(p, "qc"), order="F")}
        return out
    return init


# The +-1 double well F = (phi^2-1)^2/4 with (eps, dt) is the reference
# module's well F = alpha c^2 (c-1)^2/4 under c = (phi+1)/2 with alpha = 8,
# eps_ref = sqrt(2) eps and dt_ref = dt/2 (mu unchanged). PM_* hold that map.
PM_EPS, PM_DT = 0.03, 1e-3

STEP_CASES = {
    # C1 geometry (128^2, [0,1]^2, Neumann, eps 0.03, dt 1e-3), reference
    # double well around its spinodal point c = 0.5
    "c1_128": dict(layout="c1_128", init=noise_init(0.5, 0.05, 1), dt=1e-3, alpha=1.0,
                   eps=0.03, maxv=40, tol=1e-8, steps=2, keep_levels=(0,)),
    "single32": dict(layout="single32", init=noise_init(0.5, 0.05, 2), dt=1e-3, alpha=1.0,
                     eps=0.03, maxv=40, tol=1e-8, steps=3),
    "amr2d": dict(layout="amr2d", init=noise_init(0.5, 0.05, 3), dt=1e-3, alpha=1.0,
                  eps=0.03, maxv=40, tol=1e-8, steps=3),
    "amr2d_userbc": dict(layout="amr2d_userbc", init=noise_init(0.6, 0.05, 4), dt=2e-3,
                         alpha=2.0, eps=0.05, maxv=40, tol=1e-9, steps=2),
    "rect2d": dict(layout="rect2d", init=noise_init(0.5, 0.1, 5), dt=5e-2, alpha=1.0,
                   eps=0.3, maxv=40, tol=1e-8, steps=2),
    "amr3d": dict(layout="amr3d", init=noise_init(0.5, 0.05, 7), dt=5e-4, alpha=1.0,
                  eps=0.02 * 8, maxv=40, tol=1e-8, steps=2),
    "single16_3d": dict(layout="single16_3d", init=noise_init(0.5, 0.05, 8), dt=5e-4,
                        alpha=1.0, eps=0.1, maxv=40, tol=1e-8, steps=2),
    "periodic2d": dict(layout="periodic2d", init=noise_init(0.5, 0.05, 13), dt=1e-3,
                       alpha=1.0, eps=0.03, maxv=40, tol=1e-8, steps=3),
    "periodic3d": dict(layout="periodic3d", init=noise_init(0.5, 0.05, 14), dt=5e-4,
                       alpha=1.0, eps=0.16, maxv=40, tol=1e-8, steps=2),
    "periodic2d_qpo": dict(layout="periodic2d", init=noise_init(0.5, 0.05, 15), dt=1e-3,
                           alpha=1.0, eps=0.03, maxv=40, tol=1e-8, steps=3),
    "periodic3d_qpo": dict(layout="periodic3d", init=noise_init(0.5, 0.05, 16), dt=5e-4,
                           alpha=1.0, eps=0.16, maxv=40, tol=1e-8, steps=2),
    # fixed 6 V-cycles (tol 0) of the mapped +-1 problem, phi ~ 0.1 + U(-.05,.05)
    "plusminus_map": dict(layout="amr2d", init=noise_init(0.55, 0.025, 6), dt=PM_DT / 2,
                          alpha=8.0, eps=math.sqrt(2.0) * PM_EPS, maxv=6, tol=0.0, steps=1),
}
