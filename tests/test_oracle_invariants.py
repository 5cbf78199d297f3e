"""Closed-form checks of the oracle, independent of the reference fixtures.

Each check is chosen so that a plausible restatement mistake (a dropped
term, a wrong index or sign, a transposed operand) breaks it.
"""

import numpy as np
import pytest

import orc
from bsam_amd.layout import Layout
from helpers import layout_amr2d, layout_amr2d_tiles, layout_amr3d


def centers(p, ndims, shape):
    """Cell-centre coordinates over the ghost-extended index range."""
    ax = []
    for d in range(ndims):
        idx = np.arange(shape[d]) - 1 + 1  # index i from 0 .. mx+1
        ax.append(p.xlower[d] + (idx - 0.5) * p.dx)
    return np.meshgrid(*ax, indexing="ij")


def test_interpolation_weights_sum_to_one():
    a = (-1 / 5, 2 / 3, 8 / 15)
    b = (5 / 32, 15 / 16, -3 / 32)
    c = (9 / 16, 3 / 16, 3 / 16, 1 / 16)
    assert abs(sum(a) - 1) < 1e-15 and abs(sum(b) - 1) < 1e-15 and abs(sum(c) - 1) < 1e-15
    # quadratic Lagrange through parameters {-1, 0.5, 1.5} evaluated at -0.5
    t, x = -0.5, (-1.0, 0.5, 1.5)
    lag = [np.prod([(t - x[m]) / (x[k] - x[m]) for m in range(3) if m != k]) for k in range(3)]
    np.testing.assert_allclose(lag, (a[2], a[1], a[0]), rtol=1e-14)
    # tangential quadratic through {-1, 0, 1} at +1/4
    t, x = 0.25, (1.0, 0.0, -1.0)
    lag = [np.prod([(t - x[m]) / (x[k] - x[m]) for m in range(3) if m != k]) for k in range(3)]
    np.testing.assert_allclose(lag, b, rtol=1e-14)


@pytest.mark.parametrize("mk", [layout_amr2d, layout_amr2d_tiles])
def test_set_ghost_reproduces_linear_fields(mk):
    """Coarse-fine interpolation is exact for linear data at every ghost that
    a sibling or the interpolation fills (physical faces excepted)."""
    lay = mk()
    o = orc.Oracle(lay, dt=1e-3, eps=0.03)
    lin = lambda X, Y: 0.3 + 1.7 * X - 0.9 * Y
    for level in range(1, lay.finestlevel + 1):
        # fresh linear data everywhere (parent ghosts included), then one level
        for n, p in enumerate(o.patches):
            q = o.field(n, "q")
            X, Y = centers(p, 2, q.shape)
            q[:, :, 0, 0] = lin(X, Y)
            q[:, :, 0, 1] = 2 * lin(X, Y)
        o.set_ghost(level, 0)
        check_level(o, lay, level, lin)


def check_level(o, lay, level, lin):
    for n, p in enumerate(o.patches):
        if p.level != level:
            continue
        q = o.field(n, "q")
        X, Y = centers(p, 2, q.shape)
        want = lin(X, Y)
        # ghost ring cells not on a physical face
        mask = np.zeros(q.shape[:2], bool)
        mask[0, :] = mask[-1, :] = mask[:, 0] = mask[:, -1] = True
        if p.mthbc[0] < 999:
            mask[0, :] = False
        if p.mthbc[1] < 999:
            mask[-1, :] = False
        if p.mthbc[2] < 999:
            mask[:, 0] = False
        if p.mthbc[3] < 999:
            mask[:, -1] = False
        np.testing.assert_allclose(q[:, :, 0, 0][mask], want[mask], atol=1e-13)
        np.testing.assert_allclose(q[:, :, 0, 1][mask], 2 * want[mask], atol=1e-13)


def test_restriction_conserves_mass_and_constants():
    lay = layout_amr2d()
    o = orc.Oracle(lay, dt=1e-3, eps=0.03)
    rng = np.random.default_rng(1)
    for n, p in enumerate(o.patches):
        o.field(n, "q")[...] = rng.uniform(-1, 1, o.field(n, "q").shape)
    o.restrict(2)
    for n, p in enumerate(o.patches):
        if p.level != 2:
            continue
        q, qc = o.field(n, "q"), o.field(n, "qc")
        fine = q[1:-1, 1:-1, 0, :].sum(axis=(0, 1)) * p.dx ** 2
        coarse = qc[1:-1, 1:-1, 0, :].sum(axis=(0, 1)) * (2 * p.dx) ** 2
        np.testing.assert_allclose(coarse, fine, rtol=1e-13)
    # constants restrict to themselves
    for n, p in enumerate(o.patches):
        o.field(n, "q")[...] = 0.375
    o.fill_down(2)
    for n, p in enumerate(o.patches):
        assert np.all(o.field(n, "q")[1:-1, 1:-1] == 0.375)


def test_correction_prolongs_linear_corrections_exactly():
    """CorrectFine adds P1(parent q - qc); a linear coarse correction yields
    the same linear function at the fine centres, ghost ring included."""
    lay = Layout(2, (16, 16), 1.0 / 16)
    lay.add_child(0, ((5, 12), (3, 10)))
    o = orc.Oracle(lay, dt=1e-3, eps=0.03)
    idx = lay.index_map()
    root, child = idx[0], idx[1]
    lin = lambda X, Y: -0.2 + 0.8 * X + 0.45 * Y
    pr, pc = o.patches[root], o.patches[child]
    qr = o.field(root, "q")
    X, Y = centers(pr, 2, qr.shape)
    qr[:, :, 0, 0] = lin(X, Y)
    qr[:, :, 0, 1] = 0.0
    o.field(child, "qc")[...] = 0.0
    o.field(child, "q")[...] = 0.0
    o.correct(1)
    q = o.field(child, "q")
    X, Y = centers(pc, 2, q.shape)
    np.testing.assert_allclose(q[:, :, 0, 0], lin(X, Y), atol=1e-13)
    assert np.all(q[:, :, 0, 1] == 0.0)


def test_relaxed_cells_satisfy_the_linear_equation():
    """After relaxing a colour, the first (linear, Mob = 1) equation of the CH
    system holds at every relaxed cell: f1 - A1(q) = 0 to rounding."""
    lay = Layout(2, (16, 16), 1.0 / 16)
    o = orc.Oracle(lay, dt=1e-2, eps=0.05)
    rng = np.random.default_rng(3)
    n = lay.index_map()[0]
    o.field(n, "q")[...] = rng.uniform(0, 1, o.field(n, "q").shape)
    o.field(n, "f")[...] = rng.uniform(-1, 1, o.field(n, "f").shape)
    o.relax(0, 1)
    o.error_afas(0)            # fills rf = f - A(q)
    rf = o.field(n, "rf")
    i, j = np.meshgrid(np.arange(1, 17), np.arange(1, 17), indexing="ij")
    red = (i + j + 1) % 2 == 1
    assert np.abs(rf[:, :, 0, 0][red]).max() < 1e-12
    assert np.abs(rf[:, :, 0, 0][~red]).max() > 1e-3


def test_operator_on_constants():
    """A(q) for constant (phi, mu): A1 = phi, A2 = mu - Dfc(phi)."""
    lay = Layout(2, (8, 8), 1.0 / 8)
    o = orc.Oracle(lay, dt=1e-3, eps=0.03, alpha=2.0)
    n = lay.index_map()[0]
    q = o.field(n, "q")
    q[..., 0] = 0.8
    q[..., 1] = -0.3
    o.field(n, "f")[...] = 0.0
    o.error_afas(0)
    rf = o.field(n, "rf")
    np.testing.assert_allclose(-rf[..., 0], 0.8, rtol=1e-15)
    np.testing.assert_allclose(-rf[..., 1], -0.3 - 2.0 * (0.8 - 0.5) ** 3, rtol=1e-14)


@pytest.mark.parametrize("mk", [layout_amr2d, layout_amr3d])
def test_error_flags_threshold(mk):
    """SetErrFlags2D/3D (lib/bsamroutines.f90:1093-1157), which the reference's
    module keeps private: the 2^d children of a coarse cell are flagged exactly
    when the largest |qrte| over the variables is strictly above
    qtolerance / h^d, and every other flag is cleared. One coarse cell holds
    (0.9 x, -x) with x and h powers of two, so every threshold is exact."""
    lay = mk()
    o = orc.Oracle(lay, dt=1e-3, eps=0.03)
    level = lay.finestlevel
    nd = lay.ndims
    pats = [(n, p) for n, p in enumerate(o.patches) if p.level == level]
    n0, p0 = pats[0]
    x = 2.0 ** 5
    h = p0.dx
    cell = (1, 2, 1 if nd == 3 else 0)
    for tol_factor, flagged in ((0.5, True), (0.95, True), (1.0, False), (2.0, False)):
        for n, p in pats:
            o.field(n, "qrte")[...] = 0.0
            o.flags(n)[...] = 7
        o.field(n0, "qrte")[cell + (0,)] = 0.9 * x
        o.field(n0, "qrte")[cell + (1,)] = -x
        o.error_flags(level, tol_factor * x * h ** nd)
        for n, p in pats:
            want = np.zeros(lay.field_shape(p, "errorflags"), dtype=np.int32)
            if n == n0 and flagged:
                ci, cj, ck = cell
                want[2 * ci:2 * ci + 2, 2 * cj:2 * cj + 2,
                     (2 * ck if nd == 3 else 0):(2 * ck + 2 if nd == 3 else 1)] = 1
            assert np.array_equal(o.flags(n), want), (tol_factor, p.id)
