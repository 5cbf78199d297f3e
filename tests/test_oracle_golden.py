"""The C oracle against fixtures written by the reference itself.

tests/golden/*.npz hold inputs and outputs of the reference BSAM modules
(compiled unmodified from the reference sources with flang and driven by
oracle/ref_harness.f90; see tests/golden/make_golden.py). The oracle keeps
the reference's operation order, so results must agree bit for bit.
"""

import json
import os

import numpy as np
import pytest

import orc
from golden_cases import LAYOUTS, OPS_CASES, PM_DT, PM_EPS, QPO, STEP_CASES

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def run_ops(o, ops):
    """Runs ops on the oracle; returns the statistics the stats ops gave."""
    stats = []
    for op in ops:
        name, a = op[0], op[1:]
        if name == "setghost":
            o.set_ghost(*a)
        elif name == "relax":
            o.relax(*a)
        elif name == "restrict":
            o.restrict(*a)
        elif name == "coarseghost":
            o.coarse_ghost(*a)
        elif name == "coarseload":
            o.coarse_load(*a)
        elif name == "correct":
            o.correct(*a)
        elif name == "filldown":
            o.fill_down(*a)
        elif name == "vcycle":
            o.vcycle(*a)
        elif name == "truncerr":
            o.truncation_error(*a)
        elif name == "userflags":
            o.error_flags_user(a[0], a[1] * 1e-3)
        elif name == "stats":
            stats.extend(o.after_step_stats())
        else:
            raise KeyError(name)
    return stats


@pytest.mark.parametrize("case", sorted(OPS_CASES))
def test_ops_match_reference(case):
    lname, _, ops = OPS_CASES[case]
    lay = LAYOUTS[lname]()
    z = load(f"ops_{case}.npz")
    o = orc.Oracle(lay, dt=1e-3, eps=0.03, alpha=1.0, qpo=QPO.get(case))
    for n, _ in enumerate(o.patches):
        for f in ("q", "f", "qc"):
            o.field(n, f)[...] = z[f"in/{n}/{f}"]
    stats = run_ops(o, ops)
    for n, p in enumerate(o.patches):
        for f in ("q", "f", "qc", "qold", "qrte"):
            key = f"out/{n}/{f}"
            if f in ("q", "f", "qc") or key in z:
                assert np.array_equal(o.field(n, f), z[key]), (case, p.id, f)
        if f"out/{n}/errorflags" in z:
            assert np.array_equal(o.flags(n), z[f"out/{n}/errorflags"]), (case, p.id)
    if "stats" in z:
        # the reference sums in array element order, as the oracle does
        assert stats == list(z["stats"]), (stats, z["stats"])


@pytest.mark.parametrize("case", sorted(k for k in STEP_CASES if k != "plusminus_map"))
def test_time_steps_match_reference(case):
    spec = STEP_CASES[case]
    lay = LAYOUTS[spec["layout"]]()
    z = load(f"steps_{case}.npz")
    o = orc.Oracle(lay, dt=spec["dt"], eps=spec["eps"], alpha=spec["alpha"],
                   maxvcycles=spec["maxv"], qerrortol=spec["tol"], qpo=QPO.get(case))
    for n, _ in enumerate(o.patches):
        o.field(n, "q")[...] = z[f"in/{n}/q"]
    res = z["residuals"]
    got = []
    for s in range(spec["steps"]):
        got.extend(o.time_step())
    assert len(got) == len(res)
    assert np.array_equal(np.array(got), res[:, 2])
    for n, _ in enumerate(o.patches):
        key = f"out/{n}/q"
        if key in z:
            assert np.array_equal(o.field(n, "q"), z[key]), n


def test_plusminus_well_matches_mapped_reference():
    """The +-1 well (Dfc = phi^3, Dfe = phi) is the reference well under
    c = (phi+1)/2, alpha = 8, eps_ref = sqrt(2) eps, dt_ref = dt/2: six
    V-cycles of both agree to rounding (phi = 2c-1, mu equal)."""
    spec = STEP_CASES["plusminus_map"]
    lay = LAYOUTS[spec["layout"]]()
    z = load("steps_plusminus_map.npz")
    o = orc.Oracle(lay, dt=PM_DT, eps=PM_EPS, well="plusminus", maxvcycles=6, qerrortol=0.0)
    for n, _ in enumerate(o.patches):
        c = z[f"in/{n}/q"]
        o.field(n, "q")[..., 0] = 2.0 * c[..., 0] - 1.0
        o.field(n, "q")[..., 1] = c[..., 1]
    hist = o.time_step()
    assert len(hist) == 6
    for n, p in enumerate(o.patches):
        ref = z[f"out/{n}/q"]
        phi, mu = o.field(n, "q")[..., 0], o.field(n, "q")[..., 1]
        np.testing.assert_allclose(phi, 2.0 * ref[..., 0] - 1.0, rtol=0, atol=1e-12)
        np.testing.assert_allclose(mu, ref[..., 1], rtol=0, atol=1e-11 * np.abs(ref[..., 1]).max())


@pytest.mark.parametrize("lname", sorted(LAYOUTS))
def test_layout_matches_reference_tree(lname):
    """Patch order (youngest first per level), mx, mbounds, mglobal and mthbc
    equal the reference tree's (bsamroutines MakeNewGrid / InitSeed rules)."""
    listing = json.load(open(os.path.join(GOLDEN, "layouts_reference.json")))[lname]
    lay = LAYOUTS[lname]()
    pats = lay.ordered()
    assert len(pats) == len(listing)
    for p, row in zip(pats, listing):
        ngrid, level = row[0], row[1]
        mx = row[2:5]
        mb = row[5:11]      # column-major mbounds(3,2): lo1 lo2 lo3 hi1 hi2 hi3
        mg = row[11:17]
        bc = row[17:23]
        assert (p.id, p.level) == (ngrid, level)
        assert list(p.mx) == mx
        assert [p.mbounds[d][0] for d in range(3)] == mb[:3]
        assert [p.mbounds[d][1] for d in range(3)] == mb[3:]
        assert [p.mglobal[d][0] for d in range(3)] == mg[:3]
        assert [p.mglobal[d][1] for d in range(3)] == mg[3:]
        assert list(p.mthbc) == bc
