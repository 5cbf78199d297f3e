"""The reference Fortran host code drives the engine through its Fortran
binding (bsam_amd/fortran/bsam_gpu.f90, afas_gpu.f90).

tests/fortran/dropin_check is linked from the reference's own modules
(oracle/_ref, compiled from the reference sources), the binding and
libbsamgpu.so. On a CPU it checks the engine's ghost-fill programs applied to
the reference's own arrays against the reference SetGhost (host-only context);
on a GPU it checks GpuMultigridIterations against the reference
MultigridIterations time step by time step. Both must agree bit for bit.
Skipped where the reference build (oracle/_ref) is absent.
"""

import os
import subprocess

import numpy as np
import pytest

import refrun
from golden_cases import LAYOUTS

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "fortran", "_build", "dropin_check")
REF_OBJ = os.path.join(ROOT, "oracle", "_ref", "treeops.o")


def ensure_built():
    if not os.path.exists(REF_OBJ):
        pytest.skip("reference objects (oracle/_ref) not built")
    if not os.path.exists(BIN):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "fortran")], check=True)


def write_case(tmp_path, lname, seed, dt=1e-3, alpha=1.0, eps=0.03, maxv=40, tol=1e-8):
    lay = LAYOUTS[lname]()
    spec = os.path.join(tmp_path, "spec.txt")
    inp = os.path.join(tmp_path, "in.bin")
    refrun.write_spec(spec, lay, dt, alpha, eps, 2, maxv, tol, [])
    rng = np.random.default_rng(seed)
    with open(inp, "wb") as fh:
        for p in lay.ordered():
            q = np.zeros(lay.field_shape(p, "q"), order="F")
            q[..., 0] = 0.5 + rng.uniform(-0.05, 0.05, q.shape[:-1])
            for name, arr in (("q", q), ("f", np.zeros(lay.field_shape(p, "f"))),
                              ("qc", np.zeros(lay.field_shape(p, "qc")))):
                fh.write(np.asarray(arr, dtype=np.float64).tobytes(order="F"))
    return spec, inp


def run(mode, spec, inp, *extra, qpo=None):
    """Runs dropin_check; `qpo` sets the module's quasi-periodic offsets
    (read by the tree set-up from BSAM_HARNESS_QPO)."""
    env = dict(os.environ)
    env.pop("BSAM_HARNESS_QPO", None)
    if qpo is not None:
        env["BSAM_HARNESS_QPO"] = " ".join(repr(float(v)) for v in qpo)
    out = subprocess.run([BIN, mode, spec, inp, *extra], capture_output=True, text=True,
                         timeout=300, env=env)
    assert out.returncode == 0, out.stdout + out.stderr
    diff = [l for l in out.stdout.splitlines() if l.startswith("max_abs_diff")]
    assert diff, out.stdout
    return float(diff[-1].split()[1]), out.stdout


@pytest.mark.parametrize("lname", ["amr2d", "amr2d_userbc", "rect2d", "amr3d", "amr3d_userbc",
                                   "periodic2d", "periodic2d_x_userbc", "periodic3d",
                                   "periodic3d_xz"])
def test_ghost_programs_on_reference_arrays(tmp_path, lname):
    ensure_built()
    spec, inp = write_case(str(tmp_path), lname, seed=1)
    diff, _ = run("host", spec, inp)
    assert diff == 0.0


@pytest.mark.parametrize("lname", ["periodic2d", "periodic2d_x_userbc", "periodic3d",
                                   "periodic3d_xz"])
def test_ghost_programs_with_qpo_on_reference_arrays(tmp_path, lname):
    """GpuOpen hands the module's qpo to the engine: its host programs match
    the reference SetGhost with quasi-periodic offsets."""
    ensure_built()
    spec, inp = write_case(str(tmp_path), lname, seed=4)
    diff, _ = run("host", spec, inp, qpo=(0.375, -1.25))
    assert diff == 0.0


@pytest.mark.gpu
@pytest.mark.parametrize("lname", ["periodic2d", "periodic3d_xz"])
def test_gpu_multigrid_iterations_with_qpo_replace_reference(tmp_path, lname):
    if not os.path.exists(BIN):
        pytest.skip("dropin_check not built (needs the reference build)")
    spec, inp = write_case(str(tmp_path), lname, seed=5)
    diff, out = run("gpu", spec, inp, "2", qpo=(0.02, -0.001))
    assert "converged 1" in out, out
    assert diff == 0.0, out


@pytest.mark.gpu
@pytest.mark.parametrize("lname", ["amr2d", "rect2d", "amr3d", "periodic2d", "periodic3d"])
def test_gpu_multigrid_iterations_replace_reference(tmp_path, lname):
    if not os.path.exists(BIN):
        pytest.skip("dropin_check not built (needs the reference build)")
    spec, inp = write_case(str(tmp_path), lname, seed=2)
    diff, out = run("gpu", spec, inp, "2")
    assert "converged 1" in out, out
    assert diff == 0.0, out


@pytest.mark.gpu
@pytest.mark.parametrize("lname", ["amr2d", "amr2d_userbc", "amr3d", "periodic2d", "periodic3d"])
def test_gpu_truncation_error_replaces_reference(tmp_path, lname):
    """GpuRelativeTruncationError leaves the reference arrays q, qold, qc and
    qrte bit for bit as the reference calls of EstimateLevelErrors do, and
    GpuSetErrFlags' flags follow SetErrFlags on that qrte."""
    if not os.path.exists(BIN):
        pytest.skip("dropin_check not built (needs the reference build)")
    spec, inp = write_case(str(tmp_path), lname, seed=3)
    diff, out = run("trunc", spec, inp)
    assert diff == 0.0, out
    flagged = [l for l in out.splitlines() if l.startswith("flagged")]
    assert flagged and int(flagged[-1].split()[1]) > 0 and flagged[-1].split()[3] == "0", out
