"""Run the reference BSAM modules (oracle/_ref/harness) on a Layout.

Test and measurement infrastructure: the fixture generator and the tests
use it to cross-check the C oracle against the reference itself, and
bench.py's CPU-baseline leg times the reference's own time steps with it.
The harness is built from the reference sources in the build container
(oracle/Makefile) and travels as a binary. Arrays travel as raw FP64 in
Fortran order, patch by patch in ApplyOnLevel order (levels ascending,
youngest patch first).
"""

from __future__ import annotations

import os
import resource
import subprocess
import tempfile

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HARNESS = os.path.join(REPO, "oracle", "_ref", "harness")


def _big_stack():
    """The reference keeps array temporaries on the stack (a 128^3 patch's
    exceed the usual 8 MiB): run the harness with the stack limit raised to
    the hard limit."""
    _, hard = resource.getrlimit(resource.RLIMIT_STACK)
    resource.setrlimit(resource.RLIMIT_STACK, (hard, hard))


def available() -> bool:
    return os.path.exists(HARNESS)


def write_spec(path, layout, dt, alpha, eps, nsmooth, maxv, tol, ops):
    root = layout.patches[0]
    with open(path, "w") as fh:
        fh.write(f"{layout.ndims} {layout.nrvars} {layout.mbc}\n")
        fh.write(f"{dt!r} {alpha!r} {eps!r}\n".replace("e", "d"))
        fh.write(f"{nsmooth} {maxv} {tol!r}\n".replace("e", "d"))
        fh.write(f"{layout.minlevel} {layout.finestlevel}\n")
        fh.write(" ".join(str(v) for v in root.mx) + " "
                 + " ".join(repr(v).replace("e", "d") for v in root.xlower) + " "
                 + repr(root.dx).replace("e", "d") + " "
                 + " ".join(str(v) for v in root.mthbc) + "\n")
        ref = layout.refined()
        fh.write(f"{len(ref)}\n")
        for p in ref:
            mb = " ".join(f"{p.mbounds[d][0]} {p.mbounds[d][1]}" for d in range(3))
            fh.write(f"{p.id} {p.level} {p.parent_id} {mb}\n")
        fh.write(f"{len(ops)}\n")
        for op in ops:
            name, a1, a2 = (list(op) + [0, 0])[:3]
            fh.write(f"{name} {a1} {a2}\n")


def run(layout, arrays, ops, dt=1e-3, alpha=1.0, eps=0.03, nsmooth=2, maxv=20,
        tol=1e-8, keep=None, qpo=None):
    """Run ``ops`` on the reference. ``arrays`` maps patch id -> dict of
    'q', 'f', 'qc' Fortran-order arrays; ``qpo`` the quasi-periodic offsets
    (one per variable, 0 when None). Returns (outputs, residuals, listing)
    with outputs: id -> {'q','f','qc','rf','qold','qrte','errorflags'}
    (errorflags as int32)."""
    if not available():
        raise RuntimeError("reference harness not built (make -C oracle ref)")
    tmp = keep or tempfile.mkdtemp(prefix="bsamref_")
    os.makedirs(tmp, exist_ok=True)
    spec = os.path.join(tmp, "spec.txt")
    fin = os.path.join(tmp, "in.bin")
    fout = os.path.join(tmp, "out.bin")
    fres = os.path.join(tmp, "res.txt")
    write_spec(spec, layout, dt, alpha, eps, nsmooth, maxv, tol, ops)
    os.makedirs(os.path.join(tmp, "out"), exist_ok=True)   # the 'uniform' op's file
    order = layout.ordered()
    with open(fin, "wb") as fh:
        for p in order:
            a = arrays[p.id]
            for name in ("q", "f", "qc"):
                arr = np.asarray(a[name], dtype=np.float64)
                assert arr.shape == layout.field_shape(p, name), (p.id, name, arr.shape)
                fh.write(arr.tobytes(order="F"))
    env = dict(os.environ)
    env.pop("BSAM_HARNESS_QPO", None)
    if qpo is not None:
        env["BSAM_HARNESS_QPO"] = " ".join(repr(float(v)) for v in qpo)
    proc = subprocess.run([HARNESS, spec, fin, fout, fres], cwd=tmp, env=env,
                          capture_output=True, text=True, preexec_fn=_big_stack)
    if proc.returncode != 0:
        raise RuntimeError(f"harness failed: {proc.stdout}\n{proc.stderr}")
    raw = np.fromfile(fout, dtype=np.float64)
    out, pos = {}, 0
    for p in order:
        d = {}
        for name in ("q", "f", "qc", "rf", "qold", "qrte", "errorflags"):
            shape = layout.field_shape(p, name)
            n = int(np.prod(shape))
            d[name] = raw[pos:pos + n].reshape(shape, order="F").copy()
            pos += n
        d["errorflags"] = d["errorflags"].astype(np.int32)
        out[p.id] = d
    assert pos == raw.size
    res = []
    with open(fres) as fh:
        for line in fh:
            s, it, r = line.split()
            res.append((int(s), int(it), float(r)))
    # MultigridIterations prints "it = N  residual = R" per V-cycle
    # (lib/afasroutines.f90:96); the harness precedes each solve with
    # "harness step S".
    step = 0
    for line in proc.stdout.splitlines():
        tok = line.split()
        if tok[:2] == ["harness", "step"]:
            step = int(tok[2])
        elif len(tok) == 6 and tok[0] == "it" and tok[3] == "residual":
            res.append((step, int(tok[2]), float(tok[5])))
    listing = []
    with open(fout + ".patches") as fh:
        for line in fh:
            listing.append([int(v) for v in line.split()])
    return out, res, listing
