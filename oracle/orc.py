"""ctypes wrapper of the C FP64 oracle (liborc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, always as the checker, never by the product.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborc.so")

FIELDS = {"q": 0, "qold": 1, "qc": 2, "f": 3, "rf": 4, "ftmp": 5, "qrte": 6}

# Problem constants. REFERENCE_CH is the reference module's double well
# (problems/simple_cahn-hilliard/problem.f90:781-813); PLUSMINUS_CH is the
# +-1 well F = (phi^2-1)^2/4 with Dfc = phi^3, Dfe = phi.
REFERENCE_CH = dict(fc_shift=0.5, fe_den=4.0)   # fc_alpha = fe_num = alpha
PLUSMINUS_CH = dict(fc_alpha=1.0, fc_shift=0.0, fe_num=1.0, fe_den=1.0)

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run make -C oracle")
        L = C.CDLL(LIB_PATH)
        vp, ip, dp = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double)
        L.orc_create.restype = vp
        L.orc_create.argtypes = [C.c_int] * 4
        L.orc_destroy.argtypes = [vp]
        L.orc_add_patch.restype = C.c_int
        L.orc_add_patch.argtypes = [vp, C.c_int, C.c_int, C.c_int, ip, ip, ip, ip,
                                    C.c_double, dp]
        L.orc_commit.argtypes = [vp]
        L.orc_set_problem.argtypes = [vp] + [C.c_double] * 6
        L.orc_set_solver.argtypes = [vp, C.c_int, C.c_int, C.c_double]
        L.orc_set_qpo.argtypes = [vp, C.c_double, C.c_double]
        L.orc_field.restype = dp
        L.orc_field.argtypes = [vp, C.c_int, C.c_int]
        L.orc_field_len.restype = C.c_long
        L.orc_field_len.argtypes = [vp, C.c_int, C.c_int]
        L.orc_error.restype = C.c_char_p
        L.orc_error.argtypes = [vp]
        for name in ("orc_set_ghost", "orc_relax_level"):
            getattr(L, name).argtypes = [vp, C.c_int, C.c_int]
        for name in ("orc_restrict_level", "orc_coarse_ghost_level",
                     "orc_coarse_load_level", "orc_correct_level",
                     "orc_fill_down", "orc_vcycle", "orc_truncation_error"):
            getattr(L, name).argtypes = [vp, C.c_int]
        L.orc_error_flags.argtypes = [vp, C.c_int, C.c_double]
        L.orc_error_flags_user.argtypes = [vp, C.c_int, C.c_double]
        L.orc_after_step_stats.argtypes = [vp, dp]
        L.orc_error_flag_field.restype = ip
        L.orc_error_flag_field.argtypes = [vp, C.c_int]
        L.orc_error_flag_len.restype = C.c_long
        L.orc_error_flag_len.argtypes = [vp, C.c_int]
        for name in ("orc_source", "orc_update_source", "orc_copy_qold"):
            getattr(L, name).argtypes = [vp]
        L.orc_error_afas.restype = C.c_double
        L.orc_error_afas.argtypes = [vp, C.c_int]
        L.orc_multigrid_iterations.argtypes = [vp, dp, C.c_int]
        L.orc_time_step.argtypes = [vp, dp, C.c_int]
        _lib = L
    return _lib


def _ia(vals):
    return (C.c_int * len(vals))(*vals)


class Oracle:
    """An oracle hierarchy mirroring a :class:`bsam_amd.layout.Layout`."""

    def __init__(self, layout, dt, eps, alpha=None, well="reference",
                 nsmoothingpasses=2, maxvcycles=20, qerrortol=1e-8, qpo=None):
        L = lib()
        self.layout = layout
        self.h = L.orc_create(layout.ndims, layout.nrvars, layout.mbc, layout.minlevel)
        self.patches = layout.ordered()
        idx = layout.index_map()
        for p in self.patches:
            mb = [v for d in range(3) for v in p.mbounds[d]]
            mg = [v for d in range(3) for v in p.mglobal[d]]
            par = idx[p.parent_id] if p.parent_id is not None else -1
            n = L.orc_add_patch(self.h, p.id, p.level, par, _ia(p.mx), _ia(mb), _ia(mg),
                                _ia(p.mthbc), p.dx, (C.c_double * 3)(*p.xlower))
            if n < 0:
                raise RuntimeError(self.error())
        if L.orc_commit(self.h):
            raise RuntimeError(self.error())
        if well == "reference":
            a = 1.0 if alpha is None else alpha
            L.orc_set_problem(self.h, dt, eps * eps, a, 0.5, a, 4.0)
        elif well == "plusminus":
            L.orc_set_problem(self.h, dt, eps * eps, 1.0, 0.0, 1.0, 1.0)
        else:
            raise ValueError(well)
        L.orc_set_solver(self.h, nsmoothingpasses, maxvcycles, qerrortol)
        if qpo is not None:
            q = [float(v) for v in qpo] + [0.0, 0.0]
            L.orc_set_qpo(self.h, q[0], q[1])
        self.maxvcycles = maxvcycles

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def error(self):
        return lib().orc_error(self.h).decode()

    def field(self, n, name):
        """Writable Fortran-order view of field ``name`` of patch index ``n``."""
        L = lib()
        ptr = L.orc_field(self.h, n, FIELDS[name])
        cnt = L.orc_field_len(self.h, n, FIELDS[name])
        flat = np.ctypeslib.as_array(ptr, shape=(cnt,))
        shape = self.layout.field_shape(self.patches[n], name)
        return flat.reshape(shape, order="F")

    def _ck(self, rc):
        if rc < 0:
            raise RuntimeError(self.error())
        return rc

    def set_ghost(self, level, ipass):
        self._ck(lib().orc_set_ghost(self.h, level, ipass))

    def relax(self, level, color):
        self._ck(lib().orc_relax_level(self.h, level, color))

    def restrict(self, level):
        self._ck(lib().orc_restrict_level(self.h, level))

    def coarse_ghost(self, level):
        self._ck(lib().orc_coarse_ghost_level(self.h, level))

    def coarse_load(self, level):
        self._ck(lib().orc_coarse_load_level(self.h, level))

    def correct(self, level):
        self._ck(lib().orc_correct_level(self.h, level))

    def fill_down(self, field):
        self._ck(lib().orc_fill_down(self.h, field))

    def source(self):
        lib().orc_source(self.h)
        lib().orc_update_source(self.h)

    def update_source(self):
        lib().orc_update_source(self.h)

    def truncation_error(self, level):
        """EstimateLevelErrors' truncation error on ``level`` into qrte."""
        self._ck(lib().orc_truncation_error(self.h, level))

    def error_flags(self, level, qtolerance):
        """SetErrFlags on ``level`` from qrte."""
        self._ck(lib().orc_error_flags(self.h, level, qtolerance))

    def error_flags_user(self, level, qtolerance):
        """SetErrFlagsUser2D/3D on ``level`` from q."""
        self._ck(lib().orc_error_flags_user(self.h, level, qtolerance))

    def after_step_stats(self):
        """(mass, energy) of the composite grid; energy is 0 in 3D."""
        out = (C.c_double * 2)()
        self._ck(lib().orc_after_step_stats(self.h, out))
        return out[0], out[1]

    def flags(self, n):
        """Writable Fortran-order view of patch ``n``'s error flags."""
        L = lib()
        ptr = L.orc_error_flag_field(self.h, n)
        cnt = L.orc_error_flag_len(self.h, n)
        flat = np.ctypeslib.as_array(ptr, shape=(cnt,))
        return flat.reshape(self.layout.field_shape(self.patches[n], "errorflags"), order="F")

    def vcycle(self, level):
        self._ck(lib().orc_vcycle(self.h, level))

    def error_afas(self, level):
        return lib().orc_error_afas(self.h, level)

    def copy_qold(self):
        lib().orc_copy_qold(self.h)

    def multigrid_iterations(self):
        hist = (C.c_double * self.maxvcycles)()
        n = self._ck(lib().orc_multigrid_iterations(self.h, hist, self.maxvcycles))
        return list(hist[:n])

    def time_step(self):
        hist = (C.c_double * self.maxvcycles)()
        n = self._ck(lib().orc_time_step(self.h, hist, self.maxvcycles))
        return list(hist[:n])
