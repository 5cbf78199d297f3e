"""Python binding of the engine's C ABI (include/bsam_gpu.h) via ctypes.

:class:`Engine` mirrors the reference's solver interface for one patch
hierarchy: the method names follow the Fortran routines they replace
(``set_ghost`` for SetGhost, ``vcycle`` for AFASVcycle,
``multigrid_iterations`` for MultigridIterations, ...). There is no
fallback: if ``libbsamgpu.so`` is missing or no HIP device is present, the
constructor raises.
"""

from __future__ import annotations

import ctypes as C
import os
from typing import Dict, List, Optional, Sequence

import numpy as np

from .layout import Layout

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbsamgpu.so")

FIELD_F, FIELD_Q, FIELD_QOLD, FIELD_QC, FIELD_FTMP, FIELD_RF, FIELD_QRTE = 1, 2, 3, 4, 5, 6, 7
FIELDS = {"f": FIELD_F, "q": FIELD_Q, "qold": FIELD_QOLD, "qc": FIELD_QC,
          "ftmp": FIELD_FTMP, "rf": FIELD_RF, "qrte": FIELD_QRTE}
PREC_FP64, PREC_MIXED = 0, 1

# Double-well constants: the reference module (problem.f90:781-813) with
# alpha, and the +-1 well F = (phi^2-1)^2/4.
WELLS = {
    "reference": lambda alpha: (alpha, 0.5, alpha, 4.0),
    "plusminus": lambda alpha: (1.0, 0.0, 1.0, 1.0),
}


class BsamError(RuntimeError):
    pass


class Cfg(C.Structure):
    _fields_ = [("ndims", C.c_int), ("nrvars", C.c_int), ("mbc", C.c_int),
                ("minlevel", C.c_int), ("dt", C.c_double), ("eps2", C.c_double),
                ("fc_alpha", C.c_double), ("fc_shift", C.c_double),
                ("fe_num", C.c_double), ("fe_den", C.c_double),
                ("nsmoothingpasses", C.c_int), ("maxvcycles", C.c_int),
                ("qerrortol", C.c_double), ("precision", C.c_int), ("device", C.c_int),
                ("errortype", C.c_int)]


_lib = None


class RegridParams(C.Structure):
    _fields_ = [("ndims", C.c_int), ("level", C.c_int), ("ibuffer", C.c_int),
                ("minimumgridpoints", C.c_int), ("desiredfillratio", C.c_double),
                ("root_mx", C.c_int * 3), ("root_mthbc", C.c_int * 6)]


def regrid_level(ndims: int, level: int, ibuffer: int, minimumgridpoints: int,
                 desiredfillratio: float, root_mx, root_mthbc, patches, flags):
    """New subgrids of one level (BufferAndList, InflateEdgeTags, NewSubGrids)
    from its patches [(mx, mglobal)] in level-list order and their error flags
    (int32 arrays of shape mx, Fortran order; buffered in place). Returns
    [(parent index, mbounds)] in creation order."""
    L = lib()
    if ndims not in (2, 3):
        raise BsamError("regrid: ndims must be 2 or 3")
    rp = RegridParams(ndims, level, ibuffer, minimumgridpoints, desiredfillratio,
                      (C.c_int * 3)(*[int(v) for v in list(root_mx)[:3]] + [1] * (3 - len(root_mx))),
                      (C.c_int * 6)(*[int(v) for v in list(root_mthbc)[:6]] +
                                    [1] * (6 - len(root_mthbc))))
    n = len(patches)
    mx = np.zeros(3 * max(n, 1), dtype=np.int32)
    mg = np.ones(6 * max(n, 1), dtype=np.int32)
    ptrs = (C.POINTER(C.c_int) * max(n, 1))()
    keep = []
    for k, ((pmx, pmg), fl) in enumerate(zip(patches, flags)):
        for d in range(3):
            mx[3 * k + d] = pmx[d] if d < ndims else 1
            if d < ndims:
                mg[6 * k + 2 * d], mg[6 * k + 2 * d + 1] = pmg[d][0], pmg[d][1]
        if fl.dtype != np.int32 or not fl.flags.f_contiguous:
            raise ValueError("flags must be Fortran-ordered int32 arrays")
        keep.append(fl)
        ptrs[k] = fl.ctypes.data_as(C.POINTER(C.c_int))
    cap = 1
    for (pmx, _) in patches:
        c = 1
        for d in range(ndims):
            c *= pmx[d]
        cap += c
    parent = np.zeros(cap, dtype=np.int32)
    mb = np.zeros(6 * cap, dtype=np.int32)
    ng = C.c_int()
    ip = C.POINTER(C.c_int)
    rc = L.bsam_regrid_level(C.byref(rp), n, mx.ctypes.data_as(ip), mg.ctypes.data_as(ip),
                             ptrs, cap, C.byref(ng), parent.ctypes.data_as(ip),
                             mb.ctypes.data_as(ip))
    if rc != 0:
        raise BsamError(L.bsam_last_error(None).decode())
    return [(int(parent[g]), tuple((int(mb[6 * g + 2 * d]), int(mb[6 * g + 2 * d + 1]))
                                   for d in range(3))) for g in range(ng.value)]


def copy_rate(device: int = 0, nbytes: int = 1 << 30, reps: int = 20) -> float:
    """Measured device-to-device copy rate (GB/s, read plus write bytes) with
    the engine's copy kernel on buffers of nbytes."""
    g = C.c_double()
    if lib().bsam_copy_rate(device, nbytes, reps, C.byref(g)) != 0:
        raise BsamError(lib().bsam_last_error(None).decode())
    return g.value


def lib():
    """Load libbsamgpu.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise BsamError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        vp, ip, dp = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double)
        L.bsam_ctx_create.restype = vp
        L.bsam_ctx_create.argtypes = [C.POINTER(Cfg)]
        L.bsam_ctx_destroy.argtypes = [vp]
        L.bsam_last_error.restype = C.c_char_p
        L.bsam_last_error.argtypes = [vp]
        L.bsam_hierarchy_begin.argtypes = [vp]
        L.bsam_add_patch.argtypes = [vp, C.c_int, C.c_int, C.c_int, ip, ip, ip, ip,
                                     C.c_double, dp]
        L.bsam_hierarchy_commit.argtypes = [vp]
        L.bsam_num_patches.argtypes = [vp]
        L.bsam_footprint_bytes.restype = C.c_longlong
        L.bsam_footprint_bytes.argtypes = [vp]
        L.bsam_tune_get.argtypes = [vp, C.c_char_p, dp]
        L.bsam_profile_kernel.argtypes = [vp, C.c_int, C.c_int, dp, dp]
        L.bsam_field_len.restype = C.c_long
        L.bsam_field_len.argtypes = [vp, C.c_int, C.c_int]
        L.bsam_upload.argtypes = [vp, C.c_int, C.c_int, dp, C.c_long]
        L.bsam_download.argtypes = [vp, C.c_int, C.c_int, dp, C.c_long]
        for n in ("bsam_set_ghost", "bsam_relax"):
            getattr(L, n).argtypes = [vp, C.c_int, C.c_int]
        for n in ("bsam_restrict_solution", "bsam_coarse_ghost_points", "bsam_coarse_load",
                  "bsam_correct_fine", "bsam_fill_down", "bsam_vcycle",
                  "bsam_truncation_error"):
            getattr(L, n).argtypes = [vp, C.c_int]
        L.bsam_error_flags.argtypes = [vp, C.c_int, C.c_double]
        L.bsam_download_flags.argtypes = [vp, C.c_int, ip, C.c_long]
        L.bsam_error_flags_user.argtypes = [vp, C.c_int, C.c_double]
        L.bsam_copy_levels.argtypes = [vp, vp, C.c_int]
        L.bsam_solve_status.argtypes = [vp]
        L.bsam_format_rows.restype = C.c_long
        L.bsam_format_rows.argtypes = [C.POINTER(C.c_double), C.c_long, C.c_int, C.c_long,
                                       C.c_long, C.c_int, C.c_char_p, C.c_long]
        L.bsam_init_fields.argtypes = [vp, C.c_int]
        L.bsam_transfer_values.argtypes = [vp, vp, C.c_int]
        L.bsam_overlay_level.argtypes = [vp, vp, C.c_int]
        L.bsam_set_dt.argtypes = [vp, C.c_double]
        L.bsam_set_qpo.argtypes = [vp, dp, C.c_int]
        L.bsam_regrid_level.argtypes = [C.POINTER(RegridParams), C.c_int, ip, ip,
                                        C.POINTER(ip), C.c_int, ip, ip, ip]
        L.bsam_after_step_stats.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        for n in ("bsam_source", "bsam_copy_qold", "bsam_sync"):
            getattr(L, n).argtypes = [vp]
        L.bsam_error_afas.argtypes = [vp, C.c_int, dp]
        L.bsam_set_groups.argtypes = [vp, C.c_int, ip, dp]
        L.bsam_num_groups.argtypes = [vp]
        L.bsam_group_result.argtypes = [vp, C.c_int, ip, dp, C.c_int, ip]
        L.bsam_after_step_stats_group.argtypes = [vp, C.c_int, dp, dp]
        L.bsam_source_norm.argtypes = [vp, C.c_int, dp]
        L.bsam_copy_rate.argtypes = [C.c_int, C.c_longlong, C.c_int, dp]
        for n in ("bsam_multigrid_iterations", "bsam_time_step"):
            getattr(L, n).argtypes = [vp, ip, dp, C.c_int, ip]
        L.bsam_tune.argtypes = [vp, C.c_char_p, C.c_double]
        L.bsam_profile_begin.argtypes = [vp, C.c_uint]
        L.bsam_profile_end.argtypes = [vp]
        L.bsam_profile_get.argtypes = [vp, C.c_int, C.POINTER(C.c_longlong),
                                       C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.bsam_ghost_program_size.argtypes = [vp, C.c_int, C.c_int]
        L.bsam_ghost_program_dst.argtypes = [vp, C.c_int, C.c_int, ip, C.c_int]
        L.bsam_ghost_program_apply_host.argtypes = [vp, C.c_int, C.c_int,
                                                    C.POINTER(C.POINTER(C.c_double))]
        L.bsam_fused_ghost_size.argtypes = [vp, C.c_int, C.c_int, C.c_int]
        L.bsam_fused_ghost_pre_size.argtypes = [vp, C.c_int, C.c_int, C.c_int]
        L.bsam_ghost_rides_down.argtypes = [vp, C.c_int, C.c_int]
        L.bsam_fused_ghost_apply_host.argtypes = [vp, C.c_int, C.c_int, C.c_int,
                                                  C.POINTER(C.POINTER(C.c_double))]
        L.bsam_fused_ghost_part_apply_host.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int,
                                                       C.POINTER(C.POINTER(C.c_double))]
        L.bsam_post_ghost_size.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int]
        L.bsam_post_ghost_apply_host.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int,
                                                 C.POINTER(C.POINTER(C.c_double))]
        L.bsam_iso_patches.argtypes = [vp, C.c_int, ip, C.c_int]
        L.bsam_ride_loop_size.argtypes = [vp, C.c_int]
        L.bsam_ride_loop_apply_host.argtypes = [vp, C.c_int, C.POINTER(C.POINTER(C.c_double))]
        L.bsam_iso_program_apply_host.argtypes = [vp, C.c_int, C.c_int,
                                                  C.POINTER(C.POINTER(C.c_double))]
        _lib = L
    return _lib


def _ia(vals):
    return (C.c_int * len(vals))(*[int(v) for v in vals])


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Engine:
    """A device-resident patch hierarchy and its FAS solver.

    ``device=-1`` gives a host-only context (geometry and ghost programs,
    no GPU), used to check the integer maps on a CPU.
    """

    def __init__(self, layout: Layout, dt: float, eps: float, alpha: float = 1.0,
                 well: str = "reference", nsmoothingpasses: int = 2,
                 maxvcycles: int = 20, qerrortol: float = 1e-8,
                 precision: int = PREC_FP64, device: int = 0,
                 tune: Optional[Dict[str, float]] = None, errortype: int = 1,
                 qpo: Optional[Sequence[float]] = None):
        L = lib()
        fa, fs, fn, fd = WELLS[well](alpha)
        cfg = Cfg(layout.ndims, layout.nrvars, layout.mbc, layout.minlevel, dt, eps * eps,
                  fa, fs, fn, fd, nsmoothingpasses, maxvcycles, qerrortol, precision, device,
                  errortype)
        self.layout = layout
        self.maxvcycles = maxvcycles
        self.dt = float(dt)
        self.ctx = L.bsam_ctx_create(C.byref(cfg))
        if not self.ctx:
            raise BsamError(L.bsam_last_error(None).decode())
        for key, val in (tune or {}).items():
            self._ck(L.bsam_tune(self.ctx, key.encode(), float(val)))
        if qpo is not None:
            self.set_qpo(qpo)
        self.describe(layout)

    def describe(self, layout: Layout):
        """Hands the engine a hierarchy (after a regrid). Patch fields start
        at zero; a hierarchy equal to the current one keeps its plans."""
        L = lib()
        self.layout = layout
        self.patches = layout.ordered()
        idx = layout.index_map()
        self._ck(L.bsam_hierarchy_begin(self.ctx))
        for p in self.patches:
            mb = [v for d in range(3) for v in p.mbounds[d]]
            mg = [v for d in range(3) for v in p.mglobal[d]]
            par = idx[p.parent_id] if p.parent_id is not None else -1
            self._ck(L.bsam_add_patch(self.ctx, p.id, p.level, par, _ia(p.mx), _ia(mb),
                                      _ia(mg), _ia(p.mthbc), p.dx,
                                      (C.c_double * 3)(*p.xlower)))
        self._ck(L.bsam_hierarchy_commit(self.ctx))

    def close(self):
        if getattr(self, "ctx", None):
            lib().bsam_ctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        self.close()

    def _ck(self, rc):
        if rc != 0:
            raise BsamError(lib().bsam_last_error(self.ctx).decode())
        return rc

    def footprint_bytes(self) -> int:
        return int(lib().bsam_footprint_bytes(self.ctx))

    # -- fields -----------------------------------------------------------
    def upload(self, n: int, name: str, arr: np.ndarray):
        a = np.asfortranarray(arr, dtype=np.float64)
        shape = self.layout.field_shape(self.patches[n], name)
        if a.shape != shape:
            raise BsamError(f"shape {a.shape} != {shape}")
        flat = a.reshape(-1, order="F")
        self._ck(lib().bsam_upload(self.ctx, n, FIELDS[name], _dptr(flat), flat.size))

    def download(self, n: int, name: str) -> np.ndarray:
        shape = self.layout.field_shape(self.patches[n], name)
        flat = np.empty(int(np.prod(shape)), dtype=np.float64)
        self._ck(lib().bsam_download(self.ctx, n, FIELDS[name], _dptr(flat), flat.size))
        return flat.reshape(shape, order="F")

    # -- reference operations ------------------------------------------------
    def set_ghost(self, level: int, ipass: int):
        self._ck(lib().bsam_set_ghost(self.ctx, level, ipass))

    def relax(self, level: int, color: int):
        self._ck(lib().bsam_relax(self.ctx, level, color))

    def restrict(self, level: int):
        self._ck(lib().bsam_restrict_solution(self.ctx, level))

    def coarse_ghost(self, level: int):
        self._ck(lib().bsam_coarse_ghost_points(self.ctx, level))

    def coarse_load(self, level: int):
        self._ck(lib().bsam_coarse_load(self.ctx, level))

    def correct(self, level: int):
        self._ck(lib().bsam_correct_fine(self.ctx, level))

    def fill_down(self, field: int):
        """field 1 = f (sourcefield), 2 = q (solutionfield)."""
        self._ck(lib().bsam_fill_down(self.ctx, field))

    def truncation_error(self, level: int):
        """EstimateLevelErrors' truncation error on ``level`` into qrte."""
        self._ck(lib().bsam_truncation_error(self.ctx, level))

    def error_flags(self, level: int, qtolerance: float):
        """SetErrFlags (default tagging) on ``level`` from qrte."""
        self._ck(lib().bsam_error_flags(self.ctx, level, qtolerance))

    def copy_levels_from(self, src: "Engine", top: int):
        """q of levels minlevel..top from another engine with the same patches there."""
        self._ck(lib().bsam_copy_levels(self.ctx, src.ctx, top))

    def init_fields(self, level: int):
        """InitFields (MC prolongation from the parent) for the patches of `level`."""
        self._ck(lib().bsam_init_fields(self.ctx, level))

    def transfer_values_from(self, src: "Engine", level: int):
        """TransferValues: q of the cells `level`'s patches share with src's."""
        self._ck(lib().bsam_transfer_values(self.ctx, src.ctx, level))

    def set_dt(self, dt: float):
        """The time step of the problem constants from now on."""
        self._ck(lib().bsam_set_dt(self.ctx, float(dt)))
        self.dt = float(dt)

    def set_qpo(self, qpo: Sequence[float]):
        """The quasi-periodic offsets qpo(1:nrvars) (nodeinfodef.f90:101):
        added to values copied across the periodic x boundary, +qpo(n) to
        the high side, -qpo(n) to the low side (boundary.f90:441-461,
        :578-591)."""
        vals = np.ascontiguousarray(qpo, dtype=np.float64)
        self._ck(lib().bsam_set_qpo(self.ctx, _dptr(vals), int(vals.size)))

    def overlay_level_from(self, src: "Engine", level: int):
        """CopyPatchToUniformGrid over `level`: src's patches, ghosts
        included, copied into this engine's patches of that level in
        level-list order."""
        self._ck(lib().bsam_overlay_level(self.ctx, src.ctx, level))

    SOLVE_STATUS = {0: "converged", 1: "maxcycles", 2: "nonfinite", -1: None}

    def solve_status(self) -> Optional[str]:
        """How the last solve ended: converged, maxcycles or nonfinite."""
        return self.SOLVE_STATUS[int(lib().bsam_solve_status(self.ctx))]

    def error_flags_user(self, level: int, qtolerance: float):
        self._ck(lib().bsam_error_flags_user(self.ctx, level, qtolerance))

    def after_step_stats(self):
        """(mass, energy) of the composite grid after a step; energy is 0 in 3D."""
        m, e = C.c_double(), C.c_double()
        self._ck(lib().bsam_after_step_stats(self.ctx, C.byref(m), C.byref(e)))
        return m.value, e.value

    def flags(self, n: int) -> np.ndarray:
        """Patch ``n``'s error flags errorflags(1:mx) (int32, Fortran order)."""
        shape = self.layout.field_shape(self.patches[n], "errorflags")
        flat = np.empty(int(np.prod(shape)), dtype=np.int32)
        self._ck(lib().bsam_download_flags(self.ctx, n, flat.ctypes.data_as(C.POINTER(C.c_int)),
                                           flat.size))
        return flat.reshape(shape, order="F")

    def source(self):
        self._ck(lib().bsam_source(self.ctx))

    def vcycle(self, level: int):
        self._ck(lib().bsam_vcycle(self.ctx, level))

    def error_afas(self, level: int) -> float:
        r = C.c_double()
        self._ck(lib().bsam_error_afas(self.ctx, level, C.byref(r)))
        return r.value

    def source_norm(self, level: int) -> float:
        """RMS of f over the level's cells (both variables): the denominator
        of the relative residual RMS(f - A(q)) / RMS(f)."""
        r = C.c_double()
        self._ck(lib().bsam_source_norm(self.ctx, level, C.byref(r)))
        return r.value

    def copy_qold(self):
        self._ck(lib().bsam_copy_qold(self.ctx))

    def sync(self):
        self._ck(lib().bsam_sync(self.ctx))

    def _solve(self, fn):
        n, conv = C.c_int(), C.c_int()
        hist = (C.c_double * self.maxvcycles)()
        self._ck(fn(self.ctx, C.byref(n), hist, self.maxvcycles, C.byref(conv)))
        return list(hist[:n.value]), bool(conv.value)

    def multigrid_iterations(self):
        """Returns (per-cycle residuals, converged)."""
        return self._solve(lib().bsam_multigrid_iterations)

    def time_step(self):
        return self._solve(lib().bsam_time_step)

    def tune(self, key: str, value: float):
        self._ck(lib().bsam_tune(self.ctx, key.encode(), float(value)))

    def profile_kernel(self, kind: str = "relax", reps: int = 20):
        """Back-to-back timing of one kernel over the V-cycle's launch mix:
        (us per launch, algorithmic bytes per launch)."""
        us, by = C.c_double(), C.c_double()
        self._ck(lib().bsam_profile_kernel(self.ctx, self.PROF_KINDS.index(kind), reps,
                                           C.byref(us), C.byref(by)))
        return us.value, by.value

    def tune_get(self, key: str) -> float:
        """A tuning value, or a schedule property: tail_top, tail_lds_fp64,
        tail_lds_fp32, commits_reused (hand-overs that kept the committed
        hierarchy)."""
        v = C.c_double()
        self._ck(lib().bsam_tune_get(self.ctx, key.encode(), C.byref(v)))
        return v.value

    # -- kernel timing ----------------------------------------------------------
    PROF_KINDS = ("relax", "ghost", "restrict", "coarse_ghost", "coarse_load", "correct",
                  "filldown", "source", "resnorm", "tail", "relax_fused", "copy", "defect",
                  "smooth_iso")

    def profile_begin(self, kinds=("relax",)):
        mask = 0
        for k in kinds:
            mask |= 1 << self.PROF_KINDS.index(k)
        self._ck(lib().bsam_profile_begin(self.ctx, mask))

    def profile_end(self) -> Dict[str, dict]:
        """Per kernel kind: launches, summed ms and summed algorithmic bytes."""
        self._ck(lib().bsam_profile_end(self.ctx))
        out = {}
        for n, k in enumerate(self.PROF_KINDS):
            cnt, ms, by = C.c_longlong(), C.c_double(), C.c_double()
            self._ck(lib().bsam_profile_get(self.ctx, n, C.byref(cnt), C.byref(ms),
                                            C.byref(by)))
            if cnt.value:
                out[k] = {"launches": cnt.value, "ms": ms.value, "bytes": by.value}
        return out

    # -- integer maps ---------------------------------------------------------
    def ghost_program_dst(self, level: int, ipass: int) -> np.ndarray:
        L = lib()
        n = L.bsam_ghost_program_size(self.ctx, level, ipass)
        if n < 0:
            raise BsamError("bad level or ipass")
        out = np.zeros(4 * max(n, 1), dtype=np.int32)
        self._ck(L.bsam_ghost_program_dst(self.ctx, level, ipass,
                                          out.ctypes.data_as(C.POINTER(C.c_int)), n))
        return out[:4 * n].reshape(n, 4)

    def fused_ghost_size(self, level: int, ipass: int, color: int) -> int:
        """Updates of SetGhost(level, ipass) applied inside the relaxation of
        `color`; -1 when that transition is not fused."""
        return int(lib().bsam_fused_ghost_size(self.ctx, level, ipass, color))

    def ghost_rides_down(self, level: int, ipass: int) -> int:
        """1 if SetGhost(level, ipass) runs inside the level's restriction /
        fill-down launch."""
        return int(lib().bsam_ghost_rides_down(self.ctx, level, ipass))

    def fused_ghost_apply_host(self, level: int, ipass: int, color: int,
                               qs: List[np.ndarray]):
        ptrs = (C.POINTER(C.c_double) * len(qs))()
        for n, a in enumerate(qs):
            assert a.flags.f_contiguous and a.dtype == np.float64
            ptrs[n] = _dptr(a)
        self._ck(lib().bsam_fused_ghost_apply_host(self.ctx, level, ipass, color, ptrs))

    def fused_ghost_part_apply_host(self, level: int, ipass: int, color: int, part: int,
                                    qs: List[np.ndarray]):
        """Part 0 (inside the launch) or 1 (launched before it) of a fused
        transition, applied to host q arrays."""
        ptrs = (C.POINTER(C.c_double) * len(qs))()
        for n, a in enumerate(qs):
            assert a.flags.f_contiguous and a.dtype == np.float64
            ptrs[n] = _dptr(a)
        self._ck(lib().bsam_fused_ghost_part_apply_host(self.ctx, level, ipass, color, part,
                                                        ptrs))

    def post_ghost_size(self, level: int, pending: int, color: int, part: int) -> int:
        """Updates the relaxation launch of `color` (SetGhost(pending) before
        it, -1 none) applies at its end (part 0) or leaves to a launch before
        the next relaxation (part 1); -1 when it has no such plan."""
        return int(lib().bsam_post_ghost_size(self.ctx, level, pending, color, part))

    def post_ghost_apply_host(self, level: int, pending: int, color: int, part: int,
                              qs: List[np.ndarray]):
        ptrs = (C.POINTER(C.c_double) * len(qs))()
        for n, a in enumerate(qs):
            assert a.flags.f_contiguous and a.dtype == np.float64
            ptrs[n] = _dptr(a)
        self._ck(lib().bsam_post_ghost_apply_host(self.ctx, level, pending, color, part, ptrs))

    def iso_patches(self, level: int) -> List[int]:
        """Patch indices of the level's isolated patches (their smoothing runs
        in one launch out of LDS)."""
        n = int(lib().bsam_iso_patches(self.ctx, level, None, 0))
        if n < 0:
            raise BsamError("bad level")
        out = (C.c_int * max(n, 1))()
        lib().bsam_iso_patches(self.ctx, level, out, n)
        return list(out[:n])

    def iso_program_apply_host(self, level: int, ipass: int, qs: List[np.ndarray]):
        """SetGhost(level, ipass) of the isolated patches, from the records the
        smoothing kernel reads, applied to host q arrays."""
        ptrs = (C.POINTER(C.c_double) * len(qs))()
        for n, a in enumerate(qs):
            assert a.flags.f_contiguous and a.dtype == np.float64
            ptrs[n] = _dptr(a)
        self._ck(lib().bsam_iso_program_apply_host(self.ctx, level, ipass, ptrs))

    def ride_loop_size(self, level: int) -> int:
        """Updates of SetGhost(level - 1, 0) the loop's coarse loading of
        `level` carries (-1: no such plan)."""
        return int(lib().bsam_ride_loop_size(self.ctx, level))

    def ride_loop_apply_host(self, level: int, qs: List[np.ndarray]):
        ptrs = (C.POINTER(C.c_double) * len(qs))()
        for n, a in enumerate(qs):
            assert a.flags.f_contiguous and a.dtype == np.float64
            ptrs[n] = _dptr(a)
        self._ck(lib().bsam_ride_loop_apply_host(self.ctx, level, ptrs))

    def ghost_apply_host(self, level: int, ipass: int, qs: List[np.ndarray]):
        """Apply SetGhost(level, ipass)'s program to host q arrays in place."""
        ptrs = (C.POINTER(C.c_double) * len(qs))()
        for n, a in enumerate(qs):
            assert a.flags.f_contiguous and a.dtype == np.float64
            ptrs[n] = _dptr(a)
        self._ck(lib().bsam_ghost_program_apply_host(self.ctx, level, ipass, ptrs))


class BatchEngine(Engine):
    """Several independent hierarchies of one shape class (same ndims and
    minlevel) solved side by side in one engine context, one group each
    (bsam_set_groups): every level launch spans all of them, each keeps its
    own eps and residual test, and a group that has converged is left
    untouched, so each hierarchy's fields end as they would alone.

    Group g's patches keep their geometry but are moved apart in the global
    index space (x offset 4 g times the level's extent), so no two groups'
    ghost-extended boxes meet. Patches are added level by level, each level
    group by group in the hierarchy's own level-list order.
    """

    def __init__(self, layouts: List[Layout], eps: List[float], dt: float, alpha: float = 1.0,
                 well: str = "reference", nsmoothingpasses: int = 2,
                 maxvcycles: int = 20, qerrortol: float = 1e-8,
                 precision: int = PREC_FP64, device: int = 0,
                 tune: Optional[Dict[str, float]] = None, errortype: int = 1):
        if not layouts or len(layouts) != len(eps):
            raise BsamError("BatchEngine needs one eps per layout")
        nd, minl = layouts[0].ndims, layouts[0].minlevel
        for lay in layouts:
            if lay.ndims != nd or lay.minlevel != minl:
                raise BsamError("batched hierarchies must share ndims and minlevel")
        self.layouts = layouts
        self.eps = [float(e) for e in eps]
        super().__init__(layouts[0], dt, self.eps[0], alpha, well, nsmoothingpasses,
                         maxvcycles, qerrortol, precision, device, tune, errortype)

    def describe(self, layout: Layout):
        L = lib()
        lays = self.layouts
        nd = lays[0].ndims
        lo = min(l.minlevel for l in lays)
        hi = max(l.finestlevel for l in lays)
        self.index = {}          # (group, patch index in its layout) -> engine index
        self.patches = []
        groups = []
        maps = [lay.index_map() for lay in lays]
        self._ck(L.bsam_hierarchy_begin(self.ctx))
        n = 0
        for level in range(lo, hi + 1):
            for g, lay in enumerate(lays):
                root = lay.patches[0]
                # the level's extent in x and the group's offset in it
                w = root.mx[0] * 2 ** level if level >= 0 else root.mx[0] // 2 ** (-level)
                if level < 0 and root.mx[0] % 2 ** (-level):
                    raise BsamError("batched roots must halve down to their coarsest level")
                shift = 4 * g * max(w, 1)
                order = lay.ordered()
                for k, p in enumerate(order):
                    if p.level != level:
                        continue
                    mb = [v for d in range(3) for v in p.mbounds[d]]
                    mg = [v for d in range(3) for v in p.mglobal[d]]
                    mg[0] += shift
                    mg[1] += shift
                    par = (self.index[(g, maps[g][p.parent_id])]
                           if p.parent_id is not None else -1)
                    self._ck(L.bsam_add_patch(self.ctx, n, p.level, par, _ia(p.mx), _ia(mb),
                                              _ia(mg), _ia(p.mthbc), p.dx,
                                              (C.c_double * 3)(*p.xlower)))
                    self.index[(g, k)] = n
                    self.patches.append(p)
                    groups.append(g)
                    n += 1
        eps2 = (C.c_double * len(lays))(*[e * e for e in self.eps])
        self._ck(L.bsam_set_groups(self.ctx, len(lays), _ia(groups), eps2))
        self._ck(L.bsam_hierarchy_commit(self.ctx))
        self._group_layout = {self.index[key]: lays[key[0]] for key in self.index}

    def upload(self, n: int, name: str, arr: np.ndarray, group: int = 0):
        """Field `name` of patch n (index in group `group`'s own layout order)."""
        e = self.index[(group, n)]
        a = np.asfortranarray(arr, dtype=np.float64)
        lay = self.layouts[group]
        shape = lay.field_shape(lay.ordered()[n], name)
        if a.shape != shape:
            raise BsamError(f"shape {a.shape} != {shape}")
        flat = a.reshape(-1, order="F")
        self._ck(lib().bsam_upload(self.ctx, e, FIELDS[name], _dptr(flat), flat.size))

    def download(self, n: int, name: str, group: int = 0) -> np.ndarray:
        e = self.index[(group, n)]
        lay = self.layouts[group]
        shape = lay.field_shape(lay.ordered()[n], name)
        flat = np.empty(int(np.prod(shape)), dtype=np.float64)
        self._ck(lib().bsam_download(self.ctx, e, FIELDS[name], _dptr(flat), flat.size))
        return flat.reshape(shape, order="F")

    def group_results(self):
        """Per group of the last solve: (residual history, converged)."""
        out = []
        for g in range(len(self.layouts)):
            n, st = C.c_int(), C.c_int()
            hist = (C.c_double * self.maxvcycles)()
            self._ck(lib().bsam_group_result(self.ctx, g, C.byref(n), hist, self.maxvcycles,
                                             C.byref(st)))
            out.append((list(hist[:n.value]), st.value == 0))
        return out

    def time_steps(self):
        """One time step of every group: per group (history, converged)."""
        self.time_step()
        return self.group_results()

    def after_step_stats(self, group: int = 0):
        m, e = C.c_double(), C.c_double()
        self._ck(lib().bsam_after_step_stats_group(self.ctx, group, C.byref(m), C.byref(e)))
        return m.value, e.value
