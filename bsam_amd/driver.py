"""The BSAM time-step driver with its regrid every step, on the MI355X engine.

Reads the reference's input deck (rundata.dat, griddata.dat,
problemdata.dat: the namelists of BSAMSolver, lib/bsamroutines.f90:68-100,
RootInit :216-254, SetProb problem.f90:12-37) and runs TakeTimeSteps
(lib/bsamroutines.f90:582-789) for the Cahn-Hilliard problem:

* every step AMR(rootlevel) (:941-961): per level, SetGhost(level, 0), the
  error estimate (EstimateLevelErrors :963-1002: the truncation error and
  the default or the problem's own tagging, on the GPU), the clustering of
  the flags into new subgrids (bsam_regrid_level, on the host), and
  GridAdapt (:1288-1332): the new level's hierarchy is committed, the
  levels below copied over, the new patches initialised (QInit on initial
  grids, else InitFields' mass-corrected prolongation) and given the values
  of the old grids they overlap (TransferValues), all on the GPU;
* FillDown on the first step, qold = q, MultigridIterations (eqfn times);
* the after-step statistics (out/stats.dat) if asked for, and the frames
  out/mNNNNN.dat (frame 0 after the first regrid), each followed by the
  uniform-mesh output out/uNNNNN.dat if asked for (on the GPU: a context of
  one domain-wide patch per level, filled by prolongation and the level's
  patches).

The hierarchy under construction lives in spare engine contexts: two solver
contexts alternate as the hierarchy of the step (the one before the step
keeps the old grids), and two lighter ones, which never run a V-cycle and so
skip the V-cycle's fusion plans at commit, alternate for the intermediate
hierarchies of a deeper regrid.
"""

from __future__ import annotations

import json
import os
import re
import time
from typing import Dict, List, Optional

import numpy as np

from . import frames
from .engine import FIELD_Q, Engine, regrid_level
from .layout import Layout
from .problem import qinit2d, qinit3d

ERRFLAG_DEFAULT = 1
ERRFLAG_USER = 10


def read_namelist(path: str) -> Dict:
    """The first namelist group of a file as {name: value} with indexed
    entries as {name: {index: value}} (indices as ints or int tuples)."""
    text = open(path).read()
    m = re.search(r"&\w+(.*?)/", text, re.S)
    body = m.group(1) if m else ""
    out: Dict = {}
    for key, idx, val in re.findall(r"(\w+)\s*(?:\(([^)]*)\))?\s*=\s*([^\s,/]+)", body):
        key = key.lower()
        v = val.strip()
        lv = v.lower()
        if lv in (".true.", "t", ".t."):
            x = True
        elif lv in (".false.", "f", ".f."):
            x = False
        elif re.fullmatch(r"[-+]?\d+", v):
            x = int(v)
        else:
            x = float(lv.replace("d", "e"))
        if idx:
            ii = tuple(int(t) for t in idx.split(","))
            out.setdefault(key, {})[ii[0] if len(ii) == 1 else ii] = x
        else:
            out[key] = x
    return out


def _arr(v, n, default, base=1):
    """A namelist array (dict by index or scalar) as a list over base..base+n-1."""
    if isinstance(v, dict):
        return [v.get(i, default) for i in range(base, base + n)]
    if v is None:
        return [default] * n
    return [v] * n


def uniform_layout(layout: Layout):
    """WriteUniformMeshQ's uniform grids as a hierarchy: one patch over the
    whole domain on every level 0..finest of `layout` (no levels below the
    root). Returns it and the id of its finest patch."""
    nd = layout.ndims
    root = layout.patches[0]
    lay = Layout(nd, root.mx[:nd], root.dx, xlower=root.xlower, mthbc=root.mthbc,
                 minlevel=0, nrvars=layout.nrvars)
    top = 0
    for _ in range(layout.finestlevel):
        top = lay.add_child(top, [(1, lay.patches[top].mx[d]) for d in range(nd)])
    return lay, top


def uniform_mesh(u: Engine, eng: Engine, top: int) -> np.ndarray:
    """The finest uniform grid of WriteUniformMeshQ (lib/bsaminputoutput.f90:
    468-498) for `eng`'s hierarchy, built on the GPU in `u` (an engine on
    uniform_layout(eng.layout)): the root's q with its ghosts; then per level
    the MC prolongation of the level below over the domain (InitFields'
    BiLin/TriLinProlongationP1MC, ghosts included) with the level's patches
    copied over it, ghosts included, in level-list order
    (CopyPatchToUniformGrid :543-577). Returns q(0:mx+1, ..., 1:nrvars),
    Fortran order, downloaded."""
    u.overlay_level_from(eng, 0)
    for level in range(1, eng.layout.finestlevel + 1):
        u.init_fields(level)
        u.overlay_level_from(eng, level)
    n = next(n for n, p in enumerate(u.patches) if p.id == top)
    p = u.layout.patches[top]
    nd = u.layout.ndims
    return u.download(n, "q").reshape(tuple(p.mx[d] + 2 for d in range(nd)) + (u.layout.nrvars,),
                                      order="F")


class Driver:
    """TakeTimeSteps on the GPU engine, regridding every step."""

    def __init__(self, workdir: str, device: int = 0, metrics: Optional[str] = None):
        self.workdir = workdir
        # one JSON line per solve (MultigridIterations call), appended to
        # `metrics` when given: step, time, equation, V-cycles, last residual,
        # status, wall milliseconds, patches and cells of the hierarchy
        self.metrics = metrics
        rd = read_namelist(os.path.join(workdir, "rundata.dat"))
        gd = read_namelist(os.path.join(workdir, "griddata.dat"))
        pd = read_namelist(os.path.join(workdir, "problemdata.dat"))
        # BSAMSolver's defaults (lib/bsamroutines.f90:74-89) under the deck
        self.dt = float(rd.get("dt", 0.0))
        self.maxvcycles = int(rd.get("maxvcycles", 20))
        self.nsmooth = int(rd.get("nsmoothingpasses", 2))
        self.qerrortol = float(rd.get("qerrortol", 1.0e-6))
        self.outframes = int(rd.get("outframes", 1))
        self.timeiterations = int(rd.get("timeiterations", 1))
        self.stats = bool(rd.get("getafterstepstats", False))
        self.eqfn = int(rd.get("eqfn", 6))
        self.restart = bool(rd.get("restart", False))
        self.restartframe = int(rd.get("restartframe", 0))
        self.uniform = bool(rd.get("outputuniformmesh", False))
        # a restart never synchronises (BSAMSolver :133-138)
        self.sync = bool(rd.get("syncelliptic", False)) and not self.restart
        self.errortype = int(rd.get("errortype", 1))
        if self.errortype not in (1, 2):
            raise ValueError("ErrorAFAS: only errortype = 1, 2 are supported")
        if int(rd.get("updateauxfreq", 1)) != 1:
            raise NotImplementedError("updateauxfreq 1 only")
        # RootInit's defaults (:237-254)
        self.ndims = nd = int(gd.get("ndims", 2))
        self.maxlevel = int(gd.get("maxlevel", 0))
        self.minlevel = int(gd.get("minlevel", 0))
        depth = self.maxlevel + 1
        self.qtol = _arr(gd.get("qtolerance"), depth, 1.0e-6, 0)
        self.ibuffer = _arr(gd.get("ibuffer"), depth, 0, 0)
        self.mgp = _arr(gd.get("minimumgridpoints"), depth, 2, 0)
        self.fill = _arr(gd.get("desiredfillratios"), depth, 8.0e-01, 0)
        self.errflag = _arr(gd.get("errflagopt"), depth, ERRFLAG_DEFAULT, 0)
        self.mx = [int(v) for v in _arr(gd.get("mx"), nd, 1)]
        self.xlower = [float(v) for v in _arr(gd.get("xlower"), nd, 0.0)]
        self.xupper = [float(v) for v in _arr(gd.get("xupper"), nd, 1.0)]
        self.mthbc = [int(v) for v in _arr(gd.get("mthbc"), 2 * nd, 1)]
        # the quasi-periodic offsets of the periodic x copies (RootInit: 0)
        self.qpo = [float(v) for v in _arr(gd.get("qpo"), 2, 0.0)]
        if int(gd.get("mbc", 1)) != 1 or int(gd.get("maux", 0)) != 0 or int(gd.get("nrvars", 1)) != 2:
            raise NotImplementedError("mbc = 1, maux = 0 and nrvars = 2 only")
        self.alpha = float(pd.get("alpha", 1.0))
        self.eps = float(pd.get("eps", 0.0))
        # the root grid: dx = (xupper - xlower) / mx (RootInit :363-364)
        dx = (self.xupper[0] - self.xlower[0]) / float(self.mx[0])
        xl = tuple(self.xlower) + (0.0,) * (3 - nd)
        self.root_layout = Layout(nd, self.mx, dx, xlower=xl, mthbc=self.mthbc,
                                  minlevel=self.minlevel, nrvars=2)
        self.minlevel = self.root_layout.minlevel
        self.device = device
        self._dt_now = self.dt                  # dt of the pass under way
        self._engines: List[Engine] = []
        self.engine: Optional[Engine] = None   # the last step's hierarchy, after run()
        # wall time by phase: the regrid (of which the hierarchy hand-overs,
        # the tagging with its download, the clustering and the field work),
        # the solves and the frames
        self.timing = {"regrid_s": 0.0, "commit_s": 0.0, "tag_s": 0.0, "cluster_s": 0.0,
                       "fields_s": 0.0, "solve_s": 0.0, "frames_s": 0.0, "steps": 0}

    # -- engines --------------------------------------------------------------
    SOLVERS = (0, 1)     # contexts that hold a step's hierarchy and solve
    SCRATCH = (2, 3)     # contexts for intermediate hierarchies only
    UNIFORM = 4          # context of the uniform grids of the uniform-mesh output

    def _engine(self, k: int, layout: Layout) -> Engine:
        while len(self._engines) <= k:
            self._engines.append(None)
        if self._engines[k] is None:
            tune = None if k in self.SOLVERS else {"fuse_ghost": 0, "tail_cells": 0, "graph": 0}
            self._engines[k] = Engine(layout, dt=self._dt_now, eps=self.eps, alpha=self.alpha,
                                      well="reference", nsmoothingpasses=self.nsmooth,
                                      maxvcycles=self.maxvcycles, qerrortol=self.qerrortol,
                                      device=self.device, tune=tune,
                                      errortype=self.errortype, qpo=self.qpo)
        else:
            self._engines[k].describe(layout)
            if self._engines[k].dt != self._dt_now:
                self._engines[k].set_dt(self._dt_now)
        return self._engines[k]

    def close(self):
        for e in self._engines:
            if e is not None:
                e.close()
        self._engines = []

    # -- initial condition ----------------------------------------------------
    def _qinit(self, p) -> np.ndarray:
        if self.ndims == 2:
            return qinit2d(p.mx, p.dx, p.xlower)
        return qinit3d(p.mx, p.dx, p.xlower, self.eps)

    def _upload_qinit(self, eng: Engine, level: int):
        for n, p in enumerate(eng.patches):
            if p.level == level:
                q = self._qinit(p)
                eng.upload(n, "q", q)
                eng.upload(n, "qold", q)

    # -- AMR --------------------------------------------------------------------
    def _flags(self, eng: Engine, level: int):
        if self.errflag[level] == ERRFLAG_DEFAULT:
            eng.error_flags(level, self.qtol[level])
        elif self.errflag[level] == ERRFLAG_USER:
            eng.error_flags_user(level, self.qtol[level])
        else:
            raise ValueError(f"errflagopt({level}) = {self.errflag[level]} is not supported")
        out = []
        for n, p in enumerate(eng.patches):
            if p.level == level:
                out.append((p, np.asfortranarray(eng.flags(n).astype(np.int32))))
        return out

    def _amr(self, old: int, initial: bool) -> int:
        """AMR(rootlevel) from the hierarchy in engine `old`; returns the
        engine that holds the new hierarchy."""
        final = self.SOLVERS[1] if old == self.SOLVERS[0] else self.SOLVERS[0]
        cur = old
        lay = self._engines[old].layout
        nb = 0   # which spare builds next
        level = 0
        while True:
            eng = self._engines[cur]
            eng.set_ghost(level, 0)
            if level >= self.maxlevel:
                break
            # EstimateLevelErrors(level): its truncation-error part always
            # runs (the restriction onto level-1 is part of the state)
            t0 = time.perf_counter()
            eng.truncation_error(level)
            tagged = self._flags(eng, level)
            t1 = time.perf_counter()
            root = lay.patches[0]
            new = regrid_level(self.ndims, level, int(self.ibuffer[level]), int(self.mgp[level]),
                               float(self.fill[level]), root.mx, root.mthbc,
                               [(p.mx, p.mglobal) for p, _ in tagged], [f for _, f in tagged])
            t2 = time.perf_counter()
            self.timing["tag_s"] += t1 - t0
            self.timing["cluster_s"] += t2 - t1
            if not new and lay.finestlevel == level:
                nlay = None   # the hierarchy already ends at this level
            else:
                # GridAdapt(level): the levels up to `level` and the new grids
                nlay, idmap = self._truncated(lay, level)
                for k, mb in new:
                    nlay.add_child(idmap[tagged[k][0].id], mb[:self.ndims])
                # the last level's hierarchy goes straight to the solver context
                last = bool(new) and level + 1 >= self.maxlevel
                nxt = final if last else self.SCRATCH[nb % 2]
                nb += 1
                t0 = time.perf_counter()
                neng = self._engine(nxt, nlay)
                t1 = time.perf_counter()
                neng.copy_levels_from(eng, level)
                if new:
                    if initial:
                        self._upload_qinit(neng, level + 1)   # InitFields' QInit
                    else:
                        neng.init_fields(level + 1)
                        neng.transfer_values_from(self._engines[old], level + 1)
                neng.sync()
                self.timing["commit_s"] += t1 - t0
                self.timing["fields_s"] += time.perf_counter() - t1
                cur, lay = nxt, nlay
            if not new:
                # AMR(level+1) on the empty level: its EstimateLevelErrors
                # still refills this level's ghosts (SetGhost(level, 0))
                if level + 1 < self.maxlevel:
                    self._engines[cur].set_ghost(level, 0)
                break
            level += 1
        if cur in self.SCRATCH:
            # the hierarchy ended early: hand it to the solver context
            lay = self._engines[cur].layout
            t0 = time.perf_counter()
            neng = self._engine(final, lay)
            self.timing["commit_s"] += time.perf_counter() - t0
            neng.copy_levels_from(self._engines[cur], lay.finestlevel)
            cur = final
        return cur

    def _truncated(self, lay: Layout, top: int):
        """The levels <= top of `lay` as a new Layout in the same creation
        order, and the map from `lay`'s patch ids to the new ones."""
        root = lay.patches[0]
        out = Layout(lay.ndims, root.mx[:lay.ndims], root.dx, xlower=root.xlower,
                     mthbc=root.mthbc, minlevel=lay.minlevel, nrvars=lay.nrvars)
        idmap = {pid: pid for pid in lay.patches if lay.patches[pid].level <= 0}
        for level in range(1, top + 1):
            for pid in lay.creation.get(level, []):
                p = lay.patches[pid]
                idmap[pid] = out.add_child(idmap[p.parent_id], p.mbounds[:lay.ndims])
        return out, idmap

    # -- frames and statistics ------------------------------------------------
    def _write_frame(self, eng: Engine, n: int, t: float):
        t0 = time.perf_counter()
        path = os.path.join(self.workdir, "out", f"m{n:05d}.dat")
        frames.write_frame(path, t, eng.layout.finestlevel,
                           frames.frame_from_engine(eng, self.xupper))
        self.timing["frames_s"] += time.perf_counter() - t0

    def _write_uniform(self, eng: Engine, t: float):
        """WriteUniformMeshQ (lib/bsaminputoutput.f90:438-541) of the current
        hierarchy, written as a frame of one level-0 patch at the finest
        level's resolution to out/uNNNNN.dat, numbered by restartframe and
        not by the frame, as the reference names it: each frame overwrites
        the previous one."""
        t0 = time.perf_counter()
        nd, finest = self.ndims, eng.layout.finestlevel
        lay, top = uniform_layout(eng.layout)
        q = uniform_mesh(self._engine(self.UNIFORM, lay), eng, top)
        mxuf = lay.patches[top].mx[:nd]
        # dx(1:ndims) = (xupper - xlower) / mx on the root, halved per level
        dx = [(self.xupper[d] - self.xlower[d]) / float(self.mx[d]) for d in range(nd)]
        for _ in range(finest):
            dx = [v / 2 for v in dx]
        fp = frames.FramePatch(0, nd, 2, tuple(dx), tuple(self.xlower[:nd]),
                               tuple(self.xupper[:nd]), tuple(mxuf),
                               tuple((1, m) for m in mxuf), q)
        frames.write_frame(os.path.join(self.workdir, "out", f"u{self.restartframe:05d}.dat"),
                           t, 0, [fp])
        self.timing["frames_s"] += time.perf_counter() - t0

    def _mesh_size(self, lay: Layout) -> int:
        """GetMeshSize (:836-858) summed over levels >= 0."""
        total = 0
        for level in range(0, lay.finestlevel + 1):
            for pid in lay.level_ids(level):
                p = lay.patches[pid]
                m = c = 1
                for d in range(self.ndims):
                    m *= p.mx[d]
                    c *= p.mx[d] // 2
                total += m if level == 0 else m - c
        return total

    # -- TakeTimeSteps ----------------------------------------------------------
    def _read_restart(self) -> float:
        """ReadQ (lib/bsaminputoutput.f90:174-398) into solver context 0: the
        frame's root (checked against the deck) and its refined patches,
        which serve the first regrid as the old grids. Returns the time."""
        path = os.path.join(self.workdir, "out", f"m{self.restartframe:05d}.dat")
        if not os.path.exists(path):
            raise FileNotFoundError(f"restart frame {path} does not exist")
        t, finest, ps = frames.read_frame(path)
        root = ps[0]
        small = 1.0e-08
        if (root.level != 0 or root.ndims != self.ndims or root.nrvars != 2 or
                tuple(root.mx) != tuple(self.mx[:self.ndims]) or
                any(root.dx[d] - self.root_layout.patches[0].dx > small for d in range(self.ndims)) or
                any(abs(root.xlower[d] - self.xlower[d]) > small for d in range(self.ndims)) or
                any(abs(root.xupper[d] - self.xupper[d]) > small for d in range(self.ndims))):
            raise ValueError("restart frame does not match the deck's root grid")
        if finest > self.maxlevel:
            raise ValueError("restart frame is finer than maxlevel")
        lay, ids = frames.layout_from_frame(ps, self.mthbc, minlevel=self.minlevel)
        eng = self._engine(0, lay)
        pos = {p.id: n for n, p in enumerate(eng.patches)}
        for i, fp in enumerate(ps):
            q = np.asfortranarray(fp.q.reshape(lay.field_shape(lay.patches[ids[i]], "q"), order="F"))
            eng.upload(pos[ids[i]], "q", q)
            if fp.level == 0:
                eng.upload(pos[ids[i]], "qold", q)   # CopyQToQold on the root
        return t

    def _log_solve(self, eng, step, t, eqn, hist, conv, ms):
        lay = eng.layout
        rec = {"step": step, "time": t, "equation": eqn, "vcycles": len(hist),
               "residual": float(hist[-1]) if len(hist) else None, "converged": bool(conv),
               "status": eng.solve_status(), "ms": ms, "patches": len(lay.ordered()),
               "cells": int(sum(lay.cells_on_level(l) for l in range(0, lay.finestlevel + 1)))}
        with open(self.metrics, "a") as fh:
            fh.write(json.dumps(rec) + "\n")

    def run(self):
        os.makedirs(os.path.join(self.workdir, "out"), exist_ok=True)
        cur = 0
        if self.restart:
            if self.restartframe >= self.outframes:
                raise ValueError("outframes must be greater than restartframe")
            starttime = self._read_restart()
            firstframe = self.restartframe + 1
            eng = self._engines[0]
        else:
            eng = self._engine(0, self.root_layout)
            self._upload_qinit(eng, 0)           # Initialq on the root
            starttime, firstframe = 0.0, 1
        eng.set_ghost(0, 0)
        itperprint = self.timeiterations // self.outframes
        currenttime = starttime
        first = True                             # TakeTimeSteps' firstamr
        initial = not self.restart               # initial grids take QInit
        frame0 = not self.restart                # outputinitialdata
        if self.sync:
            # the synchronisation pass (:620-622, :694-699, :748-756): the
            # first regrid and solve with dt = 0, which leaves c and makes mu
            # consistent with it; no frame, no statistics, no step counted
            self._dt_now = 0.0
            for e in self._engines:
                if e is not None:
                    e.set_dt(0.0)
            t0 = time.perf_counter()
            cur = self._amr(cur, initial=True)
            eng = self._engines[cur]
            eng.fill_down(FIELD_Q)
            eng.copy_qold()
            self.timing["regrid_s"] += time.perf_counter() - t0
            t0 = time.perf_counter()
            for _ in range(self.eqfn):
                eng.multigrid_iterations()
            eng.sync()
            self.timing["solve_s"] += time.perf_counter() - t0
            self._dt_now = self.dt
            for e in self._engines:
                if e is not None and e.dt != self._dt_now:
                    e.set_dt(self._dt_now)
            first = initial = False
        for n in range(firstframe, self.outframes + 1):
            for it in range(1, itperprint + 1):
                t0 = time.perf_counter()
                cur = self._amr(cur, initial=initial)
                eng = self._engines[cur]
                if first:
                    eng.fill_down(FIELD_Q)
                eng.copy_qold()
                self.timing["regrid_s"] += time.perf_counter() - t0
                if frame0:
                    self._write_frame(eng, 0, currenttime)
                    if self.uniform:
                        self._write_uniform(eng, currenttime)
                    frame0 = False
                first = initial = False
                t0 = time.perf_counter()
                step = self.timing["steps"] + 1
                step_time = starttime + float((n - firstframe) * itperprint + it) * self.dt
                for eqn in range(self.eqfn):
                    ts = time.perf_counter()
                    hist, conv = eng.multigrid_iterations()
                    if self.metrics:
                        self._log_solve(eng, step, step_time, eqn + 1, hist, conv,
                                        1e3 * (time.perf_counter() - ts))
                eng.sync()
                self.timing["solve_s"] += time.perf_counter() - t0
                self.timing["steps"] += 1
                currenttime = real_time = step_time
                if self.stats:
                    mass, energy = eng.after_step_stats()
                    with open(os.path.join(self.workdir, "out", "stats.dat"), "a") as fh:
                        fh.write(f"{real_time:25.12f}{mass:25.12f}{energy:25.12f} "
                                 f"{self._mesh_size(eng.layout):8d}\n")
            self._write_frame(eng, n, currenttime)
            if self.uniform:
                self._write_uniform(eng, currenttime)
        self.engine = eng
        return self.timing


def main(argv=None):
    """python -m bsam_amd.driver WORKDIR [--device N]: runs the deck in
    WORKDIR (rundata.dat, griddata.dat, problemdata.dat) and writes WORKDIR/out."""
    import argparse
    import json
    ap = argparse.ArgumentParser(description="BSAM time steps with regrid on the GPU engine")
    ap.add_argument("workdir")
    ap.add_argument("--device", type=int, default=0)
    ap.add_argument("--metrics", default=None,
                    help="append one JSON line per solve (step, V-cycles, residual, status, ms)")
    a = ap.parse_args(argv)
    drv = Driver(a.workdir, device=a.device, metrics=a.metrics)
    try:
        print(json.dumps(drv.run()))
    finally:
        drv.close()


if __name__ == "__main__":
    main()
