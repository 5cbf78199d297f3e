/*
 * bsam_gpu.h -- C ABI of the MI355X FAS multigrid engine for block-structured
 * AMR Cahn-Hilliard solves.
 *
 * The engine replaces the reference's adaptive FAS solve and ghost fill. The
 * host (Fortran via ISO_C_BINDING, C, C++ or Python ctypes) keeps the patch
 * tree and its arrays; it describes the hierarchy once per regrid, uploads
 * fields, and calls the entry points below, which run hand-written gfx950
 * kernels on device-resident mirrors of the fields.
 *
 * Conventions
 *  - Every function returns 0 on success and a negative code on failure;
 *    bsam_last_error(ctx) then holds a message. Nothing exits the process.
 *  - Arrays cross the ABI in the reference layout (lib/bsamstorage.f90:46-123):
 *    column-major, variable index slowest, q/qold/qc ghost-extended with lower
 *    bound 1-mbc, f/rf/ftmp interior only. 2D arrays have a third extent 1.
 *  - Patches are identified by their index in the order they were added:
 *    levels ascending, each level in ApplyOnLevel order (youngest first,
 *    lib/treeops.f90:536-572).
 *  - One host thread per context. Contexts are independent (own device
 *    memory and stream); several may share one GPU.
 */
#ifndef BSAM_GPU_H
#define BSAM_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bsam_ctx bsam_ctx;

/* Field selectors; values 1 and 2 match sourcefield/solutionfield
 * (lib/nodeinfodef.f90:64-66). */
enum {
  BSAM_FIELD_F = 1,     /* source/load function f (interior) */
  BSAM_FIELD_Q = 2,     /* solution q (ghost-extended) */
  BSAM_FIELD_QOLD = 3,  /* previous time level (ghost-extended) */
  BSAM_FIELD_QC = 4,    /* coarse-under-fine restriction (ghost-extended) */
  BSAM_FIELD_FTMP = 5,  /* time-step source part (interior) */
  BSAM_FIELD_RF = 6,    /* residual scratch (interior) */
  BSAM_FIELD_QRTE = 7   /* relative truncation error qrte(1:cmx) (interior, coarse) */
};

/* Smoothing precision. */
enum {
  BSAM_PREC_FP64 = 0,   /* plain FP64 everywhere: parity path */
  BSAM_PREC_MIXED = 1   /* FP32 smoothing inside FP64 defect correction */
};

typedef struct bsam_cfg {
  int ndims;              /* 2 or 3 */
  int nrvars;             /* 2: (phi, mu) */
  int mbc;                /* ghost layers; 1 */
  int minlevel;           /* coarsest level (<= 0) */
  /* Cahn-Hilliard problem constants (problem module hooks):
   *   Dfc(c) = fc_alpha*(c-fc_shift)^3, D2fc = 3*fc_alpha*(c-fc_shift)^2,
   *   Dfe(c) = ((c-fc_shift)*fe_num)/fe_den, Mob = 1.
   * The reference module is fc_shift=0.5, fc_alpha=fe_num=alpha, fe_den=4. */
  double dt, eps2;
  double fc_alpha, fc_shift, fe_num, fe_den;
  /* &rundata solver keys (lib/bsamroutines.f90:68-90) */
  int nsmoothingpasses;
  int maxvcycles;
  double qerrortol;
  int precision;          /* BSAM_PREC_* */
  int device;             /* HIP device ordinal; -1 = host-only context
                             (geometry and ghost programs, no device memory) */
  int errortype;          /* ErrorAFAS (lib/afasroutines.f90:108-152): 1 the
                             blended RMS of f - A(q) over both components
                             (L2Error), 2 the larger of the two components'
                             RMS (L2ComponentErrors); any other value is
                             refused, as the reference stops on it (:143-146) */
} bsam_cfg;

/* Context lifetime. Replaces the module globals of NodeInfoDef
 * (lib/nodeinfodef.f90:47-104) with an explicit configuration. */
bsam_ctx *bsam_ctx_create(const bsam_cfg *cfg);
int bsam_ctx_destroy(bsam_ctx *ctx);
const char *bsam_last_error(const bsam_ctx *ctx);

/* Hierarchy description, fed from an ApplyOnLevel walk of the reference tree
 * (treeops.f90:536-572), one patch per call, levels ascending, each level in
 * level-list order. Geometry is the nodeinfo record (nodeinfodef.f90:124-219):
 * mx[3], mbounds[3][2] and mglobal[3][2] flattened as {lo1,hi1,lo2,hi2,lo3,hi3},
 * mthbc[6], dx, xlower[3]; parent is the parent's patch index (-1 for the
 * coarsest patch). Commit builds the device patch table, the ghost-fill
 * programs and the tile lists, and allocates field storage (q, qold zeroed as
 * in AllocFields, bsamstorage.f90:116-117). A commit whose patch list equals
 * the committed one in order and geometry (ids may differ: MakeNewGrid draws
 * them at random, bsamroutines.f90:1665-1666) keeps the plans, tables,
 * storage and graphs and only zeroes the fields. */
int bsam_hierarchy_begin(bsam_ctx *ctx);
int bsam_add_patch(bsam_ctx *ctx, int id, int level, int parent, const int *mx,
                   const int *mbounds, const int *mglobal, const int *mthbc,
                   double dx, const double *xlower);
int bsam_hierarchy_commit(bsam_ctx *ctx);
/* Independent sub-problems in one hierarchy (no counterpart in the
 * reference; for serving many requests with shared launches). Called after
 * the last bsam_add_patch and before the commit: group[p] (0..ngroups-1)
 * for each patch added, eps2[g] the problem's eps^2 for group g (NULL: the
 * configuration's for all). The groups' patches must not touch one
 * another's (place each group's root apart in the global index space), and
 * on every level each group's patches must be consecutive in the level
 * list. A solve (bsam_multigrid_iterations, bsam_time_step) then runs every
 * level launch over all groups, tests each group's residual on its own
 * finest-level patches (summed in the order a solve of that group alone
 * sums them) and leaves a group untouched from the iteration its test
 * passes (or its residual is not finite), so each group's fields end as
 * they would in a context of its own. The solve's return values are the
 * largest iteration count and whether every group converged; per group:
 * bsam_group_result. Without this call the hierarchy is one group. */
int bsam_set_groups(bsam_ctx *ctx, int ngroups, const int *group, const double *eps2);
int bsam_num_groups(const bsam_ctx *ctx);
/* Group g's last solve: iterations, residual history (up to maxhist) and
 * status (BSAM_SOLVE_*). */
int bsam_group_result(bsam_ctx *ctx, int group, int *ncycles, double *hist, int maxhist,
                      int *status);
int bsam_num_patches(const bsam_ctx *ctx);
/* Device bytes held by the committed hierarchy's field pools (the quantity a
 * server admits requests by; no counterpart in the reference). */
long long bsam_footprint_bytes(const bsam_ctx *ctx);

/* Host <-> device field transfer for one patch, n doubles in the reference
 * layout. Synchronous. */
long bsam_field_len(const bsam_ctx *ctx, int patch, int field);
int bsam_upload(bsam_ctx *ctx, int patch, int field, const double *host, long n);
int bsam_download(bsam_ctx *ctx, int patch, int field, double *host, long n);

/* Solver entry points. Each replaces the reference routine it names and runs
 * on the context's stream. */
/* SetGhost(level, ipass), lib/boundary.f90:58-89; ipass 0 all, 1 red, 2 black */
int bsam_set_ghost(bsam_ctx *ctx, int level, int ipass);
/* ApplyOnLevel(level, RelaxPatch) for one colour, lib/afasroutines.f90:549-598
 * with the problem's RelaxGrid2D/3D (problem.f90:164-215, 453-518) */
int bsam_relax(bsam_ctx *ctx, int level, int color);
/* ApplyOnLevel(level, RestrictSolution), lib/afasroutines.f90:600-644 */
int bsam_restrict_solution(bsam_ctx *ctx, int level);
/* ApplyOnLevel(level, GetCoarseGhostPoints), lib/boundary.f90:1203-1269 */
int bsam_coarse_ghost_points(bsam_ctx *ctx, int level);
/* ApplyOnLevel(level, CoarseLoadingFunction), lib/afasroutines.f90:646-714 */
int bsam_coarse_load(bsam_ctx *ctx, int level);
/* ApplyOnLevel(level, CorrectFine), lib/afasroutines.f90:842-963 */
int bsam_correct_fine(bsam_ctx *ctx, int level);
/* FillDown(field), lib/afasroutines.f90:223-239; field BSAM_FIELD_F or _Q */
int bsam_fill_down(bsam_ctx *ctx, int field);
/* The truncation-error part of EstimateLevelErrors (lib/bsamroutines.f90:
 * 971-981) on a refined level: RestrictSolution, SetGhost(level - 1, 0),
 * GetCoarseGhostPoints, CopyQToQold on level and level - 1, then
 * RelativeTruncationError (lib/afasroutines.f90:716-797) into
 * BSAM_FIELD_QRTE of the level's patches. Replaces the reference's
 * ApplyOnLevel(level, RelativeTruncationError, ...) and the calls before it. */
int bsam_truncation_error(bsam_ctx *ctx, int level);
/* SetErrFlags2D/3D (lib/bsamroutines.f90:1093-1157), the default error
 * tagging, on every patch of `level` from its qrte with qtolerance(level);
 * bsam_download_flags copies a patch's errorflags(1:mx) (n = interior cells)
 * to the host. */
int bsam_error_flags(bsam_ctx *ctx, int level, double qtolerance);
int bsam_download_flags(bsam_ctx *ctx, int patch, int *host, long n);
/* ErrFlag's user tagging (errflagopt = errflaguser, lib/bsamroutines.f90:
 * 1061-1075) with the Cahn-Hilliard module's SetErrFlagsUser2D/3D
 * (problems/simple_cahn-hilliard/problem.f90:135-163, :425-452): every cell
 * of `level` whose undivided central gradient of variable 1 exceeds
 * qtolerance(level) is flagged, from the device q with the ghosts that
 * SetGhost(level, 0) leaves (AMR calls it first, bsamroutines.f90:946). */
int bsam_error_flags_user(bsam_ctx *ctx, int level, double qtolerance);
/* TakeTimeSteps' after-step statistics (lib/bsamroutines.f90:690-697:
 * AfterStepStatistics :896-939 over levels finest..0 with AfterStep2D/3D,
 * problem.f90:102-134, :398-424): the mass of variable 1 on the composite
 * grid and, in 2D, the free energy Energy2D (problem.f90:330-361) with the
 * double well whose convex and concave parts the configuration's Dfc and
 * Dfe are; *energy is 0 in 3D, as the reference leaves it. Per-patch sums
 * are reduced on the device in a fixed order and combined on the host in the
 * reference's order. Synchronises the stream. */
int bsam_after_step_stats(bsam_ctx *ctx, double *mass, double *energy);
/* The same statistics over one group's patches (bsam_set_groups), with that
 * group's eps. */
int bsam_after_step_stats_group(bsam_ctx *ctx, int group, double *mass, double *energy);
/* The same statistics without allocating device memory. No counterpart in
 * the reference: it is for callers that cannot allocate (tight serving
 * loops that report the statistics per step, harnesses that forbid
 * hipMalloc), where bsam_after_step_stats allocates its task table and task
 * sums on first use and whenever a hierarchy needs more tasks. This variant
 * keeps both in a fixed scratch array that every GPU the library runs on
 * holds until the process ends (1.25 MiB), and calls no hipMalloc/hipFree.
 * Limit: BSAM_STATS_NOALLOC_MAX_TASKS tasks, a task being a run of rows (2D)
 * or planes (3D) of about 4096 cells of one patch of levels finest..0 or of
 * its parent's section under it (14.2 M cells of 1963 3D patches take about
 * 4000); a larger hierarchy is refused with an error. Calls on one GPU share
 * the scratch; within one process they are serialised, so calls from
 * several threads or contexts on one GPU do not interleave there. Results
 * are bit-identical to bsam_after_step_stats. Synchronises the stream. */
#define BSAM_STATS_NOALLOC_MAX_TASKS 32768
int bsam_after_step_stats_noalloc(bsam_ctx *ctx, double *mass, double *energy);
/* GetSourceFunction then UpdateSourceFunction on levels finest..0,
 * lib/afasroutines.f90:57-64 */
int bsam_source(bsam_ctx *ctx);
/* AFASVcycle(level), lib/afasroutines.f90:488-521 */
int bsam_vcycle(bsam_ctx *ctx, int level);
/* ErrorAFAS(level) with the context's errortype, lib/afasroutines.f90:108-208 */
int bsam_error_afas(bsam_ctx *ctx, int level, double *residual);
/* The RMS of the source f over the cells of `level` (both variables, the
 * count ErrorAFAS's errortype 1 divides by): the denominator of the relative
 * residual RMS(f - A(q)) / RMS(f) that the mixed-precision path reports beside
 * the reference's absolute test (lib/afasroutines.f90:94-101). No counterpart
 * in the reference. */
int bsam_source_norm(bsam_ctx *ctx, int level, double *rms);
/* qold = q on every level (CopyQToQold, lib/bsamroutines.f90:568-580) */
int bsam_copy_qold(bsam_ctx *ctx);
/* MultigridIterations, lib/afasroutines.f90:38-106. Writes the number of
 * V-cycles run, up to maxhist per-cycle residuals, and whether the last
 * residual fell below qerrortol (the reference sets lconverged, :98-101).
 * With BSAM_PREC_MIXED the iterations are FP64 defect-correction steps around
 * FP32 V-cycles, with the same FP64 residual test and history. */
int bsam_multigrid_iterations(bsam_ctx *ctx, int *ncycles, double *hist,
                              int maxhist, int *converged);
/* One time step on a fixed hierarchy: SetGhost on levels 0..finest, FillDown
 * of q, qold = q, then MultigridIterations (TakeTimeSteps,
 * lib/bsamroutines.f90:648-723, without the regrid). */
int bsam_time_step(bsam_ctx *ctx, int *ncycles, double *hist, int maxhist,
                   int *converged);
/* Wait for all work queued on the context's stream. */
int bsam_sync(bsam_ctx *ctx);
/* How the last solve (bsam_multigrid_iterations / bsam_time_step) ended:
 * converged below qerrortol, stopped at maxvcycles, or stopped at the first
 * residual that is not finite (NaN or Inf; the reference iterates on to
 * maxvcycles there, lib/afasroutines.f90:69-104, with every later residual
 * also non-finite); -1 before any solve. The reference reports nothing: its
 * lconverged is set and never read. No counterpart in the reference. */
enum { BSAM_SOLVE_CONVERGED = 0, BSAM_SOLVE_MAXCYCLES = 1, BSAM_SOLVE_NONFINITE = 2 };
int bsam_solve_status(const bsam_ctx *ctx);

/* ---- regrid: host-side clustering of a level's flags -------------------
 * The integer core of the regrid of one level, on the host: ErrFlag's
 * BufferAndList for every patch (lib/bsamroutines.f90:1159-1205), the
 * inflation of edge tags into neighbouring and periodic-image patches
 * (InflateEdgeTags :1207-1280, GetPeriodicTagOffset boundary.f90:1046-1201)
 * and, per patch in level-list order, Berger-Rigoutsos clustering into new
 * subgrids (RefineGrid / NewSubGrids / GetSignatures / GridFlagRatio
 * bsamroutines.f90:1334-1643). Context-free. The level's npatch patches are
 * given in level-list order by mx[3*p..], mglobal[6*p..] (lo, hi per
 * dimension) and their errorflags(1:mx) (Fortran order, flags[p]; buffered
 * in place as the reference buffers them). On return *ngrids new subgrids in
 * creation order (MakeNewGrid order): parent[g] (index into the patch list)
 * and mbounds[6*g..] in the parent's coarse indices. Returns 0, -1 with the
 * message in bsam_last_error(NULL) (also when more than max_grids result). */
/* GridAdapt's field work (lib/bsamroutines.f90:1288-1332) between the
 * hierarchy before a regrid (src) and the one being built (dst, committed
 * with the new patches of level+1; same device):
 *   bsam_copy_levels     q of levels minlevel..top, whose patches must be
 *                        those of src in the same order (they persist);
 *   bsam_init_fields     InitFields (:1685-1772) for every patch of `level`
 *                        of dst: qc = the parent's q under it with one ring,
 *                        q = qold = BiLinProlongationP1MC / TriLinProlongation
 *                        P1MC of it (gridutilities.f90:219-279, :588-686);
 *   bsam_transfer_values TransferValues / TransferOverlap (:1774-1867): every
 *                        new patch of `level` takes the q of the interior cells
 *                        it shares with src's patches of that level (the old
 *                        grids).
 *   bsam_overlay_level   CopyPatchToUniformGrid (lib/bsaminputoutput.f90:
 *                        543-577) over a level: every src patch of `level`,
 *                        in level-list order, copies q over its ghost-extended
 *                        box (0:mx+1) into the dst patches of that level where
 *                        they overlap; a later patch wins where two boxes
 *                        meet. WriteUniformMeshQ (:438-541) is this on a dst
 *                        whose levels are each one patch over the domain,
 *                        with bsam_init_fields between levels.
 * bsam_copy_levels, bsam_transfer_values and bsam_overlay_level wait for src's
 * stream first and return when their copies are done, so src may change
 * afterwards. */
/* The time step of the problem constants (bsam_cfg.dt) from now on, for
 * every solve, truncation-error estimate and source set-up. TakeTimeSteps
 * runs the first pass of a syncelliptic run with dt = 0 and restores it
 * after (lib/bsamroutines.f90:620-622, :748-750); the module variable dt
 * changes under MultigridIterations, so the Fortran binding passes it on
 * each call. dt must be finite and >= 0. Waits for the context's stream. */
int bsam_set_dt(bsam_ctx *ctx, double dt);
/* The quasi-periodic offsets qpo(1:nrvars) (lib/nodeinfodef.f90:101; 0 by
 * default, as RootInit sets them, lib/bsamroutines.f90:241): a value copied
 * across the periodic x boundary gains +qpo(n) going to the high side and
 * -qpo(n) going to the low side (PeriodicBC, PeriodicGhostOverlap with
 * GetPeriodicOffsets' qoffset, lib/boundary.f90:417-596, :965-1044). n must be
 * nrvars. Takes effect for the following SetGhost programs. */
int bsam_set_qpo(bsam_ctx *ctx, const double *qpo, int n);
int bsam_copy_levels(bsam_ctx *dst, const bsam_ctx *src, int top);
int bsam_init_fields(bsam_ctx *ctx, int level);
int bsam_transfer_values(bsam_ctx *dst, const bsam_ctx *src, int level);
int bsam_overlay_level(bsam_ctx *dst, const bsam_ctx *src, int level);

/* The frame files' number formats (WriteQ2D / WriteQ3D,
 * lib/bsaminputoutput.f90:138-168): nrows rows of ncols values
 * (vals[r*rowstride + c*colstride]) as text, es20.12 (kind 0, 2D frames; a
 * three-digit exponent drops the letter E as Fortran does) or f25.12 (kind
 * 1, 3D frames), two blanks between values, a newline after each row. Returns
 * the number of characters written to out (no terminating zero), or -1 if
 * cap is too small. No device work. */
long bsam_format_rows(const double *vals, long nrows, int ncols, long rowstride,
                      long colstride, int kind, char *out, long cap);

typedef struct bsam_regrid_params {
  int ndims, level;
  int ibuffer;               /* ibuffer(level) */
  int minimumgridpoints;     /* minimumgridpoints(level) */
  double desiredfillratio;   /* desiredfillratios(level) */
  int root_mx[3];            /* the root patch (PeriodicSetup's offsets) */
  int root_mthbc[6];
} bsam_regrid_params;
int bsam_regrid_level(const bsam_regrid_params *rp, int npatch, const int *mx,
                      const int *mglobal, int *const *flags, int max_grids, int *ngrids,
                      int *parent, int *mbounds);

/* Tuning knobs (no counterpart in the reference):
 *   "tail_cells"   levels from minlevel up whose single patch has at most
 *                  this many cells run inside one workgroup (default 4096;
 *                  0 disables). Set before bsam_hierarchy_begin.
 *   "tail_threads" workgroup size of that kernel (64..1024, default 1024).
 *   "graph"        1 (default): replay each multigrid iteration as a
 *                  captured hipGraph; 0: launch kernels one by one.
 *   "inner_cycles" FP32 V-cycles per defect-correction step (mixed
 *                  precision; 1..8, default 1).
 *   "tile_x", "tile_y"  level-kernel tile extents (powers of two, tile_x
 *                  at least 2; 2D tile_x * tile_y = 256, default 32 x 8; 3D
 *                  the third extent fills 256 threads, default 16 x 8 x 2).
 *                  Before the hierarchy; setting either turns tile_auto off.
 *   "tile_auto"    1 (default): each commit of a device context picks the
 *                  tile shape by timing the plain relaxation under each
 *                  candidate shape on the hierarchy's levels of >= 4096
 *                  cells (the default shape is kept unless another is more
 *                  than 3 % faster); winners are cached per hierarchy shape
 *                  for the process. 0: the default shape. Before the
 *                  hierarchy.
 *   "fuse_ghost"   1 (default): apply each SetGhost inside another level
 *                  launch where the level allows (the following relaxation,
 *                  the level's restriction or fill-down, the child level's
 *                  coarse loading or correction); 0: separately, and a
 *                  hierarchy committed while it is 0 skips building those
 *                  plans (contexts that never run a V-cycle).
 *   "tail_lds"     bytes of LDS the tail kernel may hold its levels' arrays
 *                  in (-1, default: the device limit; 0: none); before the
 *                  hierarchy.
 *   "iso_min_cells" a level smooths its isolated patches (whose SetGhost
 *                  programs read no other patch of the level and which no
 *                  other patch's programs read) in one launch per smoothing,
 *                  each patch in one workgroup out of LDS, when those that
 *                  fit hold at least this many cells (default 65536; 0:
 *                  never). Needs fuse_ghost. Before the hierarchy.
 *   "cgl_fine_cells" levels of at least this many cells (default 524288;
 *                  0: all, -1: none) run the coarse ghost + loading launch
 *                  over fine tiles, each thread one fine residual summed per
 *                  coarse cell through LDS; the others over linear tiles of
 *                  the coarse box, one coarse cell per thread. Before the
 *                  hierarchy. */
int bsam_tune(bsam_ctx *ctx, const char *key, double value);
/* Reads a tuning value back, or a property of the committed schedule:
 * "tail_top" (highest level run by the tail kernel; below minlevel when
 * none), "tail_lds_fp64" and "tail_lds_fp32" (dynamic LDS bytes of the tail
 * kernels), "commits_reused" (commits that found the hierarchy already
 * committed), "tile_z", and of the last commit's tile choice
 * "tile_auto_candidates" (shapes timed; 0 when taken from the cache) and
 * "tile_auto_us" (the chosen shape's relaxation rounds), "iso_patches" and
 * "iso_cells" (isolated patches smoothed out of LDS, over all levels). No
 * counterpart in the reference. */
int bsam_tune_get(const bsam_ctx *ctx, const char *key, double *value);

/* Kernel timing. Between bsam_profile_begin and bsam_profile_end, every
 * launch of a kind whose bit is set in kind_mask is bracketed by HIP events
 * on the context's stream; bsam_profile_get then returns, per kind, the
 * number of launches, their summed duration (ms) and their summed
 * algorithmic HBM bytes (FP64, ghosts excluded: relax 48 B per cell-update
 * -- q read 16, f read 16, q written 16 per full red-black sweep of a cell --
 * so 24 B per cell of the level for one colour's half-sweep, which updates
 * half its cells; residual norm 32 B, restriction 24 B, coarse loading 40 B,
 * correction 40 B, fill-down 20 B, source 32 B per cell of the level
 * operated on; ghost update 32 B, 16 B read and 16 B written; FP32
 * relaxation and ghost updates half of their figures). */
enum {
  BSAM_PROF_RELAX = 0, BSAM_PROF_GHOST, BSAM_PROF_RESTRICT, BSAM_PROF_COARSE_GHOST,
  BSAM_PROF_COARSE_LOAD, BSAM_PROF_CORRECT, BSAM_PROF_FILLDOWN, BSAM_PROF_SOURCE,
  BSAM_PROF_RESNORM, BSAM_PROF_TAIL,
  BSAM_PROF_RELAX_FUSED,  /* relaxation launches that also apply ghost
                             updates (the preceding SetGhost program's before
                             the sweep, the next one's moved to the end):
                             RELAX's bytes plus 32 B per update, as GHOST
                             (16 B in FP32) */
  BSAM_PROF_COPY,         /* whole-pool copies (source restore, qold = q):
                             16 B per element */
  BSAM_PROF_DEFECT,       /* mixed precision: defect set-up (q, f read, qb, e, fe
                             written: 52 B per ghost-extended cell) and q += e
                             (40 B per interior cell) */
  BSAM_PROF_SMOOTH_ISO,   /* whole smoothings of isolated patches (one launch
                             per smoothing of a level): RELAX's 24 B per cell
                             for each half-sweep and GHOST's 32 B per update
                             applied, for every sweep and program it runs */
  BSAM_PROF_NKINDS
};
int bsam_profile_begin(bsam_ctx *ctx, unsigned kind_mask);
int bsam_profile_end(bsam_ctx *ctx);
int bsam_profile_get(const bsam_ctx *ctx, int kind, long long *launches, double *ms,
                     double *bytes);
/* Back-to-back timing of one kernel kind (BSAM_PROF_RELAX only): `reps`
 * rounds of the V-cycle's launch mix of that kernel (both colours on every
 * level above the tail) between two HIP events on the context's stream;
 * returns the mean duration per launch (us) and the mean algorithmic bytes
 * per launch. Changes the fields as relaxation does. */
int bsam_profile_kernel(bsam_ctx *ctx, int kind, int reps, double *us_per_launch,
                        double *bytes_per_launch);

/* Device-to-device copy rate of `device`: `reps` copies of `bytes` bytes
 * between two buffers of that size with the engine's copy kernel, timed with
 * HIP events; *gbs = 2 * bytes * reps / time (read plus write) in GB/s. The
 * measured ceiling a bandwidth-bound kernel is compared with beside the
 * chip's specification. Allocates and frees the two buffers. No counterpart
 * in the reference. */
int bsam_copy_rate(int device, long long bytes, int reps, double *gbs);

/* Introspection of the integer maps (valid on host-only contexts too).
 * Ghost-fill program of (level, ipass): number of ghost-cell updates, and
 * their destinations as q-field (patch, i, j, k) tuples, 4 ints each. */
int bsam_ghost_program_size(const bsam_ctx *ctx, int level, int ipass);
int bsam_ghost_program_dst(const bsam_ctx *ctx, int level, int ipass, int *out,
                           int maxn);
/* The updates of SetGhost(level, ipass) that the engine applies inside the
 * following relaxation of `color` ((ipass, color) = (0, 1), (1, 2) or (2, 1)):
 * their number (-1 when that transition runs as two launches), and their
 * application to host arrays. Updates the relaxation does not read and the
 * next SetGhost(level, color) rewrites are left out. */
int bsam_fused_ghost_size(const bsam_ctx *ctx, int level, int ipass, int color);
/* Of those, the updates that read a cell another tile of the relaxation
 * relaxes: they run in a launch of their own just before it (-1 when the
 * transition is not fused). */
int bsam_fused_ghost_pre_size(const bsam_ctx *ctx, int level, int ipass, int color);
/* 1 if SetGhost(level, ipass) reads none of the parent cells that the
 * level's restriction and fill-down write, so the engine runs it in the same
 * launch as those; 0 if not; -1 for bad arguments. */
int bsam_ghost_rides_down(const bsam_ctx *ctx, int level, int ipass);
/* 1 if the parent level's SetGhost(0) runs inside this level's coarse ghost +
 * loading launch (which 0), or the parent's SetGhost(2) inside its
 * correction launch (which 1); 0 if it is launched on its own; -1 for bad
 * arguments. No counterpart in the reference. */
int bsam_parent_ghost_rides(const bsam_ctx *ctx, int level, int which);
/* Inside the FP64 multigrid loop the coarse ghost + loading launch of level
 * `level` (>= 1) carries only the updates of SetGhost(level - 1, 0) that read
 * a cell changed since the loop's FillDown applied that program: level-1
 * cells under the level (the restriction), level-2 cells under level - 1 or
 * level-2 ghost cells (FillDown's next level and its SetGhost). Their number (-1 when there is no such plan) and their host
 * application, for the CPU tests. No counterpart in the reference. */
int bsam_ride_loop_size(const bsam_ctx *ctx, int level);
int bsam_ride_loop_apply_host(const bsam_ctx *ctx, int level, double *const *q_patches);
int bsam_fused_ghost_apply_host(const bsam_ctx *ctx, int level, int ipass, int color,
                                double *const *q_patches);
/* Applies on the host one part of a fused transition: part 0 the updates
 * applied inside the relaxation launch, part 1 those launched before it. */
int bsam_fused_ghost_part_apply_host(const bsam_ctx *ctx, int level, int ipass, int color,
                                     int part, double *const *q_patches);
/* The relaxation launch of `color` with SetGhost(`pending`) before it (-1:
 * none) can end with the next launch's pre updates (those of the transition
 * SetGhost(color) -> the other colour), each tile applying its share after
 * its relaxation. Part 0: the updates moved there; part 1: those that still
 * run in a launch of their own before the next relaxation. Parts 2 and 3
 * (pending 1, colour 2 only): the same split of SetGhost(2), the program after
 * the last launch of a smoothing, when that launch takes it over. Sizes (-1
 * when the launch has no such plan) and host application, for the CPU tests.
 * No counterpart in the reference. */
int bsam_post_ghost_size(const bsam_ctx *ctx, int level, int pending, int color, int part);
int bsam_post_ghost_apply_host(const bsam_ctx *ctx, int level, int pending, int color, int part,
                               double *const *q_patches);
/* Isolated patches of a level (see "iso_min_cells"): their number (-1 for
 * bad arguments) and up to maxn of their patch indices; and the application
 * of SetGhost(level, ipass) to them on the host, decoded from the records the
 * smoothing kernel reads, for the CPU tests. The level's other patches'
 * updates are those of the fused and post plans. No counterpart in the
 * reference. */
int bsam_iso_patches(const bsam_ctx *ctx, int level, int *out, int maxn);
int bsam_iso_program_apply_host(const bsam_ctx *ctx, int level, int ipass,
                                double *const *q_patches);
/* Apply a ghost-fill program on the host to per-patch q arrays (reference
 * layout, one pointer per patch): used to check the integer maps on a CPU. */
int bsam_ghost_program_apply_host(const bsam_ctx *ctx, int level, int ipass,
                                  double *const *q_patches);

#ifdef __cplusplus
}
#endif
#endif
