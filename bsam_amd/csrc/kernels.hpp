// kernels.hpp -- launchers of the gfx950 kernels (kernels.hip). `prec` selects
// the field set: PREC_SOLUTION_FP64 (the solution, reference arithmetic) or
// PREC_DEFECT_FP32 (the FP32 correction of a defect-correction step).
#pragma once
#include "common.hpp"

namespace bsam {

hipError_t launch_relax(int nd, int prec, const Dev &d, const Tile *tiles, int nt, int ipass,
                        hipStream_t s);
// A SetGhost program's updates grouped per tile (list `pre`, with per-tile
// starts), then the relaxation of one colour, then the next program's updates
// grouped per tile the same way (list `post`), in one launch. Either list may
// be empty. With stage, the tile and its halo are read through LDS
// (2 x (tx+2)(ty+2)(tz+2) values of dynamic LDS).
hipError_t launch_relax_fused(int nd, int prec, const Dev &d, const Tile *tiles, int nt,
                              int color, const GhostList &pre, const GhostList &post,
                              bool stage, hipStream_t s);
// The whole smoothing of each isolated patch iso[0..n) (IsoPatch), one
// workgroup per patch with `lds` bytes of dynamic LDS (2 x (qn + ns) values
// of the largest): npasses passes of colour 1 then 2, each colour after the
// pending SetGhost program (SetGhost(0) before the first when ghost0_first),
// and SetGhost(2) after the last when final2.
hipError_t launch_smooth_iso(int nd, int prec, const Dev &d, const IsoPatch *iso, int n,
                             const int *stage, const IsoRec *rec,
                             const unsigned long long *srec, int npasses, int ghost0_first,
                             int final2, size_t lds, hipStream_t s);
// Raises the dynamic-LDS limit of the isolated-patch smoothing to `lds` bytes.
hipError_t iso_allow_lds(size_t lds);
hipError_t launch_ghost(int prec, const Dev &d, const GhostList &ops, hipStream_t s);
// Restriction / fill-down of one level; gops is a SetGhost program of that
// level applied in the same launch (grid-stride; its reads must be disjoint
// from the cells the launch writes; empty for none).
hipError_t launch_restrict(int nd, int prec, const Dev &d, const Tile *tiles, int nt,
                           const GhostList &gops, hipStream_t s);
hipError_t launch_filldown(int nd, const Dev &d, const Tile *tiles, int nt, int field,
                           const GhostList &gops, hipStream_t s);
hipError_t launch_coarse_ghost(int nd, int prec, const Dev &d, const Tile *tiles, int nt,
                               hipStream_t s);
hipError_t launch_coarse_load(int nd, int prec, const Dev &d, const Tile *tiles, int nt,
                              hipStream_t s);
// GetCoarseGhostPoints + CoarseLoadingFunction in one launch: tiles[0..nres)
// are fine tiles (cl_lx, cl_ly, cl_lz extents) forming the coarse right-hand
// side, the rest linear tiles over the qc ring cells (ring_cell); with nres
// < 0, linear tiles over the ghost-extended coarse box, one coarse cell per
// thread. And
// CorrectFine over linear tiles of the ghost-extended fine blocks. Either may
// apply a SetGhost program of the parent level in the same launch: `rest`
// grid-stride (destinations the launch does not read) and, per tile, its
// share of `tops` before its work; the launch writes none of the program's
// sources.
hipError_t launch_coarse_ghost_load(int nd, int prec, const Dev &d, const Tile *tiles, int nt,
                                   int nres, const GhostList &rest, const GhostList &tops,
                                   hipStream_t s);
hipError_t launch_correct(int nd, int prec, const Dev &d, const Tile *tiles, int nt,
                          const GhostList &rest, const GhostList &tops, hipStream_t s);
hipError_t launch_source(int nd, const Dev &d, const Tile *tiles, int nt, int full,
                         hipStream_t s);
// RelativeTruncationError over a level's coarse interior tiles (qrte) and
// SetErrFlags from it over its interior tiles (tol = qtolerance / h^d).
hipError_t launch_trunc_error(int nd, const Dev &d, const Tile *tiles, int nt, hipStream_t s);
hipError_t launch_error_flags(int nd, const Dev &d, const Tile *tiles, int nt, double tol,
                              hipStream_t s);
// SetErrFlagsUser2D/3D on the interior tiles of a level (q ghosts set).
hipError_t launch_error_flags_user(int nd, const Dev &d, const Tile *tiles, int nt,
                                   double qtol, hipStream_t s);
// InitFields (MC prolongation from the parent) for patches[0..n).
hipError_t launch_init_fields(int nd, const Dev &d, const int *patches, int n, hipStream_t s);
// Box copies src -> dst.
hipError_t launch_box_copy(double *dst, const double *src, const BoxCopy *boxes, int n,
                           hipStream_t s);
// The sums of AfterStep2D/3D for tasks[0..n) into out[3n].
hipError_t launch_after_step(int nd, const Dev &d, const StatTask *tasks, int n, double lo,
                             double hi, double *out, hipStream_t s);
hipError_t launch_defect_setup(int nd, const Dev &d, const Tile *tiles, int nt, hipStream_t s);
hipError_t launch_apply_correction(int nd, const Dev &d, const Tile *tiles, int nt,
                                   hipStream_t s);
hipError_t launch_resnorm(int nd, const Dev &d, const Tile *tiles, int nt, double *partial,
                          int what, int comps, hipStream_t s);
hipError_t launch_finalize(const double *partial, int n, double count, double tol,
                           double *hist, int *state, int comps, hipStream_t s,
                           int ngroups = 1, const int *granges = nullptr,
                           const double *gcount = nullptr, int hstride = 0);
hipError_t launch_copy(double *dst, const double *src, long long n, hipStream_t s);
// Tail kernels: one workgroup of `threads` over levels 0..nlev-1 of the
// single-patch chain; the V-cycle with `lds` bytes of dynamic LDS
// (TailLevel::lds; 0 keeps every array in the global pools); `qpo`: some
// level's SetGhost programs add periodic offsets.
hipError_t launch_tail_vcycle(int nd, int prec, const Dev &d, const TailLevel *lv, int nlev,
                              int nsmooth, int threads, size_t lds, const TailCopy *cin,
                              int nin, const TailCopy *cout, int nout, bool qpo, hipStream_t s);
// Capacity of the tail's level table, and the static LDS the tail kernel takes.
int tail_max_levels();
size_t tail_static_lds();
hipError_t launch_tail_filldown(int nd, const Dev &d, const TailLevel *lv, int nlev, int field,
                                int threads, hipStream_t s);
// Raises the dynamic-LDS limit of the tail kernels to `lds` bytes.
hipError_t tail_allow_lds(size_t lds);

}  // namespace bsam
