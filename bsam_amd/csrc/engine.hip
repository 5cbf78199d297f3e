// engine.hip -- host side of the FAS multigrid engine: hierarchy commit,
// ghost-fill programs, level orchestration and the C ABI of bsam_gpu.h.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <climits>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "bsam_gpu.h"
#include "ghost.hpp"
#include "regrid.hpp"
#include "kernels.hpp"

using namespace bsam;

namespace {

struct Error : std::runtime_error {
  int code;
  Error(const std::string &m, int c = -1) : std::runtime_error(m), code(c) {}
};

#define HIPCK(x)                                                                    \
  do {                                                                              \
    hipError_t e_ = (x);                                                            \
    if (e_ != hipSuccess)                                                           \
      throw Error(std::string(#x) + ": " + hipGetErrorString(e_), -2);              \
  } while (0)

// Contexts may be driven from several host threads at once (one per stream).
// Stream capture forbids device-wide synchronisation and legacy-stream work
// while any stream is capturing, and hipFree / hipMalloc / stream creation may
// imply such work, so those calls and every capture hold this lock. All
// copies and fills go on the context's own (non-blocking) stream.
static std::recursive_mutex g_device_lock;
using DeviceLock = std::lock_guard<std::recursive_mutex>;

template <class T> T *dev_upload(const std::vector<T> &v, hipStream_t s) {
  if (v.empty()) return nullptr;
  T *p = nullptr;
  HIPCK(hipMalloc(&p, sizeof(T) * v.size()));
  HIPCK(hipMemcpyAsync(p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, s));
  HIPCK(hipStreamSynchronize(s));
  return p;
}

// A ghost-update list in its device records (ghost_records), viewed by the
// kernels as a GhostList.
struct DGhost {
  GhostS *s = nullptr;
  GhostOp *l = nullptr;
  int *ss = nullptr, *sl = nullptr;
  int ns = 0, nl = 0;
  GhostList view() const { return GhostList{s, l, ss, sl, ns, nl}; }
};
static DGhost upload_ghost(const std::vector<GhostOp> &ops, const std::vector<int> *start,
                           bool use_short, hipStream_t st) {
  const GhostRecords r = ghost_records(ops, start, use_short);
  DGhost d;
  d.s = dev_upload(r.s, st);
  d.l = dev_upload(r.l, st);
  d.ss = dev_upload(r.ss, st);
  d.sl = dev_upload(r.sl, st);
  d.ns = (int)r.s.size();
  d.nl = (int)r.l.size();
  return d;
}

struct LevelData {
  std::vector<int> patches;  // patch indices in level-list order
  int *d_patch_list = nullptr;  // the same on the device (regrid kernels)
  std::vector<Tile> t_int, t_ext, t_coarse, t_cext;
  Tile *d_int = nullptr, *d_ext = nullptr, *d_coarse = nullptr, *d_cext = nullptr;
  // the coarse ghost + loading launch: fine tiles (cl_lx/cl_ly/cl_lz), then
  // linear tiles over the qc ring cells (n_cgl_res: the fine tiles); linear
  // tiles (LIN_TILE cells each) over the 2x2(x2) fine blocks of the correction
  std::vector<Tile> t_cgl, t_blk;
  int n_cgl_res = 0;
  Tile *d_cgl = nullptr, *d_blk = nullptr;
  // linear tiles in device order over each patch's ghost-extended q array
  std::vector<Tile> t_qlin;
  std::vector<GhostOp> gp[3];
  DGhost d_gp[3];
  GhostOp *d_tgp[3] = {nullptr, nullptr, nullptr};  // patch-local copies (tail levels)
  // SetGhost(ipass) fused into the next relaxation, per transition
  // (ipass -> colour): 0: 0 -> 1, 1: 1 -> 2, 2: 2 -> 1
  FusedGhost fused[3];
  // SetGhost(ipass) of this level may run inside this level's restriction or
  // fill-down launch (its reads avoid the parent cells those launches write)
  bool down_ok[3] = {false, false, false};
  DGhost d_fops[3];   // per tile
  DGhost d_fpre[3];   // launched before the fused relaxation
  // the next launch's pre updates applied at the end of a relaxation launch,
  // per launch kind (post_key): the updates and, for the next launch, what
  // still runs before it
  // key 4 (kPostFinal): the last launch of a smoothing (SetGhost(1) -> 2)
  // ending with most of SetGhost(2), the program that follows the smoothing
  PostGhost post[5];
  DGhost d_pops[5];   // per tile
  DGhost d_prest[5];
  // the parent's SetGhost(0) inside this level's coarse ghost + loading
  // launch, and its SetGhost(2) (the last of post-smoothing) inside this
  // level's correction launch
  RideGhost ride_cgl, ride_corr;
  // Inside the multigrid loop (FP64, levels >= 1): the parent's SetGhost(0)
  // that the loop's FillDown applied in full, since when only the parent
  // cells under this level (this V-cycle's restriction), the grandparent
  // cells under the parent and the grandparent's ghost cells (FillDown's
  // next level and SetGhost) have changed; the ride carries only the updates
  // that read one of those, the others would write the values their
  // destinations hold.
  RideGhost ride_cgl_loop;
  struct DRide {
    DGhost rest, ops;   // grid-stride, per tile
  } d_ride_cgl, d_ride_corr, d_ride_cgl_loop;
  long long cells = 0;
  // Isolated patches (IsoPatch): their whole smoothing runs in one launch out
  // of LDS. The level's other patches keep the per-colour launches over
  // t_relax, whose fused and post plans cover only those patches' updates.
  // Without isolated patches t_relax is the interior tile list.
  std::vector<IsoPatch> iso;
  std::vector<int> iso_stage;
  std::vector<IsoRec> iso_rec;
  std::vector<unsigned long long> iso_srec;
  IsoPatch *d_iso = nullptr;
  int *d_iso_stage = nullptr;
  IsoRec *d_iso_rec = nullptr;
  unsigned long long *d_iso_srec = nullptr;
  size_t iso_lds = 0;           // dynamic LDS of the FP64 smoothing (FP32: half)
  long long iso_cells = 0;      // interior cells of the isolated patches
  long long iso_updates[3] = {0, 0, 0};   // their updates per SetGhost program
  std::vector<Tile> t_relax;
  Tile *d_relax = nullptr;
  long long relax_cells = 0;
  // per group: its interior tiles of this level (first, count; a group's
  // patches are contiguous in the level list) and its cells
  std::vector<int> granges;
  std::vector<double> gcells;
  int *d_granges = nullptr;
  double *d_gcells = nullptr;
};

// Transition index of SetGhost(ipass) followed by relaxation of `color`.
int fused_index(int ipass, int color) {
  if (ipass == 0 && color == 1) return 0;
  if (ipass == 1 && color == 2) return 1;
  if (ipass == 2 && color == 1) return 2;
  return -1;
}
const int kFusedPass[3] = {0, 1, 2}, kFusedColor[3] = {1, 2, 1};
// Relaxation launches that can end with the next launch's pre updates, by the
// SetGhost pending before them and their colour: 0: nothing pending, colour 1
// (the first of pre-smoothing); 1: SetGhost(0) -> 1; 2: SetGhost(2) -> 1;
// 3: SetGhost(1) -> 2. The next launch is SetGhost(colour) -> the other
// colour.
int post_key(int pending, int color) {
  if (pending < 0 && color == 1) return 0;
  if (pending == 0 && color == 1) return 1;
  if (pending == 2 && color == 1) return 2;
  if (pending == 1 && color == 2) return 3;
  return -1;
}
const int kPostPending[4] = {-1, 0, 2, 1}, kPostColor[4] = {1, 1, 1, 2};
constexpr int kPostFinal = 4;
// A SetGhost returned as pending may be only partly pending: kRestOnly + p
// means that the last relaxation launch applied SetGhost(p) except its final
// plan's rest, which the caller still launches.
constexpr int kRestOnly = 8;

// The SetGhost programs of a single-patch level with offsets relative to the
// patch's q origin, for the tail kernels; f1 carries the source packed with
// the nflip byte (f0 | nflip << 26). Returns false unless every update is a
// copy within the patch (physical and periodic boundaries of a level that
// covers the domain).
bool local_programs(const HPatch &h, const std::vector<GhostOp> (&gp)[3], int nv,
                    std::vector<GhostOp> (&out)[3]) {
  long long n = h.qn;
  for (int ip = 0; ip < 3; ip++) {
    out[ip].clear();
    for (GhostOp g : gp[ip]) {
      long long dst = g.dst - h.qoff, src = g.f0 - h.qoff;
      if (g.kind != 0 || dst < 0 || dst >= n || src < 0 || src >= n || src >= (1 << 26) ||
          g.nflip > 63)
        return false;
      (void)nv;
      g.dst = (int)dst;
      g.f0 = (int)src;
      g.f1 = (int)((unsigned)src | (unsigned)g.nflip << 26);
      out[ip].push_back(g);
    }
  }
  return true;
}

}  // namespace

struct bsam_ctx {
  bsam_cfg cfg{};
  std::string err;
  bool host_only = true;
  hipStream_t stream = nullptr;
  // second stream on which a level's isolated-patch smoothing runs beside the
  // launches of its other patches (forked from and joined back into `stream`
  // by the two events; both streams at the default priority)
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  int iso_side = 1;
  bool tail_qpo = false;   // a tail level's SetGhost programs add periodic offsets
  double qpo[2] = {0.0, 0.0};   // quasi-periodic offsets (bsam_set_qpo)
  bool building = false, committed = false;
  bool had_commit = false;    // hp/dp/lev/device state describe a committed hierarchy
  bool planned_fuse = false;  // that commit built the fusion and riding plans
  long long commits_reused = 0;
  std::vector<HPatch> staged;  // patches added since bsam_hierarchy_begin
  std::vector<HPatch> hp;
  std::vector<DPatch> dp;
  std::vector<LevelData> lev;
  int finest = 0;
  long long nq = 0, nc = 0, nf = 0, nr = 0, nfl = 0;
  double *q = nullptr, *qold = nullptr, *qc = nullptr, *f = nullptr, *rf = nullptr,
         *ftmp = nullptr;
  double *qrte = nullptr;   // relative truncation error (EstimateLevelErrors)
  int *eflags = nullptr;    // error flags (SetErrFlags)
  float *e = nullptr, *qce = nullptr, *fe = nullptr, *qb = nullptr;  // mixed precision
  int inner_cycles = 1;      // FP32 V-cycles per defect-correction step
  DPatch *d_P = nullptr;
  std::vector<TailLevel> h_tail;   // host copy of the tail level table (d_tail)
  double *d_partial = nullptr;
  double *d_stats = nullptr;      // task sums of the after-step statistics
  StatTask *d_stats_list = nullptr;  // their tasks
  int n_stats = 0;
  int partial_cap = 0;
  double *d_hist = nullptr;   // ngroups residual histories of hmax entries
  int *d_state = nullptr;     // 4 ints per group (k_finalize)
  int *h_state = nullptr;
  int hmax = 0;
  // independent sub-problems solved side by side (bsam_set_groups): a
  // group's patches never touch another's, each has its own eps and
  // residual test, and a group that has converged is left untouched
  int ngroups = 1;
  std::vector<double> group_eps2;          // per group (committed)
  std::vector<int> staged_group;           // per staged patch
  std::vector<double> staged_eps2;
  int staged_ngroups = 1;
  std::vector<int> group_state;            // per group after the last solve: 4 ints
  int last_status = -1;      // of the last solve: BSAM_SOLVE_* (bsam_solve_status)
  int last_iters = 0;        // iterations the last solve took
  long long qstride = 0;     // the q pool's variable stride
  std::vector<std::pair<long long, long long>> f_restore;   // f runs [a, b) an iteration rewrites
  // kernel timing with HIP events on the context's stream (bsam_profile_*)
  unsigned prof_mask = 0;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  struct ProfRec {
    int kind;
    double bytes;
    size_t ev;
  };
  std::vector<ProfRec> prof;
  double prof_ms[BSAM_PROF_NKINDS] = {0};
  double prof_bytes[BSAM_PROF_NKINDS] = {0};
  long long prof_n[BSAM_PROF_NKINDS] = {0};

  // tuning (bsam_tune): single-workgroup tail below tail_cells cells per level,
  // hipGraph replay of the multigrid iteration
  // tile lists spanning levels: interior and ghost-extended tiles of all
  // levels, interior tiles of levels 0..finest (source)
  double *f_saved = nullptr;   // f after the solve's source set-up (constant per solve)
  std::vector<Tile> all_int, all_ext, src_int;   // all_ext: linear device-order tiles
  Tile *d_all_int = nullptr, *d_all_ext = nullptr, *d_src_int = nullptr;
  long long src_cells = 0;
  long long all_int_cells = 0, all_ext_cells = 0;   // interior / ghost-extended cells, all levels
  long long tail_cells = 4096;
  int tail_threads = 1024;
  long long tail_lds_cap = -1;      // LDS bytes the tail may use (-1: device limit)
  size_t tail_lds[2] = {0, 0};      // dynamic LDS of the tail kernels per field set
  // staging tables of the tail (per field set; entries ordered by level) and,
  // per number of levels run, how many entries those levels own
  TailCopy *d_tcin[2] = {nullptr, nullptr}, *d_tcout[2] = {nullptr, nullptr};
  std::vector<int> tcin_n[2], tcout_n[2];
  int use_graph = 1;
  int fuse_ghost = 1;               // apply SetGhost inside the next relaxation
  int post_ghost = 1;               // and the next launch's pre updates at a launch's end
  int relax_lds = 0;                // relaxation stencils read through an LDS copy of the tile
  int short_ghost = 1;              // 16-byte device records for the updates that have them
  // smoothing of isolated patches in one LDS launch on levels where they hold
  // at least this many cells (0: never)
  long long iso_min_cells = 65536;
  // coarse loading over fine tiles on levels of at least this many cells
  // (-1: never)
  long long cgl_fine_cells = 524288;
  int tail_top = 0;                 // highest level run by the tail kernels
  TailLevel *d_tail = nullptr;      // levels minlevel..tail_top
  hipGraph_t graph = nullptr, graph_mixed = nullptr;
  hipGraphExec_t graph_exec = nullptr, graph_mixed_exec = nullptr;

  int tile_lx = -1, tile_ly = -1;   // tile extents (log2); -1: per-dimension default
  // tile shape chosen at commit by timing the relaxation (unless tile_x /
  // tile_y were set): the choice, its time and whether it came from the
  // process-wide cache of hierarchy shape classes
  bool tile_user = false;
  int tile_auto = 1;
  double tile_auto_us = 0.0;
  int tile_auto_cached = 0, tile_auto_candidates = 0;
  int tx() const { return 1 << tile_lx; }
  int ty() const { return 1 << tile_ly; }
  int tz() const { return nd() == 3 ? 256 >> (tile_lx + tile_ly) : 1; }
  int nd() const { return cfg.ndims; }
  int comps() const { return cfg.errortype == 2 ? 2 : 1; }   // residual norms kept apart
  int ntail() const { return tail_top - cfg.minlevel + 1; }
  LevelData &L(int level) {
    if (level < cfg.minlevel || level > finest) throw Error("level out of range");
    return lev[level - cfg.minlevel];
  }
  Dev dev() const {
    Dev d;
    d.q = q; d.qold = qold; d.qc = qc; d.f = f; d.rf = rf; d.ftmp = ftmp;
    d.e = e; d.qce = qce; d.fe = fe; d.qb = qb;
    d.qrte = qrte; d.eflags = eflags;
    d.P = d_P;
    d.gstate = d_state;
    d.prm = Problem{cfg.dt, cfg.eps2, cfg.fc_alpha, cfg.fc_shift, cfg.fe_num, cfg.fe_den,
                    {qpo[0], qpo[1]}};
    d.nrvars = cfg.nrvars;
    d.qvs = qstride;
    d.tile_lx = tile_lx;
    d.tile_ly = tile_ly;
    d.one_group = ngroups == 1;
    return d;
  }
  void need_device() const {
    if (host_only) throw Error("host-only context: no device");
    if (!committed) throw Error("hierarchy not committed");
  }
  void free_device();
  void drop_graphs() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (graph_mixed_exec) (void)hipGraphExecDestroy(graph_mixed_exec);
    if (graph_mixed) (void)hipGraphDestroy(graph_mixed);
    graph_exec = graph_mixed_exec = nullptr;
    graph = graph_mixed = nullptr;
  }
  ~bsam_ctx() { free_device(); }
};

void bsam_ctx::free_device() {
  if (host_only) return;
  DeviceLock lock(g_device_lock);
  // Frees and forgets: free_device runs again at hierarchy_begin and in the
  // destructor, and a second hipFree of a stale pointer would fail (leaving
  // a sticky error) or free another context's reuse of the address.
  auto F = [](auto *&p) {
    if (p) (void)hipFree((void *)p);
    p = nullptr;
  };
  auto FG = [&](DGhost &g) {
    F(g.s);
    F(g.l);
    F(g.ss);
    F(g.sl);
    g.ns = g.nl = 0;
  };
  F(q); F(qold); F(qc); F(f); F(rf); F(ftmp); F(d_P); F(d_partial); F(d_hist);
  F(e); F(qce); F(fe); F(qb);
  e = qce = fe = qb = nullptr;
  F(d_state);
  F(f_saved);
  F(d_stats);
  F(d_stats_list);
  n_stats = 0;
  F(qrte);
  F(eflags);
  F(d_all_int);
  F(d_all_ext);
  F(d_src_int);
  F(d_tail);
  d_tail = nullptr;
  h_tail.clear();
  for (int st = 0; st < 2; st++) {
    F(d_tcin[st]);
    F(d_tcout[st]);
  }
  drop_graphs();
  q = qold = qc = f = rf = ftmp = nullptr;
  d_P = nullptr; d_partial = nullptr; d_hist = nullptr; d_state = nullptr;
  for (auto &l : lev) {
    F(l.d_int); F(l.d_ext); F(l.d_coarse); F(l.d_cext); F(l.d_patch_list);
    F(l.d_cgl); F(l.d_blk);
    F(l.d_granges); F(l.d_gcells);
    F(l.d_iso); F(l.d_iso_stage); F(l.d_iso_rec); F(l.d_iso_srec); F(l.d_relax);
    for (int p = 0; p < 3; p++) {
      FG(l.d_gp[p]);
      F(l.d_tgp[p]);
      FG(l.d_fops[p]);
      FG(l.d_fpre[p]);
    }
    for (int k = 0; k < 5; k++) {
      FG(l.d_pops[k]);
      FG(l.d_prest[k]);
    }
    for (LevelData::DRide *r : {&l.d_ride_cgl, &l.d_ride_corr, &l.d_ride_cgl_loop}) {
      FG(r->rest);
      FG(r->ops);
    }
  }
  if (h_state) (void)hipHostFree(h_state);
  h_state = nullptr;
  for (auto e : ev_pool) (void)hipEventDestroy(e);
  ev_pool.clear();
  ev_used = 0;
  prof.clear();
}

// Brackets one launch of `kind` with events when that kind is profiled.
struct ProfScope {
  bsam_ctx *c;
  size_t ev = (size_t)-1;
  ProfScope(bsam_ctx *cc, int kind, double bytes) : c(cc) {
    if (!(c->prof_mask & (1u << kind))) return;
    if (c->ev_used + 2 > c->ev_pool.size()) {
      DeviceLock lock(g_device_lock);
      for (int n = 0; n < 256; n++) {
        hipEvent_t e;
        HIPCK(hipEventCreate(&e));
        c->ev_pool.push_back(e);
      }
    }
    ev = c->ev_used;
    c->ev_used += 2;
    HIPCK(hipEventRecord(c->ev_pool[ev], c->stream));
    c->prof.push_back({kind, bytes, ev});
  }
  ~ProfScope() {
    if (ev != (size_t)-1) (void)hipEventRecord(c->ev_pool[ev + 1], c->stream);
  }
};

static thread_local std::string g_err;

#define API_BEGIN(ctx)                     \
  if (!(ctx)) {                            \
    g_err = "null context";                \
    return -1;                             \
  }                                        \
  try {
#define API_END(ctx)                       \
  }                                        \
  catch (const Error &e) {                 \
    (ctx)->err = e.what();                 \
    return e.code;                         \
  }                                        \
  catch (const std::exception &e) {        \
    (ctx)->err = e.what();                 \
    return -1;                             \
  }                                        \
  return 0;

// ---- hierarchy commit --------------------------------------------------------
static void build_tiles(bsam_ctx *c, LevelData &L, int level) {
  int nd = c->nd();
  int TX = c->tx(), TY = c->ty(), TZ = c->tz();
  auto cover = [&](std::vector<Tile> &out, int p, int lo, int ex, int ey, int ez) {
    // tiles over [lo, lo+ex) x [lo, lo+ey) x [lo, lo+ez) (z fixed 0 in 2D)
    for (int k = (nd == 3 ? lo : 0); k < (nd == 3 ? lo + ez : 1); k += TZ)
      for (int j = lo; j < lo + ey; j += TY)
        for (int i = lo; i < lo + ex; i += TX) out.push_back(Tile{p, i, j, k});
  };
  auto lin = [&](std::vector<Tile> &out, int p, long long n) {
    for (long long e = 0; e < n; e += LIN_TILE) out.push_back(Tile{p, (int)e, 0, 0});
  };
  // the coarse loading runs over fine tiles on levels of at least
  // cgl_fine_cells cells, else over linear tiles of the coarse box
  long long level_cells = 0;
  for (int p : L.patches) {
    const HPatch &h = c->hp[p];
    level_cells += (long long)h.mx[0] * h.mx[1] * (nd == 3 ? h.mx[2] : 1);
  }
  const bool fine_cgl = c->cgl_fine_cells >= 0 && level_cells >= c->cgl_fine_cells;
  for (int p : L.patches) {
    const HPatch &h = c->hp[p];
    cover(L.t_int, p, 1, h.mx[0], h.mx[1], h.mx[2]);
    cover(L.t_ext, p, 0, h.mx[0] + 2, h.mx[1] + 2, h.mx[2] + 2);
    lin(L.t_qlin, p, split_of(h.mx[0] + 2, h.mx[1] + 2).pl * (nd == 3 ? h.mx[2] + 2 : 1));
    if (level > c->cfg.minlevel) {
      cover(L.t_coarse, p, 1, h.cmx[0], h.cmx[1], h.cmx[2]);
      cover(L.t_cext, p, 0, h.cmx[0] + 2, h.cmx[1] + 2, h.cmx[2] + 2);
      if (fine_cgl) {
        const int rx = 1 << cl_lx(nd), ry = 1 << cl_ly(nd), rz = 1 << cl_lz(nd);
        for (int k = (nd == 3 ? 1 : 0); k < (nd == 3 ? 1 + h.mx[2] : 1); k += rz)
          for (int j = 1; j < 1 + h.mx[1]; j += ry)
            for (int i = 1; i < 1 + h.mx[0]; i += rx) L.t_cgl.push_back(Tile{p, i, j, k});
      } else {
        lin(L.t_cgl, p,
            (long long)(h.cmx[0] + 2) * (h.cmx[1] + 2) * (nd == 3 ? h.cmx[2] + 2 : 1));
      }
      lin(L.t_blk, p, (long long)(h.cmx[0] + 1) * (h.cmx[1] + 1) * (nd == 3 ? h.cmx[2] + 1 : 1));
    }
    L.cells += (long long)h.mx[0] * h.mx[1] * (nd == 3 ? h.mx[2] : 1);
  }
  // the coarse loading's ring tiles follow its fine tiles
  L.n_cgl_res = fine_cgl ? (int)L.t_cgl.size() : -1;
  if (fine_cgl && level > c->cfg.minlevel)
    for (int p : L.patches) {
      const HPatch &h = c->hp[p];
      lin(L.t_cgl, p, ring_count(nd, h.cmx[0], h.cmx[1], nd == 3 ? h.cmx[2] : 1));
    }
}

// Places the tail levels' arrays in LDS within `budget` bytes (coarsest level
// first, per level q, qc, qb, f; what does not fit stays in the global pools)
// and uploads the tail level table. Returns false if the runtime refuses the
// resulting dynamic LDS size.
static bool plan_tail(bsam_ctx *c, long long budget, hipStream_t s) {
  const int minl = c->cfg.minlevel, nv = c->cfg.nrvars;
  c->tail_qpo = false;
  std::vector<TailLevel> tl;
  long long used[2] = {0, 0};
  bool open[2] = {true, true};   // no level has failed to fit yet
  std::vector<TailCopy> cin[2], cout[2];
  std::vector<int> nin[2], nout[2];
  for (int l = c->cfg.minlevel; l <= c->tail_top; l++) {
    LevelData &L = c->lev[l - minl];
    const DPatch &D = c->dp[L.patches[0]];
    TailLevel t;
    t.patch = L.patches[0];
    t.P[0] = t.P[1] = D;
    for (int ip = 0; ip < 3; ip++) {
      t.ngp[ip] = (int)L.gp[ip].size();
      t.gp[ip] = L.d_tgp[ip];
      for (const GhostOp &g : L.gp[ip])
        if (qpo_code(g.nflip)) c->tail_qpo = true;
    }
    for (int set = 0; set < 2; set++) {
      long long sz = set == 0 ? 8 : 4;
      long long bytes[4] = {D.qn * nv * sz, D.cvs * nv * sz, D.fvs * nv * sz,
                            set == 1 ? D.qn * sz : 0};
      for (int a = 0; a < 4; a++) t.lds[set][a] = -1;
      bool usable = set == 0 || c->cfg.precision == BSAM_PREC_MIXED;
      // a level is in LDS when q, qc (and the FP32 base point) all fit, its f
      // as well if room remains; levels in LDS form a prefix of the chain
      long long core = ((bytes[0] + 15) & ~15LL) + ((bytes[1] + 15) & ~15LL) +
                       ((bytes[3] + 15) & ~15LL);
      bool fits = usable && open[set] && used[set] + core <= budget;
      if (!fits) open[set] = false;
      for (int a : {0, 1, 3, 2}) {
        if (!fits) break;
        long long b = (bytes[a] + 15) & ~15LL;
        if (b == 0) continue;
        if (a == 2 && used[set] + b > budget) continue;   // f stays in the pool
        t.lds[set][a] = (int)used[set];
        used[set] += b;
        // staging entry: pool address of the array
        char *pool = nullptr;
        long long off = 0;
        if (set == 0) {
          pool = a == 0 ? (char *)c->q : a == 1 ? (char *)c->qc : (char *)c->f;
        } else {
          pool = a == 0 ? (char *)c->e : a == 1 ? (char *)c->qce : a == 2 ? (char *)c->fe
                                                                          : (char *)c->qb;
        }
        off = a == 0 || a == 3 ? D.q : a == 1 ? D.qc : D.f;
        // staging chunks of one wave's width; q's variables lie the pool's
        // stride apart and are packed one after the other in LDS
        const int nvar = a == 0 ? nv : 1;
        const int n = (int)(bytes[a] / sz / nvar);
        for (int v = 0; v < nvar; v++) {
          for (int e = 0; e < n; e += 64) {
            TailCopy tc{pool + (off + v * D.qvs + e) * sz,
                        t.lds[set][a] + (int)((v * n + e) * sz), std::min(64, n - e)};
            cin[set].push_back(tc);
            if (a != 3) cout[set].push_back(tc);
          }
        }
      }
      if (t.lds[set][0] >= 0) t.P[set].qvs = D.qn;   // q packed in LDS
      nin[set].push_back((int)cin[set].size());
      nout[set].push_back((int)cout[set].size());
    }
    tl.push_back(t);
  }
  size_t need = (size_t)std::max(used[0], used[1]);
  if (need > 0 && tail_allow_lds(need) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  c->tail_lds[0] = (size_t)used[0];
  c->tail_lds[1] = (size_t)used[1];
  if (c->d_tail) (void)hipFree(c->d_tail);
  c->d_tail = dev_upload(tl, s);
  c->h_tail = tl;
  for (int set = 0; set < 2; set++) {
    if (c->d_tcin[set]) (void)hipFree(c->d_tcin[set]);
    if (c->d_tcout[set]) (void)hipFree(c->d_tcout[set]);
    c->d_tcin[set] = dev_upload(cin[set], s);
    c->d_tcout[set] = dev_upload(cout[set], s);
    c->tcin_n[set] = nin[set];
    c->tcout_n[set] = nout[set];
  }
  return true;
}

// Every field pool starts at zero after a commit.
static void zero_pools(bsam_ctx *c) {
  hipStream_t s = c->stream;
  HIPCK(hipMemsetAsync(c->qrte, 0, sizeof(double) * c->nr, s));
  HIPCK(hipMemsetAsync(c->eflags, 0, sizeof(int) * c->nfl, s));
  HIPCK(hipMemsetAsync(c->q, 0, sizeof(double) * c->nq, s));
  HIPCK(hipMemsetAsync(c->qold, 0, sizeof(double) * c->nq, s));
  HIPCK(hipMemsetAsync(c->qc, 0, sizeof(double) * c->nc, s));
  HIPCK(hipMemsetAsync(c->f, 0, sizeof(double) * c->nf, s));
  HIPCK(hipMemsetAsync(c->rf, 0, sizeof(double) * c->nf, s));
  HIPCK(hipMemsetAsync(c->ftmp, 0, sizeof(double) * c->nf, s));
  HIPCK(hipMemsetAsync(c->f_saved, 0, sizeof(double) * c->nf, s));
  if (c->cfg.precision == BSAM_PREC_MIXED) {
    HIPCK(hipMemsetAsync(c->e, 0, sizeof(float) * c->nq, s));
    HIPCK(hipMemsetAsync(c->qb, 0, sizeof(float) * c->nq, s));
    HIPCK(hipMemsetAsync(c->qce, 0, sizeof(float) * c->nc, s));
    HIPCK(hipMemsetAsync(c->fe, 0, sizeof(float) * c->nf, s));
  }
}

// True if two patch lists describe the same hierarchy (ids aside): same
// order, levels, parents and geometry.
static bool same_geometry(const std::vector<HPatch> &a, const std::vector<HPatch> &b) {
  if (a.size() != b.size()) return false;
  for (size_t i = 0; i < a.size(); i++) {
    const HPatch &x = a[i], &y = b[i];
    if (x.level != y.level || x.parent != y.parent || x.dx != y.dx || x.group != y.group)
      return false;
    for (int d = 0; d < 3; d++) {
      if (x.mx[d] != y.mx[d] || x.xlower[d] != y.xlower[d]) return false;
      for (int e = 0; e < 2; e++)
        if (x.mb[d][e] != y.mb[d][e] || x.mg[d][e] != y.mg[d][e]) return false;
    }
    for (int f = 0; f < 6; f++)
      if (x.mthbc[f] != y.mthbc[f]) return false;
  }
  return true;
}

// The tile shape of the level kernels, chosen per hierarchy: every candidate
// shape times both colours of the plain FP64 relaxation (the bulk of a
// V-cycle's launches) on each level of at least 4096 cells at or above the
// root, three rounds between HIP events on the context's stream, over
// scratch pools of the hierarchy's size. The first candidate is the default
// shape; another replaces it only if more than 3 % faster, since the timing
// leaves out the ghost updates the fused relaxations apply, whose cost also
// depends on the shape (a full-solve sweep of the 3D workload ranks the
// default first, 1.3 % ahead of the relaxation-only winner). Tile shape changes
// which workgroup relaxes a cell, never its arithmetic. Winners are kept per
// shape class (dimension and the patch extents of every level) for the
// process, so recommitting a hierarchy of the same shape does not time again.
static std::mutex g_tile_cache_mutex;
static std::map<std::string, std::pair<int, int>> g_tile_cache;

static void choose_tiles(bsam_ctx *c) {
  const int nd = c->nd(), minl = c->cfg.minlevel;
  std::vector<std::vector<int>> levels;   // patches of each timed level
  std::string key = std::to_string(nd) + ":";
  for (int l = 0; l <= c->finest; l++) {
    std::vector<int> ps;
    long long cells = 0;
    std::vector<std::array<int, 3>> ext;
    for (int p = 0; p < (int)c->hp.size(); p++)
      if (c->hp[p].level == l) {
        ps.push_back(p);
        const HPatch &h = c->hp[p];
        cells += (long long)h.mx[0] * h.mx[1] * (nd == 3 ? h.mx[2] : 1);
        ext.push_back({h.mx[0], h.mx[1], h.mx[2]});
      }
    std::sort(ext.begin(), ext.end());
    key += "L" + std::to_string(l);
    for (auto &e : ext) key += "," + std::to_string(e[0]) + "x" + std::to_string(e[1]) + "x" +
                                 std::to_string(e[2]);
    if (cells >= 4096) levels.push_back(ps);
  }
  (void)minl;
  if (levels.empty()) return;
  {
    std::lock_guard<std::mutex> g(g_tile_cache_mutex);
    auto it = g_tile_cache.find(key);
    if (it != g_tile_cache.end()) {
      c->tile_lx = it->second.first;
      c->tile_ly = it->second.second;
      c->tile_auto_cached = 1;
      return;
    }
  }
  // candidate shapes (log2 extents): 256 threads' cells per tile
  std::vector<std::pair<int, int>> cand;
  if (nd == 2) cand = {{5, 3}, {6, 2}, {4, 4}, {7, 1}, {3, 5}};
  else cand = {{4, 3}, {3, 3}, {4, 2}, {5, 2}, {4, 4}, {5, 3}, {3, 2}, {5, 1}};
  double *q = nullptr, *f = nullptr;
  DPatch *dP = nullptr;
  Tile *dt = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipStream_t s = c->stream;
  double best = 1e300;
  int bl = c->tile_lx, bly = c->tile_ly;
  hipError_t err = hipSuccess;
  {
    DeviceLock lock(g_device_lock);
    err = hipMalloc(&q, sizeof(double) * c->nq);
    if (err == hipSuccess) err = hipMalloc(&f, sizeof(double) * c->nf);
    if (err == hipSuccess) err = hipEventCreate(&e0);
    if (err == hipSuccess) err = hipEventCreate(&e1);
  }
  if (err == hipSuccess) err = hipMemsetAsync(q, 0, sizeof(double) * c->nq, s);
  if (err == hipSuccess) err = hipMemsetAsync(f, 0, sizeof(double) * c->nf, s);
  if (err == hipSuccess) {
    DeviceLock lock(g_device_lock);
    dP = dev_upload(c->dp, s);
  }
  for (auto &cd : cand) {
    if (err != hipSuccess) break;
    const int lx = cd.first, ly = cd.second;
    const int TX = 1 << lx, TY = 1 << ly, TZ = nd == 3 ? 256 >> (lx + ly) : 1;
    std::vector<Tile> tiles;
    std::vector<std::pair<int, int>> ranges;   // per timed level: first tile, count
    for (auto &ps : levels) {
      const int first = (int)tiles.size();
      for (int p : ps) {
        const HPatch &h = c->hp[p];
        for (int k = (nd == 3 ? 1 : 0); k < (nd == 3 ? 1 + h.mx[2] : 1); k += TZ)
          for (int j = 1; j < 1 + h.mx[1]; j += TY)
            for (int i = 1; i < 1 + h.mx[0]; i += TX) tiles.push_back(Tile{p, i, j, k});
      }
      ranges.push_back({first, (int)tiles.size() - first});
    }
    {
      DeviceLock lock(g_device_lock);
      dt = dev_upload(tiles, s);
    }
    Dev d = c->dev();
    d.q = q;
    d.f = f;
    d.P = dP;
    d.gstate = nullptr;
    d.tile_lx = lx;
    d.tile_ly = ly;
    for (int r = 0; r < 4 && err == hipSuccess; r++) {   // round 0 warms up
      if (r == 1) err = hipEventRecord(e0, s);
      for (auto &rg : ranges)
        for (int color = 1; color <= 2 && err == hipSuccess; color++)
          err = launch_relax(nd, PREC_SOLUTION_FP64, d, dt + rg.first, rg.second, color, s);
    }
    if (err == hipSuccess) err = hipEventRecord(e1, s);
    if (err == hipSuccess) err = hipEventSynchronize(e1);
    float ms = 0.f;
    if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
    {
      DeviceLock lock(g_device_lock);
      (void)hipFree(dt);
      dt = nullptr;
    }
    c->tile_auto_candidates++;
    const double margin = c->tile_auto_candidates == 1 ? 1.0 : 0.97;
    if (err == hipSuccess && ms / 3.0 < best * margin) {
      best = ms / 3.0;
      bl = lx;
      bly = ly;
    }
  }
  {
    DeviceLock lock(g_device_lock);
    (void)hipStreamSynchronize(s);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (q) (void)hipFree(q);
    if (f) (void)hipFree(f);
    if (dP) (void)hipFree(dP);
  }
  HIPCK(err);
  c->tile_lx = bl;
  c->tile_ly = bly;
  c->tile_auto_us = best * 1e3;
  std::lock_guard<std::mutex> g(g_tile_cache_mutex);
  g_tile_cache[key] = {bl, bly};
}

// The sources an update reads (q-pool offsets of variable 1): fine f0 (the
// copy source of kind 0), f1, and one, three or four coarse values.
static int ghost_sources(const GhostOp &g, long long *s) {
  s[0] = g.f0;
  if (g.kind == 0) return 1;
  s[1] = g.f1;
  const int nc = g.kind == 1 ? 1 : g.kind == 2 ? 3 : 4;
  for (int k = 0; k < nc; k++) s[2 + k] = g.c[k];
  return 2 + nc;
}

// LDS of one workgroup on gfx950, the bound of an isolated patch's arrays.
constexpr long long kIsoLdsBytes = 160 * 1024;

// Finds the level's isolated patches: no update of the level's three
// SetGhost programs with its destination in the patch reads another patch of
// the level, and no update with its destination in another patch reads the
// patch (sources on the parent level are read, not written, while the level
// is smoothed). Those whose ghost-extended q and read parent cells fit the
// LDS of one workgroup in FP64 (both variables) become IsoPatch records with
// their programs in LDS indices, if together they hold at least
// iso_min_cells cells. Sets t_relax to the other patches' interior tiles and
// prog[ip] to the updates of program ip with their destination in those
// patches (gp[ip] itself when no patch is isolated).
static void plan_iso(bsam_ctx *c, LevelData &L, std::vector<GhostOp> (&prog)[3]) {
  L.iso.clear();
  L.iso_stage.clear();
  L.iso_rec.clear();
  L.iso_srec.clear();
  L.iso_lds = 0;
  L.iso_cells = 0;
  for (long long &u : L.iso_updates) u = 0;
  L.t_relax = L.t_int;
  L.relax_cells = L.cells;
  for (int ip = 0; ip < 3; ip++) prog[ip] = L.gp[ip];
  const std::vector<int> &P = L.patches;
  if (c->iso_min_cells <= 0 || !c->fuse_ghost || c->cfg.nrvars != 2 || P.empty()) return;
  // the level's patches take ascending, contiguous ranges of the q pool
  auto owner = [&](long long off) -> int {
    size_t lo = 0, hi = P.size();
    while (hi - lo > 1) {
      const size_t mid = (lo + hi) / 2;
      if (c->hp[P[mid]].qoff <= off) lo = mid;
      else hi = mid;
    }
    const HPatch &h = c->hp[P[lo]];
    return off >= h.qoff && off < h.qoff + h.qn ? (int)lo : -1;
  };
  std::vector<char> coupled(P.size(), 0);
  long long src[6];
  for (int ip = 0; ip < 3; ip++)
    for (const GhostOp &g : L.gp[ip]) {
      const int a = owner(g.dst);
      if (a < 0) return;   // not a program of this level: no isolated patches
      const int n = ghost_sources(g, src);
      for (int k = 0; k < n; k++) {
        const int b = owner(src[k]);
        if (b >= 0 && b != a) coupled[(size_t)a] = coupled[(size_t)b] = 1;
      }
    }
  // per candidate: its updates per program and its staged (parent) sources
  struct Cand {
    std::vector<const GhostOp *> ops[3];
    std::unordered_map<long long, int> stage;
    std::vector<long long> order;
  };
  std::vector<Cand> cand(P.size());
  for (int ip = 0; ip < 3; ip++)
    for (const GhostOp &g : L.gp[ip]) {
      const int a = owner(g.dst);
      if (coupled[(size_t)a]) continue;
      Cand &C = cand[(size_t)a];
      C.ops[ip].push_back(&g);
      const HPatch &h = c->hp[P[(size_t)a]];
      const int n = ghost_sources(g, src);
      for (int k = 0; k < n; k++)
        if ((src[k] < h.qoff || src[k] >= h.qoff + h.qn) && !C.stage.count(src[k])) {
          C.stage[src[k]] = (int)C.order.size();
          C.order.push_back(src[k]);
        }
    }
  std::vector<char> iso(P.size(), 0);
  long long cells = 0;
  for (size_t a = 0; a < P.size(); a++) {
    if (coupled[a]) continue;
    const HPatch &h = c->hp[P[a]];
    const long long w = h.qn + (long long)cand[a].order.size();
    if (w > 65535 || 2 * w * (long long)sizeof(double) > kIsoLdsBytes) continue;
    iso[a] = 1;
    cells += (long long)h.mx[0] * h.mx[1] * (c->nd() == 3 ? h.mx[2] : 1);
  }
  if (cells < c->iso_min_cells) return;
  for (size_t a = 0; a < P.size(); a++) {
    if (!iso[a]) continue;
    const HPatch &h = c->hp[P[a]];
    const Cand &C = cand[a];
    IsoPatch I{};
    I.patch = P[a];
    I.ns = (int)C.order.size();
    I.stage0 = (int)L.iso_stage.size();
    for (long long o : C.order) L.iso_stage.push_back((int)o);
    auto lds_index = [&](long long off) -> unsigned short {
      if (off >= h.qoff && off < h.qoff + h.qn) return (unsigned short)(off - h.qoff);
      return (unsigned short)(h.qn + C.stage.at(off));
    };
    for (int ip = 0; ip < 3; ip++) {
      I.rec[ip] = (int)L.iso_rec.size();
      I.srec[ip] = (int)L.iso_srec.size();
      for (const GhostOp *g : C.ops[ip]) {
        IsoRec r{};
        r.dst = lds_index(g->dst);
        r.kind = g->kind;
        r.nflip = g->nflip;
        const int n = ghost_sources(*g, src);
        for (int k = 0; k < n; k++) r.s[k] = lds_index(src[k]);
        unsigned long long w;
        if (iso_short(r, w)) L.iso_srec.push_back(w);
        else L.iso_rec.push_back(r);
      }
      L.iso_updates[ip] += (long long)C.ops[ip].size();
    }
    I.rec[3] = (int)L.iso_rec.size();
    I.srec[3] = (int)L.iso_srec.size();
    L.iso.push_back(I);
    L.iso_lds = std::max(L.iso_lds, (size_t)(2 * (h.qn + I.ns) * (long long)sizeof(double)));
  }
  L.iso_cells = cells;
  L.relax_cells = L.cells - cells;
  L.t_relax.clear();
  for (const Tile &t : L.t_int)
    if (!iso[(size_t)owner(c->hp[t.patch].qoff)]) L.t_relax.push_back(t);
  for (int ip = 0; ip < 3; ip++) {
    prog[ip].clear();
    for (const GhostOp &g : L.gp[ip])
      if (!iso[(size_t)owner(g.dst)]) prog[ip].push_back(g);
  }
}

static void commit(bsam_ctx *c) {
  const int nd = c->nd(), nv = c->cfg.nrvars;
  int n = (int)c->hp.size();
  if (n == 0) throw Error("empty hierarchy");
  // tile shape: 32 x 8 in 2D and 16 x 8 x 2 in 3D unless tuned (the fastest
  // of the tile sweep on the 2D and 3D AMR workloads)
  int &lx = c->tile_lx, &ly = c->tile_ly;
  if (!c->tile_user) lx = ly = -1;   // the defaults, or the chooser's pick below
  if (nd == 2) {
    if (lx < 0 && ly < 0) lx = 5, ly = 3;
    else if (lx < 0) lx = 8 - ly;
    else if (ly < 0) ly = 8 - lx;
    if (lx + ly != 8) throw Error("2D tiles must hold 256 cells (tile_x * tile_y = 256)");
  } else {
    if (lx < 0 && ly < 0) lx = 4, ly = 3;
    else if (lx < 0) lx = std::min(4, 8 - ly);
    else if (ly < 0) ly = std::min(3, 8 - lx);
    if (lx + ly > 8) throw Error("3D tiles: tile_x * tile_y must be at most 256");
  }
  // the relaxation kernels give each thread one cell of a colour in its row
  if (lx < 1) throw Error("tiles must be at least 2 cells wide (tile_x >= 2)");
  int minl = c->cfg.minlevel;
  c->finest = minl;
  for (auto &h : c->hp) c->finest = std::max(c->finest, h.level);
  c->lev.assign(c->finest - minl + 1, LevelData());
  int prev = minl - 1;
  for (int p = 0; p < n; p++) {
    HPatch &h = c->hp[p];
    if (h.level < prev) throw Error("patches must be added with levels ascending");
    prev = h.level;
    if (h.level < minl) throw Error("patch level below minlevel");
    c->lev[h.level - minl].patches.push_back(p);
    if (h.level > minl) {
      if (h.parent < 0 || h.parent >= p || c->hp[h.parent].level != h.level - 1)
        throw Error("patch " + std::to_string(p) + ": parent must be an earlier patch one level coarser");
      const HPatch &pa = c->hp[h.parent];
      for (int d = 0; d < nd; d++) {
        if (h.mx[d] != 2 * (h.mb[d][1] - h.mb[d][0] + 1))
          throw Error("patch " + std::to_string(p) + ": mx != 2*(mbounds extent)");
        if (h.mb[d][0] < 1 || h.mb[d][1] > pa.mx[d])
          throw Error("patch " + std::to_string(p) + ": mbounds outside the parent");
      }
    } else if (h.parent != -1) {
      throw Error("coarsest patch must have parent -1");
    }
    for (int d = 0; d < nd; d++)
      if (h.mx[d] < 1) throw Error("mx must be positive");
    for (int fc = 0; fc < 2 * nd; fc++) {
      int t = h.mthbc[fc];
      if (!(t == 1 || t == 2 || t == 999 || (t >= 10 && t <= 99)))
        throw Error("unsupported boundary code " + std::to_string(t));
    }
  }
  for (int l = minl; l <= c->finest; l++) {
    auto &P = c->lev[l - minl].patches;
    if (P.empty()) throw Error("level " + std::to_string(l) + " has no patches");
    // levels <= 0 hold one patch (one per group when there are several)
    if (l <= 0) {
      std::vector<int> per_group(c->ngroups, 0);
      for (int p : P) per_group[c->hp[p].group]++;
      for (int k : per_group)
        if (k != 1) throw Error("levels <= 0 hold one patch (of each group)");
    }
    // same-level interiors must not overlap (GhostOverlap copies assume it)
    for (size_t a = 0; a < P.size(); a++)
      for (size_t b = a + 1; b < P.size(); b++) {
        const HPatch &A = c->hp[P[a]], &B = c->hp[P[b]];
        bool ov = true;
        for (int d = 0; d < nd; d++)
          ov = ov && A.mg[d][0] <= B.mg[d][1] && B.mg[d][0] <= A.mg[d][1];
        if (ov) throw Error("overlapping patches on level " + std::to_string(l));
      }
  }
  // pool offsets
  c->nq = c->nc = c->nf = c->nr = c->nfl = 0;
  c->dp.resize(n);
  for (int p = 0; p < n; p++) {
    HPatch &h = c->hp[p];
    DPatch &D = c->dp[p];
    memset(&D, 0, sizeof D);
    long long ex = h.mx[0] + 2, ey = h.mx[1] + 2, ez = nd == 3 ? h.mx[2] + 2 : 1;
    long long cx = h.cmx[0] + 2, cy = h.cmx[1] + 2, cz = nd == 3 ? h.cmx[2] + 2 : 1;
    long long fn = (long long)h.mx[0] * h.mx[1] * (nd == 3 ? h.mx[2] : 1);
    D.mx = h.mx[0]; D.my = h.mx[1]; D.mz = nd == 3 ? h.mx[2] : 1;
    D.px = (int)ex; D.py = (int)ey;
    D.cmx = h.cmx[0]; D.cmy = h.cmx[1]; D.cmz = nd == 3 ? h.cmx[2] : 1;
    D.level = h.level; D.parent = h.parent; D.group = h.group;
    for (int d = 0; d < 3; d++) {
      D.mb[d][0] = h.mb[d][0];
      D.mb[d][1] = h.mb[d][1];
    }
    // device planes of rows of n cells take 2 * ceil(n / 2) cells per row;
    // q holds every patch's first variable, then every patch's second (the
    // stride is set below, once the first variables are counted)
    D.q = c->nq; D.qn = split_of(ex, ey).pl * ez;
    D.qc = c->nc; D.cvs = split_of(cx, cy).pl * cz;
    D.f = c->nf; D.fvs = split_of(h.mx[0], h.mx[1]).pl * (nd == 3 ? h.mx[2] : 1);
    {
      // qrte(1:cmx) with cmx = mx/2, or 1 for odd mx (AllocFields)
      long long r = 1;
      for (int d = 0; d < nd; d++) r *= h.mx[d] % 2 ? 1 : h.mx[d] / 2;
      D.qr = c->nr; D.rvs = r;
      D.fl = c->nfl;
      c->nr += r * nv;
      c->nfl += fn;
    }
    D.dx = h.dx;
    {
      double h2 = h.dx * h.dx;   // as RelaxGrid / Operator compute it per call
      D.dt_h2 = c->cfg.dt / h2;
      const double eps2 = c->group_eps2[(size_t)h.group];   // the group's eps^2
      D.eps2_h2 = eps2 / h2;
      D.lap_h2 = (nd == 2 ? 4.0 : 6.0) * eps2 / h2;
    }
    h.qoff = D.q; h.qn = D.qn; h.px = ex; h.py = ey;
    c->nq += D.qn;
    c->nc += D.cvs * nv;
    c->nf += D.fvs * nv;
  }
  c->qstride = c->nq;   // cells per variable over all patches
  for (DPatch &D : c->dp) D.qvs = c->qstride;
  c->nq *= nv;
  if (c->nq >= (1LL << 31)) throw Error("q pool exceeds 2^31 elements");
  // f ranges a multigrid iteration rewrites (the coarse loading writes f on
  // every level below the finest, the tail the whole coarse chain's): the f
  // of every patch below the finest level, merged into contiguous runs. The
  // finest level's f is only read.
  c->f_restore.clear();
  for (int p = 0; p < n; p++) {
    const DPatch &D = c->dp[p];
    if (D.level >= c->finest) continue;
    const long long a = D.f, b = D.f + D.fvs * nv;
    if (!c->f_restore.empty() && c->f_restore.back().second == a) c->f_restore.back().second = b;
    else c->f_restore.push_back({a, b});
  }
  if (!c->host_only && c->tile_auto && !c->tile_user) choose_tiles(c);
  Periodic per;
  if (c->finest >= 0) {
    const HPatch &root = c->hp[c->lev[0 - minl].patches[0]];
    per = periodic_setup(root, nd);
    for (int d = 0; d < nd; d++)
      if ((root.mthbc[2 * d] == 2) != (root.mthbc[2 * d + 1] == 2))
        throw Error("periodic boundaries must pair opposite faces");
    if (c->ngroups > 1)
      for (int p : c->lev[0 - minl].patches)
        for (int fc = 0; fc < 2 * nd; fc++)
          if (c->hp[p].mthbc[fc] == 2)
            throw Error("groups: periodic boundaries would join the groups' roots");
  }
  // per level l: the parent cells under level l's patches (and the same of
  // level l - 1), marked over the q pool
  std::vector<unsigned char> under, under_prev;
  for (int l = minl; l <= c->finest; l++) {
    LevelData &L = c->lev[l - minl];
    build_tiles(c, L, l);
    for (int ip = 0; ip < 3; ip++) {
      std::string msg;
      if (build_ghost_program(c->hp, L.patches, nd, l, ip, per, L.gp[ip], msg))
        throw Error("ghost program level " + std::to_string(l) + ": " + msg);
      if (c->ngroups > 1) set_op_groups(c->hp, L.gp[ip]);
    }
    // each group's patches: contiguous in the level list, so its interior
    // tiles are one range of the level's tile list
    {
      const int G = c->ngroups;
      L.granges.assign(2 * G, 0);
      L.gcells.assign(G, 0.0);
      std::vector<int> seen(G, 0);
      int prev = -1;
      for (int p : L.patches) {
        const int g = c->hp[p].group;
        if (g != prev && seen[g])
          throw Error("level " + std::to_string(l) + ": a group's patches must be contiguous");
        seen[g] = 1;
        prev = g;
      }
      for (size_t t = 0; t < L.t_int.size(); t++) {
        const int g = c->hp[L.t_int[t].patch].group;
        if (L.granges[2 * g + 1] == 0) L.granges[2 * g] = (int)t;
        L.granges[2 * g + 1]++;
      }
      for (int p : L.patches) {
        const HPatch &h = c->hp[p];
        const double cells = (double)h.mx[0] * h.mx[1] * (nd == 3 ? h.mx[2] : 1);
        L.gcells[h.group] += (c->cfg.errortype == 2 ? 1.0 : (double)nv) * cells;
      }
    }
    // the fusion and riding plans serve the V-cycle only; a context that
    // does not fuse (fuse_ghost 0) goes without them. They cover the patches
    // the smoothing relaxes per colour: all but the isolated ones.
    std::vector<GhostOp> prog[3];
    plan_iso(c, L, prog);
    auto plan_fused = [&]() {
      for (int f = 0; f < 3; f++) {
        std::string why;
        L.fused[f] = c->fuse_ghost && !L.t_relax.empty()
                         ? plan_fused_ghost(c->hp, L.t_relax, nd, c->tx(), c->ty(), c->tz(),
                                            prog[kFusedPass[f]], prog[kFusedColor[f]],
                                            kFusedColor[f], why)
                         : FusedGhost();
      }
    };
    plan_fused();
    if (!L.iso.empty() && !L.t_relax.empty() &&
        !(L.fused[0].ok && L.fused[1].ok && L.fused[2].ok)) {
      // the smoothing of the other patches must fuse every transition; else
      // no patch is isolated
      const long long keep = c->iso_min_cells;
      c->iso_min_cells = 0;
      plan_iso(c, L, prog);
      c->iso_min_cells = keep;
      plan_fused();
    }
    for (int k = 0; k < 4 && c->fuse_ghost; k++) {
      L.post[k] = PostGhost();
      const int color = kPostColor[k], pend = kPostPending[k];
      const int fb = pend >= 0 ? fused_index(pend, color) : -1;
      const int fn = fused_index(color, 3 - color);
      if (!L.fused[fn].ok || L.fused[fn].pre.empty()) continue;
      std::string why;
      L.post[k] = plan_post_ghost(c->hp, L.t_relax, nd, c->tx(), c->ty(), c->tz(), color,
                                  fb >= 0 && L.fused[fb].ok ? &L.fused[fb] : nullptr,
                                  L.fused[fn].pre, why);
    }
    L.post[kPostFinal] = PostGhost();
    if (c->fuse_ghost && !prog[2].empty() && !L.t_relax.empty()) {
      // the last launch of a smoothing relaxes colour 2 after SetGhost(1)
      const int fb = fused_index(1, 2);
      std::string why;
      L.post[kPostFinal] = plan_post_ghost(c->hp, L.t_relax, nd, c->tx(), c->ty(), c->tz(), 2,
                                           L.fused[fb].ok ? &L.fused[fb] : nullptr, prog[2],
                                           why);
    }
    if (l > minl) {
      // parent cells under this level's patches: what its restriction and
      // fill-down write
      under_prev.swap(under);
      std::vector<unsigned char> &written = under;
      written.assign((size_t)c->nq, 0);
      for (int p : L.patches) {
        const HPatch &h = c->hp[p];
        const HPatch &pa = c->hp[h.parent];
        for (int k = (nd == 3 ? h.mb[2][0] : 0); k <= (nd == 3 ? h.mb[2][1] : 0); k++)
          for (int j = h.mb[1][0]; j <= h.mb[1][1]; j++)
            for (int i = h.mb[0][0]; i <= h.mb[0][1]; i++)
              written[(size_t)(pa.qoff + qcell(pa.px, pa.py, i, j, k))] = 1;
      }
      for (int ip = 0; ip < 3; ip++) L.down_ok[ip] = program_reads_avoid(L.gp[ip], written);
      const LevelData &PL = c->lev[l - 1 - minl];
      // only the reads of cells the parent's program writes matter to a ride
      std::vector<int> at;
      if (c->fuse_ghost && destination_index(PL.gp[0], at))
        L.ride_cgl = plan_ride(PL.gp[0],
                               coarse_ghost_load_reads(c->hp, L.t_cgl, L.n_cgl_res, nd, &at));
      L.ride_cgl_loop = RideGhost();
      if (c->fuse_ghost && l - 1 >= 0) {
        // cells changed since the loop's FillDown applied SetGhost(l - 1, 0):
        // level l - 1 under level l (this V-cycle's restriction: `written`),
        // and what FillDown wrote after that program, level l - 2 under
        // l - 1 (the previous level's mask) and, when l - 2 >= 0, level l - 2's
        // ghost cells (its SetGhost(l - 2, 0))
        std::vector<unsigned char> later;
        if (l - 2 >= minl) later = under_prev;
        if (l - 2 >= 0)
          for (const GhostOp &g : c->lev[l - 2 - minl].gp[0]) {
            if ((size_t)g.dst >= later.size()) later.resize((size_t)c->nq, 0);
            later[(size_t)g.dst] = 1;
          }
        std::vector<GhostOp> sub;
        long long src[6];
        auto hit = [&](const std::vector<unsigned char> &m, long long o) {
          return o >= 0 && o < (long long)m.size() && m[(size_t)o];
        };
        for (const GhostOp &g : PL.gp[0]) {
          const int n = ghost_sources(g, src);
          bool changed = false;
          for (int k = 0; k < n && !changed; k++)
            changed = hit(written, src[k]) || hit(later, src[k]);
          if (changed) sub.push_back(g);
        }
        if (destination_index(sub, at))
          L.ride_cgl_loop = plan_ride(sub, coarse_ghost_load_reads(c->hp, L.t_cgl, L.n_cgl_res,
                                                                   nd, &at));
      }
      if (c->fuse_ghost && destination_index(PL.gp[2], at))
        L.ride_corr = plan_ride(PL.gp[2], correct_reads(c->hp, L.t_blk, nd, &at));
    }
  }
  if (c->host_only) return;
  // device allocation
  DeviceLock lock(g_device_lock);
  c->free_device();
  hipStream_t s = c->stream;
  HIPCK(hipMalloc(&c->q, sizeof(double) * c->nq));
  HIPCK(hipMalloc(&c->qold, sizeof(double) * c->nq));
  HIPCK(hipMalloc(&c->qc, sizeof(double) * c->nc));
  HIPCK(hipMalloc(&c->f, sizeof(double) * c->nf));
  HIPCK(hipMalloc(&c->rf, sizeof(double) * c->nf));
  HIPCK(hipMalloc(&c->ftmp, sizeof(double) * c->nf));
  HIPCK(hipMalloc(&c->f_saved, sizeof(double) * c->nf));
  HIPCK(hipMalloc(&c->qrte, sizeof(double) * c->nr));
  HIPCK(hipMalloc(&c->eflags, sizeof(int) * c->nfl));
  if (c->cfg.precision == BSAM_PREC_MIXED) {
    HIPCK(hipMalloc(&c->e, sizeof(float) * c->nq));
    HIPCK(hipMalloc(&c->qb, sizeof(float) * c->nq));
    HIPCK(hipMalloc(&c->qce, sizeof(float) * c->nc));
    HIPCK(hipMalloc(&c->fe, sizeof(float) * c->nf));
  }
  zero_pools(c);
  c->d_P = dev_upload(c->dp, s);
  size_t maxt = 1, iso_lds = 0;
  for (auto &L : c->lev) {
    L.d_int = dev_upload(L.t_int, s);
    L.d_ext = dev_upload(L.t_ext, s);
    L.d_coarse = dev_upload(L.t_coarse, s);
    L.d_cext = dev_upload(L.t_cext, s);
    L.d_cgl = dev_upload(L.t_cgl, s);
    L.d_blk = dev_upload(L.t_blk, s);
    const bool sh = c->short_ghost != 0;
    for (int ip = 0; ip < 3; ip++) L.d_gp[ip] = upload_ghost(L.gp[ip], nullptr, sh, s);
    for (int f = 0; f < 3; f++)
      if (L.fused[f].ok) {
        L.d_fops[f] = upload_ghost(L.fused[f].ops, &L.fused[f].start, sh, s);
        L.d_fpre[f] = upload_ghost(L.fused[f].pre, nullptr, sh, s);
      }
    for (int k = 0; k < 5; k++)
      if (L.post[k].ok) {
        L.d_pops[k] = upload_ghost(L.post[k].ops, &L.post[k].start, sh, s);
        L.d_prest[k] = upload_ghost(L.post[k].rest, nullptr, sh, s);
      }
    auto up_ride = [&](const RideGhost &r, LevelData::DRide &d) {
      if (!r.ok) return;
      d.rest = upload_ghost(r.rest, nullptr, sh, s);
      d.ops = upload_ghost(r.ops, &r.start, sh, s);
    };
    up_ride(L.ride_cgl, L.d_ride_cgl);
    up_ride(L.ride_corr, L.d_ride_corr);
    up_ride(L.ride_cgl_loop, L.d_ride_cgl_loop);
    L.d_relax = dev_upload(L.t_relax, s);
    if (!L.iso.empty()) {
      L.d_iso = dev_upload(L.iso, s);
      L.d_iso_stage = dev_upload(L.iso_stage, s);
      L.d_iso_rec = dev_upload(L.iso_rec, s);
      L.d_iso_srec = dev_upload(L.iso_srec, s);
      iso_lds = std::max(iso_lds, L.iso_lds);
    }
    maxt = std::max(maxt, L.t_int.size());
  }
  if (iso_lds > 0) {
    HIPCK(iso_allow_lds(iso_lds));
    if (!c->side) {
      HIPCK(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
      HIPCK(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
      HIPCK(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
    }
  }
  // tail: the longest run of single-patch levels from minlevel whose cell
  // counts stay within tail_cells (levels <= 0 hold one patch each) and whose
  // ghost programs are patch-local copies
  c->tail_top = c->cfg.minlevel - 1;
  std::vector<std::vector<GhostOp>> tprog;
  for (int l = c->cfg.minlevel; l <= std::min(0, c->finest); l++) {
    LevelData &L = c->lev[l - minl];
    std::vector<GhostOp> loc[3];
    if (L.cells > c->tail_cells || L.patches.size() != 1 || l - minl >= tail_max_levels() ||
        !local_programs(c->hp[L.patches[0]], L.gp, nv, loc))
      break;
    c->tail_top = l;
    for (int ip = 0; ip < 3; ip++) L.d_tgp[ip] = dev_upload(loc[ip], s);
  }
  c->tail_lds[0] = c->tail_lds[1] = 0;
  if (c->ntail() > 0) {
    // LDS placement, coarsest level first and per level q, qc, qb, f, while
    // the budget lasts; the rest stays in the global pools
    int maxlds = 0;
    HIPCK(hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock,
                                c->cfg.device));
    long long avail = maxlds - (long long)tail_static_lds();
    long long budget = c->tail_lds_cap < 0 ? avail : std::min<long long>(c->tail_lds_cap, avail);
    // if the runtime refuses the planned dynamic LDS, plan again within the
    // default 64 KiB, then without LDS
    for (long long b : {budget, std::min<long long>(budget, 65536), 0LL}) {
      if (plan_tail(c, b, s)) break;
    }
  }
  c->all_int.clear();
  c->all_ext.clear();
  c->src_int.clear();
  c->src_cells = 0;
  c->all_int_cells = c->all_ext_cells = 0;
  for (const DPatch &P : c->dp)
    c->all_ext_cells += (long long)(P.mx + 2) * (P.my + 2) * (c->nd() == 3 ? P.mz + 2 : 1);
  for (int l = c->cfg.minlevel; l <= c->finest; l++) {
    LevelData &L = c->lev[l - minl];
    c->all_int_cells += L.cells;
    c->all_int.insert(c->all_int.end(), L.t_int.begin(), L.t_int.end());
    c->all_ext.insert(c->all_ext.end(), L.t_qlin.begin(), L.t_qlin.end());
    if (l >= 0) {
      c->src_int.insert(c->src_int.end(), L.t_int.begin(), L.t_int.end());
      c->src_cells += L.cells;
    }
  }
  c->d_all_int = dev_upload(c->all_int, s);
  c->d_all_ext = dev_upload(c->all_ext, s);
  c->d_src_int = dev_upload(c->src_int, s);
  c->partial_cap = (int)maxt;
  HIPCK(hipMalloc(&c->d_partial, sizeof(double) * 2 * maxt));   // two components (errortype 2)
  const int hmax = std::max(c->cfg.maxvcycles, 1) + 1;
  c->hmax = hmax;
  const int G = c->ngroups;
  HIPCK(hipMalloc(&c->d_hist, sizeof(double) * hmax * G));
  HIPCK(hipMalloc(&c->d_state, sizeof(int) * 4 * G));
  HIPCK(hipMemsetAsync(c->d_state, 0, sizeof(int) * 4 * G, s));
  HIPCK(hipHostMalloc(&c->h_state, sizeof(int) * 4 * G));
  for (int l = minl; l <= c->finest; l++) {
    LevelData &L = c->lev[l - minl];
    L.d_granges = dev_upload(L.granges, s);
    L.d_gcells = dev_upload(L.gcells, s);
  }
  HIPCK(hipStreamSynchronize(s));
}

// ---- level operations ------------------------------------------------------------
static void op_set_ghost(bsam_ctx *c, int level, int ipass, int prec = PREC_SOLUTION_FP64) {
  LevelData &L = c->L(level);
  if (ipass < 0 || ipass > 2) throw Error("ipass must be 0, 1 or 2");
  ProfScope ps(c, BSAM_PROF_GHOST,
               (prec == PREC_SOLUTION_FP64 ? 32.0 : 16.0) * L.gp[ipass].size());
  HIPCK(launch_ghost(prec, c->dev(), L.d_gp[ipass].view(), c->stream));
}
// What remains of SetGhost(2) on `level` after the last relaxation launch of
// a smoothing applied its final plan (a pending kRestOnly + 2).
static void op_set_ghost_rest(bsam_ctx *c, int level, int prec) {
  LevelData &L = c->L(level);
  const std::vector<GhostOp> &r = L.post[kPostFinal].rest;
  if (r.empty()) return;
  ProfScope ps(c, BSAM_PROF_GHOST, (prec == PREC_SOLUTION_FP64 ? 32.0 : 16.0) * r.size());
  HIPCK(launch_ghost(prec, c->dev(), L.d_prest[kPostFinal].view(), c->stream));
}
// Applies a pending SetGhost of `level` now: whole, or its rest.
static void op_apply_pending(bsam_ctx *c, int level, int pending, int prec) {
  if (pending >= kRestOnly) op_set_ghost_rest(c, level, prec);
  else if (pending >= 0) op_set_ghost(c, level, pending, prec);
}
// RelaxPatch of colour `color` over the level (with coupled_only, over the
// patches that are not isolated).
static void op_relax(bsam_ctx *c, int level, int color, int prec = PREC_SOLUTION_FP64,
                     bool coupled_only = false) {
  LevelData &L = c->L(level);
  const Tile *tiles = coupled_only ? L.d_relax : L.d_int;
  const int nt = (int)(coupled_only ? L.t_relax.size() : L.t_int.size());
  ProfScope ps(c, BSAM_PROF_RELAX, (prec == PREC_SOLUTION_FP64 ? 24.0 : 12.0) *
                                       (double)(coupled_only ? L.relax_cells : L.cells));
  if (c->relax_lds)   // the staged form of the same sweep, with no ghost updates
    HIPCK(launch_relax_fused(c->nd(), prec, c->dev(), tiles, nt, color, GhostList{},
                             GhostList{}, true, c->stream));
  else
    HIPCK(launch_relax(c->nd(), prec, c->dev(), tiles, nt, color, c->stream));
}
static LevelData &child_level(bsam_ctx *c, int level) {
  if (level <= c->cfg.minlevel) throw Error("level has no parent level");
  return c->L(level);
}
// RestrictSolution of `level`; a pending SetGhost(level, ghost) (ghost >= 0)
// runs in the same launch when its reads allow, else just before.
static void op_restrict(bsam_ctx *c, int level, int prec = PREC_SOLUTION_FP64, int ghost = -1) {
  LevelData &L = child_level(c, level);
  GhostList gops{};
  if (ghost >= kRestOnly) {
    op_set_ghost_rest(c, level, prec);
    ghost = -1;
  }
  if (ghost >= 0) {
    if (c->fuse_ghost && L.down_ok[ghost]) {
      gops = L.d_gp[ghost].view();
    } else {
      op_set_ghost(c, level, ghost, prec);
    }
  }
  ProfScope ps(c, BSAM_PROF_RESTRICT, 24.0 * L.cells);
  HIPCK(launch_restrict(c->nd(), prec, c->dev(), L.d_coarse, (int)L.t_coarse.size(), gops,
                        c->stream));
}
static void op_coarse_ghost(bsam_ctx *c, int level, int prec = PREC_SOLUTION_FP64) {
  LevelData &L = child_level(c, level);
  ProfScope ps(c, BSAM_PROF_COARSE_GHOST, 0.0);
  HIPCK(launch_coarse_ghost(c->nd(), prec, c->dev(), L.d_cext, (int)L.t_cext.size(), c->stream));
}
static void op_coarse_load(bsam_ctx *c, int level, int prec = PREC_SOLUTION_FP64) {
  LevelData &L = child_level(c, level);
  ProfScope ps(c, BSAM_PROF_COARSE_LOAD, 40.0 * L.cells);
  HIPCK(launch_coarse_load(c->nd(), prec, c->dev(), L.d_coarse, (int)L.t_coarse.size(), c->stream));
}
// The parent's SetGhost(parent_ghost), still pending, rides inside the launch
// when a plan for it exists (its SetGhost(0) in the coarse ghost + loading
// launch, its SetGhost(2) in the correction launch); otherwise it is launched
// first.
static const LevelData::DRide *parent_ride(bsam_ctx *c, LevelData &L, int parent_ghost,
                                           int level, int prec, bool correct,
                                           bool loop = false) {
  if (parent_ghost < 0) return nullptr;
  if (parent_ghost >= kRestOnly) {
    op_set_ghost_rest(c, level - 1, prec);
    return nullptr;
  }
  if (loop && !correct && parent_ghost == 0 && prec == PREC_SOLUTION_FP64 && c->fuse_ghost &&
      L.ride_cgl_loop.ok)
    return &L.d_ride_cgl_loop;
  const RideGhost &r = correct ? L.ride_corr : L.ride_cgl;
  if (c->fuse_ghost && r.ok && parent_ghost == (correct ? 2 : 0))
    return correct ? &L.d_ride_corr : &L.d_ride_cgl;
  op_set_ghost(c, level - 1, parent_ghost, prec);
  return nullptr;
}
// With `loop`, the call comes from the FP64 multigrid loop's V-cycle, whose
// FillDown has applied the parent's SetGhost(0) since the parent last changed
// outside this descent (ride_cgl_loop).
static void op_coarse_ghost_load(bsam_ctx *c, int level, int prec = PREC_SOLUTION_FP64,
                                 int parent_ghost = -1, bool loop = false) {
  LevelData &L = child_level(c, level);
  const LevelData::DRide *r = parent_ride(c, L, parent_ghost, level, prec, false, loop);
  ProfScope ps(c, BSAM_PROF_COARSE_LOAD, 40.0 * L.cells);
  HIPCK(launch_coarse_ghost_load(c->nd(), prec, c->dev(), L.d_cgl, (int)L.t_cgl.size(),
                                 L.n_cgl_res,
                                 r ? r->rest.view() : GhostList{},
                                 r ? r->ops.view() : GhostList{}, c->stream));
}
static void op_correct(bsam_ctx *c, int level, int prec = PREC_SOLUTION_FP64,
                       int parent_ghost = -1) {
  LevelData &L = child_level(c, level);
  const LevelData::DRide *r = parent_ride(c, L, parent_ghost, level, prec, true);
  ProfScope ps(c, BSAM_PROF_CORRECT, 40.0 * L.cells);
  HIPCK(launch_correct(c->nd(), prec, c->dev(), L.d_blk, (int)L.t_blk.size(),
                       r ? r->rest.view() : GhostList{}, r ? r->ops.view() : GhostList{},
                       c->stream));
}
// FillDown, afasroutines.f90:223-239. For q, each SetGhost(level - 1, 0) runs
// inside the next level's fill-down launch when its reads allow; `ghost` >= 0
// is a SetGhost(finest, ghost) still pending from the caller, handled the
// same way. `bottom` limits the levels filled to those >= bottom (the whole
// chain by default); the part below can be completed later by a call with
// ghost = -1 and top = bottom.
static void op_fill_down(bsam_ctx *c, int field, int ghost = -1, int top = INT_MAX,
                         int bottom = INT_MIN) {
  if (field != BSAM_FIELD_F && field != BSAM_FIELD_Q) throw Error("fill_down: field must be F or Q");
  int low = std::max(c->cfg.minlevel + 1, c->tail_top + 1);
  low = std::max(low, bottom + 1);
  int carry = ghost;   // SetGhost program of `level` not yet applied
  int level = std::min(top, c->finest);
  for (; level >= low; level--) {
    LevelData &L = c->L(level);
    GhostList gops{};
    if (carry >= kRestOnly) {
      op_set_ghost_rest(c, level, PREC_SOLUTION_FP64);
      carry = -1;
    }
    if (carry >= 0) {
      if (c->fuse_ghost && L.down_ok[carry]) {
        gops = L.d_gp[carry].view();
      } else {
        op_set_ghost(c, level, carry);
      }
    }
    {
      ProfScope ps(c, BSAM_PROF_FILLDOWN, 20.0 * L.cells);
      HIPCK(launch_filldown(c->nd(), c->dev(), L.d_coarse, (int)L.t_coarse.size(),
                            field == BSAM_FIELD_F ? 1 : 2, gops, c->stream));
    }
    carry = field == BSAM_FIELD_Q ? 0 : -1;
  }
  op_apply_pending(c, level, carry, PREC_SOLUTION_FP64);
  if (c->ntail() > 1 && bottom < c->tail_top && top >= c->tail_top) {
    ProfScope ps(c, BSAM_PROF_TAIL, 0.0);
    HIPCK(launch_tail_filldown(c->nd(), c->dev(), c->d_tail, c->ntail(),
                               field == BSAM_FIELD_F ? 1 : 2, c->tail_threads, c->stream));
  }
}
// GetSourceFunction / UpdateSourceFunction on levels finest..0: independent
// per cell, so one launch over the interior tiles of all those levels.
static void op_source(bsam_ctx *c, int full) {
  ProfScope ps(c, BSAM_PROF_SOURCE, 32.0 * c->src_cells);
  HIPCK(launch_source(c->nd(), c->dev(), c->d_src_int, (int)c->src_int.size(), full,
                      c->stream));
}
// The smoothing passes of AFASVcycle (afasroutines.f90:493-497, 512-516): per
// pass, relax colour 1, SetGhost(1), relax colour 2, SetGhost(2). With
// ghost0_first, a SetGhost(level, 0) precedes them. Each SetGhost is applied
// inside the following relaxation launch where the level's fused plan allows
// (its updates that read cells other tiles relax in a launch just before).
// With defer_last the final SetGhost is not launched but returned (its ipass)
// for the caller to apply. With final_post (a caller that would launch it on
// its own rather than ride it in another launch), the last launch applies the
// final SetGhost(2) where its plan allows and kRestOnly + 2 is returned.
static int level_relax(bsam_ctx *c, int level, int prec = PREC_SOLUTION_FP64,
                       bool ghost0_first = false, bool defer_last = false,
                       bool final_post = false) {
  LevelData &L = c->L(level);
  int pending = ghost0_first ? 0 : -1;   // SetGhost program not yet applied
  int moved = -1;   // post plan the previous launch ran: its rest replaces this pre
  const double sz = prec == PREC_SOLUTION_FP64 ? 1.0 : 0.5;
  // The isolated patches' whole smoothing, SetGhost(2) after it included, in
  // one launch; the launches below relax the other patches, whose plans
  // leave the isolated patches' updates out (they need the fused plans, so
  // without fusion every patch takes the per-colour launches).
  const bool iso = !L.iso.empty() && c->fuse_ghost;
  // with other patches to relax, the isolated smoothing runs on the side
  // stream beside their launches (not while it is being timed)
  const bool side = iso && c->iso_side && c->side && !L.t_relax.empty() &&
                    !(c->prof_mask & (1u << BSAM_PROF_SMOOTH_ISO));
  if (iso) {
    const int np = c->cfg.nsmoothingpasses;
    const long long ups = (ghost0_first ? L.iso_updates[0] : 0) + (np - 1) * L.iso_updates[2] +
                          np * L.iso_updates[1] + L.iso_updates[2];
    if (side) {
      HIPCK(hipEventRecord(c->ev_fork, c->stream));
      HIPCK(hipStreamWaitEvent(c->side, c->ev_fork, 0));
    }
    ProfScope ps(c, BSAM_PROF_SMOOTH_ISO, sz * (48.0 * np * L.iso_cells + 32.0 * ups));
    HIPCK(launch_smooth_iso(c->nd(), prec, c->dev(), L.d_iso, (int)L.iso.size(),
                            L.d_iso_stage, L.d_iso_rec, L.d_iso_srec, np, ghost0_first ? 1 : 0, 1,
                            prec == PREC_SOLUTION_FP64 ? L.iso_lds : L.iso_lds / 2,
                            side ? c->side : c->stream));
    if (L.t_relax.empty()) return -1;
  }
  const Tile *tiles = iso ? L.d_relax : L.d_int;
  const int nt = (int)(iso ? L.t_relax.size() : L.t_int.size());
  const double relax_bytes = 24.0 * sz * (double)(iso ? L.relax_cells : L.cells);
  for (int pass = 1; pass <= c->cfg.nsmoothingpasses; pass++)
    for (int color = 1; color <= 2; color++) {
      const bool last = pass == c->cfg.nsmoothingpasses && color == 2;
      int f = pending >= 0 ? fused_index(pending, color) : -1;
      const bool fused = f >= 0 && c->fuse_ghost && L.fused[f].ok;
      // this launch ends with the next launch's pre updates where a plan
      // exists (the next launch is SetGhost(color) -> the other colour)
      int k = (!last && c->fuse_ghost && c->post_ghost) ? post_key(pending, color) : -1;
      if (last && final_post && c->fuse_ghost && c->post_ghost && pending == 1 && color == 2)
        k = kPostFinal;
      if (k >= 0 && !L.post[k].ok) k = -1;
      const GhostList pop = k >= 0 ? L.d_pops[k].view() : GhostList{};
      if (fused) {
        const std::vector<GhostOp> &pre = moved >= 0 ? L.post[moved].rest : L.fused[f].pre;
        if (!pre.empty()) {
          ProfScope pg(c, BSAM_PROF_GHOST,
                       (prec == PREC_SOLUTION_FP64 ? 32.0 : 16.0) * pre.size());
          HIPCK(launch_ghost(prec, c->dev(),
                             (moved >= 0 ? L.d_prest[moved] : L.d_fpre[f]).view(), c->stream));
        }
        // the sweep's bytes plus 32 B per ghost update applied in the launch
        // (16 B read, 16 B written: the stand-alone ghost launch's figure)
        const double nops = (double)L.fused[f].ops.size() + (k >= 0 ? L.post[k].ops.size() : 0);
        ProfScope ps(c, BSAM_PROF_RELAX_FUSED,
                     relax_bytes + (prec == PREC_SOLUTION_FP64 ? 32.0 : 16.0) * nops);
        HIPCK(launch_relax_fused(c->nd(), prec, c->dev(), tiles, nt, color,
                                 L.d_fops[f].view(), pop, c->relax_lds != 0, c->stream));
      } else {
        // with isolated patches only the first launch of pre-smoothing comes
        // here (no SetGhost pending): every transition after it is fused
        if (iso && pending >= 0) throw Error("isolated patches need every transition fused");
        if (pending >= 0) op_set_ghost(c, level, pending, prec);
        if (k >= 0) {
          ProfScope ps(c, BSAM_PROF_RELAX_FUSED,
                       relax_bytes +
                           (prec == PREC_SOLUTION_FP64 ? 32.0 : 16.0) * L.post[k].ops.size());
          HIPCK(launch_relax_fused(c->nd(), prec, c->dev(), tiles, nt, color, GhostList{}, pop,
                                   c->relax_lds != 0, c->stream));
        } else {
          op_relax(c, level, color, prec, iso);
        }
      }
      moved = k;
      pending = color;
    }
  if (side) {
    // the side stream joins back before anything after the smoothing (the
    // SetGhost(2) below or the caller's may rewrite isolated patches' ghosts)
    HIPCK(hipEventRecord(c->ev_join, c->side));
    HIPCK(hipStreamWaitEvent(c->stream, c->ev_join, 0));
  }
  if (moved == kPostFinal) pending = kRestOnly + 2;   // SetGhost(2) all but its rest
  if (defer_last) return pending;
  op_apply_pending(c, level, pending, prec);
  return -1;
}
// AFASVcycle, afasroutines.f90:488-521. With defer_final, the last
// SetGhost of the level's post-smoothing is returned (its ipass, kRestOnly +
// its ipass when final_post let the last launch apply most of it, or -1)
// instead of launched.
static int op_vcycle(bsam_ctx *c, int level, int prec = PREC_SOLUTION_FP64,
                     bool defer_final = false, bool final_post = false, bool loop = false) {
  if (level <= c->tail_top) {
    ProfScope ps(c, BSAM_PROF_TAIL, 0.0);
    int nlev = level - c->cfg.minlevel + 1, set = prec == PREC_SOLUTION_FP64 ? 0 : 1;
    HIPCK(launch_tail_vcycle(c->nd(), prec, c->dev(), c->d_tail, nlev, c->cfg.nsmoothingpasses,
                             c->tail_threads, c->tail_lds[set], c->d_tcin[set],
                             c->tcin_n[set][nlev - 1], c->d_tcout[set],
                             c->tcout_n[set][nlev - 1], c->tail_qpo, c->stream));
    return -1;
  }
  if (level == c->cfg.minlevel)
    return level_relax(c, level, prec, false, defer_final, defer_final && final_post);
  // the restriction runs the pending SetGhost(2) inside its launch when its
  // reads allow; otherwise the last relaxation launch takes most of it
  LevelData &L = c->L(level);
  int pend = level_relax(c, level, prec, false, true, !(c->fuse_ghost && L.down_ok[2]));
  op_restrict(c, level, prec, pend);
  // SetGhost(level - 1, 0) rides inside the coarse ghost + loading launch and
  // the parent's last SetGhost of post-smoothing inside the correction launch
  // (or each is launched just before, without a plan)
  op_coarse_ghost_load(c, level, prec, 0, loop);
  int pp = op_vcycle(c, level - 1, prec, true, !(c->fuse_ghost && L.ride_corr.ok), loop);
  op_correct(c, level, prec, pp);
  return level_relax(c, level, prec, true, defer_final, defer_final && final_post);
}
static void reset_state(bsam_ctx *c) {
  HIPCK(hipMemsetAsync(c->d_state, 0, sizeof(int) * 4 * c->ngroups, c->stream));
}
// queue ErrorAFAS(level): residual norm into d_hist[state[0]]
static void queue_error(bsam_ctx *c, int level, double tol) {
  LevelData &L = c->L(level);
  int nt = (int)L.t_int.size();
  {
    ProfScope ps(c, BSAM_PROF_RESNORM, 32.0 * L.cells);
    HIPCK(launch_resnorm(c->nd(), c->dev(), L.d_int, nt, c->d_partial, 0, c->comps(),
                         c->stream));
  }
  // L2Error counts nrvars values per cell, L2ComponentErrors one
  double count = (double)(c->comps() == 2 ? L.cells : c->cfg.nrvars * L.cells);
  if (c->ngroups > 1 && tol >= 0.0)   // a solve: each group its own test
    HIPCK(launch_finalize(c->d_partial, nt, count, tol, c->d_hist, c->d_state, c->comps(),
                          c->stream, c->ngroups, L.d_granges, L.d_gcells, c->hmax));
  else
    HIPCK(launch_finalize(c->d_partial, nt, count, tol, c->d_hist, c->d_state, c->comps(),
                          c->stream));
}

// Captures `body` on the context's stream into *exec (once per hierarchy).
template <class F> static void capture(bsam_ctx *c, hipGraph_t *g, hipGraphExec_t *exec, F body) {
  if (*exec) return;
  DeviceLock lock(g_device_lock);
  HIPCK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  try {
    body();
  } catch (...) {
    hipGraph_t tmp = nullptr;
    (void)hipStreamEndCapture(c->stream, &tmp);
    if (tmp) (void)hipGraphDestroy(tmp);
    throw;
  }
  HIPCK(hipStreamEndCapture(c->stream, g));
  HIPCK(hipGraphInstantiate(exec, *g, nullptr, nullptr, 0));
}

// Runs `body` (one iteration ending in queue_error) until the residual test
// passes or maxvcycles iterations are done; returns the iteration count.
template <class F>
static int iterate(bsam_ctx *c, hipGraph_t *g, hipGraphExec_t *exec, F body, int *done_out) {
  bool graph = c->use_graph && c->prof_mask == 0;
  if (graph) capture(c, g, exec, body);
  const int G = c->ngroups;
  for (int g = 0; g < 4 * G; g++) c->h_state[g] = 0;
  // Iterations are queued back to back, as many as the previous solve took
  // less one, before the first read-back of the stop state, then one at a
  // time: the kernels leave a solve that has met its test untouched, so
  // queued iterations past convergence change nothing (their launches return
  // at once) and the result equals stopping at the test.
  int it = 0;
  while (it < c->cfg.maxvcycles) {
    const int k = it == 0 ? std::max(1, std::min(c->last_iters - 1, c->cfg.maxvcycles)) : 1;
    for (int r = 0; r < k; r++) {
      if (graph) HIPCK(hipGraphLaunch(*exec, c->stream));
      else body();
    }
    it += k;
    HIPCK(hipMemcpyAsync(c->h_state, c->d_state, sizeof(int) * 4 * G, hipMemcpyDeviceToHost,
                         c->stream));
    HIPCK(hipStreamSynchronize(c->stream));
    // a group stops once its residual is below the tolerance, or is not
    // finite (every later one is not either); the solve ends when all have
    bool all = true;
    for (int g = 0; g < G; g++) all = all && c->h_state[4 * g + 3];
    if (all) break;
  }
  c->group_state.assign(c->h_state, c->h_state + 4 * G);
  // the stopped flags only hold within a solve: entry points called after it
  // act on every group
  reset_state(c);
  bool conv = true, nonfinite = false;
  int n = 0;
  for (int g = 0; g < G; g++) {
    conv = conv && c->h_state[4 * g + 1];
    nonfinite = nonfinite || c->h_state[4 * g + 2];
    n = std::max(n, c->h_state[4 * g]);   // histories recorded: iterations run
  }
  c->last_iters = n;
  c->last_status = conv ? BSAM_SOLVE_CONVERGED
                        : (nonfinite ? BSAM_SOLVE_NONFINITE : BSAM_SOLVE_MAXCYCLES);
  *done_out = conv;
  return n;
}

int bsam_solve_status(const bsam_ctx *c) { return c ? c->last_status : -1; }

static void report(bsam_ctx *c, int n, int done, int *ncycles, double *hist, int maxhist,
                   int *converged) {
  if (ncycles) *ncycles = n;
  if (converged) *converged = done;
  if (hist && maxhist > 0) {
    int m = std::min(n, maxhist);
    HIPCK(hipMemcpyAsync(hist, c->d_hist, sizeof(double) * m, hipMemcpyDeviceToHost,
                         c->stream));
    HIPCK(hipStreamSynchronize(c->stream));
  }
}

// MultigridIterations (afasroutines.f90:38-106) in FP64.
static void multigrid_iterations(bsam_ctx *c, int *ncycles, double *hist, int maxhist,
                                 int *converged) {
  op_fill_down(c, BSAM_FIELD_Q);
  op_source(c, 1);
  op_fill_down(c, BSAM_FIELD_F);
  // The loop's source restore and FillDown(f) (afasroutines.f90:84-90) rebuild
  // f on every level from ftmp alone: UpdateSourceFunction adds 0 for this
  // problem (f = ftmp + 0 on levels >= 0) and FillDown(f) then overwrites
  // every covered cell by restriction. Their result is the same each
  // iteration, so it is computed once here and saved (the f an iteration
  // rewrites: the levels below the finest, f_restore). Inside the loop the
  // restore is dead: an iteration writes f only at parent cells under a
  // child (CoarseLoadingFunction) and on the chain below the root (the tail),
  // and the next iteration's descent rewrites each of those cells before
  // reading it (the coarse loading into level l runs before level l's
  // smoothing and before level l's own coarse loading reads f_l), while
  // ErrorAFAS reads the finest level's f, which nothing writes. The restore
  // therefore runs once, after the loop, so f ends as the reference's loop
  // leaves it.
  for (const auto &r : c->f_restore) {
    ProfScope ps(c, BSAM_PROF_COPY, 16.0 * (r.second - r.first));
    HIPCK(launch_copy(c->f_saved + r.first, c->f + r.first, r.second - r.first, c->stream));
  }
  reset_state(c);
  // one iteration of the vcycleloop (afasroutines.f90:69-104): the V-cycle,
  // FillDown(q), (source restore + FillDown(f): after the loop), the
  // residual norm
  int done = 0;
  //
  // Inside the loop FillDown(q) stops at level 0: every level below the root
  // is one patch that the next V-cycle overwrites whole (RestrictSolution
  // from the level above, then SetGhost) before anything reads it, and the
  // residual is taken on the finest level. The part below the root runs once
  // after the loop, so the fields end as the reference's do.
  int n = iterate(c, &c->graph, &c->graph_exec, [&]() {
    // the fill-down runs the pending SetGhost(2) inside its launch when its
    // reads allow; otherwise the last relaxation launch takes most of it
    int pend = op_vcycle(c, c->finest, PREC_SOLUTION_FP64, true,
                         !(c->fuse_ghost && c->L(c->finest).down_ok[2]), true);
    op_fill_down(c, BSAM_FIELD_Q, pend, INT_MAX, 0);
    queue_error(c, c->finest, c->cfg.qerrortol);
  }, &done);
  op_fill_down(c, BSAM_FIELD_Q, -1, 0, INT_MIN);
  for (const auto &r : c->f_restore) {
    ProfScope ps(c, BSAM_PROF_COPY, 16.0 * (r.second - r.first));
    HIPCK(launch_copy(c->f + r.first, c->f_saved + r.first, r.second - r.first, c->stream));
  }
  report(c, n, done, ncycles, hist, maxhist, converged);
}

// Mixed precision: FP64 defect correction around FP32 FAS V-cycles. Each
// step computes the FP64 residual r = f - A(q) on every level, solves
// A(qb + e) - A(qb) = r for the correction e with inner FP32 V-cycles (the
// same level schedule and ghost programs as the FP64 solve), adds e to q in
// FP64 and restores FP64 consistency (FillDown, ghost fill) before the FP64
// residual test, which is the FP64 path's: ErrorAFAS(finest) < qerrortol.
// The per-level set-up and correction run as one launch over the tiles of
// all levels.
static void mixed_iterations(bsam_ctx *c, int *ncycles, double *hist, int maxhist,
                             int *converged) {
  op_fill_down(c, BSAM_FIELD_Q);
  op_source(c, 1);
  op_fill_down(c, BSAM_FIELD_F);
  reset_state(c);
  int done = 0;
  int n = iterate(c, &c->graph_mixed, &c->graph_mixed_exec, [&]() {
    {
      ProfScope ps(c, BSAM_PROF_DEFECT, 52.0 * c->all_ext_cells);
      HIPCK(launch_defect_setup(c->nd(), c->dev(), c->d_all_ext, (int)c->all_ext.size(),
                                c->stream));
    }
    // the last inner cycle's final ghost fill of the correction is never read
    // (the correction is applied on interior cells, then restarts at zero)
    for (int k = 0; k < c->inner_cycles; k++) {
      const bool more = k + 1 < c->inner_cycles;
      int pend = op_vcycle(c, c->finest, PREC_DEFECT_FP32, true, more);
      if (more) op_apply_pending(c, c->finest, pend, PREC_DEFECT_FP32);
    }
    {
      ProfScope ps(c, BSAM_PROF_DEFECT, 40.0 * c->all_int_cells);
      HIPCK(launch_apply_correction(c->nd(), c->dev(), c->d_all_int, (int)c->all_int.size(),
                                    c->stream));
    }
    op_fill_down(c, BSAM_FIELD_Q);
    for (int l = 0; l <= c->finest; l++) op_set_ghost(c, l, 0);
    queue_error(c, c->finest, c->cfg.qerrortol);
  }, &done);
  report(c, n, done, ncycles, hist, maxhist, converged);
}

static void solve(bsam_ctx *c, int *ncycles, double *hist, int maxhist, int *converged) {
  if (c->cfg.precision == BSAM_PREC_MIXED)
    mixed_iterations(c, ncycles, hist, maxhist, converged);
  else
    multigrid_iterations(c, ncycles, hist, maxhist, converged);
}

// ---- C ABI ---------------------------------------------------------------------------
extern "C" {

bsam_ctx *bsam_ctx_create(const bsam_cfg *cfg) {
  if (!cfg) {
    g_err = "null config";
    return nullptr;
  }
  if (cfg->ndims != 2 && cfg->ndims != 3) {
    g_err = "ndims must be 2 or 3";
    return nullptr;
  }
  if (cfg->nrvars != 2) {
    g_err = "the Cahn-Hilliard engine needs nrvars = 2";
    return nullptr;
  }
  if (cfg->mbc != 1) {
    g_err = "only mbc = 1 is supported";
    return nullptr;
  }
  if (cfg->minlevel > 0) {
    g_err = "minlevel must be <= 0";
    return nullptr;
  }
  if (cfg->errortype != 1 && cfg->errortype != 2) {
    // the reference stops on any other value (lib/afasroutines.f90:143-146)
    g_err = "ErrorAFAS: only errortype = 1, 2 are supported";
    return nullptr;
  }
  if (cfg->precision != BSAM_PREC_FP64 && cfg->precision != BSAM_PREC_MIXED) {
    g_err = "unknown precision";
    return nullptr;
  }
  bsam_ctx *c = new (std::nothrow) bsam_ctx();
  if (!c) {
    g_err = "out of memory";
    return nullptr;
  }
  c->cfg = *cfg;
  c->host_only = cfg->device < 0;
  if (!c->host_only) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || cfg->device >= ndev) {
      g_err = "no such HIP device";
      delete c;
      return nullptr;
    }
    DeviceLock lock(g_device_lock);
    if (hipSetDevice(cfg->device) != hipSuccess ||
        hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
      g_err = "cannot create a stream on the device";
      delete c;
      return nullptr;
    }
  }
  return c;
}

int bsam_ctx_destroy(bsam_ctx *c) {
  if (!c) return 0;
  if (!c->host_only) {
    (void)hipSetDevice(c->cfg.device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    DeviceLock lock(g_device_lock);
    c->free_device();
    if (c->side) (void)hipStreamSynchronize(c->side);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    if (c->side) (void)hipStreamDestroy(c->side);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    c->stream = c->side = nullptr;
  }
  delete c;
  return 0;
}

const char *bsam_last_error(const bsam_ctx *c) { return c ? c->err.c_str() : g_err.c_str(); }

int bsam_hierarchy_begin(bsam_ctx *c) {
  API_BEGIN(c)
  if (!c->host_only) HIPCK(hipStreamSynchronize(c->stream));
  c->staged.clear();
  c->staged_group.clear();
  c->staged_eps2.clear();
  c->staged_ngroups = 1;
  c->committed = false;
  c->building = true;
  API_END(c)
}

int bsam_set_groups(bsam_ctx *c, int ngroups, const int *group, const double *eps2) {
  API_BEGIN(c)
  if (!c->building) throw Error("set_groups: call it between bsam_hierarchy_begin and commit");
  if (ngroups < 1 || ngroups > 32767) throw Error("set_groups: ngroups must be 1..32767");
  if (!group) throw Error("set_groups: null group array");
  c->staged_group.assign(group, group + c->staged.size());
  for (int g : c->staged_group)
    if (g < 0 || g >= ngroups) throw Error("set_groups: group index out of range");
  c->staged_eps2.clear();
  if (eps2) {
    c->staged_eps2.assign(eps2, eps2 + ngroups);
    for (double e : c->staged_eps2)
      if (!std::isfinite(e) || e < 0.0) throw Error("set_groups: eps^2 must be finite and >= 0");
  }
  c->staged_ngroups = ngroups;
  API_END(c)
}

int bsam_add_patch(bsam_ctx *c, int id, int level, int parent, const int *mx,
                   const int *mbounds, const int *mglobal, const int *mthbc, double dx,
                   const double *xlower) {
  API_BEGIN(c)
  if (!c->building) throw Error("call bsam_hierarchy_begin first");
  if (!mx || !mbounds || !mglobal || !mthbc) throw Error("null geometry");
  HPatch h;
  int nd = c->nd();
  h.id = id;
  h.level = level;
  h.parent = parent;
  for (int d = 0; d < 3; d++) {
    h.mx[d] = d < nd ? mx[d] : 1;
    h.mb[d][0] = d < nd ? mbounds[2 * d] : 1;
    h.mb[d][1] = d < nd ? mbounds[2 * d + 1] : 1;
    h.mg[d][0] = d < nd ? mglobal[2 * d] : 1;
    h.mg[d][1] = d < nd ? mglobal[2 * d + 1] : 1;
    h.cmx[d] = d < nd ? (h.mx[d] % 2 == 1 ? 1 : h.mx[d] / 2) : 1;
    h.xlower[d] = (xlower && d < nd) ? xlower[d] : 0.0;
  }
  for (int fc = 0; fc < 6; fc++) h.mthbc[fc] = fc < 2 * nd ? mthbc[fc] : 1;
  h.dx = dx;
  c->staged.push_back(h);
  API_END(c)
}

int bsam_hierarchy_commit(bsam_ctx *c) {
  API_BEGIN(c)
  if (!c->building) throw Error("call bsam_hierarchy_begin first");
  if (!c->host_only) HIPCK(hipSetDevice(c->cfg.device));
  // groups (one unless bsam_set_groups staged several, for these patches)
  const int G = c->staged_ngroups;
  if (G > 1 && c->staged_group.size() != c->staged.size())
    throw Error("set_groups was called before every patch was added");
  for (size_t i = 0; i < c->staged.size(); i++)
    c->staged[i].group = G > 1 ? c->staged_group[i] : 0;
  std::vector<double> eps2(G, c->cfg.eps2);
  if (!c->staged_eps2.empty()) eps2 = c->staged_eps2;
  if (c->had_commit && c->planned_fuse == (c->fuse_ghost != 0) && G == c->ngroups &&
      eps2 == c->group_eps2 && same_geometry(c->hp, c->staged)) {
    // The regrid produced the hierarchy already committed (the reference
    // regrids every step, mostly to the same patches): keep its plans,
    // tables, pools and graphs, and zero the pools as a fresh commit does.
    for (size_t i = 0; i < c->hp.size(); i++) c->hp[i].id = c->staged[i].id;
    if (!c->host_only) zero_pools(c);
    c->commits_reused++;
  } else {
    c->had_commit = false;
    c->free_device();
    c->ngroups = G;
    c->group_eps2 = eps2;
    c->hp = std::move(c->staged);
    c->dp.clear();
    c->lev.clear();
    commit(c);
    c->had_commit = true;
    c->planned_fuse = c->fuse_ghost != 0;
  }
  c->staged.clear();
  c->building = false;
  c->committed = true;
  API_END(c)
}

int bsam_num_patches(const bsam_ctx *c) { return c ? (int)c->hp.size() : -1; }

// Device bytes of a committed hierarchy's field pools (q, qold, qc, f, rf,
// ftmp, the saved f, qrte and the error flags, plus the FP32 pools in mixed
// precision), excluding small tables.
long long bsam_footprint_bytes(const bsam_ctx *c) {
  if (!c || !c->committed) return -1;
  long long b = 8 * (2 * c->nq + c->nc + 4 * c->nf + c->nr) + 4 * c->nfl;
  if (c->cfg.precision == BSAM_PREC_MIXED) b += 4 * (2 * c->nq + c->nc + c->nf);
  return b;
}

// A patch's field: its pool and offset, its length in the reference layout,
// and its plane shape (rows of rx cells, ry rows per plane, `planes` planes
// per variable, nvar variables `vstride` apart on the device; rx = 0 for
// arrays the device keeps in the reference layout). Device planes take
// split_of(rx, ry).pl cells.
struct FieldShape {
  long long rx = 0, ry = 0, planes = 0, nvar = 1, vstride = 0;
};
static void field_ptr(bsam_ctx *c, int patch, int field, double **base, long long *off,
                      long long *n, FieldShape *shape = nullptr) {
  if (patch < 0 || patch >= (int)c->hp.size()) throw Error("patch index out of range");
  const DPatch &D = c->dp[patch];
  const int nv = c->cfg.nrvars, nd = c->nd();
  FieldShape fs;
  switch (field) {
    case BSAM_FIELD_Q:
    case BSAM_FIELD_QOLD:
      *base = field == BSAM_FIELD_Q ? c->q : c->qold;
      *off = D.q;
      fs = FieldShape{D.px, D.py, nd == 3 ? D.mz + 2 : 1, nv, D.qvs};
      break;
    case BSAM_FIELD_QC:
      *base = c->qc;
      *off = D.qc;
      fs = FieldShape{D.cmx + 2, D.cmy + 2, nd == 3 ? D.cmz + 2 : 1, nv, D.cvs};
      break;
    case BSAM_FIELD_F:
    case BSAM_FIELD_FTMP:
    case BSAM_FIELD_RF:
      *base = field == BSAM_FIELD_F ? c->f : field == BSAM_FIELD_FTMP ? c->ftmp : c->rf;
      *off = D.f;
      fs = FieldShape{D.mx, D.my, D.mz, nv, D.fvs};
      break;
    case BSAM_FIELD_QRTE: *base = c->qrte; *off = D.qr; *n = D.rvs * nv; break;
    default: throw Error("unknown field");
  }
  if (fs.rx > 0) *n = fs.rx * fs.ry * fs.planes * fs.nvar;
  if (shape) *shape = fs;
}

long bsam_field_len(const bsam_ctx *cc, int patch, int field) {
  bsam_ctx *c = const_cast<bsam_ctx *>(cc);
  if (!c || !c->committed) return -1;
  try {
    double *b;
    long long off, n;
    field_ptr(c, patch, field, &b, &off, &n);
    return (long)n;
  } catch (...) {
    return -1;
  }
}

int bsam_upload(bsam_ctx *c, int patch, int field, const double *host, long n) {
  API_BEGIN(c)
  c->need_device();
  double *b;
  long long off, len;
  FieldShape fs;
  field_ptr(c, patch, field, &b, &off, &len, &fs);
  if (n != len) throw Error("upload: size mismatch");
  if (!host && n > 0) throw Error("upload: null host array");
  if (fs.rx > 0) {   // reference planes into the device's split planes, per variable
    const long long dn = split_of(fs.rx, fs.ry).pl * fs.planes, hn = fs.rx * fs.ry * fs.planes;
    std::vector<double> tmp((size_t)(dn * fs.nvar), 0.0);
    for (long long v = 0; v < fs.nvar; v++) {
      to_device(host + v * hn, tmp.data() + v * dn, fs.rx, fs.ry, fs.planes);
      HIPCK(hipMemcpyAsync(b + off + v * fs.vstride, tmp.data() + v * dn, sizeof(double) * dn,
                           hipMemcpyHostToDevice, c->stream));
    }
    HIPCK(hipStreamSynchronize(c->stream));
  } else {
    HIPCK(hipMemcpyAsync(b + off, host, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    HIPCK(hipStreamSynchronize(c->stream));
  }
  API_END(c)
}

int bsam_download(bsam_ctx *c, int patch, int field, double *host, long n) {
  API_BEGIN(c)
  c->need_device();
  double *b;
  long long off, len;
  FieldShape fs;
  field_ptr(c, patch, field, &b, &off, &len, &fs);
  if (n != len) throw Error("download: size mismatch");
  if (!host && n > 0) throw Error("download: null host array");
  if (fs.rx > 0) {   // the device's split planes back into reference planes
    const long long dn = split_of(fs.rx, fs.ry).pl * fs.planes, hn = fs.rx * fs.ry * fs.planes;
    std::vector<double> tmp((size_t)(dn * fs.nvar));
    for (long long v = 0; v < fs.nvar; v++)
      HIPCK(hipMemcpyAsync(tmp.data() + v * dn, b + off + v * fs.vstride, sizeof(double) * dn,
                           hipMemcpyDeviceToHost, c->stream));
    HIPCK(hipStreamSynchronize(c->stream));
    for (long long v = 0; v < fs.nvar; v++)
      from_device(tmp.data() + v * dn, host + v * hn, fs.rx, fs.ry, fs.planes);
  } else {
    HIPCK(hipMemcpyAsync(host, b + off, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    HIPCK(hipStreamSynchronize(c->stream));
  }
  API_END(c)
}

int bsam_truncation_error(bsam_ctx *c, int level) {
  API_BEGIN(c)
  c->need_device();
  if (level <= c->cfg.minlevel || level > c->finest) throw Error("truncation_error: bad level");
  // EstimateLevelErrors' truncation error (bsamroutines.f90:971-981)
  op_restrict(c, level);
  op_set_ghost(c, level - 1, 0);
  LevelData &L = c->L(level);
  HIPCK(launch_coarse_ghost(c->nd(), PREC_SOLUTION_FP64, c->dev(), L.d_cext,
                            (int)L.t_cext.size(), c->stream));
  // CopyQToQold on level and level - 1: the patches of consecutive levels
  // occupy one contiguous range of the pool (patches are added by level)
  long long beg = LLONG_MAX, end = 0;
  for (int l = level - 1; l <= level; l++)
    for (int p : c->L(l).patches) {
      beg = std::min(beg, c->dp[p].q);
      end = std::max(end, c->dp[p].q + c->dp[p].qn);
    }
  for (int v = 0; v < c->cfg.nrvars; v++)   // each variable's range of the pool
    HIPCK(launch_copy(c->qold + beg + v * c->qstride, c->q + beg + v * c->qstride, end - beg,
                      c->stream));
  HIPCK(launch_trunc_error(c->nd(), c->dev(), L.d_coarse, (int)L.t_coarse.size(), c->stream));
  API_END(c)
}

int bsam_error_flags(bsam_ctx *c, int level, double qtolerance) {
  API_BEGIN(c)
  c->need_device();
  if (level < c->cfg.minlevel || level > c->finest) throw Error("error_flags: bad level");
  LevelData &L = c->L(level);
  // tol = qtolerance(level)/h/h (/h in 3D), as SetErrFlags2D/3D evaluate it
  const double h = c->hp[L.patches.front()].dx;
  double tol = qtolerance / h / h;
  if (c->nd() == 3) tol = tol / h;
  HIPCK(launch_error_flags(c->nd(), c->dev(), L.d_int, (int)L.t_int.size(), tol, c->stream));
  API_END(c)
}

int bsam_error_flags_user(bsam_ctx *c, int level, double qtolerance) {
  API_BEGIN(c)
  c->need_device();
  if (level < c->cfg.minlevel || level > c->finest) throw Error("error_flags_user: bad level");
  LevelData &L = c->L(level);
  HIPCK(launch_error_flags_user(c->nd(), c->dev(), L.d_int, (int)L.t_int.size(), qtolerance,
                                c->stream));
  API_END(c)
}

// ---- after-step statistics ------------------------------------------------------
// Tasks of about 4096 cells: runs of rows (2D) or planes (3D) of every patch
// of levels finest..0 (level-list order) and of its parent section. `first`
// gets, per patch of the list and part, the index of its first task.
static void stats_tasks(bsam_ctx *c, std::vector<int> &list, std::vector<StatTask> &tasks,
                        std::vector<int> &first, int group = -1) {
  const int nd = c->nd();
  list.clear();
  tasks.clear();
  for (int l = c->finest; l >= 0; l--)
    for (int p : c->L(l).patches)
      if (group < 0 || c->hp[p].group == group) list.push_back(p);
  first.assign(list.size() * 2 + 1, 0);
  for (size_t n = 0; n < list.size(); n++) {
    const HPatch &h = c->hp[list[n]];
    for (int part = 0; part < 2; part++) {
      first[2 * n + part] = (int)tasks.size();
      if (part == 1 && h.level == 0) continue;
      int m1 = part ? h.mx[0] / 2 : h.mx[0], m2 = part ? h.mx[1] / 2 : h.mx[1];
      int m3 = nd == 3 ? (part ? h.mx[2] / 2 : h.mx[2]) : 1;
      int rows = nd == 3 ? m3 : m2, per = nd == 3 ? m1 * m2 : m1;
      int step = std::max(1, 4096 / std::max(1, per));
      for (int r = 1; r <= rows; r += step)
        tasks.push_back(StatTask{list[n], part, r, std::min(rows + 1, r + step)});
    }
  }
  first[2 * list.size()] = (int)tasks.size();
}

// Runs the task sums on the device with the task table at d_tasks and the
// sums (3 per task) at d_out, both device memory of the context's device,
// then combines them on the host: AfterStep2D/3D's accumulation, patch by
// patch in the reference's order.
static void stats_run(bsam_ctx *c, const std::vector<int> &list,
                      const std::vector<StatTask> &tasks, const std::vector<int> &first,
                      StatTask *d_tasks, double *d_out, double *mass, double *energy,
                      double eps2) {
  const int nd = c->nd();
  HIPCK(hipMemcpyAsync(d_tasks, tasks.data(), sizeof(StatTask) * tasks.size(),
                       hipMemcpyHostToDevice, c->stream));
  // minima of the double well whose convex and concave parts are Dfc and Dfe
  const double w = std::sqrt(c->cfg.fe_num / c->cfg.fe_den / c->cfg.fc_alpha);
  const double lo = c->cfg.fc_shift - w, hi = c->cfg.fc_shift + w;
  HIPCK(launch_after_step(nd, c->dev(), d_tasks, (int)tasks.size(), lo, hi, d_out, c->stream));
  std::vector<double> t(3 * tasks.size());
  HIPCK(hipMemcpyAsync(t.data(), d_out, sizeof(double) * t.size(), hipMemcpyDeviceToHost,
                       c->stream));
  HIPCK(hipStreamSynchronize(c->stream));
  // per patch and part: the task sums in row order
  auto part_sums = [&](size_t n, int part, double s[3]) {
    s[0] = s[1] = s[2] = 0.0;
    for (int k = first[2 * n + part]; k < first[2 * n + part + 1]; k++)
      for (int v = 0; v < 3; v++) s[v] += t[3 * (size_t)k + v];
  };
  double m = 0.0, e = 0.0;
  for (size_t n = 0; n < list.size(); n++) {
    const HPatch &h = c->hp[list[n]];
    double a[3], b[3];
    part_sums(n, 0, a);
    part_sums(n, 1, b);
    const double hh = h.dx, ch = hh * 2.0;
    if (nd == 2) {
      const double h2 = hh * hh, ch2 = ch * ch;
      const double ef = h2 * (a[1] + eps2 / h2 / 4.0 * a[2]);
      if (h.level == 0) {
        m = m + h2 * a[0];
        e = e + ef;
      } else {
        const double ec = ch2 * (b[1] + eps2 / ch2 / 4.0 * b[2]);
        m = m + h2 * a[0] - ch2 * b[0];
        e = e + ef - ec;
      }
    } else {
      const double h3 = hh * hh * hh, ch3 = ch * ch * ch;
      m = h.level == 0 ? m + h3 * a[0] : m + h3 * a[0] - ch3 * b[0];
    }
  }
  *mass = m;
  *energy = e;
}

int bsam_after_step_stats(bsam_ctx *c, double *mass, double *energy) {
  API_BEGIN(c)
  c->need_device();
  if (!mass || !energy) throw Error("null output pointer");
  std::vector<int> list, first;
  std::vector<StatTask> tasks;
  stats_tasks(c, list, tasks, first);
  if (c->n_stats < (int)tasks.size()) {
    HIPCK(hipStreamSynchronize(c->stream));
    DeviceLock lock(g_device_lock);
    if (c->d_stats) (void)hipFree(c->d_stats);
    if (c->d_stats_list) (void)hipFree(c->d_stats_list);
    c->d_stats = nullptr;
    c->d_stats_list = nullptr;
    c->n_stats = 0;
    HIPCK(hipMalloc(&c->d_stats, sizeof(double) * 3 * tasks.size()));
    HIPCK(hipMalloc(&c->d_stats_list, sizeof(StatTask) * tasks.size()));
    c->n_stats = (int)tasks.size();
  }
  stats_run(c, list, tasks, first, c->d_stats_list, c->d_stats, mass, energy, c->cfg.eps2);
  API_END(c)
}

int bsam_after_step_stats_group(bsam_ctx *c, int group, double *mass, double *energy) {
  API_BEGIN(c)
  c->need_device();
  if (!mass || !energy) throw Error("null output pointer");
  if (group < 0 || group >= c->ngroups) throw Error("after_step_stats_group: bad group");
  std::vector<int> list, first;
  std::vector<StatTask> tasks;
  stats_tasks(c, list, tasks, first, group);
  if (tasks.empty()) throw Error("after_step_stats_group: the group has no patch at level >= 0");
  if (c->n_stats < (int)tasks.size()) {
    HIPCK(hipStreamSynchronize(c->stream));
    DeviceLock lock(g_device_lock);
    if (c->d_stats) (void)hipFree(c->d_stats);
    if (c->d_stats_list) (void)hipFree(c->d_stats_list);
    c->d_stats = nullptr;
    c->d_stats_list = nullptr;
    c->n_stats = 0;
    HIPCK(hipMalloc(&c->d_stats, sizeof(double) * 3 * tasks.size()));
    HIPCK(hipMalloc(&c->d_stats_list, sizeof(StatTask) * tasks.size()));
    c->n_stats = (int)tasks.size();
  }
  stats_run(c, list, tasks, first, c->d_stats_list, c->d_stats, mass, energy,
            c->group_eps2[(size_t)group]);
  API_END(c)
}

int bsam_num_groups(const bsam_ctx *c) { return c ? c->ngroups : -1; }

int bsam_group_result(bsam_ctx *c, int group, int *ncycles, double *hist, int maxhist,
                      int *status) {
  API_BEGIN(c)
  c->need_device();
  if (group < 0 || group >= c->ngroups) throw Error("group_result: bad group");
  if (c->group_state.size() != (size_t)(4 * c->ngroups)) throw Error("group_result: no solve yet");
  const int *st = &c->group_state[(size_t)(4 * group)];
  if (ncycles) *ncycles = st[0];
  if (status)
    *status = st[1] ? BSAM_SOLVE_CONVERGED : (st[2] ? BSAM_SOLVE_NONFINITE : BSAM_SOLVE_MAXCYCLES);
  if (hist && maxhist > 0) {
    const int m = std::min(st[0], maxhist);
    HIPCK(hipMemcpyAsync(hist, c->d_hist + (size_t)group * c->hmax, sizeof(double) * m,
                         hipMemcpyDeviceToHost, c->stream));
    HIPCK(hipStreamSynchronize(c->stream));
  }
  API_END(c)
}

// Allocation-free variant: the task table and the task sums live in
// `scratch`, a file-scope device array that every GPU the library runs on
// holds (its own copy, at its own address) until the process ends. Host code
// never uses the array's name as a device address: per device, the first
// call resolves the array's address with hipGetSymbolAddress (that device
// current) and keeps it in g_scratch_ptr. Nothing in the scratch needs
// clearing: every call writes the whole task table it uses and the kernel
// writes every sum the call reads back. Calls on one GPU share the array;
// g_scratch_mutex serialises them within this process (each process has its
// own copy).
constexpr int kStatsMaxTasks = BSAM_STATS_NOALLOC_MAX_TASKS;
constexpr size_t kScratchTaskBytes = sizeof(StatTask) * (size_t)kStatsMaxTasks;
constexpr size_t kScratchBytes = kScratchTaskBytes + sizeof(double) * 3 * (size_t)kStatsMaxTasks;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
static void *g_scratch_ptr[kMaxDevices] = {nullptr};
static std::mutex g_scratch_mutex;

int bsam_after_step_stats_noalloc(bsam_ctx *c, double *mass, double *energy) {
  API_BEGIN(c)
  c->need_device();
  if (!mass || !energy) throw Error("null output pointer");
  const int dev = c->cfg.device;
  if (dev < 0 || dev >= kMaxDevices) throw Error("after_step_stats_noalloc: device ordinal out of range");
  std::vector<int> list, first;
  std::vector<StatTask> tasks;
  stats_tasks(c, list, tasks, first);
  if ((long long)tasks.size() > kStatsMaxTasks)
    throw Error("after_step_stats_noalloc: the hierarchy needs " + std::to_string(tasks.size()) +
                " statistics tasks, more than the " + std::to_string(kStatsMaxTasks) +
                " its fixed scratch holds (use bsam_after_step_stats)");
  std::lock_guard<std::mutex> guard(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    HIPCK(hipSetDevice(dev));
    void *ptr = nullptr;
    HIPCK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) throw Error("after_step_stats_noalloc: no device address for the scratch array");
    g_scratch_ptr[dev] = ptr;
  }
  unsigned char *base = static_cast<unsigned char *>(g_scratch_ptr[dev]);
  stats_run(c, list, tasks, first, reinterpret_cast<StatTask *>(base),
            reinterpret_cast<double *>(base + kScratchTaskBytes), mass, energy, c->cfg.eps2);
  API_END(c)
}

// Same geometry (level, parent, mx, mbounds, mglobal, mthbc, dx, xlower) for
// the patches of levels minlevel..top of two contexts, in the same order.
static bool same_levels(const bsam_ctx *a, const bsam_ctx *b, int top) {
  if (a->cfg.minlevel != b->cfg.minlevel || a->cfg.ndims != b->cfg.ndims ||
      a->cfg.nrvars != b->cfg.nrvars)
    return false;
  std::vector<HPatch> x, y;
  for (const HPatch &h : a->hp)
    if (h.level <= top) x.push_back(h);
  for (const HPatch &h : b->hp)
    if (h.level <= top) y.push_back(h);
  return same_geometry(x, y);
}

int bsam_copy_levels(bsam_ctx *dst, const bsam_ctx *src, int top) {
  API_BEGIN(dst)
  dst->need_device();
  if (!src || src->host_only || !src->committed) throw Error("copy_levels: bad source context");
  if (src->cfg.device != dst->cfg.device) throw Error("copy_levels: contexts on different devices");
  if (top < dst->cfg.minlevel || top > dst->finest || top > src->finest)
    throw Error("copy_levels: bad level");
  if (!same_levels(dst, src, top)) throw Error("copy_levels: hierarchies differ below the level");
  // levels minlevel..top occupy the same leading range of each variable's
  // part of both q pools
  long long end = 0;
  for (size_t p = 0; p < dst->hp.size(); p++)
    if (dst->hp[p].level <= top) end = std::max(end, dst->dp[p].q + dst->dp[p].qn);
  HIPCK(hipStreamSynchronize(src->stream));
  for (int v = 0; v < dst->cfg.nrvars; v++)
    HIPCK(hipMemcpyAsync(dst->q + v * dst->qstride, src->q + v * src->qstride,
                         sizeof(double) * end, hipMemcpyDeviceToDevice, dst->stream));
  // complete before returning: the source may be re-committed or freed next
  HIPCK(hipStreamSynchronize(dst->stream));
  API_END(dst)
}

int bsam_init_fields(bsam_ctx *c, int level) {
  API_BEGIN(c)
  c->need_device();
  if (level <= c->cfg.minlevel || level > c->finest || level < 1)
    throw Error("init_fields: bad level");
  LevelData &L = c->L(level);
  for (int p : L.patches)
    for (int d = 0; d < c->nd(); d++)
      if (c->hp[p].mx[d] % 2) throw Error("init_fields: refined patches have even mx");
  if (!L.d_patch_list) L.d_patch_list = dev_upload(L.patches, c->stream);
  HIPCK(launch_init_fields(c->nd(), c->dev(), L.d_patch_list, (int)L.patches.size(), c->stream));
  API_END(c)
}

// q of every (dst, src) patch pair of `level` where their boxes overlap in
// the level's global index space: the interiors (ghosts false), or the
// ghost-extended boxes (ghosts true). Boxes are listed dst-major, src in
// level-list order; launches are split wherever a box would write a cell an
// earlier box of the same launch writes, so a later src patch wins where
// ghost-extended boxes overlap, as copies made one patch after another do.
static void copy_level_boxes(bsam_ctx *dst, const bsam_ctx *src, int level, bool ghosts,
                             const char *what) {
  dst->need_device();
  if (!src || src->host_only || !src->committed) throw Error(std::string(what) + ": bad source context");
  if (src->cfg.device != dst->cfg.device)
    throw Error(std::string(what) + ": contexts on different devices");
  if (src->cfg.ndims != dst->cfg.ndims || src->cfg.nrvars != dst->cfg.nrvars)
    throw Error(std::string(what) + ": contexts differ in ndims or nrvars");
  if (level < dst->cfg.minlevel || level > dst->finest) throw Error(std::string(what) + ": bad level");
  if (level < src->cfg.minlevel || level > src->finest) return;   // src has no patches there
  const int nd = dst->nd(), nv = dst->cfg.nrvars, g = ghosts ? 1 : 0;   // ghost layers
  std::vector<BoxCopy> boxes;
  struct Placed { int patch, lo[3], hi[3]; };
  std::vector<Placed> placed;
  for (int a : dst->L(level).patches)
    for (int b : src->lev[level - src->cfg.minlevel].patches) {
      const HPatch &t = dst->hp[a], &o = src->hp[b];
      Placed w{a, {1, 1, 1}, {1, 1, 1}};
      bool empty = false;
      for (int d = 0; d < nd; d++) {
        w.lo[d] = std::max(t.mg[d][0], o.mg[d][0]) - g;
        w.hi[d] = std::min(t.mg[d][1], o.mg[d][1]) + g;
        empty = empty || w.hi[d] < w.lo[d];
      }
      if (empty) continue;
      const DPatch &T = dst->dp[a], &O = src->dp[b];
      BoxCopy B;
      // the box's first cell in each patch's index space
      B.s = O.q;
      B.d = T.q;
      B.si = w.lo[0] - o.mg[0][0] + 1;
      B.sj = w.lo[1] - o.mg[1][0] + 1;
      B.sk = nd == 3 ? w.lo[2] - o.mg[2][0] + 1 : 0;
      B.di = w.lo[0] - t.mg[0][0] + 1;
      B.dj = w.lo[1] - t.mg[1][0] + 1;
      B.dk = nd == 3 ? w.lo[2] - t.mg[2][0] + 1 : 0;
      B.spx = O.px;
      B.spy = O.py;
      B.dpx = T.px;
      B.dpy = T.py;
      B.ssv = O.qvs;
      B.dsv = T.qvs;
      B.nx = w.hi[0] - w.lo[0] + 1;
      B.ny = w.hi[1] - w.lo[1] + 1;
      B.nz = nd == 3 ? w.hi[2] - w.lo[2] + 1 : 1;
      B.nv = nv;
      boxes.push_back(B);
      placed.push_back(w);
    }
  if (boxes.empty()) return;
  std::vector<int> cuts{0};
  for (size_t i = 1; i < placed.size(); i++)
    for (size_t j = (size_t)cuts.back(); j < i; j++) {
      bool meet = placed[i].patch == placed[j].patch;
      for (int d = 0; d < nd && meet; d++)
        meet = placed[i].lo[d] <= placed[j].hi[d] && placed[j].lo[d] <= placed[i].hi[d];
      if (meet) {
        cuts.push_back((int)i);
        break;
      }
    }
  cuts.push_back((int)boxes.size());
  HIPCK(hipStreamSynchronize(src->stream));
  BoxCopy *d_boxes = nullptr;
  {
    DeviceLock lock(g_device_lock);
    d_boxes = dev_upload(boxes, dst->stream);
  }
  hipError_t e = hipSuccess;
  for (size_t c = 0; c + 1 < cuts.size() && e == hipSuccess; c++)
    e = launch_box_copy(dst->q, src->q, d_boxes + cuts[c], cuts[c + 1] - cuts[c], dst->stream);
  (void)hipStreamSynchronize(dst->stream);
  {
    DeviceLock lock(g_device_lock);
    (void)hipFree(d_boxes);
  }
  HIPCK(e);
}

int bsam_set_dt(bsam_ctx *c, double dt) {
  API_BEGIN(c)
  if (!std::isfinite(dt) || dt < 0.0) throw Error("set_dt: dt must be finite and >= 0");
  if (dt == c->cfg.dt) return 0;
  c->cfg.dt = dt;
  for (size_t p = 0; p < c->dp.size(); p++) {
    const double h2 = c->dp[p].dx * c->dp[p].dx;   // as at commit
    c->dp[p].dt_h2 = dt / h2;
  }
  for (TailLevel &t : c->h_tail)
    if (t.patch >= 0 && (size_t)t.patch < c->dp.size())
      t.P[0].dt_h2 = t.P[1].dt_h2 = c->dp[t.patch].dt_h2;
  if (!c->host_only && c->committed) {
    // on the context's own stream (no legacy-stream work that another
    // thread's capture would forbid); the host tables stay alive until the
    // copies are done
    HIPCK(hipMemcpyAsync(c->d_P, c->dp.data(), sizeof(DPatch) * c->dp.size(),
                         hipMemcpyHostToDevice, c->stream));
    if (c->d_tail && !c->h_tail.empty())
      HIPCK(hipMemcpyAsync(c->d_tail, c->h_tail.data(), sizeof(TailLevel) * c->h_tail.size(),
                           hipMemcpyHostToDevice, c->stream));
    HIPCK(hipStreamSynchronize(c->stream));
  }
  c->drop_graphs();   // graphs hold the old problem constants as kernel arguments
  API_END(c)
}

int bsam_set_qpo(bsam_ctx *c, const double *qpo, int n) {
  API_BEGIN(c)
  if (n != c->cfg.nrvars || n > 2 || !qpo) throw Error("set_qpo: one offset per variable");
  for (int v = 0; v < n; v++) c->qpo[v] = qpo[v];
  c->drop_graphs();   // the captured launches hold the problem constants
  API_END(c)
}

int bsam_transfer_values(bsam_ctx *dst, const bsam_ctx *src, int level) {
  API_BEGIN(dst)
  copy_level_boxes(dst, src, level, false, "transfer");
  API_END(dst)
}

int bsam_overlay_level(bsam_ctx *dst, const bsam_ctx *src, int level) {
  API_BEGIN(dst)
  copy_level_boxes(dst, src, level, true, "overlay");
  API_END(dst)
}

int bsam_regrid_level(const bsam_regrid_params *rp, int npatch, const int *mx,
                      const int *mglobal, int *const *flags, int max_grids, int *ngrids,
                      int *parent, int *mbounds) {
  try {
    if (!rp || !mx || !mglobal || !flags || !ngrids || npatch < 0 || max_grids < 0)
      throw Error("regrid: null argument");
    RegridParams p;
    p.ndims = rp->ndims;
    p.level = rp->level;
    p.ibuffer = rp->ibuffer;
    p.mgp = rp->minimumgridpoints;
    p.desfill = rp->desiredfillratio;
    HPatch root;
    for (int d = 0; d < 3; d++) root.mx[d] = d < rp->ndims ? rp->root_mx[d] : 1;
    for (int f = 0; f < 6; f++) root.mthbc[f] = f < 2 * rp->ndims ? rp->root_mthbc[f] : 1;
    p.per = periodic_setup(root, rp->ndims);
    std::vector<RegridPatch> ps((size_t)npatch);
    for (int k = 0; k < npatch; k++) {
      for (int d = 0; d < 3; d++) {
        ps[k].mx[d] = d < rp->ndims ? mx[3 * k + d] : 1;
        ps[k].mg[d][0] = d < rp->ndims ? mglobal[6 * k + 2 * d] : 1;
        ps[k].mg[d][1] = d < rp->ndims ? mglobal[6 * k + 2 * d + 1] : 1;
        if (ps[k].mx[d] < 1) throw Error("regrid: mx must be positive");
      }
      if (!flags[k]) throw Error("regrid: null flags");
      ps[k].flags = flags[k];
    }
    std::vector<NewGrid> out;
    std::string msg;
    if (regrid_level(p, ps, out, msg)) throw Error(msg);
    if ((int)out.size() > max_grids) throw Error("regrid: more new subgrids than max_grids");
    *ngrids = (int)out.size();
    for (size_t g = 0; g < out.size(); g++) {
      if (parent) parent[g] = out[g].parent;
      if (mbounds)
        for (int d = 0; d < 3; d++) {
          mbounds[6 * g + 2 * d] = out[g].mb[d][0];
          mbounds[6 * g + 2 * d + 1] = out[g].mb[d][1];
        }
    }
    return 0;
  } catch (const std::exception &e) {
    g_err = e.what();
    return -1;
  }
}

int bsam_download_flags(bsam_ctx *c, int patch, int *host, long n) {
  API_BEGIN(c)
  c->need_device();
  if (patch < 0 || patch >= (int)c->hp.size()) throw Error("patch index out of range");
  const DPatch &D = c->dp[patch];
  if (n != D.fvs) throw Error("download_flags: size mismatch");
  HIPCK(hipMemcpyAsync(host, c->eflags + D.fl, sizeof(int) * n, hipMemcpyDeviceToHost,
                       c->stream));
  HIPCK(hipStreamSynchronize(c->stream));
  API_END(c)
}

int bsam_set_ghost(bsam_ctx *c, int level, int ipass) {
  API_BEGIN(c)
  c->need_device();
  op_set_ghost(c, level, ipass);
  API_END(c)
}
int bsam_relax(bsam_ctx *c, int level, int color) {
  API_BEGIN(c)
  c->need_device();
  if (color != 1 && color != 2) throw Error("color must be 1 or 2");
  op_relax(c, level, color);
  API_END(c)
}
int bsam_restrict_solution(bsam_ctx *c, int level) {
  API_BEGIN(c)
  c->need_device();
  op_restrict(c, level);
  API_END(c)
}
int bsam_coarse_ghost_points(bsam_ctx *c, int level) {
  API_BEGIN(c)
  c->need_device();
  op_coarse_ghost(c, level);
  API_END(c)
}
int bsam_coarse_load(bsam_ctx *c, int level) {
  API_BEGIN(c)
  c->need_device();
  op_coarse_load(c, level);
  API_END(c)
}
int bsam_correct_fine(bsam_ctx *c, int level) {
  API_BEGIN(c)
  c->need_device();
  op_correct(c, level);
  API_END(c)
}
int bsam_fill_down(bsam_ctx *c, int field) {
  API_BEGIN(c)
  c->need_device();
  op_fill_down(c, field);
  API_END(c)
}
int bsam_source(bsam_ctx *c) {
  API_BEGIN(c)
  c->need_device();
  op_source(c, 1);
  API_END(c)
}
int bsam_vcycle(bsam_ctx *c, int level) {
  API_BEGIN(c)
  c->need_device();
  c->L(level);
  op_vcycle(c, level);
  API_END(c)
}
int bsam_error_afas(bsam_ctx *c, int level, double *residual) {
  API_BEGIN(c)
  c->need_device();
  reset_state(c);
  queue_error(c, level, -1.0);
  double r = 0.0;
  HIPCK(hipMemcpyAsync(&r, c->d_hist, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  reset_state(c);   // a non-finite residual must not stop the next calls' kernels
  HIPCK(hipStreamSynchronize(c->stream));
  if (residual) *residual = r;
  API_END(c)
}
int bsam_source_norm(bsam_ctx *c, int level, double *rms) {
  API_BEGIN(c)
  c->need_device();
  LevelData &L = c->L(level);
  int nt = (int)L.t_int.size();
  HIPCK(launch_resnorm(c->nd(), c->dev(), L.d_int, nt, c->d_partial, 1, 1, c->stream));
  reset_state(c);
  HIPCK(launch_finalize(c->d_partial, nt, (double)(c->cfg.nrvars * L.cells), -1.0, c->d_hist,
                        c->d_state, 1, c->stream));
  double r = 0.0;
  HIPCK(hipMemcpyAsync(&r, c->d_hist, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  reset_state(c);
  HIPCK(hipStreamSynchronize(c->stream));
  if (rms) *rms = r;
  API_END(c)
}
int bsam_copy_qold(bsam_ctx *c) {
  API_BEGIN(c)
  c->need_device();
  HIPCK(launch_copy(c->qold, c->q, c->nq, c->stream));
  API_END(c)
}
int bsam_multigrid_iterations(bsam_ctx *c, int *ncycles, double *hist, int maxhist,
                              int *converged) {
  API_BEGIN(c)
  c->need_device();
  solve(c, ncycles, hist, maxhist, converged);
  API_END(c)
}
int bsam_time_step(bsam_ctx *c, int *ncycles, double *hist, int maxhist, int *converged) {
  API_BEGIN(c)
  c->need_device();
  for (int level = 0; level <= c->finest; level++) op_set_ghost(c, level, 0);
  op_fill_down(c, BSAM_FIELD_Q);
  {
    ProfScope ps(c, BSAM_PROF_COPY, 16.0 * c->nq);
    HIPCK(launch_copy(c->qold, c->q, c->nq, c->stream));
  }
  solve(c, ncycles, hist, maxhist, converged);
  API_END(c)
}
int bsam_sync(bsam_ctx *c) {
  API_BEGIN(c)
  if (!c->host_only) HIPCK(hipStreamSynchronize(c->stream));
  API_END(c)
}

int bsam_tune(bsam_ctx *c, const char *key, double value) {
  API_BEGIN(c)
  std::string k = key ? key : "";
  if (k == "tail_cells") {
    if (c->committed || c->building) throw Error("tail_cells must be set before the hierarchy");
    c->tail_cells = (long long)value;
  } else if (k == "tile_x" || k == "tile_y") {
    if (c->committed || c->building) throw Error(k + " must be set before the hierarchy");
    int v = (int)value, lg = 0;
    while ((1 << lg) < v) lg++;
    if (v < 1 || (1 << lg) != v || v > 256) throw Error(k + " must be a power of two <= 256");
    if (k == "tile_x" && v < 2) throw Error("tile_x must be at least 2");
    (k == "tile_x" ? c->tile_lx : c->tile_ly) = lg;
    c->tile_user = true;
  } else if (k == "tile_auto") {
    if (c->committed || c->building) throw Error("tile_auto must be set before the hierarchy");
    c->tile_auto = value != 0.0;
  } else if (k == "tail_lds") {
    if (c->committed || c->building) throw Error("tail_lds must be set before the hierarchy");
    c->tail_lds_cap = (long long)value;
  } else if (k == "tail_threads") {
    int t = (int)value;
    if (t < 64 || t > 1024 || t % 64) throw Error("tail_threads must be 64..1024, multiple of 64");
    c->tail_threads = t;
    c->drop_graphs();
  } else if (k == "inner_cycles") {
    int n = (int)value;
    if (n < 1 || n > 8) throw Error("inner_cycles must be 1..8");
    c->inner_cycles = n;
    c->drop_graphs();
  } else if (k == "graph") {
    c->use_graph = value != 0.0;
  } else if (k == "fuse_ghost") {
    c->fuse_ghost = value != 0.0;
    c->drop_graphs();
  } else if (k == "post_ghost") {
    c->post_ghost = value != 0.0;
    c->drop_graphs();
  } else if (k == "relax_lds") {
    c->relax_lds = value != 0.0;
    c->drop_graphs();
  } else if (k == "short_ghost") {
    if (c->committed || c->building) throw Error("short_ghost must be set before the hierarchy");
    c->short_ghost = value != 0.0;
  } else if (k == "iso_side") {
    c->iso_side = value != 0.0;
    c->drop_graphs();
  } else if (k == "cgl_fine_cells") {
    if (c->committed || c->building)
      throw Error("cgl_fine_cells must be set before the hierarchy");
    c->cgl_fine_cells = (long long)value;
  } else if (k == "iso_min_cells") {
    if (c->committed || c->building)
      throw Error("iso_min_cells must be set before the hierarchy");
    if (value < 0) throw Error("iso_min_cells must be >= 0");
    c->iso_min_cells = (long long)value;
  } else {
    throw Error("unknown tuning key " + k);
  }
  API_END(c)
}

int bsam_tune_get(const bsam_ctx *cc, const char *key, double *value) {
  bsam_ctx *c = const_cast<bsam_ctx *>(cc);
  API_BEGIN(c)
  std::string k = key ? key : "";
  if (!value) throw Error("null value pointer");
  if (k == "tail_cells") *value = (double)c->tail_cells;
  else if (k == "tail_threads") *value = c->tail_threads;
  else if (k == "tail_lds") *value = (double)c->tail_lds_cap;
  else if (k == "inner_cycles") *value = c->inner_cycles;
  else if (k == "graph") *value = c->use_graph;
  else if (k == "commits_reused") *value = (double)c->commits_reused;
  else if (k == "fuse_ghost") *value = c->fuse_ghost;
  else if (k == "post_ghost") *value = c->post_ghost;
  else if (k == "relax_lds") *value = c->relax_lds;
  else if (k == "short_ghost") *value = c->short_ghost;
  else if (k == "iso_min_cells") *value = (double)c->iso_min_cells;
  else if (k == "iso_side") *value = c->iso_side;
  else if (k == "cgl_fine_cells") *value = (double)c->cgl_fine_cells;
  else if (k == "iso_patches" || k == "iso_cells") {
    double n = 0;
    for (const LevelData &L : c->lev)
      n += k == "iso_patches" ? (double)L.iso.size() : (double)L.iso_cells;
    *value = n;
  }
  else if (k == "tile_x") *value = c->tx();
  else if (k == "tile_y") *value = c->ty();
  else if (k == "tile_z") *value = c->tz();
  else if (k == "tail_top") *value = c->tail_top;
  else if (k == "tile_auto") *value = c->tile_auto && !c->tile_user;
  else if (k == "tile_auto_us") *value = c->tile_auto_us;
  else if (k == "tile_auto_cached") *value = c->tile_auto_cached;
  else if (k == "tile_auto_candidates") *value = c->tile_auto_candidates;
  else if (k == "tail_lds_fp64") *value = (double)c->tail_lds[0];
  else if (k == "tail_lds_fp32") *value = (double)c->tail_lds[1];
  else throw Error("unknown tuning key " + k);
  API_END(c)
}

int bsam_profile_begin(bsam_ctx *c, unsigned kind_mask) {
  API_BEGIN(c)
  c->need_device();
  HIPCK(hipStreamSynchronize(c->stream));
  c->prof_mask = kind_mask;
  c->prof.clear();
  c->ev_used = 0;
  for (int k = 0; k < BSAM_PROF_NKINDS; k++) {
    c->prof_ms[k] = c->prof_bytes[k] = 0.0;
    c->prof_n[k] = 0;
  }
  API_END(c)
}

int bsam_profile_end(bsam_ctx *c) {
  API_BEGIN(c)
  c->need_device();
  HIPCK(hipStreamSynchronize(c->stream));
  for (auto &r : c->prof) {
    float ms = 0.f;
    HIPCK(hipEventElapsedTime(&ms, c->ev_pool[r.ev], c->ev_pool[r.ev + 1]));
    c->prof_ms[r.kind] += ms;
    c->prof_bytes[r.kind] += r.bytes;
    c->prof_n[r.kind] += 1;
  }
  c->prof.clear();
  c->ev_used = 0;
  c->prof_mask = 0;
  API_END(c)
}

int bsam_profile_kernel(bsam_ctx *c, int kind, int reps, double *us_per_launch,
                        double *bytes_per_launch) {
  API_BEGIN(c)
  c->need_device();
  if (kind != BSAM_PROF_RELAX) throw Error("bsam_profile_kernel: only the relax kernel");
  if (reps < 1) throw Error("reps must be positive");
  int low = std::max(c->cfg.minlevel, c->tail_top + 1);
  if (low > c->finest) throw Error("no level above the tail");
  hipEvent_t e[2];
  {
    DeviceLock lock(g_device_lock);
    HIPCK(hipEventCreate(&e[0]));
    HIPCK(hipEventCreate(&e[1]));
  }
  double bytes = 0.0;
  long long n = 0;
  hipError_t err = hipEventRecord(e[0], c->stream);
  for (int r = 0; r < reps && err == hipSuccess; r++)
    for (int l = c->finest; l >= low && err == hipSuccess; l--) {
      LevelData &L = c->L(l);
      for (int color = 1; color <= 2 && err == hipSuccess; color++) {
        err = launch_relax(c->nd(), PREC_SOLUTION_FP64, c->dev(), L.d_int, (int)L.t_int.size(),
                           color, c->stream);
        bytes += 24.0 * L.cells;
        n++;
      }
    }
  if (err == hipSuccess) err = hipEventRecord(e[1], c->stream);
  if (err == hipSuccess) err = hipEventSynchronize(e[1]);
  float ms = 0.f;
  if (err == hipSuccess) err = hipEventElapsedTime(&ms, e[0], e[1]);
  {
    DeviceLock lock(g_device_lock);
    (void)hipEventDestroy(e[0]);
    (void)hipEventDestroy(e[1]);
  }
  HIPCK(err);
  if (us_per_launch) *us_per_launch = 1e3 * ms / (double)n;
  if (bytes_per_launch) *bytes_per_launch = bytes / (double)n;
  API_END(c)
}

int bsam_copy_rate(int device, long long bytes, int reps, double *gbs) {
  if (bytes < 8 || reps < 1 || !gbs) {
    g_err = "copy_rate: bytes >= 8, reps >= 1 and an output pointer are needed";
    return -1;
  }
  double *a = nullptr, *b = nullptr;
  hipEvent_t e[2] = {nullptr, nullptr};
  hipStream_t s = nullptr;
  long long n = bytes / 8;
  hipError_t err;
  {
    DeviceLock lock(g_device_lock);
    err = hipSetDevice(device);
    if (err == hipSuccess) err = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (err == hipSuccess) err = hipMalloc(&a, 8 * n);
    if (err == hipSuccess) err = hipMalloc(&b, 8 * n);
    if (err == hipSuccess) err = hipEventCreate(&e[0]);
    if (err == hipSuccess) err = hipEventCreate(&e[1]);
  }
  if (err == hipSuccess) err = hipMemsetAsync(a, 0, 8 * n, s);
  if (err == hipSuccess) err = launch_copy(b, a, n, s);   // first touch, untimed
  if (err == hipSuccess) err = hipEventRecord(e[0], s);
  for (int r = 0; r < reps && err == hipSuccess; r++)
    err = (r & 1) ? launch_copy(a, b, n, s) : launch_copy(b, a, n, s);
  if (err == hipSuccess) err = hipEventRecord(e[1], s);
  if (err == hipSuccess) err = hipEventSynchronize(e[1]);
  float ms = 0.f;
  if (err == hipSuccess) err = hipEventElapsedTime(&ms, e[0], e[1]);
  {
    DeviceLock lock(g_device_lock);
    if (s) (void)hipStreamSynchronize(s);
    for (hipEvent_t x : e)
      if (x) (void)hipEventDestroy(x);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (s) (void)hipStreamDestroy(s);
  }
  if (err != hipSuccess) {
    g_err = std::string("copy_rate: ") + hipGetErrorString(err);
    return -2;
  }
  *gbs = 2.0 * 8.0 * (double)n * reps / (1e-3 * ms) / 1e9;
  return 0;
}

int bsam_profile_get(const bsam_ctx *cc, int kind, long long *launches, double *ms,
                     double *bytes) {
  bsam_ctx *c = const_cast<bsam_ctx *>(cc);
  API_BEGIN(c)
  if (kind < 0 || kind >= BSAM_PROF_NKINDS) throw Error("bad profile kind");
  if (launches) *launches = c->prof_n[kind];
  if (ms) *ms = c->prof_ms[kind];
  if (bytes) *bytes = c->prof_bytes[kind];
  API_END(c)
}

int bsam_ghost_program_size(const bsam_ctx *cc, int level, int ipass) {
  bsam_ctx *c = const_cast<bsam_ctx *>(cc);
  if (!c || !c->committed || ipass < 0 || ipass > 2) return -1;
  if (level < c->cfg.minlevel || level > c->finest) return -1;
  return (int)c->L(level).gp[ipass].size();
}

int bsam_ghost_program_dst(const bsam_ctx *cc, int level, int ipass, int *out, int maxn) {
  bsam_ctx *c = const_cast<bsam_ctx *>(cc);
  API_BEGIN(c)
  if (!c->committed) throw Error("hierarchy not committed");
  if (ipass < 0 || ipass > 2) throw Error("bad ipass");
  auto &g = c->L(level).gp[ipass];
  int n = std::min((int)g.size(), maxn);
  for (int e = 0; e < n; e++) {
    int p, i, j, k;
    locate_q_cell(c->hp, g[e].dst, p, i, j, k);
    out[4 * e] = p; out[4 * e + 1] = i; out[4 * e + 2] = j; out[4 * e + 3] = k;
  }
  API_END(c)
}

int bsam_ghost_program_apply_host(const bsam_ctx *cc, int level, int ipass,
                                  double *const *q_patches) {
  bsam_ctx *c = const_cast<bsam_ctx *>(cc);
  API_BEGIN(c)
  if (!c->committed) throw Error("hierarchy not committed");
  if (ipass < 0 || ipass > 2) throw Error("bad ipass");
  apply_ghost_program_host(c->hp, c->L(level).gp[ipass], c->cfg.nrvars, q_patches, c->qpo);
  API_END(c)
}

int bsam_fused_ghost_size(const bsam_ctx *cc, int level, int ipass, int color) {
  bsam_ctx *c = const_cast<bsam_ctx *>(cc);
  if (!c || !c->committed) return -1;
  if (level < c->cfg.minlevel || level > c->finest) return -1;
  int f = fused_index(ipass, color);
  if (f < 0 || !c->L(level).fused[f].ok) return -1;
  return c->L(level).fused[f].kept;
}

int bsam_fused_ghost_pre_size(const bsam_ctx *cc, int level, int ipass, int color) {
  bsam_ctx *c = const_cast<bsam_ctx *>(cc);
  if (!c || !c->committed) return -1;
  if (level < c->cfg.minlevel || level > c->finest) return -1;
  int f = fused_index(ipass, color);
  if (f < 0 || !c->L(level).fused[f].ok) return -1;
  return (int)c->L(level).fused[f].pre.size();
}

int bsam_post_ghost_size(const bsam_ctx *cc, int level, int pending, int color, int part) {
  bsam_ctx *c = const_cast<bsam_ctx *>(cc);
  if (!c || !c->committed || part < 0 || part > 3) return -1;
  if (level < c->cfg.minlevel || level > c->finest) return -1;
  const int k = part >= 2 ? (pending == 1 && color == 2 ? kPostFinal : -1)
                          : post_key(pending, color);
  if (k < 0 || !c->L(level).post[k].ok) return -1;
  const PostGhost &g = c->L(level).post[k];
  return (int)(part % 2 == 0 ? g.ops.size() : g.rest.size());
}

int bsam_parent_ghost_rides(const bsam_ctx *cc, int level, int which) {
  bsam_ctx *c = const_cast<bsam_ctx *>(cc);
  if (!c || !c->committed || which < 0 || which > 1) return -1;
  if (level <= c->cfg.minlevel || level > c->finest) return -1;
  const LevelData &L = c->L(level);
  return (which == 0 ? L.ride_cgl.ok : L.ride_corr.ok) ? 1 : 0;
}

int bsam_ride_loop_size(const bsam_ctx *cc, int level) {
  bsam_ctx *c = const_cast<bsam_ctx *>(cc);
  if (!c || !c->committed) return -1;
  if (level <= c->cfg.minlevel || level > c->finest) return -1;
  const RideGhost &r = c->L(level).ride_cgl_loop;
  if (!r.ok) return -1;
  std::vector<int> at;
  std::vector<GhostOp> all(r.rest);
  all.insert(all.end(), r.ops.begin(), r.ops.end());
  std::sort(all.begin(), all.end(), [](const GhostOp &a, const GhostOp &b) { return a.dst < b.dst; });
  int n = 0;
  for (size_t k = 0; k < all.size(); k++) n += k == 0 || all[k].dst != all[k - 1].dst;
  return n;
}

int bsam_ride_loop_apply_host(const bsam_ctx *cc, int level, double *const *q_patches) {
  bsam_ctx *c = const_cast<bsam_ctx *>(cc);
  API_BEGIN(c)
  if (!c->committed) throw Error("hierarchy not committed");
  if (level <= c->cfg.minlevel || level > c->finest) throw Error("level out of range");
  const RideGhost &r = c->L(level).ride_cgl_loop;
  if (!r.ok) throw Error("no loop ride on this level");
  // the updates applied per tile and grid-stride (a per-tile update may be
  // listed for several tiles; each copy writes the same value)
  std::vector<GhostOp> all(r.rest);
  all.insert(all.end(), r.ops.begin(), r.ops.end());
  apply_ghost_program_host(c->hp, all, c->cfg.nrvars, q_patches, c->qpo);
  API_END(c)
}

int bsam_ghost_rides_down(const bsam_ctx *cc, int level, int ipass) {
  bsam_ctx *c = const_cast<bsam_ctx *>(cc);
  if (!c || !c->committed || ipass < 0 || ipass > 2) return -1;
  if (level < c->cfg.minlevel || level > c->finest) return -1;
  return c->L(level).down_ok[ipass] ? 1 : 0;
}

int bsam_fused_ghost_part_apply_host(const bsam_ctx *cc, int level, int ipass, int color,
                                     int part, double *const *q_patches) {
  bsam_ctx *c = const_cast<bsam_ctx *>(cc);
  API_BEGIN(c)
  if (!c->committed) throw Error("hierarchy not committed");
  if (level < c->cfg.minlevel || level > c->finest) throw Error("level out of range");
  int f = fused_index(ipass, color);
  if (f < 0) throw Error("no fused form of this ipass and colour");
  const FusedGhost &g = c->L(level).fused[f];
  if (!g.ok) throw Error("this transition is not fused on this level");
  if (part != 0 && part != 1) throw Error("part must be 0 (in the launch) or 1 (before it)");
  apply_ghost_program_host(c->hp, part == 0 ? g.ops : g.pre, c->cfg.nrvars, q_patches, c->qpo);
  API_END(c)
}

int bsam_post_ghost_apply_host(const bsam_ctx *cc, int level, int pending, int color, int part,
                               double *const *q_patches) {
  bsam_ctx *c = const_cast<bsam_ctx *>(cc);
  API_BEGIN(c)
  if (!c->committed) throw Error("hierarchy not committed");
  if (level < c->cfg.minlevel || level > c->finest) throw Error("level out of range");
  if (part < 0 || part > 3) throw Error("part must be 0..3");
  const int k = part >= 2 ? (pending == 1 && color == 2 ? kPostFinal : -1)
                          : post_key(pending, color);
  if (k < 0 || !c->L(level).post[k].ok) throw Error("this launch has no post plan");
  const PostGhost &g = c->L(level).post[k];
  apply_ghost_program_host(c->hp, part % 2 == 0 ? g.ops : g.rest, c->cfg.nrvars, q_patches, c->qpo);
  API_END(c)
}

int bsam_iso_patches(const bsam_ctx *cc, int level, int *out, int maxn) {
  bsam_ctx *c = const_cast<bsam_ctx *>(cc);
  if (!c || !c->committed) return -1;
  if (level < c->cfg.minlevel || level > c->finest) return -1;
  const LevelData &L = c->L(level);
  for (int n = 0; n < (int)L.iso.size() && n < maxn && out; n++) out[n] = L.iso[(size_t)n].patch;
  return (int)L.iso.size();
}

int bsam_iso_program_apply_host(const bsam_ctx *cc, int level, int ipass,
                                double *const *q_patches) {
  bsam_ctx *c = const_cast<bsam_ctx *>(cc);
  API_BEGIN(c)
  if (!c->committed) throw Error("hierarchy not committed");
  if (level < c->cfg.minlevel || level > c->finest) throw Error("level out of range");
  if (ipass < 0 || ipass > 2) throw Error("bad ipass");
  // the records read back into q-pool updates through the same LDS index
  // map the smoothing kernel uses (x < qn: the patch's cell x, else staged
  // parent cell x - qn)
  const LevelData &L = c->L(level);
  std::vector<GhostOp> ops;
  for (const IsoPatch &I : L.iso) {
    const HPatch &h = c->hp[(size_t)I.patch];
    auto off = [&](unsigned short x) -> int {
      return x < h.qn ? (int)(h.qoff + x) : L.iso_stage[(size_t)(I.stage0 + (x - h.qn))];
    };
    std::vector<IsoRec> recs;
    for (int r = I.srec[ipass]; r < I.srec[ipass + 1]; r++)
      recs.push_back(iso_unshort(L.iso_srec[(size_t)r]));
    for (int r = I.rec[ipass]; r < I.rec[ipass + 1]; r++) recs.push_back(L.iso_rec[(size_t)r]);
    for (const IsoRec &e : recs) {
      GhostOp g{};
      g.dst = off(e.dst);
      g.kind = e.kind;
      g.nflip = e.nflip;
      g.group = (short)h.group;
      g.f0 = off(e.s[0]);
      if (e.kind != 0) {
        g.f1 = off(e.s[1]);
        const int nc = e.kind == 1 ? 1 : e.kind == 2 ? 3 : 4;
        for (int k = 0; k < nc; k++) g.c[k] = off(e.s[2 + k]);
      }
      ops.push_back(g);
    }
  }
  apply_ghost_program_host(c->hp, ops, c->cfg.nrvars, q_patches, c->qpo);
  API_END(c)
}

int bsam_fused_ghost_apply_host(const bsam_ctx *cc, int level, int ipass, int color,
                                double *const *q_patches) {
  bsam_ctx *c = const_cast<bsam_ctx *>(cc);
  API_BEGIN(c)
  if (!c->committed) throw Error("hierarchy not committed");
  int f = fused_index(ipass, color);
  if (f < 0) throw Error("no fused form of this ipass and colour");
  const FusedGhost &g = c->L(level).fused[f];
  if (!g.ok) throw Error("this transition is not fused on this level");
  // the updates launched before the relaxation and those applied inside it
  // (no update reads what another writes, so one combined application)
  std::vector<GhostOp> all(g.pre);
  all.insert(all.end(), g.ops.begin(), g.ops.end());
  apply_ghost_program_host(c->hp, all, c->cfg.nrvars, q_patches, c->qpo);
  API_END(c)
}

}  // extern "C"
