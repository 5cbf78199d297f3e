// common.hpp -- data structures shared by the host engine and the gfx950
// kernels of the FAS multigrid engine.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bsam {

// Layout of every field array on the device. Each row (x the fastest
// index) of a patch array is split by the parity of the x index: the row's
// cells of even index first, then those of odd index. With rows of n cells
// (h = ceil(n/2) per parity, a row taking 2*h cells, one padding slot when n
// is odd), cell (i, j, k) of a plane of ny rows sits at
// (i & 1)*h + (i >> 1) + 2*h*j + 2*h*ny*k. Ghost-extended arrays (q, qold,
// qc, e, qb) index rows from the ghost (i = 0..n-1); interior arrays (f,
// ftmp, rf, fe) use i - 1 for i = 1..mx. All cells of one red-black colour in
// a row share the parity of i, so a half-sweep of RelaxGrid reads and writes
// whole half-rows instead of every other element. Arrays cross the ABI in the
// reference layout and are permuted on the host (to_device / from_device).
struct Split {
  long long xo, rs, pl;   // offset of the odd half of a row, row stride, plane stride
};
__host__ __device__ __forceinline__ Split split_of(long long n, long long ny) {
  const long long h = (n + 1) >> 1;
  return Split{h, 2 * h, 2 * h * ny};
}
__host__ __device__ __forceinline__ long long scell(const Split &s, int i, int j, int k) {
  return (i & 1) * s.xo + (i >> 1) + s.rs * j + s.pl * k;
}
// Offset of ghost-extended cell (i, j, k) (lower bound 0) of an array with
// rows of px cells and py rows per plane.
__host__ __device__ __forceinline__ long long qcell(long long px, long long py, int i, int j,
                                                   int k) {
  return scell(split_of(px, py), i, j, k);
}
// The cell (i, j, k) at offset r of such an array (inverse of qcell).
inline void qcell_index(long long r, long long px, long long py, int &i, int &j, int &k) {
  const Split s = split_of(px, py);
  k = (int)(r / s.pl);
  const long long m = r % s.pl;
  j = (int)(m / s.rs);
  const long long x = m % s.rs;
  i = (int)(x < s.xo ? 2 * x : 2 * (x - s.xo) + 1);
}
// Between the reference layout (planes of n*ny cells, `planes` of them, the
// variables' planes following each other) and the device layout.
inline void to_device(const double *in, double *out, long long n, long long ny,
                      long long planes) {
  const Split s = split_of(n, ny);
  for (long long p = 0; p < planes; p++)
    for (long long j = 0; j < ny; j++)
      for (long long i = 0; i < n; i++)
        out[p * s.pl + scell(s, (int)i, (int)j, 0)] = in[p * n * ny + i + n * j];
}
inline void from_device(const double *in, double *out, long long n, long long ny,
                        long long planes) {
  const Split s = split_of(n, ny);
  for (long long p = 0; p < planes; p++)
    for (long long j = 0; j < ny; j++)
      for (long long i = 0; i < n; i++)
        out[p * n * ny + i + n * j] = in[p * s.pl + scell(s, (int)i, (int)j, 0)];
}

// Device record of one patch; offsets index the per-field pools. The q pool
// (and qold, and the FP32 e and qb, which share its offsets) holds every
// patch's first variable, then every patch's second at one pool-wide stride
// (qvs), so a cell's variables sit at p and p + qvs for every patch; the
// other pools keep a patch's variables one after the other (fvs, cvs, rvs:
// the patch's size per variable).
struct DPatch {
  int mx, my, mz;          // interior extents (mz = 1 in 2D)
  int px, py;              // q row pitch (mx+2) and rows per plane (my+2)
  int cmx, cmy, cmz;       // coarse extents (mx/2 ...), qc pitch cmx+2
  int level, parent;       // parent patch index, -1 if none
  int mb[3][2];            // mbounds in the parent's local coarse indices
  int group;               // independent sub-problem (a served request) the patch belongs to
  long long q;             // offset of q(0,0,k0,1) in the q and qold pools
  long long f;             // offset of f(1,1,1,1) in the f/rf/ftmp pools
  long long qc;            // offset of qc(0,0,k0,1) in the qc pool
  long long qvs, fvs, cvs; // variable strides of q (the pool's), f and qc
  long long qn;            // cells per variable of the patch's q
  long long qr, rvs;       // offset and variable stride of qrte(1,1,1,1) (qrte pool)
  long long fl;            // offset of errorflags(1,1,1) (error-flag pool)
  double dx;
  // operator constants of this grid, computed on the host exactly as the
  // reference does per call (h2 = dx*dx): dt/h2, eps2/h2, (2*ndims*eps2)/h2
  double dt_h2, eps2_h2, lap_h2;
};

// One ghost-cell update of a SetGhost program. Every source is a value that
// the same program never writes, so all updates of a program run
// concurrently. kind: 0 copy, 1 and 2 two fine + coarse quadratic
// interpolation with a single coarse value or a 3-point coarse interpolant,
// 3 the 3D face form (two fine + bilinear coarse, four values). nflip counts
// trailing "2 - x" boundary maps; group is the destination patch's group.
// Offsets are of the first variable in the q pool; the second variable of
// every cell is at the pool's variable stride, so the record carries none.
// 32 bytes.
//
// The nflip byte also carries the offset a periodic copy adds to the value
// it copies (PeriodicBC, PeriodicGhostOverlap, boundary.f90:417-596): bits
// 0-3 count the "2 - x" maps, bits 4-5 hold the offset added before them:
// 1 +qpo(n), 2 -qpo(n) (across the periodic x boundary), 3 +0.0 (the zero
// qoffset PeriodicGhostOverlap adds for images shifted in y or z only, which
// turns a -0.0 into +0.0), 0 none.
struct GhostOp {
  int dst;                 // q-pool offset (variable 1)
  unsigned char kind, nflip;
  short group;
  int f0, f1;              // fine (or copy) sources
  int c[4];                // parent sources
};
static_assert(sizeof(GhostOp) == 32, "ghost update record: 32 bytes");

// The short device form of a copy (kind 0: src is the source) or of a kind-1
// interpolation whose fine sources lie within 16-bit offsets of the
// destination (src is the coarse source, the fine ones are dst + d0 and
// dst + d1). 16 bytes.
struct GhostS {
  int dst, src;
  short d0, d1;
  unsigned char kind, nflip;
  short group;
};
static_assert(sizeof(GhostS) == 16, "short ghost update record: 16 bytes");

// A list of ghost updates on the device: the short records, then the others
// in the 32-byte form. A list applied per tile has per-tile starts into each
// (ss, sl: nt + 1 entries); a list applied grid-stride has none (ss null) and
// ns, nl updates. An empty list has ns = nl = 0 and null starts.
struct GhostList {
  const GhostS *s;
  const GhostOp *l;
  const int *ss, *sl;
  int ns, nl;
};

// A patch of a level whose SetGhost programs read no other patch of the
// level and whose cells no other patch's ghost update reads ("isolated"): its
// whole smoothing (every half-sweep and the SetGhost programs between them)
// runs in one workgroup out of LDS. The workgroup holds, per variable, the
// patch's ghost-extended q array (qn cells, the pool layout) followed by the
// ns parent-level cells its programs read (`stage` entries from `stage0`,
// q-pool offsets of variable 1); program p's updates are records
// rec[p] .. rec[p + 1] - 1.
struct IsoPatch {
  int patch;
  int ns, stage0;
  int rec[4];    // 16-byte records (IsoRec)
  int srec[4];   // 8-byte records (iso_short); program p: srec[p] .. srec[p + 1] - 1
};
// One update of an isolated patch's program in LDS indices of one variable
// (x < qn: the patch's cell x; x >= qn: staged parent cell x - qn): the value
// of ghost_op's kind `kind` over sources s[0], s[1] (fine) and s[2]..s[5]
// (coarse, as many as the kind takes), flipped nflip times, into cell dst.
// 16 bytes.
struct IsoRec {
  unsigned short dst;
  unsigned char kind, nflip;
  unsigned short s[6];
};
static_assert(sizeof(IsoRec) == 16, "isolated-patch update record: 16 bytes");
// The 8-byte form of a copy or a kind-1 interpolation whose indices are below
// 2^13: dst, s[0], s[1], s[2] in 13 bits each from bit 0, the kind (0 or 1)
// at bit 52, the nflip byte's six bits in bits 53-58.
constexpr int ISO_SHORT_BITS = 13;
__host__ __device__ __forceinline__ bool iso_short(const IsoRec &r, unsigned long long &out) {
  if (r.kind > 1 || r.nflip > 63) return false;
  const int n = r.kind == 0 ? 1 : 3;
  if (r.dst >= (1u << ISO_SHORT_BITS)) return false;
  out = r.dst;
  for (int k = 0; k < n; k++) {
    if (r.s[k] >= (1u << ISO_SHORT_BITS)) return false;
    out |= (unsigned long long)r.s[k] << (ISO_SHORT_BITS * (k + 1));
  }
  out |= (unsigned long long)r.kind << 52 | (unsigned long long)r.nflip << 53;
  return true;
}
__host__ __device__ __forceinline__ IsoRec iso_unshort(unsigned long long w) {
  IsoRec r{};
  const unsigned m = (1u << ISO_SHORT_BITS) - 1;
  r.dst = (unsigned short)(w & m);
  r.kind = (unsigned char)((w >> 52) & 1);
  r.nflip = (unsigned char)((w >> 53) & 63);
  for (int k = 0; k < 3; k++) r.s[k] = (unsigned short)((w >> (ISO_SHORT_BITS * (k + 1))) & m);
  return r;
}

// Tiles of the coarse loading launch. Fine tiles of 2^lx x 2^ly (x 2^lz)
// interior cells (origins 1 + multiples of the extents, so a tile holds
// whole 2^d blocks of children): 32 x 8 in 2D, 16 x 8 x 2 in 3D. The qc ring
// (GetCoarseGhostPoints) is copied by linear tiles over the ring cells of each
// patch's ghost-extended coarse box, numbered by ring_cell.
__host__ __device__ __forceinline__ int cl_lx(int nd) { return nd == 3 ? 4 : 5; }
__host__ __device__ __forceinline__ int cl_ly(int) { return 3; }
__host__ __device__ __forceinline__ int cl_lz(int nd) { return nd == 3 ? 1 : 0; }
// Ring cells of a ghost-extended coarse box of interior cx x cy (x cz): in 3D
// the planes k = 0 and k = cz + 1, then per plane k = 1..cz the rows j = 0 and
// j = cy + 1 and the cells i = 0 and i = cx + 1 of the rows between; in 2D
// the rows and end cells alone (k = 0).
__host__ __device__ __forceinline__ long long ring_count(int nd, int cx, int cy, int cz) {
  const long long X = cx + 2, Y = cy + 2, per = 2 * X + 2LL * cy;
  return nd == 3 ? 2 * X * Y + (long long)cz * per : per;
}
__host__ __device__ __forceinline__ void ring_cell(int nd, int cx, int cy, int cz, long long e,
                                                   int &i, int &j, int &k) {
  const long long X = cx + 2, Y = cy + 2, per = 2 * X + 2LL * cy;
  k = 0;
  if (nd == 3) {
    if (e < 2 * X * Y) {
      k = e < X * Y ? 0 : cz + 1;
      const long long r = e % (X * Y);
      i = (int)(r % X);
      j = (int)(r / X);
      return;
    }
    e -= 2 * X * Y;
    k = (int)(e / per) + 1;
    e %= per;
  }
  if (e < 2 * X) {
    j = e < X ? 0 : cy + 1;
    i = (int)(e % X);
  } else {
    e -= 2 * X;
    j = (int)(e >> 1) + 1;
    i = (e & 1) ? cx + 1 : 0;
  }
}

// A tile of cells of one patch: origin (i0, j0, k0) in the patch's index
// space; the tile extent is fixed per kernel.
struct Tile {
  int patch, i0, j0, k0;
};

// One level of the single-patch coarse chain, for the tail kernels: its patch
// and, per field set, a copy of the patch's record, its three SetGhost
// programs (ipass 0, 1, 2; offsets relative to the patch's q origin) and
// where its arrays live: lds[set][a] is a byte offset into the tail kernel's
// dynamic LDS or -1 for the global pool, for a = q, qc, f, qb and set 0 (FP64
// solution) or 1 (FP32 correction). P[set].qvs is the variable stride of
// that set's q where it lives: the patch's size per variable in LDS, the
// pool's stride in the pool.
struct TailLevel {
  int patch;
  int ngp[3];
  const GhostOp *gp[3];
  int lds[2][4];
  DPatch P[2];
};

// One chunk of an array the tail kernel stages between the global pools and
// LDS: n <= 64 elements (one per lane of a wave) at `glob` and at LDS byte
// offset `lds`.
struct TailCopy {
  void *glob;
  int lds, n;
};

struct Problem {
  double dt, eps2;
  double fc_alpha, fc_shift, fe_num, fe_den;
  double qpo[2];   // quasi-periodic offsets per variable (0: periodic)
};

// The parts of a ghost update's nflip byte (GhostOp): the number of "2 - x"
// maps and the periodic offset code.
__host__ __device__ __forceinline__ int flip_count(int b) { return b & 15; }
__host__ __device__ __forceinline__ int qpo_code(int b) { return (b >> 4) & 3; }
__host__ __device__ __forceinline__ int flip_byte(int nflip, int code) {
  return nflip | code << 4;
}
// The two variables' periodic offsets, passed by value: kernels read them
// from their argument block, and no address of it is taken (an address would
// copy the argument block to scratch memory).
struct Qpo {
  double a, b;
};
// A value with its periodic offset code applied: the reference's
// `q + qpo(n)`, `q - qpo(n)` or `q + 0.0`.
template <class T>
__host__ __device__ __forceinline__ T add_qpo(T val, int code, Qpo qpo, int v) {
  if (code == 0) return val;
  const T q = code == 3 ? (T)0 : (T)(v == 0 ? qpo.a : qpo.b);
  return code == 2 ? val - q : val + q;
}

// Kernel argument bundle: pool base pointers plus the patch table. The FP64
// pools hold the solution; the FP32 pools (mixed precision only) hold the
// correction e, its coarse restriction qce, its right-hand side fe and the
// FP32 copy qb of the solution at which the defect operator is linearised.
// FP32 pools share the offsets of their FP64 counterparts.
struct Dev {
  double *q, *qold, *qc, *f, *rf, *ftmp;
  float *e, *qce, *fe, *qb;
  double *qrte;   // relative truncation error per coarse cell (refined levels)
  int *eflags;    // error flags per fine cell
  const DPatch *P;
  // per group: iteration count, converged, non-finite, stopped; the kernels of
  // a solve skip the patches and ghost updates of a group that has stopped
  const int *gstate;
  Problem prm;
  long long qvs;          // the q pool's variable stride (also e and qb)
  int nrvars;
  int tile_lx, tile_ly;   // level-kernel tiles: 2^tile_lx x 2^tile_ly (x the rest of 256)
  int one_group;          // 1: every patch is in group 0
};

// One workgroup's share of the after-step statistics: rows (2D) or planes
// (3D) r0..r1-1 of a patch's interior (part 0) or of its parent's cells
// under it (part 1).
struct StatTask {
  int patch, part, r0, r1;
};

// A box of cells copied between two q pools (TransferValues): nx*ny*nz cells
// of each of nv variables; cell (i, j, k) of the box, variable v, is the
// source array's cell (si + i, sj + j, sk + k) at s + v*ssv (array of rows of
// spx cells, spy rows per plane), and likewise in the destination.
struct BoxCopy {
  long long s, d, ssv, dsv;
  int nx, ny, nz, nv;
  int si, sj, sk, spx, spy;
  int di, dj, dk, dpx, dpy;
};

// Which field set a level operation acts on.
enum { PREC_SOLUTION_FP64 = 0, PREC_DEFECT_FP32 = 1 };

// Coarse-fine interpolation weights (lib/boundary.f90:45-54).
#define BSAM_A1 (-1.0 / 5.0)
#define BSAM_A2 (2.0 / 3.0)
#define BSAM_A3 (8.0 / 15.0)
#define BSAM_B1 (5.0 / 32.0)
#define BSAM_B2 (15.0 / 16.0)
#define BSAM_B3 (-3.0 / 32.0)
#define BSAM_C1 (9.0 / 16.0)
#define BSAM_C2 (3.0 / 16.0)
#define BSAM_C3 (3.0 / 16.0)
#define BSAM_C4 (1.0 / 16.0)

}  // namespace bsam
