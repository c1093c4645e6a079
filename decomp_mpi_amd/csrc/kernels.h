// Launch interface between the host plan compilers and the HIP kernels.
// Plain structs of device pointers and sizes; no runtime types.
#pragma once
#include <cstdint>
#include <vector>

#include "internal.h"

namespace dm {

// ---- K1: reshape gather ---------------------------------------------------
// Every global coordinate g with a valid owner in both layouts is copied:
//   dst_slot[D(g)] + offset_dst(g)  <-  src_slot[S(g)] + offset_src(g)
// Coordinates are walked as 2D tiles over (a, b) = (dst fastest axis, src
// fastest axis) so that reads and writes are both coalesced; the remaining
// axes are batch.  a == b walks 1D runs along that axis.
struct ReshapeLaunch {
  DevLayout src, dst;
  int ndim;
  int L[DM_MAX_DIMS];
  int a, b;
  int nbatch;
  int batch_axis[DM_MAX_DIMS];
  int tiles_a, tiles_b;
  long long nbatch_total;      // product of the batch axes' lengths
  long long nblocks;
  const char *src_base;   // slot 0 pointer (already offset by the caller's pointer)
  char *dst_base;
  i64 src_slot;           // elements per rank slot
  i64 dst_slot;
  int elem_bytes;
  int nontemporal;            // 1: non-temporal hint on the element accesses
};
// Launchers return 0 or the HIP error code of the launch (hipError_t).
int launch_reshape(const ReshapeLaunch &p, void *stream);

// Stage s of a line of N points: nr = N / radix butterflies per line, ns =
// product of the earlier radices (Stockham span), magics for both.
struct StageInfo {
  unsigned long long m_nr;   // magic for N / radix
  unsigned long long m_ns;   // magic for Ns
  int nr, ns;
};

// ---- K2: batched 1D FFT pass -----------------------------------------------
// Lines along axis t of an ndim array.  With equal strides on every rank
// (uniform) point g lives at element offset sum_d abs[d][g_d] from the base
// pointer; otherwise at R*slot + sum_d strides[R][d]*off[d][g_d] with owner
// R = sum_d rank[d][g_d].  A tile holds T consecutive lines along axis u
// (u < 0: one line per tile) and goes through LDS.
struct FftSide {
  int uniform;
  int proc_used;
  i64 slot;                    // elements per rank slot
  i64 aff;                     // uniform side whose offset along t is x * aff (0: see jump)
  i64 jump, rstride;           // aff 0, jump != 0: runs of rmask + 1 points, offset
  int rsh, rmask;              //   (x >> rsh) * jump + (x & rmask) * rstride; else the table
  const int *rank[DM_MAX_DIMS];
  const int *off[DM_MAX_DIMS];
  const i64 *strides;          // [proc_used * ndim]
};

struct FftLaunch {
  FftSide sin, sout;           // non-uniform addressing (used when !uniform)
  int ndim;
  int L[DM_MAX_DIMS];
  int t, u;
  int T;
  int N;
  int nstage;
  int radix[16];
  int in_row, out_row;        // 1: walk memory along t; 0: along u
  int nbatch;
  int batch_axis[DM_MAX_DIMS];
  int tiles_u;
  int merge_v, merge_k;        // merge_k > 1: a tile holds all L[u] lines for merge_k consecutive
                               //   values of batch axis merge_v (tiles_u 1; merge_k divides L[v])
  long long nblocks;
  const i64 *off_in[DM_MAX_DIMS];
  const i64 *off_out[DM_MAX_DIMS];
  const int *ord_in;          // row mode: t coordinates in input address order (or null)
  const int *ord_out;         // row mode: t coordinates in output address order (or null)
  const int *ord_u;           // tile line l holds u coordinate ord_u[u0 + l] (or u0 + l)
  const i64 *line_in_tab;     // uniform sides: base offset of line l of tile b at [b * T + l]
  const i64 *line_out_tab;    //   (null: computed from the per-axis tables)
  const void *twiddle;        // stage-ordered roots (N 16-byte entries, twiddle_image): stage with
                              //   Ns earlier points and radix R holds w^(jj*r*N/(Ns*R)) at
                              //   (Ns-1) + (r-1)*Ns + jj
  int inverse;                // 0 forward, 1 backward (conjugate roots)
  int dbl;                    // 1: complex double, 0: complex float
  const void *in;
  void *out;
  long long tile_base;        // this launch runs tiles [tile_base, tile_base + nblocks)
  int pass_index;             // host bookkeeping: the pass this launch belongs to
  int stream_in, stream_out;  // 1: non-temporal hint on the data loads / stores
  int ctw;                    // complex float: 1 compensated twiddle products, 0 the float part only
};
// Radices with a stage kernel of their own; any other (a prime above 13)
// runs as a direct p-point sum.
inline bool has_radix_kernel(int r) {
  return r == 2 || r == 3 || r == 4 || r == 5 || r == 7 || r == 8 || r == 11 || r == 13 || r == 16;
}
// Tiles of a whole pass: tiles along u times the batch, the merged axis
// counting in blocks of merge_k.
inline long long fft_pass_tiles(const FftLaunch &p) {
  long long n = p.tiles_u;
  for (int k = 0; k < p.nbatch; k++) {
    const int d = p.batch_axis[k];
    n *= (p.merge_k > 1 && d == p.merge_v) ? p.L[d] / p.merge_k : p.L[d];
  }
  return n;
}
// nblocks == 0: the launch covers every tile of the pass.
int launch_fft_pass(const FftLaunch &p, void *stream);
// Largest points per tile the FFT pass kernel holds for an element size:
// row walks a half tile, column walks the full one, and complex-double
// column walks at a large line stride ("wide") a double one (16 lines of
// 512 points, 256-byte rows).
int fft_tile_capacity(int dbl, int column, int wide = 0);

// ---- K4: twiddle multiply over every rank's local block -------------------
// Rank r's block: count positions per axis at its memory strides; the
// factor of axis d at local position p is tw[d][first + p] (complex double,
// (re, im) pairs), multiplied outer axis to inner; zero products skip the
// slot; conj multiplies by the conjugate.  A workgroup of `block` threads
// takes one chunk of one row (the last axis), or block / count whole rows
// when the rows are shorter; max_blocks: the most workgroups any rank needs.
struct ScaleLaunch {
  void *data;                  // slot 0 of the buffer (pass precision)
  i64 slot;                    // elements per rank slot
  int ndim;
  int nranks;
  const int *rank_tab;         // [nranks][ndim][2]: first table row, count
  const i64 *strides;          // [nranks][ndim]
  const double *tw[DM_MAX_DIMS];
  const double *rows;          // per rank, per row: the outer axes' factors' product
  const long long *row_off;    // [nranks]: rank r's first entry in rows
  long long max_elems;         // largest per-rank block
  long long max_blocks;        // most row chunks of any rank
  int block;                   // threads per workgroup (a multiple of 64, at most 256)
  int dbl, conj;
};
int launch_scale(const ScaleLaunch &p, void *stream);

// ---- K3: batched local FFT pass (many transforms, one axis of each) --------
// A tile holds nl (<= T) lines of N points of one transform: point x of line
// l at element base + l * (row ? N : 1) + x * ps of that transform's array.
// row 2: a plane tile -- the last two axes of one transform at once, nl = T
// rows of N contiguous points (schedule `sched`), then its N columns of T
// points (schedule `sched2`, tile height N) in the same LDS tile.
struct BatchTile {
  long long base;
  long long ps;
  int xform;
  int sched;
  int nl;
  int row;
  int sched2;
};
// Per (line length, tile height, radix order): stages and twiddles.
struct BatchSched {
  int N, T, nstage, tw_off;    // tw_off: start of the stage-ordered roots
  int radix[16];
  unsigned long long m_n, m_t;
  StageInfo st[16];
};
// A segment: the pass over one axis of one transform, described by its
// shape instead of a tile list (the request loop's launches).  The lines of
// N points have point stride `inner`.  Rows (inner 1): `outer` lines of N
// contiguous points, tiles of T consecutive lines.  Columns: for each of
// `outer` coordinates, inner lines at consecutive offsets, `ctiles` tiles of
// T of them.  Tile q of the segment is computed, not looked up.  The
// schedule (one BatchSched, roots at tw_off 0 of `twiddle`) is a device-
// resident entry shared by every transform with that line length and tile
// height.
struct BatchSeg {
  const BatchSched *sched;
  const void *twiddle;
  void *data;
  long long end;               // launch tiles [previous end, end) are this segment's
  long long inner;             // point stride of the axis
  int outer;                   // rows: lines; columns: outer coordinates
  int ctiles;                  // columns: tiles per outer coordinate
  int N, T;                    // the schedule's line length and tile height
};
#define DM_BATCH_MAX_SEG 16
struct BatchLaunch {
  const BatchTile *tiles;
  long long ntiles;
  const BatchSched *sched;
  void *const *data;           // device array: transform i's data pointer
  const void *twiddle;         // all schedules' roots in the pass precision
  int dbl, inverse;
  int primes;                  // some schedule has a radix 5, 7, 11 or 13
  int long_lines;              // some schedule's lines are too long for twiddles in LDS
  int plane;                   // plane tiles (row 2) only: 2 within the line tile's points, 1 larger
  int any;                     // some schedule has a direct-sum stage (line tiles only)
  long long grid;              // persistent workgroups (tile loop inside)
  size_t lds;
  int nseg;                    // > 0: tiles come from seg[] (tiles, sched, data, twiddle unused)
  BatchSeg seg[DM_BATCH_MAX_SEG];
};
int launch_batch_pass(const BatchLaunch &p, void *stream);
int launch_copy16(const void *src, void *dst, long long n16, void *stream);
// Host tables of a batch (batch.cpp): tiles[pass] (last axis first),
// schedules, their stage-ordered roots as complex doubles, LDS bytes.
struct BatchHost {
  std::vector<std::vector<BatchTile>> tiles;
  // plane tiles (run before tiles[0]): [0] planes within the line tile's
  // points (the line kernel's shape), [1] larger ones
  std::vector<BatchTile> planes[2];
  size_t plane_lds[2] = {0, 0};
  int plane_primes[2] = {0, 0};
  std::vector<BatchSched> sched;
  std::vector<double> tw;
  size_t lds = 0;
  int primes = 0;
  int long_lines = 0;
  int any = 0;  // some line schedule has a direct-sum stage (a prime above 13)
};
bool build_batch_host(int count, int ndim, const int *shapes, int dbl, int tile_choice, int radix_order,
                      BatchHost *out);
// One schedule: lines of N points (radices 2..16, 3, 5, 7, 11, 13 by
// butterfly kernels, larger primes by direct sums) in tiles of T lines,
// tw_off 0; tw: its N stage-ordered roots as (re, im) doubles, then the p
// roots of each direct-sum stage in stage order.
void make_batch_sched(int N, int T, int radix_order, BatchSched *out, std::vector<double> *tw);
// Whether a schedule has a direct-sum stage.
bool batch_sched_any(const BatchSched &S);
int batch_tile_capacity(int dbl);
int batch_tile_height(int N, long long inner, int dbl, int tile_choice);
// LDS of a tile of T lines of N points, with the schedule's twiddles when
// batch_tw_in_lds(N) (a launch copies them only if that holds for all its
// schedules: BatchLaunch::long_lines 0).
size_t batch_lds_bytes(int dbl, int T, int N);
bool batch_tw_in_lds(int N);
long long batch_resident_groups(int dbl, int primes, size_t lds);
// Plane tiles (BatchTile row 2): the most points a plane may hold, the LDS
// bytes of a plane of T rows of N points (0 when it does not fit: too many
// points, a side above the LDS-twiddle limit, or over 160 KiB), and the
// resident workgroups of the large-plane kernel (planes above the line
// tile's points) for a launch's LDS.
int batch_plane_capacity(int dbl);
size_t batch_plane_lds_bytes(int dbl, int T, int N);
long long batch_plane_resident_groups(size_t lds);

}  // namespace dm
