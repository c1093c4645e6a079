// Split (four-step) parallel FFT: planner in the reference's terms, executed
// as a device pass program.
//
// Reference: split_fft.c (planner 36-77, 249-614, 747-788; executor 622-744;
// search 78-247), dimscale.c, processor_grid.c.
//
// Execution model.  The forward transform maps the start decomposition (in
// dat) to the output decomposition (work[0:extent), B layout with frequency k
// of axis d at B-position k / B_d + A_d * (k % B_d)).  Mathematically this is
// the multi-dimensional DFT with both ends given by layouts; the device
// program runs one batched FFT pass per axis whose loads and stores go
// through the layouts' per-axis offset tables, so the reference's
// start->A reshape, twiddle multiply and A->B reshape need no passes of
// their own.  Intermediate data live in the output layout itself (any
// bijection is a valid intermediate layout), which keeps every pass
// coalesced: the start layout's fastest axis is transformed first (forward)
// or last (backward), the other passes walk rows of consecutive output
// positions.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <numeric>
#include <string>

#include "internal.h"
#include "kernels.h"
#include "fft_program.h"
#include <functional>
#include "block_decomp.h"
#include "interleave_decomp.h"
#include "factor_reverse_decomp.h"
#include "processor_grid.h"
#include "logger.h"

namespace dm {
bool enqueue_reshape(dm_comm_t comm, ReshapePlan plan, const void *src, void *dst, void *stream);
}

// ---------------------------------------------------------------- processor grid
// processor_grid.c:24-97 (same enumeration order, score and tie-break).

static void search_grid(GridResult best, int ndim, int mydim, int *grid, const int *tasks, int max_proc,
                        int proc, int load, int mn, int mx) {
  int newproc = proc;
  for (grid[mydim] = 1; grid[mydim] <= max_proc && grid[mydim] <= tasks[mydim] && newproc <= max_proc;
       grid[mydim]++) {
    newproc = proc * grid[mydim];
    if (newproc > max_proc) continue;
    int lo = tasks[mydim] / grid[mydim];
    int hi = (tasks[mydim] + grid[mydim] - 1) / grid[mydim];
    int newload = load * hi;
    if (mn < lo) lo = mn;
    if (mx > hi) hi = mx;
    if (mydim < ndim - 1) {
      search_grid(best, ndim, mydim + 1, grid, tasks, max_proc, newproc, newload, lo, hi);
    } else {
      double ratio = (double)hi / (double)lo;
      if (newload <= best->load || best->load == 0) {
        if (best->load == 0 || newload < best->load || ratio < best->asym) {
          best->procs = newproc;
          best->load = newload;
          best->asym = ratio;
          for (int i = 0; i < ndim; i++) best->length[i] = grid[i];
        }
      }
    }
  }
}

extern "C" int *makeProcessorGrid(MPI_Comm comm, int ndim, int tasks[]) {
  struct grid_result gr;
  std::vector<int> len(ndim, 1);
  gr.ndim = ndim;
  gr.length = len.data();
  gr.procs = 1;
  gr.load = 1;
  gr.asym = 10000.0;
  int mx = 1;
  for (int i = 0; i < ndim; i++) {
    gr.load *= tasks[i];
    mx = std::max(mx, tasks[i]);
  }
  int *grid = (int *)calloc(ndim, sizeof(int));
  search_grid(&gr, ndim, 0, grid, tasks, comm->world->nranks, 1, 1, mx, 0);
  for (int i = 0; i < ndim; i++) grid[i] = gr.length[i];
  return grid;
}

// split_fft.c:890-926 (any gcd algorithm gives the same value)
extern "C" unsigned int gcd(unsigned int u, unsigned int v) {
  while (v != 0) {
    unsigned int t = u % v;
    u = v;
    v = t;
  }
  return u;
}

// ---------------------------------------------------------------- decomps

extern "C" Decomp makeInterleavedDecomp(int ndim, MPI_Comm comm, int A[], int B[], int grid[]) {
  std::vector<DimDecomp> ax(ndim);
  for (int i = 0; i < ndim; i++) {
    DimDecomp blocked = makeBlockDecomp(A[i], grid[i]);
    ax[i] = makeInterleaveDecomp(blocked, B[i]);
    freeDimDecomp(blocked);  // the interleave holds its own clone
  }
  return makePackedDecomp(comm, ndim, ax.data());
}

extern "C" Decomp makeBlockedDecomp(int ndim, MPI_Comm comm, int A[], int B[], int grid[]) {
  std::vector<DimDecomp> ax(ndim);
  for (int i = 0; i < ndim; i++) {
    ax[i] = makeBlockDecompWithFactor(A[i] * B[i], A[i], grid[i]);
    if (ax[i] == nullptr) {
      for (int j = 0; j < i; j++) freeDimDecomp(ax[j]);
      return nullptr;
    }
  }
  return makePackedDecomp(comm, ndim, ax.data());
}

// ---------------------------------------------------------------- dimscale
// dimscale.c:7-67: multiply every slot by the product of its per-axis
// factors (walked outer axis to inner); a zero partial product marks an
// unused slot and skips it.  sign +1 conjugates the product.

static void dim_scale(int dim, Decomp d, _Complex double factor, _Complex double **factors,
                      _Complex double *data, long long offset, int sign) {
  AxisDecomp a = d->dims[dim];
  int count = a->max_local + 1;
  for (int pos = 0; pos < count; pos++) {
    long long o = offset + (long long)pos * a->memory_stride;
    _Complex double f = factor;
    if (factors != nullptr && factors[dim] != nullptr) f = factor * factors[dim][pos];
    if (f == 0.0) continue;
    if (dim == d->ndim - 1) {
      if (sign == 1) f = __builtin_conj(f);
      data[o] = f * data[o];
    } else {
      dim_scale(dim + 1, d, f, factors, data, o, sign);
    }
  }
}

extern "C" void decomp_scale(Decomp decomp, _Complex double **factors, _Complex double *data, int sign) {
  dim_scale(0, decomp, 1.0, factors, data, 0, sign);
}

// split_fft.c:862-889 walks every slot of the local bounding box.  Here the
// visitor also receives the slot's global coordinates (-1 on axes where the
// slot holds no coordinate), which the reference leaves unset.
static void dim_visit(int dim, int *gpos, Decomp d, data_visitor vis, _Complex double *data, long long offset,
                      const std::vector<std::vector<int>> &inv) {
  AxisDecomp a = d->dims[dim];
  int count = a->max_local + 1;
  for (int pos = 0; pos < count; pos++) {
    gpos[dim] = inv[dim][pos];
    long long o = offset + (long long)pos * a->memory_stride;
    if (dim == d->ndim - 1) vis(d->ndim, gpos, data + o);
    else dim_visit(dim + 1, gpos, d, vis, data, o, inv);
  }
}

extern "C" void visitData(Decomp d, data_visitor vis, _Complex double *dat) {
  std::vector<std::vector<int>> inv(d->ndim);
  for (int i = 0; i < d->ndim; i++) {
    AxisDecomp a = d->dims[i];
    inv[i].assign(a->max_local + 1, -1);
    for (int g = 0; g < a->decomp->coord_size; g++)
      if (isLocal(a, g)) {
        int o = a->decomp->type_offset(a->decomp, g);
        if (o >= 0 && o <= a->max_local) inv[i][o] = g;
      }
  }
  std::vector<int> gpos(d->ndim, 0);
  dim_visit(0, gpos.data(), d, vis, dat, 0, inv);
}

// ---------------------------------------------------------------- device program

struct dm_fft_program {
  int inverse = 0;
  std::vector<dm::FftLaunch> passes;
  void *mem = nullptr;  // device tables
};

// The reference's phase order on the device: reshape (K1), stage DFTs over
// sub-axes (K2), twiddle multiply (K4), reshape, stage DFTs.
struct dm_phased_program {
  bool compiled = false;
  dm_fft_program st1_fwd, st1_bwd, st2_fwd, st2_bwd;  // stage passes, in place
  void *tw_mem = nullptr;   // twiddle tables and the per-rank table of K4
  bool has_twiddle = false;
  dm::ScaleLaunch scale{};
  ReshapePlan a_to_b = nullptr, b_to_a = nullptr;  // owned: plans whose A_to_B the reference leaves NULL
};

struct dm_split_ext {
  int dbl = 1;
  int ndim = 0;
  std::vector<int> A, B, L;
  Decomp start = nullptr;   // caller's (aliased)
  Decomp out = nullptr;     // B_flipped, or A for single-stage plans (owned by the plan)
  bool use_first = false;   // single stage: output in the A decomposition
  bool compiled = false;
  dm_fft_program fwd, bwd;
  bool phased = false;      // run in the reference's phase order
  bool fused_ok = true;     // every axis fits one pass (else phased only)
  bool custom_out = false;  // out is the caller's (dm_split_set_output)
  int ctw = 1;              // complex float: compensated twiddle products
  dm_phased_program ph;
};

static void program_release(dm_fft_program &p) {
  if (p.mem) dm::device_free(p.mem);
  p.mem = nullptr;
  p.passes.clear();
}

static void phased_release(dm_phased_program &ph) {
  program_release(ph.st1_fwd);
  program_release(ph.st1_bwd);
  program_release(ph.st2_fwd);
  program_release(ph.st2_bwd);
  if (ph.tw_mem) dm::device_free(ph.tw_mem);
  ph.tw_mem = nullptr;
  freeReshapePlan(ph.a_to_b);
  freeReshapePlan(ph.b_to_a);
  ph.a_to_b = ph.b_to_a = nullptr;
  ph.compiled = false;
}

// split_fft.c:373-393: w = exp(-2 pi i y l / L) with the quarter turns exact.
static void twiddle_value(int y, int l, int L, double *re, double *im) {
  long long rem = ((long long)y * l) % L;
  if (rem == 0) *re = 1.0, *im = 0.0;
  else if (4 * rem == L) *re = 0.0, *im = -1.0;
  else if (2 * rem == L) *re = -1.0, *im = 0.0;
  else if (4 * rem == 3LL * L) *re = -0.0, *im = 1.0;
  else {
    double angle = -(2.0 * M_PI) * ((double)((long long)y * l) / (double)L);
    *re = cos(angle);
    *im = sin(angle);
  }
}

// A dense intermediate layout: axes in `order` (fastest first), the slowest
// axis split into q equal blocks, one per rank slot (q the largest divisor
// of its length not above P).  False when a rank's block would not fit the
// region of `region` elements per slot.
static bool dense_layout(const std::vector<int> &L, const std::vector<int> &order, int P, i64 region,
                         dm::HostLayout *out) {
  const int nd = (int)L.size();
  const int slow = order.back();
  int q = 1;
  for (int c = std::min(P, L[slow]); c >= 1; c--)
    if (L[slow] % c == 0) {
      q = c;
      break;
    }
  i64 total = 1;
  for (int l : L) total *= l;
  const i64 per = total / q;
  if (per > region) return false;
  dm::HostLayout h;
  h.ndim = nd;
  h.L = L;
  h.proc_used = q;
  h.rank_part.assign(nd, std::vector<int>());
  h.off.assign(nd, std::vector<int>());
  h.m.assign(nd, 0);
  const int blk = L[slow] / q;
  i64 stride = 1;
  for (int k = 0; k < nd; k++) {
    int d = order[k];
    h.m[d] = stride;
    stride *= (d == slow) ? blk : L[d];
  }
  for (int d = 0; d < nd; d++) {
    h.rank_part[d].assign(L[d], 0);
    h.off[d].resize(L[d]);
    for (int g = 0; g < L[d]; g++) {
      if (d == slow) {
        h.rank_part[d][g] = g / blk;
        h.off[d][g] = g % blk;
      } else {
        h.off[d][g] = g;
      }
    }
  }
  h.strides.clear();
  for (int r = 0; r < q; r++) h.strides.insert(h.strides.end(), h.m.begin(), h.m.end());
  h.uniform = true;
  h.slot_extent.assign(q, per);
  h.fastest = order[0];
  h.offsets_ok = true;
  *out = h;
  return true;
}

// Build one direction's passes.  Forward: dat (start) -> ... -> W0
// (output); backward: W0 (output) -> ... -> dat (start).  Each direction
// orders its own passes the same way: the axis fastest in memory on the
// side it reads first, then the others from fast to slow on the side it
// writes last.  With dense intermediates (below) every pass then reads
// whole contiguous lines and writes rows of T elements at the stride of a
// line: the forward's and the backward's passes have the same access
// pattern.  Without them, the intermediates stay in the output layout and
// run in place (every tile reads all of its points before writing them,
// tiles are disjoint, and such a pass reads and writes one layout).
static bool compile_direction(SplitPlan plan, bool inverse, const dm::Side &S, const dm::Side &O, void *dat,
                              void *w0, void *w1, dm_fft_program *prog) {
  dm_split_ext *x = plan->ext;
  int nd = x->ndim;
  std::vector<int> axes;
  for (int d = 0; d < nd; d++)
    if (x->L[d] > 1) axes.push_back(d);
  if (axes.empty()) axes.push_back(0);
  static const int dense_env = [] {
    const char *e = getenv("DM_DENSE_INTERMEDIATES");
    return e ? atoi(e) : 1;
  }();
  const int n = (int)axes.size();
  bool dense = dense_env != 0 && n >= 2;
  // pass order: dense -- the first-read side's fastest axis, then the rest
  // by the last-written side's strides; otherwise the forward order
  // (start's fastest axis first, the rest by output strides), reversed for
  // the backward
  auto make_order = [&](const dm::Side &first, const dm::Side &last) {
    const int f = first.h.fastest;
    std::vector<int> ord;
    if (std::find(axes.begin(), axes.end(), f) != axes.end()) ord.push_back(f);
    std::vector<int> rest;
    for (int d : axes)
      if (d != f) rest.push_back(d);
    std::sort(rest.begin(), rest.end(), [&](int a, int b) { return last.h.m[a] < last.h.m[b]; });
    ord.insert(ord.end(), rest.begin(), rest.end());
    return ord;
  };
  auto fallback_order = [&]() {
    std::vector<int> ord = make_order(S, O);
    if (inverse) std::reverse(ord.begin(), ord.end());
    return ord;
  };
  const dm::Side &last = inverse ? S : O;
  std::vector<int> order = dense ? (inverse ? make_order(O, S) : make_order(S, O)) : fallback_order();
  // Dense intermediates: after the pass over axis order[i] the data are
  // stored with the next pass's axis fastest and this pass's axis next, so
  // the next pass reads whole lines; they alternate between the workspace
  // halves, ending in W0 (forward) or leaving W0 first (backward).
  std::vector<dm::Side> mids;
  if (dense) {
    const i64 region = (i64)plan->extent;
    for (int i = 0; i + 1 < n && dense; i++) {
      int cur = order[i], next = order[i + 1];
      std::vector<int> ord = {next, cur};
      std::vector<int> others;
      for (int d = 0; d < nd; d++)
        if (d != cur && d != next) others.push_back(d);
      std::sort(others.begin(), others.end(), [&](int a, int b) { return last.h.m[a] < last.h.m[b]; });
      ord.insert(ord.end(), others.begin(), others.end());
      dm::HostLayout h;
      if (!dense_layout(x->L, ord, plan->comm->world->nranks, region, &h)) dense = false;
      else mids.push_back(dm::make_side(h, O.slot));
    }
    if (!dense) {  // a block did not fit: the in-place scheme
      mids.clear();
      order = fallback_order();
    }
  }
  dm::TableBuilder tb;
  std::vector<dm::PendingPass> pp(n);
  void *prev = nullptr;
  for (int i = 0; i < n; i++) {
    void *bin, *bout;
    const dm::Side *si, *so;
    if (dense) {
      if (!inverse) {
        // intermediate k in W1 when n - 2 - k is even, else W0, so that the
        // last pass writes W0 from the other half
        auto mid_buf = [&](int k) { return ((n - 2 - k) % 2 == 0) ? w1 : w0; };
        bin = (i == 0) ? dat : prev;
        si = (i == 0) ? &S : &mids[i - 1];
        bout = (i == n - 1) ? w0 : mid_buf(i);
        so = (i == n - 1) ? &O : &mids[i];
      } else {
        // intermediate k in W1 when k is even, else W0: the first pass
        // leaves W0, which holds its input
        bin = (i == 0) ? w0 : prev;
        si = (i == 0) ? &O : &mids[i - 1];
        bout = (i == n - 1) ? dat : ((i % 2 == 0) ? w1 : w0);
        so = (i == n - 1) ? &S : &mids[i];
      }
    } else if (!inverse) {
      bout = w0;
      bin = (i == 0) ? dat : prev;
      si = (i == 0) ? &S : &O;
      so = &O;
    } else {
      bin = (i == 0) ? w0 : prev;
      bout = (i == n - 1) ? dat : w0;
      si = &O;
      so = (i == n - 1) ? &S : &O;
    }
    if (!dm::plan_pass(plan->comm, *si, *so, order[i], x->dbl, inverse, bin, bout, tb, &pp[i], x->ctw)) return false;
    prev = bout;
  }
  prog->mem = dm::finalize_passes(plan->comm, tb, pp, &prog->passes);
  prog->inverse = inverse ? 1 : 0;
  if (prog->mem == nullptr) return false;
  for (size_t i = 0; i < prog->passes.size(); i++) prog->passes[i].pass_index = (int)i;
  return true;
}

// The plan's dat and workspace as registered buffers, checked against the
// decompositions they hold.
static bool bind_buffers(SplitPlan plan, dm::Buffer *bd, size_t *od, dm::Buffer *bw, size_t *ow) {
  dm_split_ext *x = plan->ext;
  dm_comm_t comm = plan->comm;
  size_t es = x->dbl ? 16 : 8;
  if (!dm::resolve(plan->src, bd, od) || !dm::resolve(plan->work, bw, ow)) {
    raise_error(comm, "split FFT: dat and workspace must point into distributed device buffers "
                      "(dm_malloc/dm_register)");
    return false;
  }
  if (*od % es || *ow % es || bd->slot % es || bw->slot % es) {
    raise_error(comm, "split FFT: buffers must be aligned to the element size");
    return false;
  }
  dm::HostLayout hs = dm::layout_from_decomp(x->start);
  dm::HostLayout ho = dm::layout_from_decomp(x->out);
  if (bd->nranks < hs.proc_used || bw->nranks < ho.proc_used) {
    raise_error(comm, "split FFT: buffers have fewer rank slots than the decompositions use");
    return false;
  }
  size_t ext = (size_t)plan->extent;
  if (*ow + 2 * ext * es > bw->slot) {
    raise_error(comm, "split FFT: workspace slot of %zu bytes cannot hold 2 x %zu elements", bw->slot - *ow, ext);
    return false;
  }
  if (!dm::layout_usable(comm, "start", hs, (bd->slot - *od) / es) || !dm::layout_usable(comm, "output", ho, ext))
    return false;
  return true;
}

static bool compile_plan(SplitPlan plan) {
  dm_split_ext *x = plan->ext;
  if (x->compiled) return true;
  if (!x->fused_ok) {
    raise_error(plan->comm, "split FFT: an axis is longer than one pass tile; this plan runs in phase order only");
    return false;
  }
  dm_comm_t comm = plan->comm;
  size_t es = x->dbl ? 16 : 8;
  dm::Buffer bd, bw;
  size_t od, ow;
  if (!bind_buffers(plan, &bd, &od, &bw, &ow)) return false;
  size_t ext = (size_t)plan->extent;
  dm::HostLayout hs = dm::layout_from_decomp(x->start);
  dm::HostLayout ho = dm::layout_from_decomp(x->out);
  dm::Side S = dm::make_side(hs, (i64)(bd.slot / es));
  dm::Side O = dm::make_side(ho, (i64)(bw.slot / es));
  char *dat = bd.base + od;
  char *w0 = bw.base + ow;
  char *w1 = w0 + ext * es;
  if (!compile_direction(plan, false, S, O, dat, w0, w1, &x->fwd)) return false;
  if (!compile_direction(plan, true, S, O, dat, w0, w1, &x->bwd)) return false;
  plan->plan_a = &x->fwd;
  plan->plan_b_back = &x->bwd;
  x->compiled = true;
  return true;
}

// ---------------------------------------------------------------- phase order
// The reference's runSplitPlan phases (split_fft.c:622-682), each its own
// device work: RESHAPE_1 start -> A (K1), FFT_1 the B_d-point DFTs along
// the z sub-axes of the A layout (K2, in place), TWIDDLE the multiply by
// the plan's twiddle tables (K4), RESHAPE_2 A -> B (K1), FFT_2 the
// A_d-point DFTs along the y sub-axes of the B layout (K2, in place).  The
// result lands in work[0:extent) in the B layout, as the fused program's.
// The backward runs the exact inverse in reverse order (FFT_2, RESHAPE_2
// B -> A, conjugate TWIDDLE, FFT_1, RESHAPE_1 A -> start).  Each phase is
// timed into plan->times under its own name.

// Stage passes, in place on buf: for each axis d in `axes` the DFT along
// one factor of d -- the z part (B_d points) for stage 1, the y part (A_d
// points) for stage 2 -- on a view of the layout in which only axis d is
// split, so that a tile's lines can run along a whole other axis (long
// contiguous rows, many lines per tile).  Fastest axis first.
static bool compile_stage(SplitPlan plan, const dm::HostLayout &h, std::vector<int> axes, bool stage1,
                          bool inverse, void *buf, i64 slot, dm_fft_program *prog) {
  dm_split_ext *x = plan->ext;
  prog->passes.clear();
  prog->mem = nullptr;
  prog->inverse = inverse ? 1 : 0;
  if (axes.empty()) return true;
  std::stable_sort(axes.begin(), axes.end(), [&](int a, int b) { return h.m[a] < h.m[b]; });
  dm::TableBuilder tb;
  std::vector<dm::PendingPass> pp(axes.size());
  for (size_t i = 0; i < axes.size(); i++) {
    const int d = axes[i];
    std::vector<int> f(x->ndim, 1);
    f[d] = x->A[d];
    dm::HostLayout sh;
    std::string why;
    if (!dm::split_layout(h, f, &sh, &why)) {
      raise_error(plan->comm, "split FFT: axis %d of a stage layout: %s", d, why.c_str());
      return false;
    }
    dm::Side side = dm::make_side(sh, slot);
    const int sub = stage1 ? 2 * d + 1 : 2 * d;
    if (!dm::plan_pass(plan->comm, side, side, sub, x->dbl, inverse, buf, buf, tb, &pp[i], x->ctw)) return false;
  }
  prog->mem = dm::finalize_passes(plan->comm, tb, pp, &prog->passes);
  return prog->mem != nullptr;
}

static bool compile_phased(SplitPlan plan) {
  dm_split_ext *x = plan->ext;
  dm_phased_program &ph = x->ph;
  if (ph.compiled) return true;
  dm_comm_t comm = plan->comm;
  const int nd = x->ndim;
  const size_t es = x->dbl ? 16 : 8;
  dm::Buffer bd, bw;
  size_t od, ow;
  if (!bind_buffers(plan, &bd, &od, &bw, &ow)) return false;
  char *w0 = bw.base + ow;
  char *w1 = w0 + (size_t)plan->extent * es;
  const bool two = !x->use_first;
  dm::HostLayout hA = dm::layout_from_decomp(plan->A);
  if (!dm::layout_usable(comm, "A", hA, (size_t)plan->extent)) return false;
  const i64 wslot = (i64)(bw.slot / es);
  std::vector<int> z_axes, y_axes;
  for (int d = 0; d < nd; d++) {
    if (x->B[d] > 1) z_axes.push_back(d);
    if (x->A[d] > 1) y_axes.push_back(d);
  }
  void *st1_buf = two ? w1 : w0;
  if (!compile_stage(plan, hA, z_axes, true, false, st1_buf, wslot, &ph.st1_fwd) ||
      !compile_stage(plan, hA, z_axes, true, true, st1_buf, wslot, &ph.st1_bwd))
    return false;
  if (two) {
    dm::HostLayout hB = dm::layout_from_decomp(plan->B);
    if (!dm::layout_usable(comm, "B", hB, (size_t)plan->extent)) return false;
    if (!compile_stage(plan, hB, y_axes, false, false, w0, wslot, &ph.st2_fwd) ||
        !compile_stage(plan, hB, y_axes, false, true, w0, wslot, &ph.st2_bwd))
      return false;
    if (plan->A_to_B == nullptr) {  // single second stage: the reference has no A -> B plan
      MPI_Datatype base = x->dbl ? &dm_type_c_double_complex : &dm_type_c_float_complex;
      ph.a_to_b = makeReshapePlan(comm, base, plan->A, plan->B);
      ph.b_to_a = makeReshapePlan(comm, base, plan->B, plan->A);
      if (ph.a_to_b == nullptr || ph.b_to_a == nullptr) return false;
    }
  }
  // twiddle tables of every rank position, rows in position order, and the
  // per-rank (first row, count) table with the A layout's strides
  ph.has_twiddle = two && plan->twiddle != nullptr;
  if (ph.has_twiddle) {
    std::vector<std::vector<double>> tabs(nd);
    std::vector<std::vector<int>> first(nd);
    for (int d = 0; d < nd; d++) {
      AxisDecomp ax = plan->A->dims[d];
      const int ps = ax->decomp->proc_size;
      first[d].assign(ps + 1, 0);
      for (int p = 0; p < ps; p++) first[d][p + 1] = first[d][p] + ax->ext->max_local_pos[p] + 1;
      tabs[d].assign(2 * (size_t)first[d][ps], 0.0);
      for (int g = 0; g < x->L[d]; g++) {
        const int p = ax->ext->rank_of[g], local = ax->ext->off_of[g];
        if (p < 0 || p >= ps || local < 0 || local > ax->ext->max_local_pos[p]) continue;
        double re, im;
        twiddle_value(g % x->A[d], g / x->A[d], x->L[d], &re, &im);
        tabs[d][2 * (size_t)(first[d][p] + local)] = re;
        tabs[d][2 * (size_t)(first[d][p] + local) + 1] = im;
      }
    }
    const int nr = plan->A->proc_used;
    std::vector<int> rt((size_t)nr * nd * 2);
    long long max_elems = 0;
    for (int r = 0; r < nr; r++) {
      long long n = 1;
      for (int d = 0; d < nd; d++) {
        AxisDecomp ax = plan->A->dims[d];
        const int pos = (r / ax->rank_stride) % ax->decomp->proc_size;
        rt[((size_t)r * nd + d) * 2] = first[d][pos];
        rt[((size_t)r * nd + d) * 2 + 1] = ax->ext->max_local_pos[pos] + 1;
        n *= ax->ext->max_local_pos[pos] + 1;
      }
      max_elems = std::max(max_elems, n);
    }
    // 256-thread workgroups: one chunk of one row (the last axis) each, or
    // several whole rows when the rows are shorter (the kernel's mapping)
    const int block = 256;
    long long max_blocks = 0;
    for (int r = 0; r < nr; r++) {
      long long rows = 1;
      for (int d = 0; d < nd - 1; d++) rows *= rt[((size_t)r * nd + d) * 2 + 1];
      const long long inner = rt[((size_t)r * nd + nd - 1) * 2 + 1];
      if (rows <= 0 || inner <= 0) continue;  // a rank with an empty block
      const long long need = inner >= block ? rows * ((inner + block - 1) / block)
                                            : (rows + block / inner - 1) / (block / inner);
      max_blocks = std::max(max_blocks, need);
    }
    // per rank, the product of the outer axes' factors for every row (the
    // last axis's positions share it), formed outer axis first as the
    // reference's recursion forms it
    std::vector<double> rows_tab;
    std::vector<long long> row_off(nr, 0);
    for (int r = 0; r < nr; r++) {
      long long rows = 1;
      for (int d = 0; d < nd - 1; d++) rows *= rt[((size_t)r * nd + d) * 2 + 1];
      if (rt[((size_t)r * nd + nd - 1) * 2 + 1] <= 0) rows = 0;
      row_off[r] = (long long)(rows_tab.size() / 2);
      for (long long row = 0; row < rows; row++) {
        long long rest = row;
        int loc[DM_MAX_DIMS];
        for (int d = nd - 2; d >= 0; d--) {
          const int c = rt[((size_t)r * nd + d) * 2 + 1];
          loc[d] = (int)(rest % c);
          rest /= c;
        }
        double fr = 1.0, fi = 0.0;
        for (int d = 0; d < nd - 1; d++) {
          const double *t = &tabs[d][2 * (size_t)(rt[((size_t)r * nd + d) * 2] + loc[d])];
          const double nr_ = fr * t[0] - fi * t[1];
          fi = fr * t[1] + fi * t[0];
          fr = nr_;
        }
        rows_tab.push_back(fr);
        rows_tab.push_back(fi);
      }
    }
    if (rows_tab.empty()) rows_tab.assign(2, 0.0);
    dm::TableBuilder tb;
    std::vector<size_t> o_tab(nd);
    for (int d = 0; d < nd; d++) o_tab[d] = tb.put(tabs[d].data(), tabs[d].size() * sizeof(double));
    const size_t o_rows = tb.put(rows_tab.data(), rows_tab.size() * sizeof(double));
    const size_t o_roff = tb.put(row_off.data(), row_off.size() * sizeof(long long));
    const size_t o_rt = tb.put(rt.data(), rt.size() * sizeof(int));
    const size_t o_st = tb.put(hA.strides.data(), hA.strides.size() * sizeof(i64));
    ph.tw_mem = dm::device_alloc(tb.host.size());
    if (ph.tw_mem == nullptr) {
      raise_error(comm, "split FFT: device allocation of %zu bytes of twiddle tables failed", tb.host.size());
      return false;
    }
    dm::h2d(ph.tw_mem, tb.host.data(), tb.host.size());
    char *b = (char *)ph.tw_mem;
    dm::ScaleLaunch &L = ph.scale;
    L = dm::ScaleLaunch{};
    L.data = w1;
    L.slot = (i64)(bw.slot / es);
    L.ndim = nd;
    L.nranks = nr;
    L.rank_tab = (const int *)(b + o_rt);
    L.strides = (const i64 *)(b + o_st);
    for (int d = 0; d < nd; d++) L.tw[d] = (const double *)(b + o_tab[d]);
    L.rows = (const double *)(b + o_rows);
    L.row_off = (const long long *)(b + o_roff);
    L.max_elems = max_elems;
    L.max_blocks = max_blocks;
    L.block = block;
    L.dbl = x->dbl;
  }
  ph.compiled = true;
  return true;
}

// Run one direction in phase order; times[] gets each phase's HIP-event
// time when the world is synchronous.
static void run_phased(SplitPlan plan, int direction) {
  dm_split_ext *x = plan->ext;
  dm_phased_program &ph = x->ph;
  dm_world *w = plan->comm->world;
  void *s = dm::world_stream(w);
  const bool timed = !w->async;
  const size_t es = x->dbl ? 16 : 8;
  const bool two = !x->use_first;
  char *w0 = (char *)plan->work;
  char *w1 = w0 + (size_t)plan->extent * es;
  ReshapePlan a_to_b = plan->A_to_B ? plan->A_to_B : ph.a_to_b;
  ReshapePlan b_to_a = plan->B_to_A ? plan->B_to_A : ph.b_to_a;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (timed) {
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
  }
  bool ok = true;
  double total = 0;
  auto phase = [&](int timer, const std::function<bool()> &work) {
    if (!ok) return;
    if (timed) (void)hipEventRecord(e0, (hipStream_t)s);
    ok = work();
    if (!timed || !ok) return;
    (void)hipEventRecord(e1, (hipStream_t)s);
    (void)hipEventSynchronize(e1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    plan->times[timer] += ms * 1e-3;
    total += ms * 1e-3;
  };
  auto passes = [&](dm_fft_program &prog) {
    return [&]() {
      for (auto &L : prog.passes)
        if (!dm::launch_ok(plan->comm, dm::launch_fft_pass(L, s), "FFT pass launch")) return false;
      return true;
    };
  };
  auto scale = [&](int conj) {
    return [&, conj]() {
      dm::ScaleLaunch L = ph.scale;
      L.conj = conj;
      return dm::launch_ok(plan->comm, dm::launch_scale(L, s), "twiddle kernel launch");
    };
  };
  void *st1_buf = two ? w1 : w0;
  if (direction == FFTW_FORWARD) {
    phase(RESHAPE_1, [&]() { return dm::enqueue_reshape(plan->comm, plan->start_to_A, plan->src, st1_buf, s); });
    phase(FFT_1, passes(ph.st1_fwd));
    if (ph.has_twiddle) phase(TWIDDLE, scale(0));
    if (two) {
      phase(RESHAPE_2, [&]() { return dm::enqueue_reshape(plan->comm, a_to_b, w1, w0, s); });
      phase(FFT_2, passes(ph.st2_fwd));
    }
  } else {
    if (two) {
      phase(FFT_2, passes(ph.st2_bwd));
      phase(RESHAPE_2, [&]() { return dm::enqueue_reshape(plan->comm, b_to_a, w0, w1, s); });
    }
    if (ph.has_twiddle) phase(TWIDDLE, scale(1));
    phase(FFT_1, passes(ph.st1_bwd));
    phase(RESHAPE_1, [&]() { return dm::enqueue_reshape(plan->comm, plan->A_to_start, st1_buf, plan->src, s); });
  }
  if (timed) {
    plan->times[ALL] += total;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
}

// ---------------------------------------------------------------- planner

// Why the calling thread's last plan was rejected for an unsupported axis
// (added to makeSplitPlan's "Cannot generate plan" message).
static thread_local std::string t_rejection;

static SplitPlan build_plan(int dbl, int flags, int use_first, int use_second, Decomp start, int ndim, int A[],
                            int B[], int L[], int grid_a[], int grid_b[], void *dat, int length, void *workspace) {
  if ((flags & FORCE_USE_FIRST) && !use_first) return nullptr;
  MPI_Comm comm = start->comm;
  if (ndim != start->ndim) {
    raise_error(comm, "buildParallelPlan: ndim %d differs from the start decomposition's %d", ndim, start->ndim);
    return nullptr;
  }
  for (int i = 0; i < ndim; i++)
    if (A[i] < 1 || B[i] < 1 || A[i] * B[i] != L[i] || L[i] != sizeDimDecomp(start->dims[i]->decomp)) {
      raise_error(comm, "Invalid factor");
      return nullptr;
    }
  // The fused passes transform whole axes; the phase order transforms the
  // A_d- and B_d-point factors of each axis.  An axis longer than one pass
  // tile makes a phase-order plan; an axis neither form can do rejects the
  // plan here, as the reference rejects plans it cannot build
  // (split_fft.c:252-254,304-308): a logged reason and NULL.
  bool fused_ok = true;
  for (int i = 0; i < ndim; i++) {
    std::string why, why_a, why_b;
    if (L[i] <= 1 || dm::axis_supported(L[i], dbl, &why)) continue;
    fused_ok = false;
    if (!dm::axis_supported(A[i], dbl, &why_a) || !dm::axis_supported(B[i], dbl, &why_b)) {
      logger(comm, "buildParallelPlan: axis %d: %s\n", i, why.c_str());
      t_rejection = "axis " + std::to_string(i) + ": " + why + ", and its factors " + std::to_string(A[i]) + " x " +
                    std::to_string(B[i]) + " are not both within one pass";
      return nullptr;
    }
  }
  SplitPlan plan = (SplitPlan)calloc(1, sizeof(struct split_fft_plan));
  plan->comm = comm;
  plan->src = (fftw_complex *)dat;
  plan->work = (fftw_complex *)workspace;
  dm_split_ext *x = new dm_split_ext;
  plan->ext = x;
  x->dbl = dbl;
  x->ndim = ndim;
  x->A.assign(A, A + ndim);
  x->B.assign(B, B + ndim);
  x->L.assign(L, L + ndim);
  x->start = start;
  x->use_first = use_first != 0;
  x->fused_ok = fused_ok;
  x->phased = !fused_ok || (flags & RUN_PHASE_ORDER) != 0;

  // first decomposition keeps the long-distance factor local (split_fft.c:281)
  plan->A = makeInterleavedDecomp(ndim, comm, A, B, grid_a);
  if (plan->A == nullptr) {
    freeSplitPlan(plan);
    return nullptr;
  }
  long long ext = 0;
  for (long long e : plan->A->ext->extent_of) ext = std::max(ext, e);
  if (!use_first) {
    plan->B = makeBlockedDecomp(ndim, comm, A, B, grid_b);
    if (plan->B == nullptr) {
      freeSplitPlan(plan);
      return nullptr;
    }
    for (long long e : plan->B->ext->extent_of) ext = std::max(ext, e);
  }
  plan->extent = (int)ext;
  if (2 * ext > length) {
    logger(comm, "Not enough workspace need=%lld have=%d\n", 2 * ext, length);
    freeSplitPlan(plan);
    return nullptr;
  }
  if ((flags & PREFER_LARGE_FFT) && !use_first) {
    // split_fft.c:309-334: reject a common factor in the two blockings
    for (int i = 0; i < ndim; i++) {
      InterleaveDecomp ad = (InterleaveDecomp)plan->A->dims[i]->decomp;
      unsigned a_red = (unsigned)((BlockDecomp)ad->nested)->blocksize;
      unsigned b_red = (unsigned)(((BlockDecomp)plan->B->dims[i]->decomp)->blocksize / A[i]);
      if (gcd(a_red, b_red) > 1) {
        freeSplitPlan(plan);
        return nullptr;
      }
    }
  }
  MPI_Datatype base = dbl ? &dm_type_c_double_complex : &dm_type_c_float_complex;
  plan->start_to_A = makeReshapePlan(comm, base, start, plan->A);
  plan->A_to_start = makeReshapePlan(comm, base, plan->A, start);
  if (plan->start_to_A == nullptr || plan->A_to_start == nullptr) {
    freeSplitPlan(plan);
    return nullptr;
  }
  if (!(use_first || use_second)) {
    // twiddle tables per axis, indexed by the view rank's local A offset
    // (split_fft.c:348-420): w = exp(-2 pi i y l / L), y = g % A, l = g / A.
    int non_trivial = 0;
    plan->twiddle = (_Complex double **)calloc(ndim, sizeof(_Complex double *));
    for (int i = 0; i < ndim; i++) {
      AxisDecomp ax = plan->A->dims[i];
      plan->twiddle[i] = (_Complex double *)calloc(ax->max_local + 1, sizeof(_Complex double));
      for (int g = 0; g < L[i]; g++) {
        if (!isLocal(ax, g)) continue;
        int local = ax->decomp->type_offset(ax->decomp, g);
        double re, im;
        twiddle_value(g % A[i], g / A[i], L[i], &re, &im);
        if (re != 1.0 && im != 0.0) non_trivial = 1;  // as split_fft.c:396
        __real__ plan->twiddle[i][local] = re;
        __imag__ plan->twiddle[i][local] = im;
      }
    }
    if (!non_trivial) {
      for (int i = 0; i < ndim; i++) free(plan->twiddle[i]);
      free(plan->twiddle);
      plan->twiddle = nullptr;
    }
    plan->A_to_B = makeReshapePlan(comm, base, plan->A, plan->B);
    plan->B_to_A = makeReshapePlan(comm, base, plan->B, plan->A);
  }
  if (!use_first) {
    // frequency k of axis i sits at B-position k / B_i + A_i * (k % B_i):
    // FactorReverse with the factors in that order (the reference passes
    // (A_i, B_i), which is the inverse map whenever A_i != B_i).
    std::vector<AxisDecomp> fl(ndim);
    for (int i = 0; i < ndim; i++) fl[i] = makeFactorReverseAxisDecomp(B[i], A[i], plan->B->dims[i]);
    plan->B_flipped = makeDecomp(ndim, fl.data());
    x->out = plan->B_flipped;
  } else {
    x->out = plan->A;
  }
  plan->A_factors = (int *)calloc(ndim, sizeof(int));
  for (int i = 0; i < ndim; i++) plan->A_factors[i] = A[i];
  plan->times = (double *)calloc(NUM_TIMERS, sizeof(double));
  return plan;
}

static SplitPlan propose(int dbl, int flags, Decomp start, int *initial_factor, void *dat, int length,
                         void *workspace) {
  if (start == nullptr || initial_factor == nullptr) {
    raise_error(start ? start->comm : nullptr, "Invalid factor");
    return nullptr;
  }
  int ndim = start->ndim;
  int use_first = 1, use_second = 1;
  std::vector<int> A(ndim), B(ndim), L(ndim);
  for (int i = 0; i < ndim; i++) {
    B[i] = initial_factor[i];
    use_second = use_second && B[i] == 1;
    L[i] = sizeDimDecomp(start->dims[i]->decomp);
    if (B[i] < 1 || L[i] % B[i] != 0) {
      raise_error(start->comm, "Invalid factor");
      return nullptr;
    }
    A[i] = L[i] / B[i];
    use_first = use_first && A[i] == 1;
  }
  if (use_second && (flags & PREFER_STAGE1)) return nullptr;
  int *ga = makeProcessorGrid(start->comm, ndim, A.data());
  int *gb = makeProcessorGrid(start->comm, ndim, B.data());
  SplitPlan p = build_plan(dbl, flags, use_first, use_second, start, ndim, A.data(), B.data(), L.data(), ga, gb,
                           dat, length, workspace);
  free(ga);
  free(gb);
  return p;
}

static SplitPlan split_plan(int dbl, int slab, Decomp start, void *dat, int length, void *workspace) {
  int ndim = start->ndim;
  std::vector<int> B(ndim);
  for (int i = 0; i < ndim; i++) {
    int LL = sizeDimDecomp(start->dims[i]->decomp);
    int AA = 1, BB = LL;
    if (slab == 1) {
      if (i == ndim - 1) {
        BB = 1;
        AA = LL;
      }
    } else {
      // split_fft.c:590-601: largest divisor not above sqrt(L)
      for (int q = 1; q < LL; q++)
        if (LL % q == 0) {
          int p = LL / q;
          if (q > AA && q <= p) {
            AA = q;
            BB = p;
          }
        }
    }
    B[i] = BB;
  }
  t_rejection.clear();
  SplitPlan r = propose(dbl, NO_FLAGS, start, B.data(), dat, length, workspace);
  if (r == nullptr) {
    if (t_rejection.empty()) raise_error(start->comm, "Cannot generate plan");
    else raise_error(start->comm, "Cannot generate plan (%s)", t_rejection.c_str());
  }
  return r;
}

// ---------------------------------------------------------------- autotuner
// The reference times every divisor tuple on every grid (split_fft.c:78-247).
// The device program's cost barely depends on the split (one pass per axis
// whatever A and B are); what differs is the output layout.  Candidates are
// ranked by a traffic model (bytes moved per pass, with a penalty for output
// rows shorter than a 128-byte line), and the best few are timed.

static double model_cost(SplitPlan p) {
  dm_split_ext *x = p->ext;
  double N = 1;
  for (int l : x->L) N *= l;
  double es = x->dbl ? 16.0 : 8.0;
  // every pass reads and writes the array once
  double passes = 0;
  for (int l : x->L) passes += l > 1;
  double cost = passes * 2.0 * N * es;
  // contiguous run of the output's fastest axis on one rank
  dm::HostLayout ho = dm::layout_from_decomp(x->out);
  int f = ho.fastest;
  int nloc = 0;
  for (int g = 0; g < ho.L[f]; g++) nloc += ho.rank_part[f][g] == 0;
  double run = std::max(1, nloc) * es;
  if (run < 128.0) cost *= 128.0 / run;
  // the backward's first pass reads whole lines of that axis: a line split
  // over ranks reads as many short pieces from as many slots (measured at
  // 512^3: 1.0 ms against 0.73 ms for a local line)
  if (ho.L[f] > 1) cost *= 1.0 + 0.3 * (1.0 - (double)nloc / ho.L[f]);
  // fewer, larger rank blocks keep the workspace compact
  cost *= 1.0 + 1e-3 * p->extent / std::max(1.0, N);
  return cost;
}

// Milliseconds of a forward + backward step of plan p (the backward
// overwrites dat: the caller saves it).
static double time_plan(SplitPlan p, int reps) {
  if (p->ext->phased ? !compile_phased(p) : !compile_plan(p)) return 1e30;
  void *s = dm::world_stream(p->comm->world);
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  bool ok = true;
  auto run = [&]() {
    if (p->ext->phased) {  // synchronous phases: the loop below times them whole
      run_phased(p, FFTW_FORWARD);
      run_phased(p, FFTW_BACKWARD);
      ok = !dm::error_pending();
      return;
    }
    for (auto *prog : {&p->ext->fwd, &p->ext->bwd})
      for (auto &L : prog->passes)
        if (ok) ok = dm::launch_ok(p->comm, dm::launch_fft_pass(L, s), "FFT pass launch");
  };
  run();
  (void)hipEventRecord(e0, (hipStream_t)s);
  for (int i = 0; i < reps && ok; i++) run();
  (void)hipEventRecord(e1, (hipStream_t)s);
  (void)hipEventSynchronize(e1);
  float ms = 0;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return ok ? ms / reps : 1e30;
}

static SplitPlan autotune_search(int dbl, Decomp start, void *dat, int length, void *workspace);

// Collective in a dm_spmd_run session (as the reference's search is, with
// its barriers and all-reduces): rank 0 searches and times, every rank then
// plans the chosen factors with its own view.
static SplitPlan autotune(int dbl, Decomp start, void *dat, int length, void *workspace) {
  if (dm::spmd_rank() < 0) return autotune_search(dbl, start, dat, length, workspace);
  static std::vector<int> chosen;  // rank 0 writes it between the rendezvous
  std::vector<long long> sig = {dbl, length, dm::decomp_fingerprint(start)};
  dm::sig_pointer(&sig, dat);
  dm::sig_pointer(&sig, workspace);
  if (!dm::spmd_collective_checked(sig, "autotuneParallelPlan", [&] {
        chosen.clear();
        SplitPlan best = autotune_search(dbl, start, dat, length, workspace);
        if (best == nullptr) return;
        for (int i = 0; i < best->ext->ndim; i++) chosen.push_back(best->ext->B[i]);
        freeSplitPlan(best);
      }))
    return nullptr;
  std::vector<int> B = chosen;
  dm::spmd_barrier_internal();  // every rank has copied the choice before the next search may write it
  if (B.empty()) return nullptr;
  return propose(dbl, PREFER_STAGE1, start, B.data(), dat, length, workspace);
}

static SplitPlan autotune_search(int dbl, Decomp start, void *dat, int length, void *workspace) {
  int ndim = start->ndim;
  std::vector<int> L(ndim);
  for (int i = 0; i < ndim; i++) L[i] = sizeDimDecomp(start->dims[i]->decomp);
  std::vector<std::vector<int>> divs(ndim);
  for (int i = 0; i < ndim; i++)
    for (int b = 1; b <= L[i]; b++)
      if (L[i] % b == 0) divs[i].push_back(b);
  struct Cand {
    double cost;
    SplitPlan plan;
  };
  std::vector<Cand> cands;
  std::vector<int> B(ndim, 0);
  // candidates that cannot be planned are skipped: errors are recorded, not
  // raised, on this thread while the candidates are built
  std::unique_ptr<dm::ScopedErrorMode> quiet(new dm::ScopedErrorMode(1));
  std::function<void(int)> rec = [&](int d) {
    if (d == ndim) {
      bool all_one = true;
      for (int b : B) all_one = all_one && b == 1;
      if (all_one) return;  // PREFER_STAGE1: a single second stage is the first-stage plan
      SplitPlan p = propose(dbl, PREFER_STAGE1, start, B.data(), dat, length, workspace);
      dm_clear_error();
      if (p == nullptr) return;
      cands.push_back({model_cost(p), p});
      return;
    }
    for (int b : divs[d]) {
      B[d] = b;
      rec(d + 1);
    }
  };
  // bound the enumeration: beyond a few thousand tuples only balanced splits
  long long total = 1;
  for (auto &v : divs) total *= (long long)v.size();
  if (total <= 4096) {
    rec(0);
  } else {
    for (int i = 0; i < ndim; i++) {
      std::vector<int> keep;
      for (int b : divs[i])
        if ((long long)b * b >= L[i] / 16 && (long long)b * b <= 16LL * L[i]) keep.push_back(b);
      divs[i] = keep;
    }
    rec(0);
  }
  quiet.reset();
  if (cands.empty()) return nullptr;
  std::stable_sort(cands.begin(), cands.end(), [](const Cand &a, const Cand &b) { return a.cost < b.cost; });
  size_t keep = std::min<size_t>(cands.size(), 6);
  SplitPlan best = cands[0].plan;
  if (dm::device_available() && keep > 1) {
    // The candidates' backward passes write dat: keep a copy of the
    // caller's data (every rank slot) and put it back afterwards.  Without
    // room for the copy only the model decides.
    dm::Buffer bd;
    size_t od;
    void *save = nullptr;
    size_t bytes = 0;
    if (dm::resolve(dat, &bd, &od)) {
      bytes = bd.slot * (size_t)bd.nranks;
      save = dm::device_alloc(bytes);
    }
    if (save != nullptr && hipMemcpy(save, bd.base, bytes, hipMemcpyDeviceToDevice) == hipSuccess) {
      double bt = 1e30;
      for (size_t i = 0; i < keep; i++) {
        double t = time_plan(cands[i].plan, 2);
        if (t < bt) {
          bt = t;
          best = cands[i].plan;
        }
      }
      if (hipMemcpy(bd.base, save, bytes, hipMemcpyDeviceToDevice) != hipSuccess)
        raise_error(start->comm, "autotune: restoring the input after the timing runs failed");
    }
    dm::device_free(save);
  }
  for (auto &c : cands)
    if (c.plan != best) freeSplitPlan(c.plan);
  return best;
}

// ---------------------------------------------------------------- executor

static void run_program(SplitPlan plan, dm_fft_program &prog, int t_first, int t_rest) {
  dm_world *w = plan->comm->world;
  void *s = dm::world_stream(w);
  size_t n = prog.passes.size();
  bool timed = !w->async;
  std::vector<hipEvent_t> ev;
  if (timed) {
    ev.resize(n + 1);
    for (auto &e : ev) (void)hipEventCreate(&e);
    (void)hipEventRecord(ev[0], (hipStream_t)s);
  }
  bool ok = true;
  for (size_t i = 0; i < n && ok; i++) {
    ok = dm::launch_ok(plan->comm, dm::launch_fft_pass(prog.passes[i], s), "FFT pass launch");
    if (timed) (void)hipEventRecord(ev[i + 1], (hipStream_t)s);
  }
  if (!timed) return;
  if (!ok) {
    (void)hipStreamSynchronize((hipStream_t)s);
    for (auto &e : ev) (void)hipEventDestroy(e);
    return;
  }
  (void)hipEventSynchronize(ev[n]);
  double total = 0;
  for (size_t i = 0; i < n; i++) {
    float ms = 0;
    (void)hipEventElapsedTime(&ms, ev[i], ev[i + 1]);
    plan->times[prog.passes[i].pass_index == 0 ? t_first : t_rest] += ms * 1e-3;
    total += ms * 1e-3;
  }
  plan->times[ALL] += total;
  for (auto &e : ev) (void)hipEventDestroy(e);
}

// split_fft.c:622-744.  Forward: start layout in dat -> output layout in
// work.  Backward: the exact (unnormalised) inverse, output layout in work ->
// start layout in dat (the reference's backward path runs A_to_start before
// its stage-1 inverse and builds that inverse only under MAKE_REVERSE).
static void run_split(SplitPlan plan, int direction);

// Collective like runReshapePlan: in a dm_spmd_run session rank 0's plan
// (built on slot-0 pointers) runs for every rank.
extern "C" void runSplitPlan(SplitPlan plan, int direction) {
  if (dm::spmd_rank() >= 0) {
    std::vector<long long> sig = {direction};
    if (plan != nullptr && plan->ext != nullptr) {
      dm_split_ext *x = plan->ext;
      sig.insert(sig.end(), x->A.begin(), x->A.end());
      sig.push_back(x->dbl);
      sig.push_back(x->phased);
      sig.push_back(plan->extent);
      sig.push_back(dm::decomp_fingerprint(x->start));
      sig.push_back(dm::decomp_fingerprint(x->out));
      dm::sig_pointer(&sig, plan->src);
      dm::sig_pointer(&sig, plan->work);
    }
    dm::spmd_collective_checked(sig, "runSplitPlan", [&] { run_split(plan, direction); });
    return;
  }
  run_split(plan, direction);
}

static void run_split(SplitPlan plan, int direction) {
  if (plan == nullptr) {
    dm::fail(nullptr, "runSplitPlan: NULL plan");
    return;
  }
  if (direction != FFTW_FORWARD && direction != FFTW_BACKWARD) {
    raise_error(plan->comm, "runSplitPlan: direction must be FFTW_FORWARD or FFTW_BACKWARD");
    return;
  }
  dm::check_device(plan->comm);
  if (dm::error_pending() && dm::error_mode() == 1) return;
  if (plan->ext->phased) {
    if (compile_phased(plan)) run_phased(plan, direction);
    return;
  }
  if (!compile_plan(plan)) return;
  if (direction == FFTW_FORWARD) run_program(plan, plan->ext->fwd, FFT_1, FFT_2);
  else run_program(plan, plan->ext->bwd, FFT_2, FFT_1);
}

extern "C" int dm_split_set_output(SplitPlan plan, Decomp out) {
  if (plan == nullptr || plan->ext == nullptr || out == nullptr) return DM_ERR_ARG;
  dm_split_ext *x = plan->ext;
  if (compatible(out, x->start) != COMPATIBLE) {
    raise_error(plan->comm, "dm_split_set_output: the output decomposition is not compatible with the start");
    return DM_ERR_ARG;
  }
  if (!x->fused_ok) {
    raise_error(plan->comm, "dm_split_set_output: this plan runs in phase order only, which writes the B layout");
    return DM_ERR_ARG;
  }
  std::string why;
  if (!dm::layout_rank_map_ok(out, &why)) {
    raise_error(plan->comm, "dm_split_set_output: %s", why.c_str());
    return DM_ERR_ARG;
  }
  program_release(x->fwd);
  program_release(x->bwd);
  x->compiled = false;
  x->out = out;
  x->custom_out = true;
  x->phased = false;
  return DM_SUCCESS;
}

extern "C" int dm_split_set_phase_order(SplitPlan plan, int on) {
  if (plan == nullptr || plan->ext == nullptr) return DM_ERR_ARG;
  if (on && plan->ext->custom_out) {
    raise_error(plan->comm, "dm_split_set_phase_order: the phase order writes the B layout, not a chosen output");
    return DM_ERR_ARG;
  }
  if (!on && !plan->ext->fused_ok) {
    raise_error(plan->comm, "dm_split_set_phase_order: an axis is longer than one pass tile; the plan runs in "
                            "phase order only");
    return DM_ERR_ARG;
  }
  plan->ext->phased = on != 0;
  return DM_SUCCESS;
}

extern "C" int dm_split_set_twiddles(SplitPlan plan, int mode) {
  if (plan == nullptr || plan->ext == nullptr) return DM_ERR_ARG;
  if (mode != DM_TWIDDLES_COMPENSATED && mode != DM_TWIDDLES_PLAIN) {
    raise_error(plan->comm, "dm_split_set_twiddles: unknown mode %d", mode);
    return DM_ERR_ARG;
  }
  dm_split_ext *x = plan->ext;
  const int ctw = mode == DM_TWIDDLES_COMPENSATED ? 1 : 0;
  if (ctw != x->ctw) {
    // the pass descriptors carry the form: compile again on the next run
    x->ctw = ctw;
    program_release(x->fwd);
    program_release(x->bwd);
    x->compiled = false;
    phased_release(x->ph);
  }
  return DM_SUCCESS;
}

extern "C" int dm_split_twiddles(SplitPlan plan) {
  if (plan == nullptr || plan->ext == nullptr) return -1;
  return plan->ext->ctw ? DM_TWIDDLES_COMPENSATED : DM_TWIDDLES_PLAIN;
}

extern "C" int dm_split_phase_order(SplitPlan plan) {
  return (plan != nullptr && plan->ext != nullptr && plan->ext->phased) ? 1 : 0;
}

extern "C" Decomp getTransformed(SplitPlan plan) {
  if (plan != nullptr && plan->ext != nullptr && plan->ext->custom_out) return plan->ext->out;
  return plan->B_flipped;
}

extern "C" void freeSplitPlan(SplitPlan plan) {
  if (plan == nullptr) return;
  int ndim = plan->ext ? plan->ext->ndim : (plan->A ? plan->A->ndim : 0);
  free(plan->A_factors);
  freeReshapePlan(plan->start_to_A);
  freeReshapePlan(plan->A_to_start);
  freeReshapePlan(plan->A_to_B);
  freeReshapePlan(plan->B_to_A);
  if (plan->twiddle != nullptr) {
    for (int i = 0; i < ndim; i++) free(plan->twiddle[i]);
    free(plan->twiddle);
  }
  if (plan->ext) {
    program_release(plan->ext->fwd);
    program_release(plan->ext->bwd);
    phased_release(plan->ext->ph);
    delete plan->ext;
  }
  freeDecomp(plan->B_flipped);
  freeDecomp(plan->B);
  freeDecomp(plan->A);
  free(plan->times);
  free(plan);
}

// ---------------------------------------------------------------- C entry points

extern "C" SplitPlan buildParallelPlan(int flags, int use_first, int use_second, Decomp start, int ndim, int A[],
                                       int B[], int L[], int a_grid[], int b_grid[], fftw_complex *dat, int length,
                                       fftw_complex *workspace) {
  return build_plan(1, flags, use_first, use_second, start, ndim, A, B, L, a_grid, b_grid, dat, length,
                    workspace);
}
extern "C" SplitPlan proposeParallelPlan(int flags, Decomp start, int *initial_factor, fftw_complex *dat,
                                         int length, fftw_complex *workspace) {
  return propose(1, flags, start, initial_factor, dat, length, workspace);
}
extern "C" SplitPlan autotuneParallelPlan(Decomp start, fftw_complex *dat, int length, fftw_complex *workspace) {
  return autotune(1, start, dat, length, workspace);
}
extern "C" SplitPlan makeSplitPlan(int slab, Decomp start, fftw_complex *dat, int length,
                                   fftw_complex *workspace) {
  return split_plan(1, slab, start, dat, length, workspace);
}
extern "C" SplitPlan makeParallelPlan(Decomp start, int *initial_factor, fftw_complex *dat, int length,
                                      fftw_complex *workspace) {
  return propose(1, NO_FLAGS, start, initial_factor, dat, length, workspace);
}

extern "C" SplitPlan buildParallelPlanf(int flags, int use_first, int use_second, Decomp start, int ndim, int A[],
                                        int B[], int L[], int a_grid[], int b_grid[], fftwf_complex *dat,
                                        int length, fftwf_complex *workspace) {
  return build_plan(0, flags, use_first, use_second, start, ndim, A, B, L, a_grid, b_grid, dat, length,
                    workspace);
}
extern "C" SplitPlan proposeParallelPlanf(int flags, Decomp start, int *initial_factor, fftwf_complex *dat,
                                          int length, fftwf_complex *workspace) {
  return propose(0, flags, start, initial_factor, dat, length, workspace);
}
extern "C" SplitPlan autotuneParallelPlanf(Decomp start, fftwf_complex *dat, int length,
                                           fftwf_complex *workspace) {
  return autotune(0, start, dat, length, workspace);
}
extern "C" SplitPlan makeSplitPlanf(int slab, Decomp start, fftwf_complex *dat, int length,
                                    fftwf_complex *workspace) {
  return split_plan(0, slab, start, dat, length, workspace);
}
extern "C" SplitPlan makeParallelPlanf(Decomp start, int *initial_factor, fftwf_complex *dat, int length,
                                       fftwf_complex *workspace) {
  return propose(0, NO_FLAGS, start, initial_factor, dat, length, workspace);
}
