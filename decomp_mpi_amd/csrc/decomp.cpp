// AxisDecomp / Decomp: the reference's per-rank objects plus the global view
// (every virtual rank's positions, local counts and memory strides).
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "internal.h"
#include "factor_reverse_decomp.h"
#include "logger.h"

static inline int position_of(int rank, int rank_stride, int proc_size) {
  return (rank / rank_stride) % proc_size;
}

// reshape.c:47-89.  The DimDecomp is evaluated once at every coordinate
// (the reference scans it once per rank); the view rank's fields are filled
// from that scan.  Takes ownership of d.
extern "C" AxisDecomp makeAxisDecomp(MPI_Comm comm, DimDecomp d, int stride, int rank_stride) {
  if (comm == nullptr || d == nullptr) {
    dm::fail(comm, "makeAxisDecomp: NULL argument");
    return nullptr;
  }
  int size = comm->world->nranks;
  if (rank_stride < 1 || (long long)rank_stride * d->proc_size > size) {
    raise_error(comm, "Illegal rank_stride");
    return nullptr;
  }
  AxisDecomp a = (AxisDecomp)calloc(1, sizeof(struct axis_decomp));
  a->decomp = d;
  a->comm = comm;
  a->memory_stride = stride;
  a->rank_stride = rank_stride;
  a->my_rank_in_dimension = position_of(comm->rank, rank_stride, d->proc_size);

  dm_axis_ext *e = new dm_axis_ext;
  e->world = comm->world;
  int n = d->coord_size;
  e->rank_of.resize(n);
  e->off_of.resize(n);
  e->n_local_pos.assign(d->proc_size, 0);
  e->max_local_pos.assign(d->proc_size, 0);
  for (int g = 0; g < n; g++) {
    int r = d->rank_offset(d, g);
    int o = d->type_offset(d, g);
    e->rank_of[g] = r;
    e->off_of[g] = o;
    if (r >= 0 && r < d->proc_size) {
      e->n_local_pos[r]++;
      e->max_local_pos[r] = std::max(e->max_local_pos[r], o);
    }
  }
  e->stride_of_rank.assign(size, stride);
  a->ext = e;

  int me = a->my_rank_in_dimension;
  a->n_local = e->n_local_pos[me];
  a->max_local = e->max_local_pos[me];
  a->index = makeIndex();
  for (int g = 0; g < n; g++)
    if (e->rank_of[g] == me) addIndex(a->index, e->off_of[g]);
  return a;
}

extern "C" void freeAxisDecomp(AxisDecomp a) {
  if (a == nullptr) return;
  freeDimDecomp(a->decomp);
  freeIndex(a->index);
  delete a->ext;
  free(a);
}

extern "C" void describeAxisDecomp(FILE *out, AxisDecomp d) {
  fprintf(out, "AxisDecomp: mem_stride=%d rank_stride=%d ", d->memory_stride, d->rank_stride);
  describeDimDecomp(out, d->decomp);
}

extern "C" int isLocal(AxisDecomp d, int global) {
  return d->my_rank_in_dimension == d->decomp->rank_offset(d->decomp, global);
}

// reshape.c:104-136: for each coordinate local to `mine`, the local offset
// goes to the list of the remote position that owns it.
extern "C" Index *makeIndexList(AxisDecomp mine, DimDecomp remote) {
  int count = remote->proc_size;
  if (mine->decomp->coord_size != remote->coord_size) {
    raise_error(mine->comm, "non-conformant sizes in makeIndexList %d!=%d",
                mine->decomp->coord_size, remote->coord_size);
    return nullptr;
  }
  Index *res = (Index *)calloc(count + 1, sizeof(Index));
  for (int i = 0; i < count; i++) res[i] = makeIndex();
  for (int g = 0; g < mine->decomp->coord_size; g++) {
    if (!isLocal(mine, g)) continue;
    int r = remote->rank_offset(remote, g);
    if (r >= 0 && r < count) addIndex(res[r], mine->decomp->type_offset(mine->decomp, g));
  }
  return res;
}

// reshape.c:137-148
extern "C" Index makeGlobalIndex(AxisDecomp mine) {
  Index res = makeIndex();
  for (int g = 0; g < mine->decomp->coord_size; g++)
    if (isLocal(mine, g)) addIndex(res, g);
  return res;
}

// reshape.c:158-172.  The file view of one axis: for every coordinate g
// this rank holds along it, ascending, a copy of src's view placed at
// g * extent * stride bytes (src resized to extent * stride, then indexed
// by the rank's global index runs).  Returns stride * coord_size, the
// stride of the next (slower) axis.
extern "C" int expandIOType(MPI_Datatype *result, MPI_Aint extent, MPI_Datatype src, AxisDecomp mine, int stride) {
  if (result == nullptr || src == nullptr || mine == nullptr) {
    raise_error(mine ? mine->comm : nullptr, "expandIOType: NULL argument");
    return 0;
  }
  std::vector<long long> inner = src->view;
  if (inner.empty()) inner.push_back(0);
  dm_datatype *t = new dm_datatype;
  t->predefined = 0;
  t->elem = src->elem ? src->elem : src->size;
  const long long step = (long long)extent * stride;
  long long top = 0;
  for (int g = 0; g < mine->decomp->coord_size; g++) {
    if (!isLocal(mine, g)) continue;
    for (long long o : inner) t->view.push_back(g * step + o);
    top = (g + 1) * step;
  }
  t->count = (int)t->view.size();
  t->size = t->view.size() * t->elem;
  t->extent = top;
  t->true_lb = 0;
  t->true_extent = top;
  *result = t;
  return stride * mine->decomp->coord_size;
}

// reshape.c:494-521.  The view of a whole decomposition: the calling rank's
// elements at their global row-major positions (last axis fastest), in the
// order of its loops over its coordinates (first axis outermost), spanning
// the whole file; empty (extent 0) on ranks the decomposition does not use.
extern "C" void makeIOType(MPI_Datatype *result, MPI_Datatype src, Decomp d) {
  if (result == nullptr || src == nullptr || d == nullptr) {
    raise_error(d ? d->comm : nullptr, "makeIOType: NULL argument");
    return;
  }
  dm_aint_t lb = 0, extent = 0;
  dm_type_get_extent(src, &lb, &extent);
  dm_datatype_t cur = nullptr;
  dm_type_dup(src, &cur);
  cur->view.assign(1, 0);
  cur->elem = src->elem ? src->elem : src->size;
  int stride = 1;
  if (d->comm->rank < d->proc_used) {
    for (int i = d->ndim - 1; i >= 0; i--) {
      dm_datatype_t next = nullptr;
      stride = expandIOType(&next, extent, cur, d->dims[i], stride);
      dm_type_free(&cur);
      if (next == nullptr) return;
      cur = next;
    }
    cur->extent = (ptrdiff_t)stride * extent;
  } else {
    cur->view.clear();
    cur->count = 0;
    cur->size = 0;
    cur->extent = 0;
  }
  cur->true_extent = cur->extent;
  *result = cur;
}

extern "C" long long dm_type_view(MPI_Datatype t, long long *offsets, long long max) {
  if (t == nullptr) return -1;
  if (t->view.empty()) {
    if (t->predefined || t->count == 1) {
      if (offsets != nullptr && max > 0) offsets[0] = 0;
      return 1;
    }
    return 0;
  }
  const long long n = (long long)t->view.size();
  if (offsets != nullptr)
    for (long long i = 0; i < n && i < max; i++) offsets[i] = t->view[i];
  return n;
}

// ---------------------------------------------------------------- Decomp

static void build_decomp_ext(Decomp d) {
  int P = d->comm->world->nranks;
  dm_decomp_ext *x = new dm_decomp_ext;
  x->strides.assign((size_t)P * d->ndim, 0);
  x->extent_of.assign(P, 1);
  for (int r = 0; r < P; r++) {
    i64 ext = 1;
    for (int i = 0; i < d->ndim; i++) {
      AxisDecomp a = d->dims[i];
      i64 m = a->ext->stride_of_rank[r];
      x->strides[(size_t)r * d->ndim + i] = m;
      int pos = position_of(r, a->rank_stride, a->decomp->proc_size);
      ext += m * a->ext->max_local_pos[pos];
    }
    x->extent_of[r] = ext;
  }
  x->uniform = true;
  for (int r = 1; r < d->proc_used && r < P; r++)
    for (int i = 0; i < d->ndim; i++)
      if (x->strides[(size_t)r * d->ndim + i] != x->strides[i]) x->uniform = false;
  d->ext = x;
}

// reshape.c:174-217
extern "C" Decomp makeDecomp(int ndim, AxisDecomp dims[]) {
  if (ndim < 1 || dims == nullptr) {
    dm::fail(nullptr, "makeDecomp: need at least one axis");
    return nullptr;
  }
  for (int i = 0; i < ndim; i++)
    if (dims[i] == nullptr) {
      dm::fail(nullptr, "makeDecomp: axis %d is NULL", i);
      return nullptr;
    }
  MPI_Comm comm = dims[0]->comm;
  for (int i = 1; i < ndim; i++)
    if (dims[i]->comm->world != comm->world) {
      raise_error(comm, "Different communicators in Decomp");
      return nullptr;
    }
  long long max = 0;
  for (int i = 0; i < ndim; i++) max += (long long)dims[i]->rank_stride * (dims[i]->decomp->proc_size - 1);
  int size = comm->world->nranks;
  if (max >= size) {
    raise_error(comm, "Maximum processor rank out of range max=%lld size=%d", max, size);
    return nullptr;
  }
  for (int i = 0; i < ndim; i++)
    for (int j = i + 1; j < ndim; j++) {
      AxisDecomp a = dims[i], b = dims[j];
      // size-1 axes (single/local) may share a rank stride
      if (a->decomp->proc_size > 1 && b->decomp->proc_size > 1 && a->rank_stride > 1 &&
          b->rank_stride > 1 && a->rank_stride == b->rank_stride) {
        raise_error(comm,
                    "duplicate rank strides dim=%d size=%d rank_stride=%d, dim=%d size=%d rank_stride=%d",
                    i, a->decomp->proc_size, a->rank_stride, j, b->decomp->proc_size, b->rank_stride);
        return nullptr;
      }
    }
  Decomp d = (Decomp)calloc(1, sizeof(struct decomp));
  d->ndim = ndim;
  d->comm = comm;
  d->proc_used = (int)max + 1;
  d->dims = (AxisDecomp *)malloc(ndim * sizeof(AxisDecomp));
  for (int i = 0; i < ndim; i++) d->dims[i] = dims[i];
  build_decomp_ext(d);
  return d;
}

// reshape.c:218-241: last axis fastest; each rank's strides follow its own
// local extents.
extern "C" Decomp makePackedDecomp(MPI_Comm comm, int ndim, DimDecomp dims[]) {
  if (comm == nullptr || ndim < 1) {
    dm::fail(comm, "makePackedDecomp: bad arguments");
    return nullptr;
  }
  std::vector<AxisDecomp> ax(ndim);
  int rank_stride = 1, mem_stride = 1;
  for (int i = ndim - 1; i >= 0; i--) {
    ax[i] = makeAxisDecomp(comm, dims[i], mem_stride, rank_stride);
    if (ax[i] == nullptr) {
      for (int j = i + 1; j < ndim; j++) freeAxisDecomp(ax[j]);
      return nullptr;
    }
    mem_stride *= (ax[i]->max_local + 1);
    rank_stride *= dims[i]->proc_size;
  }
  int P = comm->world->nranks;
  for (int r = 0; r < P; r++) {
    i64 m = 1;
    for (int i = ndim - 1; i >= 0; i--) {
      AxisDecomp a = ax[i];
      a->ext->stride_of_rank[r] = m;
      int pos = position_of(r, a->rank_stride, a->decomp->proc_size);
      m *= (a->ext->max_local_pos[pos] + 1);
    }
  }
  Decomp d = makeDecomp(ndim, ax.data());
  if (d == nullptr)
    for (int i = 0; i < ndim; i++) freeAxisDecomp(ax[i]);
  return d;
}

extern "C" void freeDecomp(Decomp d) {
  if (d == nullptr) return;
  for (int i = 0; i < d->ndim; i++) freeAxisDecomp(d->dims[i]);
  free(d->dims);
  delete d->ext;
  free(d);
}

extern "C" void describeDeomp(FILE *out, Decomp d) {
  fprintf(out, "Decomp ndim=%d procs_used=%d\n", d->ndim, d->proc_used);
  for (int i = 0; i < d->ndim; i++) {
    describeAxisDecomp(out, d->dims[i]);
    fprintf(out, "\n");
  }
}

// reshape.c:258-279
extern "C" int compatible(Decomp a, Decomp b) {
  if (a->ndim != b->ndim) return NOT_COMPATIBLE;
  if (a->comm->world != b->comm->world) return NOT_COMPATIBLE;
  for (int i = 0; i < a->ndim; i++)
    if (sizeDimDecomp(a->dims[i]->decomp) != sizeDimDecomp(b->dims[i]->decomp)) return NOT_COMPATIBLE;
  return COMPATIBLE;
}

extern "C" int equivalent(Decomp a, Decomp b) {
  if (compatible(a, b) != COMPATIBLE || a->proc_used != b->proc_used) return 0;
  for (int i = 0; i < a->ndim; i++) {
    AxisDecomp x = a->dims[i], y = b->dims[i];
    if (x->rank_stride != y->rank_stride && x->decomp->proc_size > 1) return 0;
    if (x->decomp->proc_size != y->decomp->proc_size) return 0;
    if (x->ext->rank_of != y->ext->rank_of || x->ext->off_of != y->ext->off_of) return 0;
  }
  return a->ext->strides == b->ext->strides ? 1 : 0;
}

// reshape.c:281-295
extern "C" int myRank(Decomp d) {
  int r = 0;
  for (int i = 0; i < d->ndim; i++) r += d->dims[i]->my_rank_in_dimension * d->dims[i]->rank_stride;
  return r;
}

extern "C" int isCoordLocal(Decomp d, int dim, int pos) { return isLocal(d->dims[dim], pos); }

extern "C" int areCoordsLocal(Decomp d, int *pos) {
  for (int i = 0; i < d->ndim; i++)
    if (!isLocal(d->dims[i], pos[i])) return 0;
  return 1;
}

// reshape.c:308-315
extern "C" int getOffset(Decomp d, int *pos) {
  int off = 0;
  for (int i = 0; i < d->ndim; i++) {
    AxisDecomp a = d->dims[i];
    off += a->memory_stride * a->decomp->type_offset(a->decomp, pos[i]);
  }
  return off;
}

// reshape.c:317-324
extern "C" int extent(Decomp d) {
  int size = 1;
  for (int i = 0; i < d->ndim; i++) size += d->dims[i]->memory_stride * d->dims[i]->max_local;
  return size;
}

// factor_reverse_decomp.c:56-58.  The new axis keeps orig's per-rank
// memory strides.
extern "C" AxisDecomp makeFactorReverseAxisDecomp(int A, int B, AxisDecomp orig) {
  DimDecomp f = makeFactorReverseDecomp(A, B, orig->decomp);
  if (f == nullptr) return nullptr;
  AxisDecomp a = makeAxisDecomp(orig->comm, f, orig->memory_stride, orig->rank_stride);
  if (a != nullptr) a->ext->stride_of_rank = orig->ext->stride_of_rank;
  return a;
}

// ---------------------------------------------------------------- layouts

namespace dm {

bool layout_rank_map_ok(Decomp d, std::string *why) {
  long long prod = 1;
  for (int i = 0; i < d->ndim; i++) prod *= d->dims[i]->decomp->proc_size;
  if (prod != d->proc_used) {
    if (why) *why = "product of axis process counts differs from proc_used";
    return false;
  }
  for (int r = 0; r < d->proc_used; r++) {
    long long s = 0;
    for (int i = 0; i < d->ndim; i++) {
      AxisDecomp a = d->dims[i];
      s += (long long)a->rank_stride * position_of(r, a->rank_stride, a->decomp->proc_size);
    }
    if (s != r) {
      if (why) *why = "rank strides do not enumerate the ranks one-to-one";
      return false;
    }
  }
  return true;
}

// FNV-1a over the global description of d: axes (sizes, rank strides, the
// owner and offset of every coordinate) and every rank's memory strides.
long long decomp_fingerprint(Decomp d) {
  unsigned long long h = 1469598103934665603ULL;
  auto mix = [&](long long v) {
    for (int k = 0; k < 8; k++) {
      h ^= (unsigned long long)((v >> (8 * k)) & 0xff);
      h *= 1099511628211ULL;
    }
  };
  if (d == nullptr) return 0;
  mix(d->ndim);
  mix(d->proc_used);
  for (int i = 0; i < d->ndim; i++) {
    AxisDecomp a = d->dims[i];
    mix(a->decomp->coord_size);
    mix(a->decomp->proc_size);
    mix(a->rank_stride);
    for (int v : a->ext->rank_of) mix(v);
    for (int v : a->ext->off_of) mix(v);
  }
  for (i64 v : d->ext->strides) mix(v);
  return (long long)h;
}

HostLayout layout_from_decomp(Decomp d) {
  HostLayout h;
  h.ndim = d->ndim;
  h.proc_used = d->proc_used;
  h.L.resize(d->ndim);
  h.rank_part.resize(d->ndim);
  h.off.resize(d->ndim);
  for (int i = 0; i < d->ndim; i++) {
    AxisDecomp a = d->dims[i];
    int n = a->decomp->coord_size, P = a->decomp->proc_size;
    h.L[i] = n;
    h.rank_part[i].resize(n);
    h.off[i] = a->ext->off_of;
    for (int g = 0; g < n; g++) {
      int r = a->ext->rank_of[g];
      // A coordinate outside the axis' process range has no owner.  The
      // sentinel exceeds any proc_used, and DM_MAX_DIMS of them still fit
      // in an int, so the kernels' owner sums cannot wrap.
      h.rank_part[i][g] = (r >= 0 && r < P) ? r * a->rank_stride : DM_NO_OWNER;
      if (r >= 0 && r < P && h.off[i][g] < 0) h.offsets_ok = false;
    }
  }
  h.strides.assign((size_t)d->proc_used * d->ndim, 0);
  h.slot_extent.assign(d->proc_used, 0);
  for (int r = 0; r < d->proc_used; r++) {
    for (int i = 0; i < d->ndim; i++) {
      i64 m = d->ext->strides[(size_t)r * d->ndim + i];
      h.strides[(size_t)r * d->ndim + i] = m;
      if (m < 0) h.offsets_ok = false;
    }
    h.slot_extent[r] = d->ext->extent_of[r];
  }
  h.uniform = d->ext->uniform;
  h.m.assign(h.strides.begin(), h.strides.begin() + d->ndim);
  // fastest axis: smallest stride among axes longer than one; ties -> later axis
  int best = d->ndim - 1;
  i64 bs = -1;
  for (int i = d->ndim - 1; i >= 0; i--) {
    if (h.L[i] <= 1) continue;
    if (bs < 0 || h.m[i] < bs) {
      bs = h.m[i];
      best = i;
    }
  }
  h.fastest = best;
  return h;
}

}  // namespace dm
