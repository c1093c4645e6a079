// Reshape plans: message bookkeeping in the reference's terms (send/recv
// lists per rank) and the device gather program that actually moves data.
//
// Reference: makeReshapePlan / runReshapePlan / checkReshapePlanBounds
// (reshape.c:326-567) and the typelist expansion (typelist.c:8-119).  The
// reference pairs the k-th element of a send datatype with the k-th element
// of the matching receive datatype; both walk the same global coordinates
// in the same axis order, so the element that lands at a destination slot
// is the one with the same global coordinate.  The device program states
// that pairing directly over global coordinates.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "internal.h"
#include "kernels.h"
#include "logger.h"

struct dm_reshape_program {
  dm::HostLayout src, dst;    // element unit: `word` bytes
  dm::DeviceLayout dsrc, ddst;
  bool uploaded = false;
  int word = 1;               // bytes per moved unit
  int words = 1;              // units per element
  int elem = 1;               // element extent in bytes
  int a = 0, b = 0;
  std::vector<int> batch;
};

// ---------------------------------------------------------------- TypeEntryList
// typelist.c:8-71

extern "C" TypeEntryList makeTypeEntryList(void) {
  TypeEntryList t = (TypeEntryList)calloc(1, sizeof(struct type_entry_list));
  t->size = INITIAL_TYPELIST_SIZE;
  t->list = (struct type_entry *)calloc(t->size, sizeof(struct type_entry));
  return t;
}

extern "C" void commitTypeEntryList(TypeEntryList) {}

extern "C" void freeTypeEntryList(TypeEntryList t) {
  if (t == nullptr) return;
  for (int i = 0; i < t->count; i++) dm_type_free(&t->list[i].type);
  free(t->list);
  free(t);
}

extern "C" void addTypeEntry(TypeEntryList t, int rank, MPI_Datatype type) {
  if (t->count == t->size) {
    t->size *= 2;
    t->list = (struct type_entry *)realloc(t->list, t->size * sizeof(struct type_entry));
  }
  t->list[t->count].rank = rank;
  dm_type_dup(type, &t->list[t->count].type);
  t->count++;
}

extern "C" void printTypeEntryList(FILE *out, TypeEntryList t) {
  for (int i = 0; i < t->count; i++) fprintf(out, " [%d](%zu)", t->list[i].rank, t->list[i].type->size);
}

// One message as the reference's nested datatype would describe it:
// elements, true lower bound and true extent in bytes.
struct MsgShape {
  long long count = 1;
  long long lb = 0;
  long long hi = 0;
};

// Expand every entry by one axis (typelist.c:74-119): for each position of
// the remote axis that holds some of my coordinates, a new entry with rank
// `rank + pos * remote.rank_stride`, kept when below max_proc.
extern "C" TypeEntryList expandList(int max_proc, MPI_Aint extent, TypeEntryList list, AxisDecomp mine,
                                    AxisDecomp remote) {
  TypeEntryList res = makeTypeEntryList();
  Index *idx = makeIndexList(mine, remote->decomp);
  if (idx == nullptr) return res;
  for (int e = 0; e < list->count; e++) {
    for (int q = 0; q < remote->decomp->proc_size; q++) {
      Index ix = idx[q];
      if (ix->count == 0) continue;
      int nr = list->list[e].rank + q * remote->rank_stride;
      if (nr >= max_proc) continue;
      long long n = 0;
      int lo = ix->offsets[0], hi = ix->offsets[0];
      for (int r = 0; r < ix->count; r++) {
        n += ix->counts[r];
        lo = std::min(lo, ix->offsets[r]);
        hi = std::max(hi, ix->offsets[r] + ix->counts[r] - 1);
      }
      dm_datatype_t old = list->list[e].type;
      dm_datatype nt;
      nt.size = old->size * (size_t)n;
      nt.predefined = 0;
      nt.count = old->count * (int)n;
      long long step = (long long)extent * mine->memory_stride;
      nt.true_lb = old->true_lb + step * lo;
      nt.true_extent = old->true_extent + step * (hi - lo);
      nt.extent = nt.true_extent;
      addTypeEntry(res, nr, &nt);
    }
  }
  freeIndexList(idx);
  return res;
}

// ---------------------------------------------------------------- plans

// reshape.c:330-350: axes in ascending source memory stride, starting from C
// order (last axis first).
static void sort_order(int ndim, int *order, Decomp src) {
  for (int i = 0; i < ndim; i++) order[i] = ndim - 1 - i;
  for (int i = 0; i < ndim - 1; i++)
    for (int j = i + 1; j < ndim; j++)
      if (src->dims[order[j]]->memory_stride < src->dims[order[i]]->memory_stride) std::swap(order[i], order[j]);
}

// The send (is_send) or receive list of rank `me`, as the reference builds it.
static TypeEntryList message_list(Decomp mine, Decomp remote, const int *order, MPI_Datatype base, int me,
                                  bool is_send) {
  (void)is_send;
  if (me >= mine->proc_used) return nullptr;
  // The lists are per rank: evaluate them with axis decomps seen from `me`.
  dm_comm_t view = dm_comm_view(mine->comm, me);
  TypeEntryList cur = makeTypeEntryList();
  addTypeEntry(cur, 0, base);
  MPI_Aint lb, ext;
  dm_type_get_extent(base, &lb, &ext);
  for (int j = 0; j < mine->ndim; j++) {
    int i = order[j];
    AxisDecomp m = mine->dims[i];
    struct axis_decomp tmp = *m;
    tmp.comm = view;
    tmp.my_rank_in_dimension = (me / m->rank_stride) % m->decomp->proc_size;
    tmp.memory_stride = (int)mine->ext->strides[(size_t)me * mine->ndim + i];
    TypeEntryList nxt = expandList(remote->proc_used, ext, cur, &tmp, remote->dims[i]);
    freeTypeEntryList(cur);
    cur = nxt;
  }
  return cur;
}

static int count_rank(TypeEntryList t, int me) {
  int n = 0;
  if (t != nullptr)
    for (int i = 0; i < t->count; i++) n += t->list[i].rank == me;
  return n;
}

static void extend_by_words(dm::HostLayout &h, int words) {
  if (words == 1) return;
  int nd = h.ndim;
  h.L.push_back(words);
  h.rank_part.push_back(std::vector<int>(words, 0));
  std::vector<int> o(words);
  for (int i = 0; i < words; i++) o[i] = i;
  h.off.push_back(o);
  std::vector<i64> s;
  for (int r = 0; r < h.proc_used; r++) {
    for (int d = 0; d < nd; d++) s.push_back(h.strides[(size_t)r * nd + d] * words);
    s.push_back(1);
  }
  h.strides = s;
  for (auto &m : h.m) m *= words;
  h.m.push_back(1);
  for (auto &e : h.slot_extent) e *= words;
  h.ndim = nd + 1;
}

// Axis d (n = Q * P points) as two axes: z = g / P stays at d (Q points) and
// y = g % P is appended (P points), when both the owner contribution and the
// local offset of g split as a part of z plus a part of y that is zero at
// y = 0 -- for example a block axis (runs of P or a multiple of P) or an
// interleave of period P.  Returns false (h untouched) otherwise.
static bool separable(const dm::HostLayout &h, int d, int P) {
  const int n = h.L[d];
  const std::vector<int> &rp = h.rank_part[d], &of = h.off[d];
  for (int g = 0; g < n; g++) {
    const int z = g / P, y = g % P;
    if (rp[g] >= (1 << 27)) return false;  // a coordinate without owner
    if (rp[g] != rp[z * P] + rp[y] - rp[0] || of[g] != of[z * P] + of[y] - of[0]) return false;
    if (rp[y] < rp[0] || of[y] < of[0]) return false;
  }
  return true;
}

static void split_axis(dm::HostLayout &h, int d, int P) {
  const int n = h.L[d], Q = n / P, nd = h.ndim;
  std::vector<int> rz(Q), oz(Q), ry(P), oy(P);
  for (int z = 0; z < Q; z++) {
    rz[z] = h.rank_part[d][z * P];
    oz[z] = h.off[d][z * P];
  }
  for (int y = 0; y < P; y++) {
    ry[y] = h.rank_part[d][y] - h.rank_part[d][0];
    oy[y] = h.off[d][y] - h.off[d][0];
  }
  h.L[d] = Q;
  h.rank_part[d] = rz;
  h.off[d] = oz;
  h.L.push_back(P);
  h.rank_part.push_back(ry);
  h.off.push_back(oy);
  std::vector<i64> s;
  for (int r = 0; r < h.proc_used; r++) {
    for (int e = 0; e < nd; e++) s.push_back(h.strides[(size_t)r * nd + e]);
    s.push_back(h.strides[(size_t)r * nd + d]);
  }
  h.strides = s;
  h.m.push_back(h.m[d]);
  h.ndim = nd + 1;
}

// Whether consecutive coordinates of axis d step by one local offset within
// one owner (the axis is read or written in whole runs).
static bool unit_steps(const dm::HostLayout &h, int d, int from, int step, int count) {
  for (int k = 1; k < count; k++) {
    const int g = from + k * step, p = g - step;
    if (h.rank_part[d][g] != h.rank_part[d][p] || h.off[d][g] != h.off[d][p] + 1) return false;
  }
  return true;
}

// Both sides share their fastest axis d but not its order along it (for
// example a block axis against an interleaved one): the gather would walk
// 1D runs with scattered stores.  Split d into (z, y) with one side
// contiguous along y and the other along z, so the reshape becomes a 2D
// tile transpose over (z, y).  Among the periods P that separate both sides
// that way, the one with the squarest (z, y) is taken.
static bool split_shared_axis(dm_reshape_program *pg) {
  dm::HostLayout &S = pg->src, &D = pg->dst;
  const int d = S.fastest;
  if (D.fastest != d || S.ndim + 2 > DM_MAX_DIMS) return false;
  const int n = S.L[d];
  if (n < 4 || (unit_steps(S, d, 0, 1, n) && unit_steps(D, d, 0, 1, n))) return false;
  int best = 0;
  bool src_y = true;
  for (int P = 2; P <= n / 2; P++) {
    if (n % P != 0 || !separable(S, d, P) || !separable(D, d, P)) continue;
    // one side in runs along y, the other in runs along z
    const bool sy = unit_steps(S, d, 0, 1, P) && unit_steps(D, d, 0, P, n / P);
    const bool sz = unit_steps(D, d, 0, 1, P) && unit_steps(S, d, 0, P, n / P);
    if (!sy && !sz) continue;
    if (best == 0 || std::min(P, n / P) > std::min(best, n / best)) {
      best = P;
      src_y = sy;
    }
  }
  if (best == 0) return false;
  split_axis(S, d, best);
  split_axis(D, d, best);
  const int y = S.ndim - 1;
  S.fastest = src_y ? y : d;
  D.fastest = src_y ? d : y;
  return true;
}

// Relabel the global coordinates of axis d on both sides so that they run in
// `key`'s memory order (owner rank, then local offset).  A bijection applied
// to both tables keeps the pairing by global coordinate, and a tile over the
// relabelled axis then starts where one of `key`'s runs starts: with a
// rotated or interleaved side, global tiles would straddle its run
// boundaries and every row would touch partial cache lines.
static void relabel_axis(dm_reshape_program *pg, int d, const dm::HostLayout &key) {
  const int n = key.L[d];
  std::vector<int> perm(n);
  for (int i = 0; i < n; i++) perm[i] = i;
  const std::vector<int> &kr = key.rank_part[d], &ko = key.off[d];
  std::stable_sort(perm.begin(), perm.end(),
                   [&](int x, int y) { return kr[x] != kr[y] ? kr[x] < kr[y] : ko[x] < ko[y]; });
  for (dm::HostLayout *h : {&pg->src, &pg->dst}) {
    std::vector<int> r(n), o(n);
    for (int i = 0; i < n; i++) {
      r[i] = h->rank_part[d][perm[i]];
      o[i] = h->off[d][perm[i]];
    }
    h->rank_part[d].swap(r);
    h->off[d].swap(o);
  }
}

extern "C" ReshapePlan makeReshapePlan(MPI_Comm comm, MPI_Datatype base, Decomp src, Decomp dest) {
  if (src == nullptr || dest == nullptr || base == nullptr) {
    dm::fail(comm, "makeReshapePlan: NULL argument");
    return nullptr;
  }
  if (compatible(src, dest) == NOT_COMPATIBLE) {
    raise_error(comm, "Decomp not compatible in makeReshapePlan");
    return nullptr;
  }
  int npe = comm->world->nranks;
  if (src->proc_used < 1 || src->proc_used > npe) {
    raise_error(comm, "bad src proc count %d\n", src->proc_used);
    return nullptr;
  }
  if (dest->proc_used < 1 || dest->proc_used > npe) {
    raise_error(comm, "bad dest proc count %d\n", dest->proc_used);
    return nullptr;
  }
  std::string why;
  if (!dm::layout_rank_map_ok(src, &why) || !dm::layout_rank_map_ok(dest, &why)) {
    raise_error(comm, "makeReshapePlan: %s", why.c_str());
    return nullptr;
  }
  int ndim = src->ndim;
  if (ndim + 1 > DM_MAX_DIMS) {
    raise_error(comm, "makeReshapePlan: at most %d axes are supported", DM_MAX_DIMS - 1);
    return nullptr;
  }
  MPI_Aint lb, ext;
  dm_type_get_extent(base, &lb, &ext);
  if (ext <= 0) {
    raise_error(comm, "makeReshapePlan: element extent must be positive");
    return nullptr;
  }

  ReshapePlan plan = (ReshapePlan)calloc(1, sizeof(struct reshape_plan));
  plan->comm = src->comm;
  plan->src = src;
  plan->dest = dest;
  plan->order = (int *)calloc(ndim, sizeof(int));
  sort_order(ndim, plan->order, src);

  int me = comm->rank;
  plan->sends = message_list(src, dest, plan->order, base, me, true);
  plan->recvs = message_list(dest, src, plan->order, base, me, false);
  // reshape.c:421-443: the self message must appear on both sides.
  if (count_rank(plan->sends, me) != count_rank(plan->recvs, me)) {
    raise_error(comm, "Self send-recv don't match %d != %d\n", count_rank(plan->sends, me),
                count_rank(plan->recvs, me));
    freeReshapePlan(plan);
    return nullptr;
  }

  dm_reshape_program *pg = new dm_reshape_program;
  pg->elem = (int)ext;
  int w = 16;
  while (ext % w != 0) w >>= 1;
  pg->word = w;
  pg->words = (int)(ext / w);
  pg->src = dm::layout_from_decomp(src);
  pg->dst = dm::layout_from_decomp(dest);
  if (!pg->src.offsets_ok || !pg->dst.offsets_ok) {
    delete pg;
    raise_error(comm, "makeReshapePlan: a decomposition has a negative local offset or memory stride");
    freeReshapePlan(plan);
    return nullptr;
  }
  const bool split = getenv("DM_NO_SPLIT") == nullptr && split_shared_axis(pg);
  const int nd_el = pg->src.ndim;  // element axes (a split adds one)
  extend_by_words(pg->src, pg->words);
  extend_by_words(pg->dst, pg->words);
  pg->a = pg->dst.fastest;
  pg->b = pg->src.fastest;
  // (a split keeps its two halves as the tile axes, also over words)
  if (!split && pg->words > 1) {  // tile over the element axes, not the word axis
    auto pick = [&](const dm::HostLayout &h) {
      int best = 0;
      i64 bs = -1;
      for (int d = nd_el - 1; d >= 0; d--)
        if (h.L[d] > 1 && (bs < 0 || h.m[d] < bs)) bs = h.m[d], best = d;
      return best;
    };
    pg->a = pick(pg->dst);
    pg->b = pick(pg->src);
  }
  // writes run along a, reads along b
  if (getenv("DM_NO_RELABEL") == nullptr) {
    relabel_axis(pg, pg->a, pg->dst);
    if (pg->b != pg->a) relabel_axis(pg, pg->b, pg->src);
  }
  std::vector<int> rest;
  for (int d = 0; d < pg->dst.ndim; d++)
    if (d != pg->a && d != pg->b) rest.push_back(d);
  // slowest destination axes first so consecutive blocks stay close in memory
  std::stable_sort(rest.begin(), rest.end(),
                   [&](int x, int y) { return pg->dst.m[x] > pg->dst.m[y]; });
  pg->batch = rest;
  plan->prog = pg;
  return plan;
}

extern "C" void freeReshapePlan(ReshapePlan plan) {
  if (plan == nullptr) return;
  freeTypeEntryList(plan->sends);
  freeTypeEntryList(plan->recvs);
  free(plan->order);
  if (plan->prog != nullptr) {
    plan->prog->dsrc.release();
    plan->prog->ddst.release();
    delete plan->prog;
  }
  free(plan);
}

namespace {
struct Span {
  dm::Buffer buf;
  size_t off;
};
}  // namespace

// Check one side of a reshape against the buffer it runs on.
static bool fits(dm_comm_t comm, const char *what, const void *p, const dm::HostLayout &h, int word,
                 Span *sp) {
  if (!dm::resolve(p, &sp->buf, &sp->off)) {
    raise_error(comm, "runReshapePlan: %s pointer is not inside a distributed buffer (dm_malloc/dm_register)",
                what);
    return false;
  }
  if (sp->buf.nranks < h.proc_used) {
    raise_error(comm, "runReshapePlan: %s buffer has %d rank slots, the decomposition uses %d", what,
                sp->buf.nranks, h.proc_used);
    return false;
  }
  if (sp->off % word != 0 || sp->buf.slot % word != 0) {
    raise_error(comm, "runReshapePlan: %s pointer or slot not aligned to %d bytes", what, word);
    return false;
  }
  for (int r = 0; r < h.proc_used; r++) {
    if (sp->off + (size_t)h.slot_extent[r] * word > sp->buf.slot) {
      raise_error(comm, "runReshapePlan: %s data of rank %d (%lld units) overruns its slot", what, r,
                  (long long)h.slot_extent[r]);
      return false;
    }
  }
  return true;
}

namespace dm {
// Enqueue the gather on `stream` (no synchronisation).
bool enqueue_reshape(dm_comm_t comm, ReshapePlan plan, const void *src, void *dst, void *stream) {
  dm_reshape_program *pg = plan->prog;
  Span s, d;
  if (!fits(comm, "source", src, pg->src, pg->word, &s)) return false;
  if (!fits(comm, "destination", dst, pg->dst, pg->word, &d)) return false;
  if (s.buf.base == d.buf.base) {
    size_t se = 0, de = 0;
    for (auto e : pg->src.slot_extent) se = std::max(se, (size_t)e * pg->word);
    for (auto e : pg->dst.slot_extent) de = std::max(de, (size_t)e * pg->word);
    if (s.off < d.off + de && d.off < s.off + se) {
      raise_error(comm, "runReshapePlan: source and destination overlap");
      return false;
    }
  }
  // Nothing owned on either side: nothing to move.  (The kernels read
  // element 0 of the source region in place of absent elements, which is in
  // bounds whenever some rank's source extent is positive.)
  size_t src_max = 0, dst_max = 0;
  for (auto e : pg->src.slot_extent) src_max = std::max(src_max, (size_t)e);
  for (auto e : pg->dst.slot_extent) dst_max = std::max(dst_max, (size_t)e);
  if (src_max == 0 || dst_max == 0) return true;
  if (!pg->uploaded) {
    pg->dsrc.upload(pg->src);
    pg->ddst.upload(pg->dst);
    pg->uploaded = true;
  }
  ReshapeLaunch L{};
  L.src = pg->dsrc.dev;
  L.dst = pg->ddst.dev;
  L.ndim = pg->dst.ndim;
  for (int i = 0; i < L.ndim; i++) L.L[i] = pg->dst.L[i];
  L.a = pg->a;
  L.b = pg->b;
  L.nbatch = (int)pg->batch.size();
  for (int i = 0; i < L.nbatch; i++) L.batch_axis[i] = pg->batch[i];
  L.src_base = s.buf.base + s.off;
  L.dst_base = d.buf.base + d.off;
  L.src_slot = (i64)(s.buf.slot / pg->word);
  L.dst_slot = (i64)(d.buf.slot / pg->word);
  L.elem_bytes = pg->word;
  // non-temporal hint: 11-20% off the 512^3 int64 hops (each element is
  // read once and written once)
  static const bool nont = getenv("DM_NO_NONTEMPORAL") != nullptr;
  L.nontemporal = nont ? 0 : 1;
  return launch_ok(comm, launch_reshape(L, stream), "reshape kernel launch");
}
}  // namespace dm

static void run_reshape(MPI_Comm comm, ReshapePlan plan, void *src, void *dest);

// reshape.c:461-492.  Collective: inside a dm_spmd_run session every rank
// calls it with pointers into its own slots; rank 0's call (whose slot
// pointers are the buffers' slot-0 pointers) moves every rank's share.
extern "C" void runReshapePlan(MPI_Comm comm, ReshapePlan plan, void *src, void *dest) {
  if (dm::spmd_rank() >= 0) {
    std::vector<long long> sig;
    sig.push_back(plan ? dm::decomp_fingerprint(plan->src) : 0);
    sig.push_back(plan ? dm::decomp_fingerprint(plan->dest) : 0);
    dm::sig_pointer(&sig, src);
    dm::sig_pointer(&sig, dest);
    dm::spmd_collective_checked(sig, "runReshapePlan", [&] { run_reshape(comm, plan, src, dest); });
    return;
  }
  run_reshape(comm, plan, src, dest);
}

static void run_reshape(MPI_Comm comm, ReshapePlan plan, void *src, void *dest) {
  if (plan == nullptr) {
    dm::fail(comm, "runReshapePlan: NULL plan");
    return;
  }
  dm::check_device(comm);
  if (dm::error_pending() && dm::error_mode() == 1) return;
  void *stream = dm::world_stream(comm->world);
  if (dm::enqueue_reshape(comm, plan, src, dest, stream)) dm::world_finish(comm->world);
}

// reshape.c:523-567, for every rank's messages; returns 1 on the first
// violation after printing the reference's message.
// reshape.c checks the calling process's own message lists.  Inside a
// dm_spmd_run session each rank passes its own lengths, so only the view's
// rank is checked; outside one, the call stands for every rank and each
// rank's messages are checked against the same lengths.
extern "C" int checkReshapePlanBounds(ReshapePlan plan, int type_size, int src_len, int dest_len) {
  int P = plan->comm->world->nranks;
  const bool own = dm::spmd_rank() >= 0;
  for (int me = own ? plan->comm->rank : 0; me < (own ? plan->comm->rank + 1 : P); me++) {
    TypeEntryList sends = message_list(plan->src, plan->dest, plan->order, &dm_type_byte, me, true);
    TypeEntryList recvs = message_list(plan->dest, plan->src, plan->order, &dm_type_byte, me, false);
    // message_list with a 1-byte base counts elements; scale by type_size
    auto check = [&](TypeEntryList t, const char *kind, int len) -> int {
      if (t == nullptr) return 0;
      long long total = 0;
      for (int i = 0; i < t->count; i++) {
        long long size = (long long)t->list[i].type->size * type_size;
        long long lb = t->list[i].type->true_lb * type_size;
        long long ex = (t->list[i].type->true_extent - 1) * type_size + type_size;
        total += size;
        if (lb < 0) {
          fprintf(stderr, "Type %s[%d] has lb=%lld below start of array\n", kind, i, lb);
          return 1;
        }
        if (lb + ex > (long long)type_size * len) {
          fprintf(stderr, "Type %s[%d] extends beyond end of array top=%lld len=%lld\n", kind, i, lb + ex,
                  (long long)type_size * len);
          return 1;
        }
      }
      if (total > (long long)type_size * len) {
        fprintf(stderr, "More data %s than fits in array %lld\n", kind[0] == 'r' ? "read" : "sent", total);
        return 1;
      }
      return 0;
    };
    int bad = check(recvs, "recv", dest_len);
    if (!bad) bad = check(sends, "send", src_len);
    freeTypeEntryList(sends);
    freeTypeEntryList(recvs);
    if (bad) return 1;
  }
  return 0;
}

// ---------------------------------------------------------------- device tables

namespace dm {

void DeviceLayout::upload(const HostLayout &h) {
  release();
  size_t bytes = 0;
  auto take = [&](size_t n) {
    size_t at = bytes;
    bytes += (n + 255) & ~(size_t)255;
    return at;
  };
  std::vector<size_t> at_rank(h.ndim), at_off(h.ndim), at_addr(h.ndim);
  for (int d = 0; d < h.ndim; d++) {
    at_rank[d] = take(h.L[d] * sizeof(int));
    at_off[d] = take(h.L[d] * sizeof(int));
    at_addr[d] = take(h.L[d] * sizeof(i64));
  }
  size_t at_str = take(h.strides.size() * sizeof(i64));
  std::vector<char> host(bytes, 0);
  for (int d = 0; d < h.ndim; d++) {
    memcpy(&host[at_rank[d]], h.rank_part[d].data(), h.L[d] * sizeof(int));
    memcpy(&host[at_off[d]], h.off[d].data(), h.L[d] * sizeof(int));
    i64 *ad = (i64 *)&host[at_addr[d]];
    for (int g = 0; g < h.L[d]; g++) ad[g] = h.m[d] * h.off[d][g];
  }
  memcpy(&host[at_str], h.strides.data(), h.strides.size() * sizeof(i64));
  mem = device_alloc(bytes);
  if (mem == nullptr) {
    fail(nullptr, "device allocation of %zu bytes for layout tables failed", bytes);
    return;
  }
  h2d(mem, host.data(), bytes);
  this->bytes = bytes;
  char *b = (char *)mem;
  dev.ndim = h.ndim;
  dev.proc_used = h.proc_used;
  dev.uniform = h.uniform ? 1 : 0;
  for (int d = 0; d < h.ndim; d++) {
    dev.rank[d] = (const int *)(b + at_rank[d]);
    dev.off[d] = (const int *)(b + at_off[d]);
    dev.addr[d] = (const i64 *)(b + at_addr[d]);
  }
  dev.strides = (const i64 *)(b + at_str);
}

void DeviceLayout::release() {
  if (mem != nullptr) device_free(mem);
  mem = nullptr;
  bytes = 0;
}

}  // namespace dm
