// Request loop with continuous batching (dm_serve.h).
//
// State of a request: queued -> admitted (device buffer from the pool,
// within the HBM budget) -> launched (part of a batch on a stream) ->
// finished (the stream's event for that batch has completed).  Each poll
// retires finished batches, admits queued requests in arrival order while
// they fit, and gives every idle stream a batch: the oldest admitted
// request plus the following admitted requests of the same precision,
// direction and rank, up to max_batch.  A batch is: table upload (from a
// pinned staging area into the stream's device arena), input copies, one
// batched pass launch per axis (dm_batch kernels), output copies, event.
// A request's transform depends only on its own shape (tile heights and
// radix schedules are functions of the line length), so its result is the
// same alone and in any mix.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <thread>
#include <vector>

#include "dm_batch.h"
#include "dm_serve.h"
#include "fft_program.h"
#include "internal.h"
#include "kernels.h"

namespace {

using Clock = std::chrono::steady_clock;

enum State { QUEUED, ADMITTED, LAUNCHED, FINISHED };

struct Req {
  int ndim = 0;
  int shape[3] = {1, 1, 1};
  int dbl = 0, inv = 0;
  const void *in = nullptr;
  void *out = nullptr;
  size_t bytes = 0, cls = 0;
  void *dev = nullptr;
  State state = QUEUED;
  double t_sub = 0, t_launch = 0, t_done = 0;
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  bool busy = false;
  std::vector<long long> reqs;
  void *arena = nullptr;   // device tables of the batch in flight
  size_t arena_cap = 0;
  char *stage = nullptr;   // pinned host image of those tables
  size_t stage_cap = 0;
};

// Buffer size classes: powers of two up to 2 MiB, then multiples of 2 MiB.
size_t size_class(size_t b) {
  const size_t step = (size_t)2 << 20;
  if (b >= step) return (b + step - 1) / step * step;
  size_t c = 4096;
  while (c < b) c <<= 1;
  return c;
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// A refused request is the caller's to handle (submit returns -1): the
// reason goes to dm_last_error without the abort of error mode 0.
void reject(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
void reject(const char *fmt, ...) {
  char msg[256];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(msg, sizeof msg, fmt, ap);
  va_end(ap);
  dm::ScopedErrorMode record(1);
  raise_error(nullptr, "%s", msg);
}

}  // namespace

struct dm_server {
  size_t budget = 0, in_use = 0, peak = 0, held = 0;
  int max_batch = 1;
  std::vector<Slot> slots;
  std::vector<Req> reqs;
  std::deque<long long> queue, ready;
  std::map<size_t, std::vector<void *>> pool;  // free device buffers by class
  Clock::time_point t0 = Clock::now();
  long long batches = 0, max_seen = 0;
  long long unfinished = 0;

  double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

static void *pool_get(DmServer s, size_t cls) {
  auto &v = s->pool[cls];
  if (!v.empty()) {
    void *p = v.back();
    v.pop_back();
    return p;
  }
  // keep the device memory held within the budget: drop idle buffers of
  // other classes first
  for (auto it = s->pool.begin(); s->held + cls > s->budget && it != s->pool.end(); ++it)
    while (!it->second.empty() && s->held + cls > s->budget) {
      dm::device_free(it->second.back());
      it->second.pop_back();
      s->held -= it->first;
    }
  void *p = dm::device_alloc(cls);
  if (p != nullptr) s->held += cls;
  return p;
}

static void pool_put(DmServer s, void *p, size_t cls) { s->pool[cls].push_back(p); }

extern "C" DmServer dm_server_create(size_t budget, int streams, int max_batch) {
  if (!dm::device_available()) {
    raise_error(nullptr, "dm_server_create: no HIP device is available");
    return nullptr;
  }
  if (budget == 0 || streams < 1 || streams > 16 || max_batch < 1) {
    raise_error(nullptr, "dm_server_create: need budget > 0, 1 <= streams <= 16, max_batch >= 1");
    return nullptr;
  }
  DmServer s = new dm_server;
  s->budget = budget;
  s->max_batch = max_batch;
  s->slots.resize(streams);
  for (auto &sl : s->slots) {
    // default priority, as every stream of this library
    if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess) {
      dm_server_destroy(s);
      raise_error(nullptr, "dm_server_create: stream or event creation failed");
      return nullptr;
    }
  }
  return s;
}

extern "C" long long dm_server_submit(DmServer s, int ndim, const int *shape, int precision, int direction,
                                      const void *in, void *out) {
  if (s == nullptr || shape == nullptr || in == nullptr || out == nullptr || ndim < 1 || ndim > 3 ||
      (direction != -1 && direction != 1)) {
    reject("dm_server_submit: bad arguments");
    return -1;
  }
  Req r;
  r.ndim = ndim;
  r.dbl = precision ? 1 : 0;
  r.inv = direction == 1 ? 1 : 0;
  size_t n = 1;
  std::vector<int> rad;
  for (int d = 0; d < ndim; d++) {
    r.shape[d] = shape[d];
    if (shape[d] < 1 || shape[d] > dm::batch_tile_capacity(r.dbl) || !dm::choose_radices(shape[d], &rad) ||
        rad.size() > 16) {
      reject("dm_server_submit: axis length %d is not supported", shape[d]);
      return -1;
    }
    n *= (size_t)shape[d];
  }
  r.in = in;
  r.out = out;
  r.bytes = n * (r.dbl ? 16 : 8);
  r.cls = size_class(r.bytes);
  if (r.cls > s->budget) {
    reject("dm_server_submit: request of %zu bytes exceeds the budget", r.bytes);
    return -1;
  }
  r.t_sub = s->now();
  long long id = (long long)s->reqs.size();
  s->reqs.push_back(r);
  s->queue.push_back(id);
  s->unfinished++;
  return id;
}

static void retire(DmServer s) {
  for (auto &sl : s->slots) {
    if (!sl.busy || hipEventQuery(sl.done) != hipSuccess) continue;
    double t = s->now();
    for (long long id : sl.reqs) {
      Req &r = s->reqs[id];
      r.state = FINISHED;
      r.t_done = t;
      pool_put(s, r.dev, r.cls);
      r.dev = nullptr;
      s->in_use -= r.cls;
      s->unfinished--;
    }
    sl.reqs.clear();
    sl.busy = false;
  }
}

static void admit(DmServer s) {
  while (!s->queue.empty()) {
    Req &r = s->reqs[s->queue.front()];
    if (s->in_use + r.cls > s->budget) break;  // arrival order: no overtaking
    r.dev = pool_get(s, r.cls);
    if (r.dev == nullptr) break;  // device full: wait for buffers to come back
    s->in_use += r.cls;
    s->peak = std::max(s->peak, s->in_use);
    r.state = ADMITTED;
    s->ready.push_back(s->queue.front());
    s->queue.pop_front();
  }
}

static bool grow(Slot &sl, size_t bytes) {
  if (bytes <= sl.arena_cap && bytes <= sl.stage_cap) return true;
  size_t cap = std::max(bytes, std::max(sl.arena_cap, sl.stage_cap) * 2);
  dm::device_free(sl.arena);
  if (sl.stage) (void)hipHostFree(sl.stage);
  sl.arena = dm::device_alloc(cap);
  void *h = nullptr;
  if (hipHostMalloc(&h, cap, hipHostMallocDefault) != hipSuccess) h = nullptr;
  sl.stage = (char *)h;
  sl.arena_cap = sl.stage_cap = (sl.arena && sl.stage) ? cap : 0;
  return sl.arena_cap >= bytes;
}

static bool launch(DmServer s, Slot &sl) {
  // the oldest admitted request and the admitted ones that batch with it
  const Req &head = s->reqs[s->ready.front()];
  std::vector<long long> ids;
  for (auto it = s->ready.begin(); it != s->ready.end() && (int)ids.size() < s->max_batch;) {
    const Req &r = s->reqs[*it];
    if (r.dbl == head.dbl && r.inv == head.inv && r.ndim == head.ndim) {
      ids.push_back(*it);
      it = s->ready.erase(it);
    } else {
      ++it;
    }
  }
  const int nd = head.ndim, dbl = head.dbl, inv = head.inv;
  std::vector<int> shapes;
  std::vector<void *> ptrs;
  for (long long id : ids) {
    const Req &r = s->reqs[id];
    shapes.insert(shapes.end(), r.shape, r.shape + nd);
    ptrs.push_back(r.dev);
  }
  dm::BatchHost h;
  if (!dm::build_batch_host((int)ids.size(), nd, shapes.data(), dbl, DM_BATCH_TILE_FULL,
                            DM_BATCH_RADIX_LARGE_FIRST, &h))
    return false;
  // table image: pointers | schedules | roots | tiles of each pass
  size_t o_ptr = 0, o_sched = align256(ptrs.size() * sizeof(void *));
  size_t o_tw = o_sched + align256(h.sched.size() * sizeof(dm::BatchSched));
  size_t tw_bytes = (h.tw.size() / 2) * (dbl ? 16 : 8);
  std::vector<size_t> o_tiles(nd);
  size_t at = o_tw + align256(tw_bytes);
  for (int p = 0; p < nd; p++) {
    o_tiles[p] = at;
    at += align256(h.tiles[p].size() * sizeof(dm::BatchTile));
  }
  if (!grow(sl, at)) return false;
  memcpy(sl.stage + o_ptr, ptrs.data(), ptrs.size() * sizeof(void *));
  memcpy(sl.stage + o_sched, h.sched.data(), h.sched.size() * sizeof(dm::BatchSched));
  if (dbl) {
    memcpy(sl.stage + o_tw, h.tw.data(), tw_bytes);
  } else {
    float *f = (float *)(sl.stage + o_tw);
    for (size_t i = 0; i < h.tw.size(); i++) f[i] = (float)h.tw[i];
  }
  for (int p = 0; p < nd; p++)
    memcpy(sl.stage + o_tiles[p], h.tiles[p].data(), h.tiles[p].size() * sizeof(dm::BatchTile));
  char *arena = (char *)sl.arena;
  hipStream_t st = sl.stream;
  bool ok = hipMemcpyAsync(arena, sl.stage, at, hipMemcpyHostToDevice, st) == hipSuccess;
  for (long long id : ids) {
    const Req &r = s->reqs[id];
    ok = ok && hipMemcpyAsync(r.dev, r.in, r.bytes, hipMemcpyHostToDevice, st) == hipSuccess;
  }
  long long resident = dm::batch_resident_groups(dbl, h.primes, h.lds);
  for (int p = 0; p < nd; p++) {
    dm::BatchLaunch L{};
    L.tiles = (const dm::BatchTile *)(arena + o_tiles[p]);
    L.ntiles = (long long)h.tiles[p].size();
    L.sched = (const dm::BatchSched *)(arena + o_sched);
    L.data = (void *const *)(arena + o_ptr);
    L.twiddle = arena + o_tw;
    L.dbl = dbl;
    L.inverse = inv;
    L.primes = h.primes;
    L.grid = std::min<long long>(L.ntiles, resident);
    L.lds = h.lds;
    dm::launch_batch_pass(L, st);
  }
  for (long long id : ids) {
    const Req &r = s->reqs[id];
    ok = ok && hipMemcpyAsync(r.out, r.dev, r.bytes, hipMemcpyDeviceToHost, st) == hipSuccess;
  }
  ok = ok && hipEventRecord(sl.done, st) == hipSuccess;
  double t = s->now();
  for (long long id : ids) {
    s->reqs[id].state = LAUNCHED;
    s->reqs[id].t_launch = t;
  }
  sl.reqs = ids;
  sl.busy = true;
  s->batches++;
  s->max_seen = std::max<long long>(s->max_seen, (long long)ids.size());
  if (!ok) raise_error(nullptr, "dm_server: a copy or event call failed");
  return ok;
}

extern "C" int dm_server_poll(DmServer s) {
  if (s == nullptr) return -1;
  retire(s);
  admit(s);
  for (auto &sl : s->slots) {
    if (s->ready.empty()) break;
    if (!sl.busy && !launch(s, sl)) return -1;
  }
  return (int)s->unfinished;
}

extern "C" int dm_server_wait(DmServer s, long long id) {
  if (s == nullptr || id < 0 || id >= (long long)s->reqs.size()) return 1;
  while (s->reqs[id].state != FINISHED) {
    if (dm_server_poll(s) < 0) return 1;
    if (s->reqs[id].state != FINISHED) std::this_thread::yield();
  }
  return 0;
}

extern "C" int dm_server_drain(DmServer s) {
  if (s == nullptr) return 1;
  while (s->unfinished > 0) {
    if (dm_server_poll(s) < 0) return 1;
    if (s->unfinished > 0) std::this_thread::yield();
  }
  return 0;
}

extern "C" int dm_server_times(DmServer s, long long id, double *submitted, double *launched, double *finished) {
  if (s == nullptr || id < 0 || id >= (long long)s->reqs.size()) return 1;
  const Req &r = s->reqs[id];
  if (submitted) *submitted = r.t_sub;
  if (launched) *launched = r.t_launch;
  if (finished) *finished = r.t_done;
  return 0;
}

extern "C" int dm_server_stats(DmServer s, long long *batches, long long *max_batch_seen, size_t *peak_bytes,
                               size_t *pool_bytes) {
  if (s == nullptr) return 1;
  if (batches) *batches = s->batches;
  if (max_batch_seen) *max_batch_seen = s->max_seen;
  if (peak_bytes) *peak_bytes = s->peak;
  if (pool_bytes) *pool_bytes = s->held;
  return 0;
}

extern "C" void dm_server_destroy(DmServer s) {
  if (s == nullptr) return;
  for (auto &sl : s->slots)
    if (sl.stream) (void)hipStreamSynchronize(sl.stream);
  for (auto &sl : s->slots) {
    if (sl.done) (void)hipEventDestroy(sl.done);
    if (sl.stream) (void)hipStreamDestroy(sl.stream);
    dm::device_free(sl.arena);
    if (sl.stage) (void)hipHostFree(sl.stage);
  }
  for (auto &r : s->reqs)
    if (r.dev) dm::device_free(r.dev);
  for (auto &kv : s->pool)
    for (void *p : kv.second) dm::device_free(p);
  delete s;
}

extern "C" void *dm_host_alloc(size_t bytes) {
  void *p = nullptr;
  if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) return nullptr;
  return p;
}

extern "C" void dm_host_free(void *p) {
  if (p) (void)hipHostFree(p);
}
