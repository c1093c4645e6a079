// Request loop with continuous batching (dm_serve.h).
//
// State of a request: queued -> admitted (device buffer from the pool,
// within the HBM budget) -> launched (part of a batch on a stream) ->
// finished (the stream's last event for that batch has completed).  Each
// poll retires finished batches, admits queued requests in arrival order
// while they fit, and gives every idle stream a batch: the oldest admitted
// request plus the following admitted requests of the same precision,
// direction and rank, up to max_batch.
//
// Plans are device-resident: a transform's plan is, per axis, a radix
// schedule with its twiddles for the line length and tile height of that
// axis; the first request needing a (length, tile height, precision) builds
// and uploads that schedule once and every later request reuses it (least
// recently used entries go when the cache is full).  Tiles are not listed:
// the kernel computes a tile's lines from the transform's shape.  A batch is
// then: input copies, one batched pass launch per axis whose kernel argument
// lists each request's schedule, shape and data as a segment (nothing is
// built or uploaded per batch), output copies.  Four timing events per batch
// give each request's copy-in, kernel and copy-out time beside its queue and
// host time.
//
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
#include <tuple>
#include <vector>

#include "dm_batch.h"
#include "dm_serve.h"
#include "fft_program.h"
#include "internal.h"
#include "kernels.h"

namespace {

using Clock = std::chrono::steady_clock;

enum State { QUEUED, ADMITTED, LAUNCHED, FINISHED };

// One radix schedule on the device: its BatchSched (tw_off 0) and roots.
struct Sched {
  void *dev = nullptr;
  size_t bytes = 0;
  const dm::BatchSched *sched = nullptr;
  const void *tw = nullptr;
  size_t lds = 0;
  int primes = 0;
  int any = 0;              // a direct-sum stage (a prime above 13)
  int long_lines = 0;       // lines too long for twiddles in LDS
  int refs = 0;             // requests admitted with this schedule and not yet finished
  long long last_use = 0;
};
typedef std::tuple<int, int, int> SchedKey;  // N, T, dbl

// A request's plan: one segment (less data and tile end) per axis pass.
struct ReqPlan {
  int npass = 0;
  Sched *sched[3] = {nullptr, nullptr, nullptr};
  dm::BatchSeg seg[3];
  long long ntiles[3] = {0, 0, 0};
  size_t lds = 0;
  int primes = 0;
  int any = 0;
  int long_lines = 0;
};

struct Req {
  int ndim = 0;
  int shape[3] = {1, 1, 1};
  int dbl = 0, inv = 0;
  const void *in = nullptr;
  void *out = nullptr;
  size_t bytes = 0, cls = 0;
  void *dev = nullptr;
  ReqPlan plan;
  bool planned = false;
  State state = QUEUED;
  double t_sub = 0, t_launch = 0, t_done = 0;
  double host_ms = 0, h2d_ms = 0, kern_ms = 0, d2h_ms = 0;
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};  // start, inputs in, kernels done, outputs out
  bool busy = false;
  std::vector<long long> reqs;
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

constexpr size_t kSchedCacheEntries = 1024;
constexpr size_t kSchedCacheBytes = (size_t)64 << 20;

}  // namespace

struct dm_server {
  size_t budget = 0, in_use = 0, peak = 0, held = 0;
  int max_batch = 1;
  std::vector<Slot> slots;
  std::vector<Req> reqs;
  std::deque<long long> queue, ready;
  std::map<size_t, std::vector<void *>> pool;  // free device buffers by class
  std::map<SchedKey, Sched *> scheds;
  size_t sched_bytes = 0;
  long long sched_builds = 0, sched_clock = 0;
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

static void sched_free(Sched *p) {
  dm::device_free(p->dev);
  delete p;
}

// Drop least recently used schedules that no request holds until the cache
// is within its entry and byte limits.
static void sched_trim(DmServer s) {
  while (s->scheds.size() > kSchedCacheEntries || s->sched_bytes > kSchedCacheBytes) {
    auto victim = s->scheds.end();
    for (auto it = s->scheds.begin(); it != s->scheds.end(); ++it)
      if (it->second->refs == 0 && (victim == s->scheds.end() || it->second->last_use < victim->second->last_use))
        victim = it;
    if (victim == s->scheds.end()) return;  // every cached schedule is in use
    s->sched_bytes -= victim->second->bytes;
    sched_free(victim->second);
    s->scheds.erase(victim);
  }
}

// The device schedule of lines of N points in tiles of T lines, with a
// reference taken for the caller: from the cache, or built and uploaded
// once.  Null (after an error message) if the device refuses.
static Sched *sched_for(DmServer s, int N, int T, int dbl) {
  SchedKey key(N, T, dbl);
  auto it = s->scheds.find(key);
  if (it != s->scheds.end()) {
    it->second->last_use = ++s->sched_clock;
    it->second->refs++;
    return it->second;
  }
  // the schedule build_batch_host gives lines of N points in tiles of T
  dm::BatchSched one{};
  std::vector<double> tw;
  dm::make_batch_sched(N, T, DM_BATCH_RADIX_LARGE_FIRST, &one, &tw);
  const size_t ntw = tw.size() / 2;  // N stage roots, then any direct-sum stages' roots
  const size_t tw_bytes = ntw * 16;
  const size_t o_tw = align256(sizeof one);
  std::vector<char> img(o_tw + tw_bytes, 0);
  memcpy(img.data(), &one, sizeof one);
  dm::twiddle_image(tw.data(), ntw, dbl, img.data() + o_tw);
  Sched *p = new Sched;
  p->bytes = img.size();
  p->dev = dm::device_alloc(p->bytes);
  if (p->dev == nullptr || hipMemcpy(p->dev, img.data(), p->bytes, hipMemcpyHostToDevice) != hipSuccess) {
    dm::device_free(p->dev);
    delete p;
    raise_error(nullptr, "dm_server: device allocation or upload of a schedule failed");
    return nullptr;
  }
  p->sched = (const dm::BatchSched *)p->dev;
  p->tw = (char *)p->dev + o_tw;
  p->lds = dm::batch_lds_bytes(dbl, T, N);
  p->long_lines = dm::batch_tw_in_lds(N) ? 0 : 1;
  for (int k = 0; k < one.nstage; k++)
    if (one.radix[k] == 5 || one.radix[k] == 7 || one.radix[k] == 11 || one.radix[k] == 13) p->primes = 1;
  p->any = dm::batch_sched_any(one) ? 1 : 0;
  p->last_use = ++s->sched_clock;
  p->refs = 1;  // held by the caller, so the trim below keeps it
  s->scheds[key] = p;
  s->sched_bytes += p->bytes;
  s->sched_builds++;
  sched_trim(s);
  return p;
}

static void plan_release(DmServer s, Req &r) {
  for (int d = 0; d < r.plan.npass; d++)
    if (r.plan.sched[d]) r.plan.sched[d]->refs--;
  r.plan = ReqPlan();
  r.planned = false;
  (void)s;
}

// A request's plan: for the pass over each axis (last axis first, as the
// batched transforms run), the schedule and the segment geometry, with the
// tile height rule of build_batch_host (as many lines as the tile holds; a
// column walk never more than its contiguous run).
static bool plan_request(DmServer s, Req &r) {
  ReqPlan &pl = r.plan;
  pl = ReqPlan();
  pl.npass = r.ndim;
  for (int k = 0; k < r.ndim; k++) {
    const int d = r.ndim - 1 - k;
    long long inner = 1, outer = 1;
    for (int e = d + 1; e < r.ndim; e++) inner *= r.shape[e];
    for (int e = 0; e < d; e++) outer *= r.shape[e];
    const int N = r.shape[d];
    const int T = dm::batch_tile_height(N, inner, r.dbl, DM_BATCH_TILE_FULL);
    Sched *sc = sched_for(s, N, T, r.dbl);
    if (sc == nullptr) {
      plan_release(s, r);
      return false;
    }
    pl.sched[k] = sc;
    dm::BatchSeg &g = pl.seg[k];
    g = dm::BatchSeg{};
    g.sched = sc->sched;
    g.twiddle = sc->tw;
    g.inner = inner;
    g.outer = (int)outer;
    g.N = N;
    g.T = T;
    if (inner == 1) {
      g.ctiles = 0;
      pl.ntiles[k] = (outer + T - 1) / T;
    } else {
      g.ctiles = (int)((inner + T - 1) / T);
      pl.ntiles[k] = outer * g.ctiles;
    }
    pl.lds = std::max(pl.lds, sc->lds);
    pl.primes = pl.primes || sc->primes;
    pl.any = pl.any || sc->any;
    pl.long_lines = pl.long_lines || sc->long_lines;
  }
  r.planned = true;
  return true;
}

extern "C" DmServer dm_server_create(size_t budget, int streams, int max_batch) {
  if (!dm::device_available()) {
    raise_error(nullptr, "dm_server_create: no HIP device is available");
    return nullptr;
  }
  if (budget == 0 || streams < 1 || streams > 16 || max_batch < 1 || max_batch > DM_BATCH_MAX_SEG) {
    raise_error(nullptr, "dm_server_create: need budget > 0, 1 <= streams <= 16, 1 <= max_batch <= %d",
                DM_BATCH_MAX_SEG);
    return nullptr;
  }
  DmServer s = new dm_server;
  s->budget = budget;
  s->max_batch = max_batch;
  s->slots.resize(streams);
  for (auto &sl : s->slots) {
    // default priority, as every stream of this library
    bool ok = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) == hipSuccess;
    for (auto &e : sl.ev) ok = ok && hipEventCreate(&e) == hipSuccess;
    if (!ok) {
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
    if (shape[d] < 1 || shape[d] > dm::batch_tile_capacity(r.dbl) || !dm::choose_radices(shape[d], &rad, true) ||
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

static float elapsed(hipEvent_t a, hipEvent_t b) {
  float ms = 0;
  return hipEventElapsedTime(&ms, a, b) == hipSuccess ? ms : 0.0f;
}

static void retire(DmServer s) {
  for (auto &sl : s->slots) {
    if (!sl.busy || hipEventQuery(sl.ev[3]) != hipSuccess) continue;
    double t = s->now();
    const float h2d = elapsed(sl.ev[0], sl.ev[1]), kern = elapsed(sl.ev[1], sl.ev[2]),
                d2h = elapsed(sl.ev[2], sl.ev[3]);
    for (long long id : sl.reqs) {
      Req &r = s->reqs[id];
      r.state = FINISHED;
      r.t_done = t;
      r.h2d_ms = h2d;
      r.kern_ms = kern;
      r.d2h_ms = d2h;
      pool_put(s, r.dev, r.cls);
      r.dev = nullptr;
      plan_release(s, r);
      s->in_use -= r.cls;
      s->unfinished--;
    }
    sl.reqs.clear();
    sl.busy = false;
  }
  sched_trim(s);
}

static void admit(DmServer s) {
  while (!s->queue.empty()) {
    Req &r = s->reqs[s->queue.front()];
    if (s->in_use + r.cls > s->budget) break;  // arrival order: no overtaking
    if (!plan_request(s, r)) break;
    r.dev = pool_get(s, r.cls);
    if (r.dev == nullptr) {  // device full: wait for buffers to come back
      plan_release(s, r);
      break;
    }
    s->in_use += r.cls;
    s->peak = std::max(s->peak, s->in_use);
    r.state = ADMITTED;
    s->ready.push_back(s->queue.front());
    s->queue.pop_front();
  }
}

static bool launch(DmServer s, Slot &sl) {
  const double t_host = s->now();
  // the oldest admitted request and the admitted ones that batch with it;
  // they leave the ready list only once the batch is enqueued
  const Req &head = s->reqs[s->ready.front()];
  std::vector<long long> ids;
  for (auto it = s->ready.begin(); it != s->ready.end() && (int)ids.size() < s->max_batch; ++it) {
    const Req &r = s->reqs[*it];
    if (r.dbl == head.dbl && r.inv == head.inv && r.ndim == head.ndim) ids.push_back(*it);
  }
  const int nd = head.ndim, dbl = head.dbl, inv = head.inv;
  hipStream_t st = sl.stream;
  bool ok = hipEventRecord(sl.ev[0], st) == hipSuccess;
  for (long long id : ids) {
    const Req &r = s->reqs[id];
    ok = ok && hipMemcpyAsync(r.dev, r.in, r.bytes, hipMemcpyHostToDevice, st) == hipSuccess;
  }
  ok = ok && hipEventRecord(sl.ev[1], st) == hipSuccess;
  size_t lds = 0;
  int primes = 0, long_lines = 0, any = 0;
  for (long long id : ids) {
    lds = std::max(lds, s->reqs[id].plan.lds);
    primes = primes || s->reqs[id].plan.primes;
    any = any || s->reqs[id].plan.any;
    long_lines = long_lines || s->reqs[id].plan.long_lines;
  }
  for (int d = 0; d < nd && ok; d++) {
    dm::BatchLaunch L{};
    L.dbl = dbl;
    L.inverse = inv;
    L.primes = primes;
    L.long_lines = long_lines;
    L.any = any;
    L.lds = lds;
    L.nseg = (int)ids.size();
    long long end = 0;
    for (size_t k = 0; k < ids.size(); k++) {
      const Req &r = s->reqs[ids[k]];
      end += r.plan.ntiles[d];
      L.seg[k] = r.plan.seg[d];
      L.seg[k].data = r.dev;
      L.seg[k].end = end;
    }
    L.ntiles = end;
    L.grid = end;  // one workgroup per tile
    ok = dm::launch_ok(nullptr, dm::launch_batch_pass(L, st), "dm_server: batched pass launch");
  }
  ok = ok && hipEventRecord(sl.ev[2], st) == hipSuccess;
  for (long long id : ids) {
    const Req &r = s->reqs[id];
    ok = ok && hipMemcpyAsync(r.out, r.dev, r.bytes, hipMemcpyDeviceToHost, st) == hipSuccess;
  }
  ok = ok && hipEventRecord(sl.ev[3], st) == hipSuccess;
  if (!ok) {
    // the requests stay admitted (ready) with their buffers; the caller sees
    // the failure through poll's -1
    (void)hipStreamSynchronize(st);
    raise_error(nullptr, "dm_server: a copy, launch or event call failed");
    return false;
  }
  for (long long id : ids) s->ready.erase(std::find(s->ready.begin(), s->ready.end(), id));
  const double t = s->now();
  for (long long id : ids) {
    Req &r = s->reqs[id];
    r.state = LAUNCHED;
    r.t_launch = t_host;  // the batch's first event is recorded right after
    r.host_ms = (t - t_host) * 1e3;
  }
  sl.reqs = ids;
  sl.busy = true;
  s->batches++;
  s->max_seen = std::max<long long>(s->max_seen, (long long)ids.size());
  return true;
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

extern "C" int dm_server_breakdown(DmServer s, long long id, double *ms) {
  if (s == nullptr || ms == nullptr || id < 0 || id >= (long long)s->reqs.size()) return 1;
  const Req &r = s->reqs[id];
  if (r.state != FINISHED) return 1;
  ms[0] = (r.t_launch - r.t_sub) * 1e3;
  ms[1] = r.host_ms;
  ms[2] = r.h2d_ms;
  ms[3] = r.kern_ms;
  ms[4] = r.d2h_ms;
  ms[5] = (r.t_done - r.t_launch) * 1e3 - (r.h2d_ms + r.kern_ms + r.d2h_ms);
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

extern "C" int dm_server_plan_stats(DmServer s, long long *cached, long long *builds, size_t *bytes) {
  if (s == nullptr) return 1;
  if (cached) *cached = (long long)s->scheds.size();
  if (builds) *builds = s->sched_builds;
  if (bytes) *bytes = s->sched_bytes;
  return 0;
}

extern "C" void dm_server_destroy(DmServer s) {
  if (s == nullptr) return;
  for (auto &sl : s->slots)
    if (sl.stream) (void)hipStreamSynchronize(sl.stream);
  for (auto &sl : s->slots) {
    for (auto &e : sl.ev)
      if (e) (void)hipEventDestroy(e);
    if (sl.stream) (void)hipStreamDestroy(sl.stream);
  }
  for (auto &r : s->reqs)
    if (r.dev) dm::device_free(r.dev);
  for (auto &kv : s->pool)
    for (void *p : kv.second) dm::device_free(p);
  for (auto &kv : s->scheds) sched_free(kv.second);
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
