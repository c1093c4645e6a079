// One host thread per rank (dm_spmd.h): a session runs a driver body on P
// threads; the collective calls rendezvous and rank 0 launches for all.
#include <condition_variable>
#include <mutex>
#include <thread>
#include <vector>

#include "dm_spmd.h"
#include "internal.h"

namespace {

struct Session {
  int P = 0;
  dm_world *world = nullptr;
  std::mutex m;
  std::condition_variable cv;
  int waiting = 0;
  long long generation = 0;
  long long sum = 0, result = 0;  // all-reduce scratch
  double dmax = 0, dresult = 0;
  void *shared = nullptr;         // collective allocation
  std::vector<std::vector<long long>> sigs;  // collective arguments, per rank
  bool agree = true;
};

thread_local Session *tls_session = nullptr;
thread_local int tls_rank = -1;

// Generation barrier: the last arrival releases everyone; `last` runs under
// the lock by that arrival before the release.
template <typename F>
void arrive(Session *s, F last) {
  std::unique_lock<std::mutex> lk(s->m);
  long long gen = s->generation;
  if (++s->waiting == s->P) {
    last();
    s->waiting = 0;
    s->generation++;
    s->cv.notify_all();
  } else {
    s->cv.wait(lk, [&] { return s->generation != gen; });
  }
}

Session *session_of(MPI_Comm me) {
  if (tls_session == nullptr || me == nullptr || me->world != tls_session->world) return nullptr;
  return tls_session;
}

}  // namespace

namespace dm {
int spmd_rank() { return tls_session ? tls_rank : -1; }

void spmd_collective(const std::function<void()> &fn) {
  Session *s = tls_session;
  arrive(s, [] {});
  if (tls_rank == 0) fn();
  arrive(s, [] {});
}

void spmd_barrier_internal() {
  if (tls_session != nullptr) arrive(tls_session, [] {});
}

bool spmd_collective_checked(const std::vector<long long> &sig, const char *what, const std::function<void()> &fn) {
  Session *s = tls_session;
  {
    std::lock_guard<std::mutex> lk(s->m);
    s->sigs[tls_rank] = sig;
  }
  arrive(s, [s] {
    s->agree = true;
    for (int r = 1; r < s->P; r++) s->agree = s->agree && s->sigs[r] == s->sigs[0];
  });
  const bool agree = s->agree;
  if (agree && tls_rank == 0) fn();
  arrive(s, [] {});
  if (!agree)
    raise_error(s->world->views[tls_rank],
                "%s: the ranks' arguments differ (each rank passes the same plan and its own slot of the same "
                "buffers at the same offset)",
                what);
  return agree;
}
}  // namespace dm

extern "C" int dm_spmd_run(MPI_Comm world, void (*body)(MPI_Comm me, void *arg), void *arg) {
  if (world == nullptr || body == nullptr) return DM_ERR_ARG;
  if (tls_session != nullptr) {
    raise_error(world, "dm_spmd_run: sessions do not nest");
    return DM_ERR_ARG;
  }
  Session s;
  s.P = world->world->nranks;
  s.world = world->world;
  s.sigs.resize(s.P);
  std::vector<std::thread> threads;
  threads.reserve(s.P);
  for (int r = 0; r < s.P; r++)
    threads.emplace_back([&s, r, body, arg] {
      tls_session = &s;
      tls_rank = r;
      body(s.world->views[r], arg);
      tls_session = nullptr;
      tls_rank = -1;
    });
  for (auto &t : threads) t.join();
  return DM_SUCCESS;
}

extern "C" void dm_spmd_barrier(MPI_Comm me) {
  Session *s = session_of(me);
  if (s) arrive(s, [] {});
}

extern "C" long long dm_spmd_allreduce_sum(MPI_Comm me, long long value) {
  Session *s = session_of(me);
  if (s == nullptr) return value;
  arrive(s, [s] { s->sum = 0; });
  {
    std::lock_guard<std::mutex> lk(s->m);
    s->sum += value;
  }
  arrive(s, [s] { s->result = s->sum; });
  long long r = s->result;
  arrive(s, [] {});  // nobody resets sum before everyone has read result
  return r;
}

extern "C" double dm_spmd_allreduce_max(MPI_Comm me, double value) {
  Session *s = session_of(me);
  if (s == nullptr) return value;
  arrive(s, [s] { s->dmax = -1e308; });
  {
    std::lock_guard<std::mutex> lk(s->m);
    if (value > s->dmax) s->dmax = value;
  }
  arrive(s, [s] { s->dresult = s->dmax; });
  double r = s->dresult;
  arrive(s, [] {});
  return r;
}

extern "C" void *dm_spmd_malloc(MPI_Comm me, size_t bytes_per_rank) {
  Session *s = session_of(me);
  if (s == nullptr) {
    raise_error(me, "dm_spmd_malloc: not inside dm_spmd_run");
    return nullptr;
  }
  arrive(s, [s, bytes_per_rank] { s->shared = dm_malloc(s->world->views[0], bytes_per_rank); });
  void *base = s->shared;
  arrive(s, [] {});
  if (base == nullptr) return nullptr;
  return (char *)base + (size_t)tls_rank * dm_slot_bytes(base);
}

extern "C" void dm_spmd_free(MPI_Comm me, void *mine) {
  Session *s = session_of(me);
  if (s == nullptr) return;
  arrive(s, [] {});
  if (tls_rank == 0 && mine != nullptr) dm_free(mine);  // rank 0's slot is the buffer's base
  arrive(s, [] {});
}

static bool own_slot(MPI_Comm me, const void *mine, size_t offset, size_t bytes, dm::Buffer *b, size_t *off) {
  int r;
  if (!dm::resolve_rank(mine, b, &r, off) || r != me->rank || *off + offset + bytes > b->slot) return false;
  *off += offset;
  return true;
}

extern "C" int dm_spmd_put(MPI_Comm me, void *mine, size_t offset, const void *host, size_t bytes) {
  dm::Buffer b;
  size_t off;
  if (me == nullptr || !own_slot(me, mine, offset, bytes, &b, &off)) return DM_ERR_ARG;
  return dm_put(b.base + off, me->rank, host, bytes);
}

extern "C" int dm_spmd_get(MPI_Comm me, void *host, const void *mine, size_t offset, size_t bytes) {
  dm::Buffer b;
  size_t off;
  if (me == nullptr || !own_slot(me, mine, offset, bytes, &b, &off)) return DM_ERR_ARG;
  return dm_get(host, b.base + off, me->rank, bytes);
}
