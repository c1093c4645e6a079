// Virtual-rank world, element types, errors, logging, timer and the
// distributed-buffer registry.
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <atomic>
#include <cstring>
#include <map>
#include <mutex>
#include <time.h>

#include "internal.h"
#include "logger.h"
#include "timer.h"

// ---------------------------------------------------------------- worlds

extern "C" dm_comm_t dm_world_create(int nranks) {
  if (nranks < 1 || nranks > DM_MAX_RANKS) {
    fprintf(stderr, "dm_world_create: nranks must be in [1, %d] (got %d)\n", DM_MAX_RANKS, nranks);
    return nullptr;
  }
  dm_world *w = new dm_world;
  w->nranks = nranks;
  w->views.resize(nranks);
  for (int r = 0; r < nranks; r++) w->views[r] = new dm_comm{w, r};
  return w->views[0];
}

extern "C" dm_comm_t dm_comm_view(dm_comm_t comm, int rank) {
  if (comm == nullptr || rank < 0 || rank >= comm->world->nranks) return nullptr;
  return comm->world->views[rank];
}

extern "C" void dm_world_free(dm_comm_t comm) {
  if (comm == nullptr) return;
  dm_world *w = comm->world;
  if (w->own_stream && w->stream != nullptr) dm::destroy_stream(w->stream);
  for (dm_comm *c : w->views) delete c;
  delete w;
}

extern "C" int dm_comm_rank(dm_comm_t comm, int *rank) {
  if (comm == nullptr) return DM_ERR_ARG;
  *rank = comm->rank;
  return DM_SUCCESS;
}

extern "C" int dm_comm_size(dm_comm_t comm, int *size) {
  if (comm == nullptr) return DM_ERR_ARG;
  *size = comm->world->nranks;
  return DM_SUCCESS;
}

extern "C" void dm_abort(dm_comm_t comm, int code) {
  fprintf(stderr, "%d: abort with code %d\n", comm ? comm->rank : -1, code);
  fflush(stderr);
  abort();
}

// ---------------------------------------------------------------- types

#define DM_BASE_TYPE(name, bytes) \
  struct dm_datatype name = {bytes, bytes, 1, 1, 0, bytes, {}, bytes}
DM_BASE_TYPE(dm_type_byte, 1);
DM_BASE_TYPE(dm_type_char, 1);
DM_BASE_TYPE(dm_type_int, 4);
DM_BASE_TYPE(dm_type_long, 8);
DM_BASE_TYPE(dm_type_long_long, 8);
DM_BASE_TYPE(dm_type_float, 4);
DM_BASE_TYPE(dm_type_double, 8);
DM_BASE_TYPE(dm_type_c_float_complex, 8);
DM_BASE_TYPE(dm_type_c_double_complex, 16);

extern "C" int dm_type_contiguous(int count, dm_datatype_t old, dm_datatype_t *result) {
  if (old == nullptr || count < 0 || result == nullptr) return DM_ERR_ARG;
  dm_datatype *t = new dm_datatype;
  t->size = old->size * (size_t)count;
  t->extent = old->extent * count;
  t->predefined = 0;
  t->count = 1;
  t->true_lb = 0;
  t->true_extent = t->extent;
  t->elem = t->size;
  *result = t;
  return DM_SUCCESS;
}

extern "C" int dm_type_dup(dm_datatype_t old, dm_datatype_t *result) {
  if (old == nullptr || result == nullptr) return DM_ERR_ARG;
  dm_datatype *t = new dm_datatype(*old);
  t->predefined = 0;
  *result = t;
  return DM_SUCCESS;
}

extern "C" int dm_type_commit(dm_datatype_t *type) {
  return (type == nullptr || *type == nullptr) ? DM_ERR_ARG : DM_SUCCESS;
}

extern "C" int dm_type_free(dm_datatype_t *type) {
  if (type == nullptr || *type == nullptr) return DM_ERR_ARG;
  if (!(*type)->predefined) delete *type;
  *type = nullptr;
  return DM_SUCCESS;
}

extern "C" int dm_type_get_extent(dm_datatype_t type, dm_aint_t *lb, dm_aint_t *extent) {
  if (type == nullptr) return DM_ERR_ARG;
  if (lb) *lb = 0;
  if (extent) *extent = type->extent;
  return DM_SUCCESS;
}

extern "C" int dm_type_size(dm_datatype_t type, int *size) {
  if (type == nullptr) return DM_ERR_ARG;
  *size = (int)type->size;
  return DM_SUCCESS;
}

// ---------------------------------------------------------------- errors

// The error mode is a process-wide default that a thread may override for a
// scope (the autotuner does, around candidates that may not plan); the last
// error and the pending flag belong to the thread that raised them, so the
// rank threads of a dm_spmd_run session never write one shared string.
static std::atomic<int> g_error_mode{0};
static thread_local int t_error_mode = -1;  // -1: follow the process default
static thread_local std::string t_last_error;
static thread_local bool t_error_pending = false;

extern "C" void dm_set_error_mode(int mode) {
  g_error_mode.store(mode);
  t_error_mode = -1;
}
extern "C" const char *dm_last_error(void) { return t_last_error.c_str(); }
extern "C" void dm_clear_error(void) {
  t_last_error.clear();
  t_error_pending = false;
}

namespace dm {
bool error_pending() { return t_error_pending; }
int error_mode() { return t_error_mode >= 0 ? t_error_mode : g_error_mode.load(); }

ScopedErrorMode::ScopedErrorMode(int mode) : saved_(t_error_mode) { t_error_mode = mode; }
ScopedErrorMode::~ScopedErrorMode() { t_error_mode = saved_; }

static void vraise(dm_comm_t comm, const char *fmt, va_list ap) {
  char msg[1024];
  vsnprintf(msg, sizeof msg, fmt, ap);
  char line[1200];
  // Same text as the reference's raise_error (reshape.c:18).
  snprintf(line, sizeof line, "%d: Error in reshape:%s\n", comm ? comm->rank : 0, msg);
  if (error_mode() == 1) {
    t_last_error = line;
    t_error_pending = true;
    return;
  }
  fputs(line, stderr);
  fflush(stderr);
  abort();
}

void fail(dm_comm_t comm, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vraise(comm, fmt, ap);
  va_end(ap);
}
}  // namespace dm

extern "C" void raise_error(MPI_Comm comm, const char *msg, ...) {
  va_list ap;
  va_start(ap, msg);
  dm::vraise(comm, msg, ap);
  va_end(ap);
}

// ---------------------------------------------------------------- logging

static bool debug_enabled() {
  static const bool v = getenv("DM_DEBUG") != nullptr;
  return v;
}

extern "C" void logger(MPI_Comm comm, const char *msg, ...) {
  if (!debug_enabled()) return;
  va_list ap;
  va_start(ap, msg);
  fprintf(stdout, "%d: ", comm ? comm->rank : 0);
  vfprintf(stdout, msg, ap);
  va_end(ap);
}

extern "C" void fragment_logger(const char *msg, ...) {
  if (!debug_enabled()) return;
  va_list ap;
  va_start(ap, msg);
  vfprintf(stdout, msg, ap);
  va_end(ap);
}

extern "C" double timer(void) {
  // the first call (from any thread) sets the base and returns 0
  static std::once_flag once;
  static struct timespec base;
  bool first = false;
  std::call_once(once, [&] {
    clock_gettime(CLOCK_MONOTONIC, &base);
    first = true;
  });
  if (first) return 0.0;
  struct timespec now;
  clock_gettime(CLOCK_MONOTONIC, &now);
  return (double)(now.tv_sec - base.tv_sec) + 1e-9 * (double)(now.tv_nsec - base.tv_nsec);
}

// ---------------------------------------------------------------- buffers

static std::mutex g_buf_mutex;
static std::map<uintptr_t, dm::Buffer> g_buffers;  // keyed by base address

namespace dm {
bool resolve(const void *p, Buffer *buf, size_t *offset) {
  std::lock_guard<std::mutex> lk(g_buf_mutex);
  uintptr_t a = (uintptr_t)p;
  auto it = g_buffers.upper_bound(a);
  if (it == g_buffers.begin()) return false;
  --it;
  const Buffer &b = it->second;
  if (a < it->first || a >= it->first + b.slot) return false;
  *buf = b;
  *offset = a - it->first;
  return true;
}

void sig_pointer(std::vector<long long> *sig, const void *p) {
  Buffer b;
  int r;
  size_t off;
  if (resolve_rank(p, &b, &r, &off)) {
    sig->push_back((long long)(uintptr_t)b.base);
    sig->push_back((long long)off);
    sig->push_back(r == spmd_rank() ? 1 : 0);  // each rank passes its own slot
  } else {
    sig->push_back(-1);
    sig->push_back(-1);
    sig->push_back((long long)(uintptr_t)p);
  }
}

bool resolve_rank(const void *p, Buffer *buf, int *rank, size_t *offset) {
  std::lock_guard<std::mutex> lk(g_buf_mutex);
  uintptr_t a = (uintptr_t)p;
  auto it = g_buffers.upper_bound(a);
  if (it == g_buffers.begin()) return false;
  --it;
  const Buffer &b = it->second;
  if (a < it->first || a >= it->first + b.slot * (size_t)b.nranks) return false;
  *buf = b;
  *rank = (int)((a - it->first) / b.slot);
  *offset = (a - it->first) % b.slot;
  return true;
}
}  // namespace dm

extern "C" void *dm_malloc(MPI_Comm comm, size_t bytes_per_rank) {
  if (comm == nullptr || bytes_per_rank == 0) return nullptr;
  dm::check_device(comm);
  size_t slot = (bytes_per_rank + 255) & ~(size_t)255;
  void *p = dm::device_alloc(slot * (size_t)comm->world->nranks);
  if (p == nullptr) return nullptr;
  std::lock_guard<std::mutex> lk(g_buf_mutex);
  g_buffers[(uintptr_t)p] = dm::Buffer{(char *)p, slot, comm->world->nranks, true};
  return p;
}

extern "C" int dm_free(void *base) {
  dm::Buffer b;
  {
    std::lock_guard<std::mutex> lk(g_buf_mutex);
    auto it = g_buffers.find((uintptr_t)base);
    if (it == g_buffers.end()) return DM_ERR_ARG;
    b = it->second;
    g_buffers.erase(it);
  }
  if (b.owned) dm::device_free(b.base);
  return DM_SUCCESS;
}

extern "C" int dm_register(MPI_Comm comm, void *base, size_t slot_bytes) {
  if (comm == nullptr || base == nullptr || slot_bytes == 0) return DM_ERR_ARG;
  std::lock_guard<std::mutex> lk(g_buf_mutex);
  g_buffers[(uintptr_t)base] = dm::Buffer{(char *)base, slot_bytes, comm->world->nranks, false};
  return DM_SUCCESS;
}

extern "C" int dm_unregister(void *base) {
  std::lock_guard<std::mutex> lk(g_buf_mutex);
  auto it = g_buffers.find((uintptr_t)base);
  if (it == g_buffers.end() || it->second.owned) return DM_ERR_ARG;
  g_buffers.erase(it);
  return DM_SUCCESS;
}

extern "C" size_t dm_slot_bytes(const void *ptr) {
  dm::Buffer b;
  size_t off;
  return dm::resolve(ptr, &b, &off) ? b.slot : 0;
}

extern "C" int dm_put(void *dst, int rank, const void *host, size_t bytes) {
  dm::Buffer b;
  size_t off;
  if (!dm::resolve(dst, &b, &off) || rank < 0 || rank >= b.nranks || off + bytes > b.slot)
    return DM_ERR_ARG;
  dm::h2d(b.base + (size_t)rank * b.slot + off, host, bytes);
  return DM_SUCCESS;
}

extern "C" int dm_get(void *host, const void *src, int rank, size_t bytes) {
  dm::Buffer b;
  size_t off;
  if (!dm::resolve(src, &b, &off) || rank < 0 || rank >= b.nranks || off + bytes > b.slot)
    return DM_ERR_ARG;
  dm::d2h(host, b.base + (size_t)rank * b.slot + off, bytes);
  return DM_SUCCESS;
}

extern "C" int dm_set_async(MPI_Comm comm, int on) {
  if (comm == nullptr) return DM_ERR_ARG;
  comm->world->async = on != 0;
  return DM_SUCCESS;
}

extern "C" void *dm_get_stream(MPI_Comm comm) {
  if (comm == nullptr) return nullptr;
  return dm::world_stream(comm->world);
}

extern "C" int dm_set_stream(MPI_Comm comm, void *stream) {
  if (comm == nullptr) return DM_ERR_ARG;
  dm_world *w = comm->world;
  if (w->own_stream && w->stream != nullptr) dm::destroy_stream(w->stream);
  w->stream = stream;
  w->own_stream = false;
  return DM_SUCCESS;
}

extern "C" int dm_synchronize(MPI_Comm comm) {
  if (comm == nullptr) return DM_ERR_ARG;
  if (comm->world->stream != nullptr) dm::stream_sync(comm->world->stream);
  return DM_SUCCESS;
}
