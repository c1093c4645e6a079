// HIP runtime plumbing: device detection, allocation, copies, streams.
#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "internal.h"
#include "kernels.h"

namespace dm {

static void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess) {
    fprintf(stderr, "decomp_mpi_amd: %s failed: %s\n", what, hipGetErrorString(e));
    fflush(stderr);
    abort();
  }
}

bool device_available() {
  static const bool avail = [] {
    int n = 0;
    return hipGetDeviceCount(&n) == hipSuccess && n > 0;
  }();
  return avail;
}

void check_device(dm_comm_t comm) {
  if (!device_available()) fail(comm, "no HIP device is available");
}

void *device_alloc(size_t bytes) {
  void *p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
  return p;
}

void device_free(void *p) {
  if (p != nullptr) hip_check(hipFree(p), "hipFree");
}

void h2d(void *dst, const void *src, size_t bytes) {
  hip_check(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "hipMemcpy H2D");
}

void d2h(void *dst, const void *src, size_t bytes) {
  hip_check(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "hipMemcpy D2H");
}

void *world_stream(dm_world *w) {
  if (w->stream == nullptr) {
    hipStream_t s;
    // Default priority on purpose: every stream of this library keeps it.
    hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
    w->stream = (void *)s;
    w->own_stream = true;
  }
  return w->stream;
}

void stream_sync(void *stream) {
  hip_check(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
}

void world_finish(dm_world *w) {
  if (!w->async) stream_sync(world_stream(w));
}

bool launch_ok(dm_comm_t comm, int err, const char *what) {
  if (err == 0) return true;
  fail(comm, "%s failed: %s", what, hipGetErrorString((hipError_t)err));
  return false;
}

void destroy_stream(void *stream) {
  hipStreamSynchronize((hipStream_t)stream);
  hipStreamDestroy((hipStream_t)stream);
}

}  // namespace dm

extern "C" int dm_device_available(void) { return dm::device_available() ? 1 : 0; }

extern "C" int dm_fill(void *dst, int value, size_t bytes_per_rank) {
  dm::Buffer b;
  size_t off;
  if (!dm::resolve(dst, &b, &off) || off + bytes_per_rank > b.slot) return DM_ERR_ARG;
  for (int r = 0; r < b.nranks; r++)
    dm::hip_check(hipMemset(b.base + (size_t)r * b.slot + off, value, bytes_per_rank), "hipMemset");
  dm::hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
  return DM_SUCCESS;
}


// Plain device-to-device copy of `bytes` (a multiple of 16) on the world's
// stream, timed with HIP events; returns milliseconds (negative on error).
extern "C" double dm_copy_bench(MPI_Comm comm, void *dst, const void *src, size_t bytes, int reps) {
  if (comm == nullptr || bytes % 16 != 0 || reps < 1) return -1.0;
  dm::check_device(comm);
  hipStream_t s = (hipStream_t)dm::world_stream(comm->world);
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  bool ok = dm::launch_ok(comm, dm::launch_copy16(src, dst, (long long)(bytes / 16), s), "copy kernel launch");
  (void)hipEventRecord(e0, s);
  for (int i = 0; i < reps && ok; i++)
    ok = dm::launch_ok(comm, dm::launch_copy16(src, dst, (long long)(bytes / 16), s), "copy kernel launch");
  (void)hipEventRecord(e1, s);
  (void)hipEventSynchronize(e1);
  float ms = 0;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return ok ? ms / reps : -1.0;
}
