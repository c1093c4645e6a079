// dm_batch_run_noalloc: the batched local FFTs of dm_batch.h in one call
// that allocates no device memory.
//
// The bound form (dm_batch_bind + dm_batch_run) keeps its tables -- the
// transforms' data pointers, the radix schedules, their twiddles and the
// tile lists of the three axis passes -- in device memory the library
// allocates.  Here the same tables are built on the host for every call and
// copied into `scratch`, a fixed device array of this file, so the call
// makes no hipMalloc / hipFree and can sit in a loop that must not
// allocate.  The kernels are the bound form's, so the results are the same
// bit for bit.
//
// Each device holds its own copy of `scratch`, at its own address.  Host
// code never uses the array's name as a device address (in host code it
// names a host-side stand-in): the address comes from hipGetSymbolAddress,
// taken once per device with that device current, and kept in
// g_scratch_dev[device].  Every call on one device shares that device's
// array: calls that overlap in time on one device (different streams, or
// threads) race on it; calls on one stream are ordered by the stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "dm_batch.h"
#include "fft_program.h"
#include "internal.h"
#include "kernels.h"

namespace {

constexpr size_t kScratchBytes = DM_BATCH_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mutex;
void *g_scratch_dev[kMaxDevices];  // device address of `scratch` per device (null until set up)

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// The calling thread's current device's `scratch` address, set up on the
// first call on that device.
void *scratch_address(int *device) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    raise_error(nullptr, "dm_batch_run_noalloc: no current HIP device (or device index above %d)",
                kMaxDevices - 1);
    return nullptr;
  }
  *device = dev;
  std::lock_guard<std::mutex> lk(g_scratch_mutex);
  if (g_scratch_dev[dev] == nullptr) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
      raise_error(nullptr, "dm_batch_run_noalloc: hipGetSymbolAddress of the scratch array failed on device %d: %s",
                  dev, hipGetErrorString(e));
      return nullptr;
    }
    g_scratch_dev[dev] = p;
  }
  return g_scratch_dev[dev];
}

}  // namespace

extern "C" int dm_batch_run_noalloc(int count, int ndim, const int *shapes, int precision, int direction,
                                    void *const *data, void *stream) {
  if (count < 1 || ndim < 1 || ndim > 3 || shapes == nullptr || data == nullptr) {
    raise_error(nullptr, "dm_batch_run_noalloc: need count >= 1, 1 <= ndim <= 3, shapes and data");
    return 1;
  }
  if (direction != -1 && direction != 1) {
    raise_error(nullptr, "dm_batch_run_noalloc: direction must be FFTW_FORWARD (-1) or FFTW_BACKWARD (+1)");
    return 1;
  }
  const int dbl = precision ? 1 : 0;
  std::vector<int> rad;
  const int cap = dm::batch_tile_capacity(dbl);
  for (long long i = 0; i < (long long)count * ndim; i++) {
    const int n = shapes[i];
    if (n < 1 || n > cap || !dm::choose_radices(n, &rad, true) || rad.size() > 16) {
      raise_error(nullptr, "dm_batch_run_noalloc: axis length %d of transform %lld is not supported", n, i / ndim);
      return 1;
    }
  }
  for (int i = 0; i < count; i++)
    if (data[i] == nullptr) {
      raise_error(nullptr, "dm_batch_run_noalloc: NULL data pointer for transform %d", i);
      return 1;
    }
  // the bound form's default schedule (full tiles, large radices first, one
  // workgroup per tile)
  dm::BatchHost h;
  if (!dm::build_batch_host(count, ndim, shapes, dbl, DM_BATCH_TILE_FULL, DM_BATCH_RADIX_LARGE_FIRST, &h)) {
    raise_error(nullptr, "dm_batch_run_noalloc: unsupported axis length");
    return 1;
  }
  // table image: pointers | schedules | roots (pass precision) | tiles per pass
  const size_t o_ptr = 0;
  const size_t o_sched = align256((size_t)count * sizeof(void *));
  const size_t o_tw = o_sched + align256(h.sched.size() * sizeof(dm::BatchSched));
  const size_t tw_bytes = (h.tw.size() / 2) * 16;
  std::vector<size_t> o_tiles(ndim);
  size_t at = o_tw + align256(tw_bytes);
  for (int p = 0; p < ndim; p++) {
    o_tiles[p] = at;
    at += align256(h.tiles[p].size() * sizeof(dm::BatchTile));
  }
  if (at > kScratchBytes) {
    raise_error(nullptr, "dm_batch_run_noalloc: the batch needs %zu bytes of tables, the scratch array holds %zu",
                at, kScratchBytes);
    return 1;
  }
  int dev = -1;
  char *base = (char *)scratch_address(&dev);
  if (base == nullptr) return 1;
  std::vector<char> img(at, 0);
  memcpy(img.data() + o_ptr, data, (size_t)count * sizeof(void *));
  memcpy(img.data() + o_sched, h.sched.data(), h.sched.size() * sizeof(dm::BatchSched));
  dm::twiddle_image(h.tw.data(), h.tw.size() / 2, dbl, img.data() + o_tw);
  for (int p = 0; p < ndim; p++)
    memcpy(img.data() + o_tiles[p], h.tiles[p].data(), h.tiles[p].size() * sizeof(dm::BatchTile));
  hipStream_t s = (hipStream_t)stream;
  // Stream-ordered after the previous work on this stream (which may still
  // read the tables of an earlier call); waiting for the copy lets the host
  // image go when the call returns.
  hipError_t e = hipMemcpyAsync(base, img.data(), at, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    raise_error(nullptr, "dm_batch_run_noalloc: table copy to the scratch array failed: %s", hipGetErrorString(e));
    return 1;
  }
  for (int p = 0; p < ndim; p++) {
    dm::BatchLaunch L{};
    L.tiles = (const dm::BatchTile *)(base + o_tiles[p]);
    L.ntiles = (long long)h.tiles[p].size();
    L.sched = (const dm::BatchSched *)(base + o_sched);
    L.data = (void *const *)(base + o_ptr);
    L.twiddle = base + o_tw;
    L.dbl = dbl;
    L.inverse = direction == 1 ? 1 : 0;
    L.primes = h.primes;
    L.long_lines = h.long_lines;
    L.any = h.any;
    L.grid = L.ntiles;  // one workgroup per tile, as the bound form's default
    L.lds = h.lds;
    if (!dm::launch_ok(nullptr, dm::launch_batch_pass(L, stream), "dm_batch_run_noalloc: batched pass launch"))
      return 1;
  }
  return 0;
}
