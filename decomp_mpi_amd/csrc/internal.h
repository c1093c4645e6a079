// Internal declarations shared by the host side of libdecomp_mpi_amd.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <functional>
#include <vector>

#include "reshape.h"
#include "split_fft.h"
#include "dm_runtime.h"

#define DM_MAX_DIMS 8
#define DM_MAX_RANKS (1 << 24)
#define DM_NO_OWNER (1 << 27)  // > DM_MAX_RANKS; 8 * DM_NO_OWNER < 2^31

typedef long long i64;

// A virtual world of nranks ranks.  Device resources are created lazily so
// that descriptor and plan construction work on machines without a GPU.
struct dm_world {
  int nranks = 0;
  std::vector<dm_comm *> views;
  void *stream = nullptr;       // hipStream_t
  bool own_stream = false;
  bool async = false;
};

struct dm_comm {
  dm_world *world;
  int rank;
};

struct dm_datatype {
  size_t size;        // bytes of data
  ptrdiff_t extent;   // bytes spanned (equal to size for contiguous types)
  int predefined;     // never freed
  int count;          // elements of a message description (1 for base types)
  ptrdiff_t true_lb;  // true lower bound (bytes) of a message description
  ptrdiff_t true_extent;
  // File views (makeIOType / expandIOType): byte offsets of the base
  // elements, in order, and the base element size; empty for other types.
  std::vector<long long> view;
  size_t elem = 0;
};

// Per-axis global view: the DimDecomp evaluated at every coordinate plus
// per-position and per-virtual-rank values.
struct dm_axis_ext {
  dm_world *world = nullptr;
  std::vector<int> rank_of;          // [coord_size] rank_offset(g)
  std::vector<int> off_of;           // [coord_size] type_offset(g)
  std::vector<int> n_local_pos;      // [proc_size]
  std::vector<int> max_local_pos;    // [proc_size]
  std::vector<i64> stride_of_rank;   // [nranks] memory stride on each rank
};

struct dm_decomp_ext {
  std::vector<i64> strides;   // [nranks * ndim]
  std::vector<i64> extent_of; // [nranks] 1 + sum stride*max_local
  bool uniform = true;        // strides equal over ranks < proc_used
};

namespace dm {

// ---- errors / logging (world.cpp)
void fail(dm_comm_t comm, const char *fmt, ...);  // raise_error with formatting
bool error_pending();
int error_mode();
// Error mode of the calling thread for the lifetime of the object.
struct ScopedErrorMode {
  explicit ScopedErrorMode(int mode);
  ~ScopedErrorMode();
  ScopedErrorMode(const ScopedErrorMode &) = delete;
  ScopedErrorMode &operator=(const ScopedErrorMode &) = delete;

 private:
  int saved_;
};

// ---- device runtime (device.cpp)
bool device_available();
void check_device(dm_comm_t comm);
void *device_alloc(size_t bytes);
void device_free(void *p);
void h2d(void *dst, const void *src, size_t bytes);
void d2h(void *dst, const void *src, size_t bytes);
void *world_stream(dm_world *w);
void world_finish(dm_world *w);   // synchronize unless async
void stream_sync(void *stream);
void destroy_stream(void *stream);
// err: a launcher's return value (a hipError_t).  Raises "<what> failed:
// <HIP error>" and returns false when it is not 0.
bool launch_ok(dm_comm_t comm, int err, const char *what);

// ---- buffer registry (world.cpp)
struct Buffer {
  char *base;
  size_t slot;       // bytes per rank
  int nranks;
  bool owned;
};
// Resolve a pointer into slot 0 of a registered buffer.
bool resolve(const void *p, Buffer *buf, size_t *offset);
// Resolve a pointer into any slot: the buffer, the slot's rank, the offset.
bool resolve_rank(const void *p, Buffer *buf, int *rank, size_t *offset);
// Inside a dm_spmd_run session (spmd.cpp): the calling thread's rank (-1
// outside), and the collective rendezvous -- all ranks wait, rank 0 runs fn,
// all ranks wait again.
int spmd_rank();
void spmd_collective(const std::function<void()> &fn);
// The same, after checking that every rank's argument signature equals
// rank 0's; when they differ nothing runs and every rank raises an error.
bool spmd_collective_checked(const std::vector<long long> &sig, const char *what,
                             const std::function<void()> &fn);
// Argument signatures: a pointer as (buffer base, slot rank, offset) or
// (-1, -1, address) when it is no registered buffer's; a decomposition's
// global layout as a hash.
void sig_pointer(std::vector<long long> *sig, const void *p);
// All ranks of the calling thread's session wait for each other.
void spmd_barrier_internal();
long long decomp_fingerprint(Decomp d);

// ---- layouts (layout.cpp)
// A decomposition flattened into per-axis lookup tables over global
// coordinates: owner rank contribution and local offset, plus per-rank
// memory strides.
struct HostLayout {
  int ndim = 0;
  std::vector<int> L;
  int proc_used = 0;
  std::vector<std::vector<int>> rank_part;  // s_d * rank_d(g)
  std::vector<std::vector<int>> off;        // type_offset_d(g)
  std::vector<i64> strides;                 // [proc_used * ndim]
  bool uniform = true;
  std::vector<i64> m;                       // strides of rank 0 (uniform case)
  std::vector<i64> slot_extent;             // [proc_used]
  int fastest = 0;                          // axis with smallest stride
  bool offsets_ok = true;                   // no negative offset or stride
};
HostLayout layout_from_decomp(Decomp d);
// Check that owner ranks are a bijection onto [0, proc_used).
bool layout_rank_map_ok(Decomp d, std::string *why);

struct DevLayout {
  int ndim;
  int proc_used;
  int uniform;
  int pad_;
  const int *rank[DM_MAX_DIMS];
  const int *off[DM_MAX_DIMS];
  const i64 *addr[DM_MAX_DIMS];   // uniform: m_d * off_d(g)
  const i64 *strides;             // [proc_used * ndim]
};

// Device copy of a HostLayout's tables.
struct DeviceLayout {
  DevLayout dev{};
  void *mem = nullptr;
  size_t bytes = 0;
  void upload(const HostLayout &h);
  void release();
};

}  // namespace dm
