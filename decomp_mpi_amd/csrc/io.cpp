// The on-disk format: a distributed array as raw native-endian elements in
// global row-major order (last axis fastest), the file the reference's
// drivers write through MPI-IO with makeIOType's views (reshape.c:494-521,
// split_plan_test.c:160-200).  Here the array is gathered on the device by
// a reshape (the K1 gather) into one rank's row-major block, copied to the
// host and written with pwrite; reading runs the same steps backwards.
#include <hip/hip_runtime_api.h>

#include <cerrno>
#include <cstring>
#include <fcntl.h>
#include <unistd.h>
#include <vector>

#include "internal.h"
#include "local_decomp.h"

namespace dm {
bool enqueue_reshape(dm_comm_t comm, ReshapePlan plan, const void *src, void *dst, void *stream);
}

namespace {

// Whole file in row-major order on rank 0 of d's communicator.
Decomp file_decomp(Decomp d) {
  std::vector<AxisDecomp> ax(d->ndim);
  long long stride = 1;
  for (int i = d->ndim - 1; i >= 0; i--) {
    const int L = d->dims[i]->decomp->coord_size;
    ax[i] = makeAxisDecomp(d->comm, makeLocalDecomp(L), (int)stride, 1);
    stride *= L;
  }
  return makeDecomp(d->ndim, ax.data());
}

long long global_count(Decomp d) {
  long long n = 1;
  for (int i = 0; i < d->ndim; i++) n *= d->dims[i]->decomp->coord_size;
  return n;
}

// A device block of `bytes` registered as a distributed buffer whose rank 0
// slot is the whole block.  The file decomposition uses rank 0 only, so no
// other slot is ever addressed (the reshape checks every used rank's extent
// against its slot before it launches).
struct Staging {
  void *dev = nullptr;
  ~Staging() {
    if (dev) {
      dm_unregister(dev);
      dm::device_free(dev);
    }
  }
};

bool io_steps(const char *path, MPI_Datatype base, Decomp d, void *buf, bool write) {
  dm_comm_t comm = d->comm;
  dm::check_device(comm);
  if (dm::error_pending() && dm::error_mode() == 1) return false;
  const size_t es = base->size;
  const long long n = global_count(d);
  const size_t bytes = (size_t)n * es;
  Decomp fd = file_decomp(d);
  ReshapePlan plan = write ? makeReshapePlan(comm, base, d, fd) : makeReshapePlan(comm, base, fd, d);
  bool ok = plan != nullptr;
  Staging st;
  if (ok) {
    st.dev = dm::device_alloc(std::max<size_t>(bytes, 256));
    ok = st.dev != nullptr && dm_register(comm, st.dev, std::max<size_t>(bytes, 256)) == DM_SUCCESS;
    if (!ok) raise_error(comm, "file I/O: device staging of %zu bytes failed", bytes);
  }
  std::vector<char> host;
  int fdesc = -1;
  if (ok) {
    host.resize(bytes);
    fdesc = write ? open(path, O_WRONLY | O_CREAT | O_TRUNC, 0644) : open(path, O_RDONLY);
    if (fdesc < 0) {
      raise_error(comm, "file I/O: cannot open %s: %s", path, strerror(errno));
      ok = false;
    }
  }
  void *stream = ok ? dm::world_stream(comm->world) : nullptr;
  if (ok && write) {
    ok = dm::enqueue_reshape(comm, plan, buf, st.dev, stream);
    if (ok) {
      dm::stream_sync(stream);
      dm::d2h(host.data(), st.dev, bytes);
      for (size_t at = 0; ok && at < bytes;) {
        ssize_t w = pwrite(fdesc, host.data() + at, std::min<size_t>(bytes - at, (size_t)1 << 30), (off_t)at);
        if (w <= 0) {
          raise_error(comm, "file I/O: writing %s failed: %s", path, w < 0 ? strerror(errno) : "no progress");
          ok = false;
        }
        at += w > 0 ? (size_t)w : 0;
      }
    }
  } else if (ok) {
    for (size_t at = 0; ok && at < bytes;) {
      ssize_t r = pread(fdesc, host.data() + at, std::min<size_t>(bytes - at, (size_t)1 << 30), (off_t)at);
      if (r <= 0) {
        raise_error(comm, "file I/O: %s holds fewer than the %zu bytes of the array", path, bytes);
        ok = false;
      }
      at += r > 0 ? (size_t)r : 0;
    }
    if (ok) {
      dm::h2d(st.dev, host.data(), bytes);
      ok = dm::enqueue_reshape(comm, plan, st.dev, buf, stream);
      if (ok) dm::stream_sync(stream);
    }
  }
  if (fdesc >= 0 && close(fdesc) != 0 && ok) {
    raise_error(comm, "file I/O: closing %s failed: %s", path, strerror(errno));
    ok = false;
  }
  freeReshapePlan(plan);
  freeDecomp(fd);
  return ok;
}

int io_collective(const char *path, MPI_Datatype base, Decomp d, void *buf, bool write) {
  if (path == nullptr || base == nullptr || d == nullptr || buf == nullptr) {
    raise_error(d ? d->comm : nullptr, "file I/O: NULL argument");
    return DM_ERR_ARG;
  }
  if (dm::spmd_rank() >= 0) {
    // rank 0's call (its decomposition and slot-0 pointer) moves every
    // rank's share; every rank returns its result
    static int shared_rc;
    std::vector<long long> sig = {write ? 1 : 0, (long long)base->size, dm::decomp_fingerprint(d)};
    for (const char *c = path; *c; c++) sig.push_back(*c);
    dm::sig_pointer(&sig, buf);
    if (!dm::spmd_collective_checked(sig, write ? "dm_file_write" : "dm_file_read", [&] {
          shared_rc = io_steps(path, base, d, buf, write) ? DM_SUCCESS : DM_ERR_ARG;
        }))
      return DM_ERR_ARG;
    const int rc = shared_rc;
    dm::spmd_barrier_internal();  // every rank has read it before the next call may write it
    return rc;
  }
  return io_steps(path, base, d, buf, write) ? DM_SUCCESS : DM_ERR_ARG;
}

}  // namespace

extern "C" int dm_file_write(const char *path, MPI_Datatype base, Decomp d, const void *buf) {
  return io_collective(path, base, d, const_cast<void *>(buf), true);
}

extern "C" int dm_file_read(const char *path, MPI_Datatype base, Decomp d, void *buf) {
  return io_collective(path, base, d, buf, false);
}
