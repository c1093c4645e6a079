// Batched local FFTs (dm_batch.h): the host planner for K3.
//
// Every transform is a C-order array (last axis contiguous).  The pass over
// axis d of a transform of shape n has lines of N = n[d] points at point
// stride inner = prod(n[e], e > d), indexed by (outer, inner) coordinates:
//   * inner == 1 (the contiguous axis): a tile is T consecutive lines, each
//     a contiguous run of N points (row walk);
//   * otherwise a tile is T consecutive inner coordinates of one outer
//     coordinate, so a row of the tile is T contiguous elements (column
//     walk).
// One launch per axis covers the tiles of all transforms (T and the radix
// schedule are per line length).  A schedule's twiddles are stage ordered
// like the distributed FFT's: stage with Ns earlier points and radix R holds
// w_N^(jj r N / (Ns R)) at (Ns - 1) + (r - 1) Ns + jj.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <vector>

#include "dm_batch.h"
#include "fft_program.h"
#include "internal.h"
#include "kernels.h"

namespace {

unsigned long long magic(int d) { return (0x100000000ULL / (unsigned long long)d) + 1ULL; }

struct PassTables {
  std::vector<dm::BatchTile> tiles;
  void *dev = nullptr;
  long long grid = 0;
};

}  // namespace

struct dm_batch {
  int count = 0, ndim = 0, dbl = 1, inverse = 0;
  std::vector<int> shapes;
  int tile_choice = DM_BATCH_TILE_FULL, radix_order = DM_BATCH_RADIX_LARGE_FIRST;
  int grid_mode = DM_BATCH_GRID_TILES;
  // bound state
  std::vector<void *> data;
  void *d_ptrs = nullptr;
  void *d_sched = nullptr;
  void *d_tw = nullptr;
  size_t lds = 0;
  int primes = 0;  // some schedule has a radix 5, 7, 11 or 13
  int long_lines = 0;  // some schedule's twiddles stay in device memory
  int any = 0;         // some schedule has a direct-sum stage (a prime above 13)
  std::vector<PassTables> passes;
  // plane tiles (DM_BATCH_TILE_PLANES), run first: [0] within the line
  // tile's points, [1] larger
  PassTables planes[2];
  size_t plane_lds[2] = {0, 0};
  int plane_primes[2] = {0, 0};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  void *graph_stream = nullptr;
};

static void drop_graph(DmBatch b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->graph_stream = nullptr;
}

static void release_bound(DmBatch b) {
  drop_graph(b);
  for (auto &p : b->passes) dm::device_free(p.dev);
  b->passes.clear();
  for (auto &pl : b->planes) {
    dm::device_free(pl.dev);
    pl = PassTables();
  }
  dm::device_free(b->d_ptrs);
  dm::device_free(b->d_sched);
  dm::device_free(b->d_tw);
  b->d_ptrs = b->d_sched = b->d_tw = nullptr;
}

extern "C" DmBatch dm_batch_plan(int count, int ndim, const int *shapes, int precision, int direction) {
  if (count < 1 || ndim < 1 || ndim > 3 || shapes == nullptr) {
    raise_error(nullptr, "dm_batch_plan: need count >= 1, 1 <= ndim <= 3 and shapes");
    return nullptr;
  }
  if (direction != -1 && direction != 1) {
    raise_error(nullptr, "dm_batch_plan: direction must be FFTW_FORWARD (-1) or FFTW_BACKWARD (+1)");
    return nullptr;
  }
  std::vector<int> rad;
  const int cap = dm::batch_tile_capacity(precision ? 1 : 0);
  for (long long i = 0; i < (long long)count * ndim; i++) {
    int n = shapes[i];
    if (n < 1 || !dm::choose_radices(n, &rad, true) || rad.size() > 16 || n > cap) {
      raise_error(nullptr, "dm_batch_plan: axis length %d of transform %lld is not supported", n, i / ndim);
      return nullptr;
    }
  }
  DmBatch b = new dm_batch;
  b->count = count;
  b->ndim = ndim;
  b->dbl = precision ? 1 : 0;
  b->inverse = direction == 1 ? 1 : 0;
  b->shapes.assign(shapes, shapes + (size_t)count * ndim);
  return b;
}

namespace dm {

// Lines per tile: as many as the tile holds; a column walk never more than
// its contiguous run (so no tile is mostly padding).
int batch_tile_height(int N, long long inner, int dbl, int tile_choice) {
  const int cap = batch_tile_capacity(dbl);
  const int limit = tile_choice == DM_BATCH_TILE_HALF ? cap / 2 : cap;
  int T = std::max(1, limit / N);
  if (inner > 1) {
    // whole 128-byte rows: tiles of a column walk start on line boundaries
    // (1.19 -> 1.16 ms per 256-transform batch)
    const int row = 128 / (dbl ? 16 : 8);
    if (T > row) T = T / row * row;
    T = (int)std::min<long long>(T, inner);
  }
  return T;
}

void make_batch_sched(int N, int T, int radix_order, BatchSched *out, std::vector<double> *tw) {
  BatchSched S{};
  std::vector<int> rad;
  choose_radices(N, &rad, true);
  if (radix_order == DM_BATCH_RADIX_SMALL_FIRST) std::reverse(rad.begin(), rad.end());
  S.N = N;
  S.nstage = (int)rad.size();
  S.T = T;
  S.m_n = magic(N);
  S.m_t = magic(T);
  S.tw_off = 0;
  int ns = 1;
  std::vector<double> root = roots(N);
  tw->assign(2 * (size_t)N, 0.0);
  for (int s = 0; s < S.nstage; s++) {
    int R = rad[s];
    S.radix[s] = R;
    S.st[s].nr = N / R;
    S.st[s].ns = ns;
    S.st[s].m_nr = magic(N / R);
    S.st[s].m_ns = magic(ns);
    int stride = N / (ns * R);
    for (int r = 1; r < R; r++)
      for (int jj = 0; jj < ns; jj++) {
        size_t at = (size_t)(ns - 1) + (size_t)(r - 1) * ns + jj;
        size_t k = (size_t)jj * r * stride;
        (*tw)[2 * at] = root[2 * k];
        (*tw)[2 * at + 1] = root[2 * k + 1];
      }
    ns *= R;
  }
  // after the N stage entries: w_p^m = w^(m N / p), m < p, for each stage
  // of a prime radix without a butterfly kernel (direct sums), in stage order
  for (int s = 0; s < S.nstage; s++) {
    const int R = S.radix[s];
    if (has_radix_kernel(R)) continue;
    for (int m = 0; m < R; m++) {
      const size_t k = (size_t)m * (N / R);
      tw->push_back(root[2 * k]);
      tw->push_back(root[2 * k + 1]);
    }
  }
  *out = S;
}

bool batch_sched_any(const BatchSched &S) {
  for (int s = 0; s < S.nstage; s++)
    if (!has_radix_kernel(S.radix[s])) return true;
  return false;
}

// Host image of a batch: tiles per axis pass (last axis first), schedules
// and their stage-ordered twiddles (complex double), LDS need, and whether a
// radix 5, 7, 11 or 13 occurs.  DM_BATCH_TILE_PLANES: a transform whose
// last two axes make a plane the plane kernel holds gets plane tiles (one
// per plane) instead of line tiles for those two axes.
bool build_batch_host(int count, int ndim, const int *shapes, int dbl, int tile_choice, int radix_order,
                      BatchHost *out) {
  BatchHost &h = *out;
  h = BatchHost();
  const int cap = batch_tile_capacity(dbl);
  std::map<std::pair<int, int>, int> sched_of;  // (N, T) -> schedule
  auto has_primes = [](const BatchSched &S) {
    for (int s = 0; s < S.nstage; s++)
      if (S.radix[s] == 5 || S.radix[s] == 7 || S.radix[s] == 11 || S.radix[s] == 13) return true;
    return false;
  };
  auto schedule = [&](int N, int T) -> int {
    auto key = std::make_pair(N, T);
    auto it = sched_of.find(key);
    if (it != sched_of.end()) return it->second;
    BatchSched S{};
    std::vector<double> block;
    make_batch_sched(N, T, radix_order, &S, &block);
    S.tw_off = (int)(h.tw.size() / 2);
    h.tw.insert(h.tw.end(), block.begin(), block.end());
    int id = (int)h.sched.size();
    h.sched.push_back(S);
    sched_of[key] = id;
    return id;
  };
  // a schedule used by line tiles
  auto line_schedule = [&](int N, int T) -> int {
    int id = schedule(N, T);
    if (has_primes(h.sched[id])) h.primes = 1;
    if (batch_sched_any(h.sched[id])) h.any = 1;
    if (!batch_tw_in_lds(N)) h.long_lines = 1;
    h.lds = std::max(h.lds, batch_lds_bytes(dbl, T, N));
    return id;
  };
  h.tiles.assign(ndim, std::vector<BatchTile>());
  for (int i = 0; i < count; i++) {
    const int *n = &shapes[(size_t)i * ndim];
    int dlast = ndim;  // axes [0, dlast) get line tiles
    if (tile_choice == DM_BATCH_TILE_PLANES && ndim >= 2) {
      const int n1 = n[ndim - 2], n2 = n[ndim - 1];
      size_t plds = batch_plane_lds_bytes(dbl, n1, n2);
      std::vector<int> r1, r2;  // planes take butterfly-kernel radices only
      choose_radices(n1, &r1, true);
      choose_radices(n2, &r2, true);
      for (int r : r1) plds = has_radix_kernel(r) ? plds : 0;
      for (int r : r2) plds = has_radix_kernel(r) ? plds : 0;
      if (plds > 0 && n1 <= cap && n2 <= cap) {
        long long outer = 1;
        for (int e = 0; e < ndim - 2; e++) outer *= n[e];
        const int rows = schedule(n2, n1), cols = schedule(n1, n2);
        const int k = (long long)n1 * n2 <= cap ? 0 : 1;
        if (has_primes(h.sched[rows]) || has_primes(h.sched[cols])) h.plane_primes[k] = 1;
        h.plane_lds[k] = std::max(h.plane_lds[k], plds);
        for (long long o = 0; o < outer; o++) {
          BatchTile t{};
          t.base = o * n1 * n2;
          t.ps = 1;
          t.xform = i;
          t.sched = rows;
          t.sched2 = cols;
          t.nl = n1;
          t.row = 2;
          h.planes[k].push_back(t);
        }
        dlast = ndim - 2;
      }
    }
    for (int d = 0; d < dlast; d++) {
      long long inner = 1, outer = 1;
      for (int e = d + 1; e < ndim; e++) inner *= n[e];
      for (int e = 0; e < d; e++) outer *= n[e];
      const int N = n[d];
      if (N < 1 || N > cap) return false;
      const bool row = inner == 1;
      // lines per tile: as many as the tile holds; a column walk never more
      // than the contiguous run (so no tile is mostly padding)
      const int T = batch_tile_height(N, inner, dbl, tile_choice);
      int sid = line_schedule(N, T);
      auto &tl = h.tiles[ndim - 1 - d];  // passes run last axis first
      if (row) {
        for (long long o = 0; o < outer; o += T) {
          BatchTile t{};
          t.base = o * N;
          t.ps = 1;
          t.xform = i;
          t.sched = sid;
          t.nl = (int)std::min<long long>(T, outer - o);
          t.row = 1;
          tl.push_back(t);
        }
      } else {
        for (long long o = 0; o < outer; o++)
          for (long long c = 0; c < inner; c += T) {
            BatchTile t{};
            t.base = o * N * inner + c;
            t.ps = inner;
            t.xform = i;
            t.sched = sid;
            t.nl = (int)std::min<long long>(T, inner - c);
            t.row = 0;
            tl.push_back(t);
          }
      }
    }
  }
  // Every other launch walks its tiles in reverse transform order, so that it
  // starts on the transforms the launch before it wrote last, which are
  // still in the memory-side cache (tiles are independent: the order changes
  // no result).  Launch order: planes[0], planes[1], tiles[0], tiles[1], ...
  std::vector<std::vector<BatchTile> *> seq;
  for (auto &pl : h.planes)
    if (!pl.empty()) seq.push_back(&pl);
  for (auto &tl : h.tiles)
    if (!tl.empty()) seq.push_back(&tl);
  for (size_t k = 1; k < seq.size(); k += 2) std::reverse(seq[k]->begin(), seq[k]->end());
  return true;
}

}  // namespace dm

static bool bind_tables(DmBatch b) {
  release_bound(b);
  dm::BatchHost h;
  if (!dm::build_batch_host(b->count, b->ndim, b->shapes.data(), b->dbl, b->tile_choice, b->radix_order, &h)) {
    raise_error(nullptr, "dm_batch_bind: unsupported axis length");
    return false;
  }
  b->lds = h.lds;
  b->primes = h.primes;
  b->long_lines = h.long_lines;
  b->any = h.any;
  b->passes.assign(b->ndim, PassTables());
  for (int d = 0; d < b->ndim; d++) b->passes[d].tiles = std::move(h.tiles[d]);
  for (int k = 0; k < 2; k++) {
    b->planes[k].tiles = std::move(h.planes[k]);
    b->plane_lds[k] = h.plane_lds[k];
    b->plane_primes[k] = h.plane_primes[k];
  }
  std::vector<dm::BatchSched> &sched = h.sched;
  std::vector<double> &tw = h.tw;
  // upload
  auto up = [&](const void *src, size_t bytes) -> void * {
    void *d = dm::device_alloc(std::max<size_t>(bytes, 16));
    if (d == nullptr) return nullptr;
    if (hipMemcpy(d, src, bytes, hipMemcpyHostToDevice) != hipSuccess) {
      dm::device_free(d);
      return nullptr;
    }
    return d;
  };
  b->d_ptrs = up(b->data.data(), b->data.size() * sizeof(void *));
  b->d_sched = up(sched.data(), sched.size() * sizeof(dm::BatchSched));
  std::vector<char> twimg(tw.size() / 2 * 16);
  dm::twiddle_image(tw.data(), tw.size() / 2, b->dbl, twimg.data());
  b->d_tw = up(twimg.data(), twimg.size());
  if (!b->d_ptrs || !b->d_sched || !b->d_tw) {
    raise_error(nullptr, "dm_batch_bind: device allocation or copy failed");
    return false;
  }
  long long resident = dm::batch_resident_groups(b->dbl, b->primes, b->lds);
  for (auto &p : b->passes) {
    p.dev = up(p.tiles.data(), p.tiles.size() * sizeof(dm::BatchTile));
    if (!p.dev) {
      raise_error(nullptr, "dm_batch_bind: device allocation or copy failed");
      return false;
    }
    p.grid = b->grid_mode == DM_BATCH_GRID_PERSISTENT ? std::min<long long>((long long)p.tiles.size(), resident)
                                                      : (long long)p.tiles.size();
  }
  for (int k = 0; k < 2; k++) {
    PassTables &pl = b->planes[k];
    if (pl.tiles.empty()) continue;
    pl.dev = up(pl.tiles.data(), pl.tiles.size() * sizeof(dm::BatchTile));
    if (!pl.dev) {
      raise_error(nullptr, "dm_batch_bind: device allocation or copy failed");
      return false;
    }
    const long long res = k == 0 ? dm::batch_resident_groups(b->dbl, b->plane_primes[0], b->plane_lds[0])
                                 : dm::batch_plane_resident_groups(b->plane_lds[1]);
    pl.grid = b->grid_mode == DM_BATCH_GRID_PERSISTENT ? std::min<long long>((long long)pl.tiles.size(), res)
                                                       : (long long)pl.tiles.size();
  }
  return true;
}

extern "C" int dm_batch_bind(DmBatch b, void *const *data) {
  if (b == nullptr || data == nullptr) {
    raise_error(nullptr, "dm_batch_bind: NULL batch or data");
    return 1;
  }
  if (!dm::device_available()) {
    raise_error(nullptr, "dm_batch_bind: no HIP device is available");
    return 1;
  }
  for (int i = 0; i < b->count; i++) {
    if (data[i] == nullptr) {
      raise_error(nullptr, "dm_batch_bind: NULL data pointer for transform %d", i);
      return 1;
    }
    hipPointerAttribute_t at;
    if (hipPointerGetAttributes(&at, data[i]) != hipSuccess || at.type != hipMemoryTypeDevice) {
      (void)hipGetLastError();
      raise_error(nullptr, "dm_batch_bind: data of transform %d is not device memory", i);
      return 1;
    }
  }
  b->data.assign(data, data + b->count);
  return bind_tables(b) ? 0 : 1;
}

extern "C" int dm_batch_set_schedule(DmBatch b, int tile, int radix_order) {
  if (b == nullptr) return 1;
  b->tile_choice = (tile == DM_BATCH_TILE_HALF || tile == DM_BATCH_TILE_PLANES) ? tile : DM_BATCH_TILE_FULL;
  b->radix_order = radix_order == DM_BATCH_RADIX_SMALL_FIRST ? DM_BATCH_RADIX_SMALL_FIRST
                                                              : DM_BATCH_RADIX_LARGE_FIRST;
  if (b->data.empty()) return 0;
  return bind_tables(b) ? 0 : 1;
}

extern "C" int dm_batch_set_grid(DmBatch b, int mode) {
  if (b == nullptr) return 1;
  b->grid_mode = mode == DM_BATCH_GRID_PERSISTENT ? DM_BATCH_GRID_PERSISTENT : DM_BATCH_GRID_TILES;
  if (b->data.empty()) return 0;
  return bind_tables(b) ? 0 : 1;
}

extern "C" int dm_batch_run(DmBatch b, void *stream) {
  if (b == nullptr || b->passes.empty()) {
    raise_error(nullptr, "dm_batch_run: batch not bound (dm_batch_bind)");
    return 1;
  }
  for (int k = 0; k < 2; k++) {  // the last two axes of the plane transforms
    const PassTables &pl = b->planes[k];
    if (pl.tiles.empty()) continue;
    dm::BatchLaunch L{};
    L.tiles = (const dm::BatchTile *)pl.dev;
    L.ntiles = (long long)pl.tiles.size();
    L.sched = (const dm::BatchSched *)b->d_sched;
    L.data = (void *const *)b->d_ptrs;
    L.twiddle = b->d_tw;
    L.dbl = b->dbl;
    L.inverse = b->inverse;
    L.primes = b->plane_primes[k];
    L.plane = k == 0 ? 2 : 1;
    L.grid = pl.grid;
    L.lds = b->plane_lds[k];
    if (!dm::launch_ok(nullptr, dm::launch_batch_pass(L, stream), "dm_batch_run: plane pass launch")) return 1;
  }
  for (auto &p : b->passes) {
    dm::BatchLaunch L{};
    L.tiles = (const dm::BatchTile *)p.dev;
    L.ntiles = (long long)p.tiles.size();
    L.sched = (const dm::BatchSched *)b->d_sched;
    L.data = (void *const *)b->d_ptrs;
    L.twiddle = b->d_tw;
    L.dbl = b->dbl;
    L.inverse = b->inverse;
    L.primes = b->primes;
    L.long_lines = b->long_lines;
    L.any = b->any;
    L.grid = p.grid;
    L.lds = b->lds;
    if (!dm::launch_ok(nullptr, dm::launch_batch_pass(L, stream), "dm_batch_run: batched pass launch")) return 1;
  }
  return 0;
}

extern "C" int dm_batch_graph_launch(DmBatch b, void *stream) {
  if (b == nullptr || b->passes.empty()) {
    raise_error(nullptr, "dm_batch_graph_launch: batch not bound (dm_batch_bind)");
    return 1;
  }
  hipStream_t s = (hipStream_t)stream;
  if (b->exec == nullptr || b->graph_stream != stream) {
    drop_graph(b);
    hipStream_t cap = s;
    bool own = false;
    if (cap == nullptr) {  // capture needs a non-legacy stream
      if (hipStreamCreateWithFlags(&cap, hipStreamNonBlocking) != hipSuccess) return 1;
      own = true;
    }
    if (hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      if (own) (void)hipStreamDestroy(cap);
      return 1;
    }
    int rc = dm_batch_run(b, cap);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(cap, &g);
    if (own) (void)hipStreamDestroy(cap);
    if (rc != 0 || e != hipSuccess || g == nullptr) {
      raise_error(nullptr, "dm_batch_graph_launch: stream capture failed");
      return 1;
    }
    b->graph = g;
    if (hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0) != hipSuccess) {
      drop_graph(b);
      raise_error(nullptr, "dm_batch_graph_launch: graph instantiation failed");
      return 1;
    }
    b->graph_stream = stream;
  }
  return hipGraphLaunch(b->exec, s) == hipSuccess ? 0 : 1;
}

extern "C" double dm_batch_autotune(DmBatch b, void *stream) {
  if (b == nullptr || b->data.empty()) {
    raise_error(nullptr, "dm_batch_autotune: batch not bound (dm_batch_bind)");
    return -1.0;
  }
  hipStream_t s = (hipStream_t)stream;
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  double best = 1e30;
  int bt = b->tile_choice, br = b->radix_order, bg = b->grid_mode;
  for (int grid : {DM_BATCH_GRID_TILES, DM_BATCH_GRID_PERSISTENT})
  for (int tile : {DM_BATCH_TILE_FULL, DM_BATCH_TILE_HALF, DM_BATCH_TILE_PLANES})
    for (int order : {DM_BATCH_RADIX_LARGE_FIRST, DM_BATCH_RADIX_SMALL_FIRST}) {
      b->grid_mode = grid;
      if (dm_batch_set_schedule(b, tile, order) != 0) continue;
      dm_batch_run(b, s);  // warm-up
      (void)hipEventRecord(e0, s);
      const int reps = 3;
      for (int r = 0; r < reps; r++) dm_batch_run(b, s);
      (void)hipEventRecord(e1, s);
      (void)hipEventSynchronize(e1);
      float ms = 0;
      (void)hipEventElapsedTime(&ms, e0, e1);
      if (ms / reps < best) {
        best = ms / reps;
        bt = tile;
        br = order;
        bg = grid;
      }
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  b->grid_mode = bg;
  if (dm_batch_set_schedule(b, bt, br) != 0) return -1.0;
  return best;
}

extern "C" int dm_batch_grid(DmBatch b) { return b == nullptr ? -1 : b->grid_mode; }

extern "C" int dm_batch_info(DmBatch b, long long *tiles, long long *grid, int *tile_choice, int *radix_order) {
  if (b == nullptr) return 1;
  for (size_t i = 0; i < b->passes.size(); i++) {
    const long long np = i == 0 ? (long long)(b->planes[0].tiles.size() + b->planes[1].tiles.size()) : 0;
    if (tiles) tiles[i] = (long long)b->passes[i].tiles.size() + np;
    if (grid) grid[i] = b->passes[i].grid + (i == 0 ? b->planes[0].grid + b->planes[1].grid : 0);
  }
  if (tile_choice) *tile_choice = b->tile_choice;
  if (radix_order) *radix_order = b->radix_order;
  return 0;
}

extern "C" double dm_batch_bytes(DmBatch b) {
  if (b == nullptr) return 0;
  double total = 0;
  for (int i = 0; i < b->count; i++) {
    double n = 1;
    for (int d = 0; d < b->ndim; d++) n *= b->shapes[(size_t)i * b->ndim + d];
    total += 2.0 * b->ndim * n * (b->dbl ? 16 : 8);
  }
  return total;
}

extern "C" double dm_batch_bytes_moved(DmBatch b) {
  if (b == nullptr) return 0;
  double total = 0;
  for (int i = 0; i < b->count; i++) {
    const int *n = &b->shapes[(size_t)i * b->ndim];
    double pts = 1;
    for (int d = 0; d < b->ndim; d++) pts *= n[d];
    int passes = b->ndim;
    // the plane rule of build_batch_host
    if (b->tile_choice == DM_BATCH_TILE_PLANES && b->ndim >= 2) {
      const int n1 = n[b->ndim - 2], n2 = n[b->ndim - 1], cap = dm::batch_tile_capacity(b->dbl);
      std::vector<int> r1, r2;
      dm::choose_radices(n1, &r1, true);
      dm::choose_radices(n2, &r2, true);
      bool kern = true;
      for (int r : r1) kern = kern && dm::has_radix_kernel(r);
      for (int r : r2) kern = kern && dm::has_radix_kernel(r);
      if (kern && dm::batch_plane_lds_bytes(b->dbl, n1, n2) > 0 && n1 <= cap && n2 <= cap) passes -= 1;
    }
    total += 2.0 * passes * pts * (b->dbl ? 16 : 8);
  }
  return total;
}

extern "C" void dm_batch_free(DmBatch b) {
  if (b == nullptr) return;
  release_bound(b);
  delete b;
}
