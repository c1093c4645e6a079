// K1: decomposition-driven reshape as a device gather.
//
// The reference moves each rank's share as MPI messages built from nested
// indexed datatypes (reshape.c:461-492, typelist.c:74-106).  Here every
// virtual rank's local array is a slot of one HBM buffer, so a reshape is
// one launch: each global coordinate is read from its source owner's slot
// and written to its destination owner's slot.  Placement is given by
// per-axis lookup tables (owner contribution and local offset per global
// coordinate) plus per-rank memory strides.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace dm {

namespace {

constexpr int TILE = 32;      // tile edge (elements)
constexpr int ROWS = 8;       // thread rows per tile pass
constexpr int NTHREADS = TILE * ROWS;
// Batch slices per 2D-tile workgroup: 2 (16 KiB of 8-byte elements; 4
// measured 5% slower, 1 20%), 1 for 16-byte elements (the same 16 KiB:
// 2.09 / 2.05 -> 2.00 / 1.96 ms per step for the phase-order reshapes).
constexpr int tile_slices(int bytes) { return bytes >= 16 ? 1 : 2; }

template <int BYTES>
struct Elem;
template <> struct Elem<1> { typedef uint8_t T; };
template <> struct Elem<2> { typedef uint16_t T; };
template <> struct Elem<4> { typedef uint32_t T; };
template <> struct Elem<8> { typedef uint64_t T; };
template <> struct Elem<16> { typedef unsigned int T __attribute__((ext_vector_type(4))); };

// Element accesses, with the non-temporal hint when NT (each element is
// read once and written once per launch).
template <bool NT, typename E>
__device__ __forceinline__ E read_elem(const E *at) {
  if constexpr (NT) return __builtin_nontemporal_load(at);
  else return *at;
}
template <bool NT, typename E>
__device__ __forceinline__ void write_elem(E *at, E v) {
  if constexpr (NT) __builtin_nontemporal_store(v, at);
  else *at = v;
}

// Per-block partial sums of the batch axes for one layout.
struct Partial {
  int rank;
  i64 addr;                   // uniform strides: sum of m_d * off_d
  int off[DM_MAX_DIMS];       // non-uniform: the batch axes' local offsets
};

// Batch part of a workgroup's coordinates: batch index bidx decoded over
// the batch axes (last batch axis fastest) straight into the sums, so no
// coordinate array is indexed at run time.
template <bool UNIFORM>
__device__ __forceinline__ void partial_init(const DevLayout &Lt, const ReshapeLaunch &p, long long bidx,
                                             Partial &q) {
  q.rank = 0;
  q.addr = 0;
  for (int k = p.nbatch - 1; k >= 0; k--) {
    int d = p.batch_axis[k];
    int g = (int)(bidx % p.L[d]);
    bidx /= p.L[d];
    q.rank += Lt.rank[d][g];
    if (UNIFORM) q.addr += Lt.addr[d][g];
    else q.off[k] = Lt.off[d][g];
  }
}

// Element offset (in elements, from slot 0) of the coordinate whose batch
// part is q and whose tile axes take values ga (axis a) and gb (axis b);
// returns -1 when the owner rank is outside the layout's ranks.
template <bool UNIFORM>
__device__ __forceinline__ i64 locate(const DevLayout &Lt, const ReshapeLaunch &p, const Partial &q,
                                      int a, int ga, int b, int gb, i64 slot) {
  int R = q.rank + Lt.rank[a][ga] + (b != a ? Lt.rank[b][gb] : 0);
  if (R >= Lt.proc_used) return -1;
  i64 off;
  if (UNIFORM) {
    off = q.addr + Lt.addr[a][ga] + (b != a ? Lt.addr[b][gb] : 0);
  } else {
    const i64 *m = Lt.strides + (i64)R * Lt.ndim;
    off = m[a] * Lt.off[a][ga];
    if (b != a) off += m[b] * Lt.off[b][gb];
    for (int k = 0; k < p.nbatch; k++) off += m[p.batch_axis[k]] * q.off[k];
  }
  return (i64)R * slot + off;
}

// 2D tile: read rows along b (source-contiguous), write rows along a.
// A workgroup moves NB tiles that differ only in the batch coordinate (so
// the tile-axis tables are shared and more bytes are in flight per
// workgroup).  Each side runs in three branch-free phases: every address of
// the thread's elements (table loads only), then every data load (clamped
// to element 0 of the region when the element is absent), then the LDS (or
// global) writes.  A load queue that completes in order would otherwise
// stall each data load behind the next element's table loads.
template <int BYTES, bool US, bool UD, int NB, bool NT>
__global__ __launch_bounds__(NTHREADS) void reshape_tile2d(ReshapeLaunch p) {
  typedef typename Elem<BYTES>::T E;
  constexpr int RPT = TILE / ROWS;
  __shared__ E tile[NB][TILE][TILE + 1];
  __shared__ unsigned char have[NB][TILE][TILE + 1];

  long long bid = blockIdx.x;
  int ta = (int)(bid % p.tiles_a);
  bid /= p.tiles_a;
  int tb = (int)(bid % p.tiles_b);
  bid /= p.tiles_b;
  Partial qs[NB], qd[NB];
  bool live[NB];
#pragma unroll
  for (int n = 0; n < NB; n++) {
    long long b = bid * NB + n;
    live[n] = b < p.nbatch_total;
    partial_init<US>(p.src, p, live[n] ? b : 0, qs[n]);
    partial_init<UD>(p.dst, p, live[n] ? b : 0, qd[n]);
  }

  const E *src = reinterpret_cast<const E *>(p.src_base);
  E *dst = reinterpret_cast<E *>(p.dst_base);
  const int La = p.L[p.a], Lb = p.L[p.b];
  int a0 = ta * TILE, b0 = tb * TILE;
  int tx = threadIdx.x % TILE, ty = threadIdx.x / TILE;

  // load: consecutive threads walk axis b
  i64 s[NB][RPT];
  if constexpr (US) {
    // uniform source: the tile axes' table entries are loaded once (one per
    // column, one per row) and shared by the NB slices
    const int gbc = min(b0 + tx, Lb - 1);
    const int rb = p.src.rank[p.b][gbc];
    const i64 ab = p.src.addr[p.b][gbc];
    int ra[RPT];
    i64 aa[RPT];
#pragma unroll
    for (int k = 0; k < RPT; k++) {
      const int gac = min(a0 + ty + k * ROWS, La - 1);
      ra[k] = p.src.rank[p.a][gac];
      aa[k] = p.src.addr[p.a][gac];
    }
#pragma unroll
    for (int n = 0; n < NB; n++)
#pragma unroll
      for (int k = 0; k < RPT; k++) {
        const bool in = live[n] && a0 + ty + k * ROWS < La && b0 + tx < Lb;
        const int R = qs[n].rank + ra[k] + rb;
        s[n][k] = (in && R < p.src.proc_used) ? (i64)R * p.src_slot + qs[n].addr + aa[k] + ab : -1;
      }
  } else {
#pragma unroll
    for (int n = 0; n < NB; n++)
#pragma unroll
      for (int k = 0; k < RPT; k++) {
        int ga = a0 + ty + k * ROWS, gb = b0 + tx;
        bool in = live[n] && ga < La && gb < Lb;
        s[n][k] = locate<US>(p.src, p, qs[n], p.a, in ? ga : 0, p.b, in ? gb : 0, p.src_slot);
        if (!in) s[n][k] = -1;
      }
  }
  E v[NB][RPT];
#pragma unroll
  for (int n = 0; n < NB; n++)
#pragma unroll
    for (int k = 0; k < RPT; k++) v[n][k] = read_elem<NT>(src + (s[n][k] >= 0 ? s[n][k] : 0));
#pragma unroll
  for (int n = 0; n < NB; n++)
#pragma unroll
    for (int k = 0; k < RPT; k++) {
      tile[n][ty + k * ROWS][tx] = v[n][k];
      have[n][ty + k * ROWS][tx] = s[n][k] >= 0;
    }
  __syncthreads();
  // store: consecutive threads walk axis a
  i64 d[NB][RPT];
  if constexpr (UD) {
    const int gac = min(a0 + tx, La - 1);
    const int ra = p.dst.rank[p.a][gac];
    const i64 aa = p.dst.addr[p.a][gac];
    int rb[RPT];
    i64 ab[RPT];
#pragma unroll
    for (int k = 0; k < RPT; k++) {
      const int gbc = min(b0 + ty + k * ROWS, Lb - 1);
      rb[k] = p.dst.rank[p.b][gbc];
      ab[k] = p.dst.addr[p.b][gbc];
    }
#pragma unroll
    for (int n = 0; n < NB; n++)
#pragma unroll
      for (int k = 0; k < RPT; k++) {
        const bool in = live[n] && a0 + tx < La && b0 + ty + k * ROWS < Lb && have[n][tx][ty + k * ROWS];
        const int R = qd[n].rank + ra + rb[k];
        d[n][k] = (in && R < p.dst.proc_used) ? (i64)R * p.dst_slot + qd[n].addr + aa + ab[k] : -1;
      }
  } else {
#pragma unroll
    for (int n = 0; n < NB; n++)
#pragma unroll
      for (int k = 0; k < RPT; k++) {
        int ga = a0 + tx, gb = b0 + ty + k * ROWS;
        bool in = live[n] && ga < La && gb < Lb;
        d[n][k] = locate<UD>(p.dst, p, qd[n], p.a, in ? ga : 0, p.b, in ? gb : 0, p.dst_slot);
        if (!in || !have[n][tx][ty + k * ROWS]) d[n][k] = -1;
      }
  }
#pragma unroll
  for (int n = 0; n < NB; n++)
#pragma unroll
    for (int k = 0; k < RPT; k++)
      if (d[n][k] >= 0) write_elem<NT>(dst + d[n][k], tile[n][tx][ty + k * ROWS]);
}

// 1D runs along axis a (same fastest axis on both sides), phased as above.
template <int BYTES, bool US, bool UD, bool NT>
__global__ __launch_bounds__(NTHREADS) void reshape_run1d(ReshapeLaunch p) {
  typedef typename Elem<BYTES>::T E;
  constexpr int PER = 4;
  long long bid = blockIdx.x;
  int ta = (int)(bid % p.tiles_a);
  bid /= p.tiles_a;
  Partial qs, qd;
  partial_init<US>(p.src, p, bid, qs);
  partial_init<UD>(p.dst, p, bid, qd);
  const E *src = reinterpret_cast<const E *>(p.src_base);
  E *dst = reinterpret_cast<E *>(p.dst_base);
  const int La = p.L[p.a];
  int base = ta * NTHREADS * PER;
  i64 s[PER], d[PER];
#pragma unroll
  for (int k = 0; k < PER; k++) {
    int ga = base + k * NTHREADS + threadIdx.x;
    bool in = ga < La;
    int g = in ? ga : 0;
    s[k] = locate<US>(p.src, p, qs, p.a, g, p.a, g, p.src_slot);
    d[k] = locate<UD>(p.dst, p, qd, p.a, g, p.a, g, p.dst_slot);
    if (!in || s[k] < 0 || d[k] < 0) s[k] = d[k] = -1;
  }
  E v[PER];
#pragma unroll
  for (int k = 0; k < PER; k++) v[k] = read_elem<NT>(src + (s[k] >= 0 ? s[k] : 0));
#pragma unroll
  for (int k = 0; k < PER; k++)
    if (d[k] >= 0) write_elem<NT>(dst + d[k], v[k]);
}

template <int BYTES, bool US, bool UD, bool NT>
hipError_t launch_hinted(const ReshapeLaunch &p, hipStream_t s) {
  if (p.a != p.b)
    hipLaunchKernelGGL((reshape_tile2d<BYTES, US, UD, tile_slices(BYTES), NT>), dim3((unsigned)p.nblocks),
                       dim3(NTHREADS), 0, s, p);
  else
    hipLaunchKernelGGL((reshape_run1d<BYTES, US, UD, NT>), dim3((unsigned)p.nblocks), dim3(NTHREADS), 0, s, p);
  return hipGetLastError();
}

template <int BYTES, bool US, bool UD>
hipError_t launch_typed(const ReshapeLaunch &p, hipStream_t s) {
  if (p.nblocks <= 0) return hipSuccess;
  if (p.nontemporal) return launch_hinted<BYTES, US, UD, true>(p, s);
  return launch_hinted<BYTES, US, UD, false>(p, s);
}

template <int BYTES>
hipError_t launch_bytes(const ReshapeLaunch &p, hipStream_t s) {
  bool us = p.src.uniform, ud = p.dst.uniform;
  if (us && ud) return launch_typed<BYTES, true, true>(p, s);
  if (us) return launch_typed<BYTES, true, false>(p, s);
  if (ud) return launch_typed<BYTES, false, true>(p, s);
  return launch_typed<BYTES, false, false>(p, s);
}

}  // namespace

// Tile counts and grid size are filled here so the host only chooses axes.
int launch_reshape(const ReshapeLaunch &p0, void *stream) {
  ReshapeLaunch p = p0;
  hipStream_t s = (hipStream_t)stream;
  long long batch = 1;
  for (int k = 0; k < p.nbatch; k++) batch *= p.L[p.batch_axis[k]];
  p.nbatch_total = batch;
  if (p.a != p.b) {
    p.tiles_a = (p.L[p.a] + TILE - 1) / TILE;
    p.tiles_b = (p.L[p.b] + TILE - 1) / TILE;
    const int slices = tile_slices(p.elem_bytes);
    p.nblocks = (long long)p.tiles_a * p.tiles_b * ((batch + slices - 1) / slices);
  } else {
    p.tiles_a = (p.L[p.a] + NTHREADS * 4 - 1) / (NTHREADS * 4);
    p.tiles_b = 1;
    p.nblocks = (long long)p.tiles_a * batch;
  }
  switch (p.elem_bytes) {
    case 1: return (int)launch_bytes<1>(p, s);
    case 2: return (int)launch_bytes<2>(p, s);
    case 4: return (int)launch_bytes<4>(p, s);
    case 8: return (int)launch_bytes<8>(p, s);
    case 16: return (int)launch_bytes<16>(p, s);
    default: return (int)hipErrorInvalidValue;  // other sizes are split into words by the caller
  }
}

}  // namespace dm

// ---- plain copy, the calibration reference for bandwidth and counters ------
namespace dm {
namespace {
// One 16-byte element per thread, 4 KiB per workgroup, non-temporal: the
// fastest plain copy measured here (6.4 TB/s for 2 GiB; 16 KiB per
// workgroup without the hint: 5.6-5.8 TB/s).
__global__ __launch_bounds__(256) void copy16_kernel(const uint4 *__restrict__ src, uint4 *__restrict__ dst,
                                                     long long n) {
  typedef unsigned int v4 __attribute__((ext_vector_type(4)));
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) {
    const v4 v = __builtin_nontemporal_load(reinterpret_cast<const v4 *>(src) + i);
    __builtin_nontemporal_store(v, reinterpret_cast<v4 *>(dst) + i);
  }
}
}  // namespace

int launch_copy16(const void *src, void *dst, long long n16, void *stream) {
  long long blocks = (n16 + 255) / 256;
  if (blocks <= 0) return 0;
  hipLaunchKernelGGL(copy16_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                     (const uint4 *)src, (uint4 *)dst, n16);
  return (int)hipGetLastError();
}
}  // namespace dm
