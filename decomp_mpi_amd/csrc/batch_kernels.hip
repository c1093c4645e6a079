// K3: one axis pass of many independent local FFTs in one persistent launch.
//
// Each tile is nl lines of N points of one transform, read into an LDS tile
// (coalesced: along the line for the contiguous axis, across the tile's
// consecutive lines otherwise), transformed by the Stockham stages of its
// schedule, and written back in place.  Tiles of different transforms (and
// line lengths) share the launch; the workgroups loop over the tile list.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "fft_device.h"
#include "kernels.h"

namespace dm {
namespace {

using namespace fftdev;

template <typename R>
using gptr = __attribute__((address_space(1))) cx<R> *;  // global memory

constexpr int BCAP_D = 4096;
constexpr int BCAP_F = 4096;  // 8 points per thread in both precisions
constexpr int BNTH = 512;
// Plane tiles: a workgroup holds a whole plane of the last two axes in LDS.
// Planes of up to the line tile's points run in the line kernel's shape (8
// points per thread, 512 threads); complex-float planes of up to 8192
// points at 16 points per thread on 512 threads (two workgroups per CU by
// LDS).  (A 16384-point plane on 1024 threads, one workgroup per CU, took
// 2.7 times as long per byte as a line pass: no memory work overlaps its
// stages.)
constexpr int PCAP_F = 8192;
// Twiddles (16 bytes per root) go to LDS after the tile when every schedule
// of the launch has lines of at most BTW_LDS_N points; otherwise the launch
// reads them from device memory.
constexpr int BTW_LDS_N = 256;
__host__ __device__ __forceinline__ size_t batch_tw_at(int T, int N, int es) {
  return ((size_t)T * (size_t)lds_line(N) * es + 15) & ~(size_t)15;
}
template <typename R>
using lds_twg = __attribute__((address_space(3))) typename tw_global<R>::type;


// Element access (the non-temporal hint measured slower here: a pass reads
// and writes in place, and a batch partly stays in the memory-side cache
// between passes; with the hint the complex-float kernel also spills at 64
// VGPRs).
template <typename R>
__device__ __forceinline__ cx<R> gload(gptr<R> data, long long i) {
  return {data[i].x, data[i].y};
}
template <typename R>
__device__ __forceinline__ void gstore(gptr<R> data, long long i, cx<R> v) {
  data[i].x = v.x;
  data[i].y = v.y;
}

// Butterfly bi of a stage with nr butterflies per line: row walks put
// consecutive butterflies of a line on consecutive threads, column walks
// consecutive lines (so that each butterfly input is one coalesced access).
__device__ __forceinline__ void bfly(int bi, int row, int nr, unsigned long long m_nr, int T,
                                     unsigned long long m_t, int &l, int &j) {
  if (row) {
    l = fdiv(bi, m_nr);
    j = bi - l * nr;
  } else {
    j = fdiv(bi, m_t);
    l = bi - j * T;
  }
}

// First stage on the loaded values: whole butterflies per thread (points
// j + nr r of line l), no twiddles, outputs to their Stockham places.
template <typename R, bool INV, int RADIX, int CAP, int NTH>
__device__ __forceinline__ void load_first(cx<R> *buf, const BatchSched &S, const BatchTile &d,
                                           gptr<R> data, long long lstride, int ls) {
  constexpr int K = (CAP / NTH) / RADIX;
  const StageInfo &st = S.st[0];
  const int nr = st.nr, nb = S.T * nr;
  long long a[K][RADIX];
  int lk[K], jk[K];
  bool ok[K];
#pragma unroll
  for (int q = 0; q < K; q++) {
    int bi = threadIdx.x + q * NTH;
    bfly(min(bi, nb - 1), d.row, nr, st.m_nr, S.T, S.m_t, lk[q], jk[q]);
    ok[q] = bi < nb && lk[q] < d.nl;
#pragma unroll
    for (int r = 0; r < RADIX; r++)
      a[q][r] = ok[q] ? (long long)lk[q] * lstride + (long long)(jk[q] + r * nr) * d.ps : 0;
  }
  cx<R> v[K][RADIX];
#pragma unroll
  for (int q = 0; q < K; q++)
#pragma unroll
    for (int r = 0; r < RADIX; r++) v[q][r] = gload(data, a[q][r]);
#pragma unroll
  for (int q = 0; q < K; q++) {
    if (!ok[q])
#pragma unroll
      for (int r = 0; r < RADIX; r++) v[q][r] = {0, 0};
    dft<R, INV, RADIX>(v[q]);
    if ((int)threadIdx.x + q * NTH < nb) {
      const int base = lk[q] * ls, idx = jk[q] * RADIX;
#pragma unroll
      for (int r = 0; r < RADIX; r++) buf[base + lds_pad(idx + r)] = v[q][r];
    }
  }
}

// Last stage straight to memory: butterfly inputs from LDS, twiddles
// (stage-ordered: Ns = nr in the last stage), transform, store.
template <typename R, bool INV, int RADIX, int CAP, int NTH, typename TW>
__device__ __forceinline__ void last_store(const cx<R> *buf, const BatchSched &S, const BatchTile &d,
                                           gptr<R> data, long long lstride, int ls, TW tw) {
  constexpr int K = (CAP / NTH) / RADIX;
  const StageInfo &st = S.st[S.nstage - 1];
  const int nr = st.nr, nb = S.T * nr;
#pragma unroll
  for (int q = 0; q < K; q++) {
    int bi = threadIdx.x + q * NTH;
    int l, j;
    bfly(min(bi, nb - 1), d.row, nr, st.m_nr, S.T, S.m_t, l, j);
    const bool ok = bi < nb && l < d.nl;
    cx<R> v[RADIX];
#pragma unroll
    for (int r = 0; r < RADIX; r++) v[r] = buf[l * ls + lds_pad(j + r * nr)];
#pragma unroll
    for (int r = 1; r < RADIX; r++) v[r] = twmul<INV, (sizeof(R) == 4)>(v[r], tw, (nr - 1) + (r - 1) * nr + j);
    dft<R, INV, RADIX>(v);
    if (ok)
#pragma unroll
      for (int r = 0; r < RADIX; r++) {
        const long long at = (long long)l * lstride + (long long)(j + r * nr) * d.ps;
        gstore(data, at, v[r]);
      }
  }
}

// PRIMES: schedules with radix 5, 7, 11 or 13 (their butterflies need more
// registers, so launches without them use a lighter kernel).  Waves per
// SIMD: complex float fits 64 VGPRs without spills, so 8 waves and four
// 512-thread workgroups per CU keep more loads in flight (81 VGPRs allowed
// only two); complex double spills below 104, so it stays at 4.
// TWL: the schedules' twiddles are copied to LDS per tile (every line of
// the launch at most BTW_LDS_N points).
// PLANE: every tile is a plane tile (row 2): the rows' stages, then the
// columns' stages on the same LDS tile, then one row-order store.  Held to
// 64 VGPRs the complex-float form spills (16 bytes per lane, 28% more bytes
// written per launch), so its planes of up to the line tile's points get 6
// waves per SIMD instead of 8; larger planes, and complex double, 4.
// ANY: some schedule has a direct-sum stage (a prime above 13; line tiles
// only), a kernel of its own so that the others keep their registers.
template <typename R, bool INV, int CAP, int NTH, bool PRIMES, bool TWL, bool PLANE, bool ANY>
__global__ __launch_bounds__(NTH, (PLANE && sizeof(R) == 4 && CAP / NTH == 8) ? 6 : ((sizeof(R) == 4 && CAP / NTH == 8) ? 8 : 4)) void batch_pass_kernel(BatchLaunch p) {
  extern __shared__ char smem_raw[];
  cx<R> *buf = reinterpret_cast<cx<R> *>(smem_raw);
  constexpr int V = CAP / NTH;
  constexpr bool CT = sizeof(R) == 4;  // complex float: compensated twiddle products
  for (long long t = blockIdx.x; t < p.ntiles; t += gridDim.x) {
    // the tile, its schedule, roots and transform: from the launch's tables,
    // or from the segment (one transform's own tables) that holds tile t
    BatchTile d;
    const BatchSched *sched;
    typedef typename tw_global<R>::type TwE;
    const TwE *twall;
    void *xdata;
    if (!PLANE && p.nseg > 0) {
      int sg = 0;
      while (sg + 1 < p.nseg && t >= p.seg[sg].end) sg++;
      const BatchSeg &g = p.seg[sg];
      const long long q = t - (sg > 0 ? p.seg[sg - 1].end : 0);
      if (g.inner == 1) {  // rows: T consecutive lines of N contiguous points
        const long long o = q * g.T;
        d.base = o * g.N;
        d.ps = 1;
        d.nl = (int)min((long long)g.T, (long long)g.outer - o);
        d.row = 1;
      } else {             // columns: T consecutive inner offsets of one outer coordinate
        const long long o = q / g.ctiles;
        const long long c = (q - o * g.ctiles) * g.T;
        d.base = o * g.N * g.inner + c;
        d.ps = g.inner;
        d.nl = (int)min((long long)g.T, g.inner - c);
        d.row = 0;
      }
      d.xform = 0;
      d.sched = 0;
      d.sched2 = 0;
      sched = g.sched;
      twall = reinterpret_cast<const TwE *>(g.twiddle);
      xdata = g.data;
    } else {
      d = p.tiles[t];
      sched = p.sched;
      twall = reinterpret_cast<const TwE *>(p.twiddle);
      xdata = p.data[d.xform];
    }
    const BatchSched &S = sched[d.sched];
    const int N = S.N, T = S.T, nl = d.nl;
    const int tot = T * N;
    const int ls = lds_line(N);
    const long long lstride = d.row ? (long long)N : 1LL;
    // the transform's pointer comes from memory, so its address space is
    // stated: through a generic pointer every access would be a flat one
    gptr<R> data = (gptr<R>)(xdata) + d.base;
    const TwE *tw = twall + S.tw_off;
    TwE *twl = reinterpret_cast<TwE *>(smem_raw + batch_tw_at(T, N, sizeof(cx<R>)));
    if constexpr (TWL)
      for (int i = threadIdx.x; i < N; i += NTH) twl[i] = tw[i];
    if constexpr (PLANE && TWL) {  // the columns' roots after the rows'
      const BatchSched &S2 = sched[d.sched2];
      const TwE *tw2 = twall + S2.tw_off;
      for (int i = threadIdx.x; i < S2.N; i += NTH) twl[N + i] = tw2[i];
    }
    // Fused first / last stages: a power-of-two radix dividing the points per
    // thread (whole butterflies per thread); the last one needs two stages.
    const int r0 = S.radix[0], rl = S.radix[S.nstage - 1];
    // (fusing row walks too measured faster than loading their short
    // per-butterfly runs through the plain element walk: 1.27 vs 1.31 ms
    // per 256-transform batch)
    const bool fuse_in = (r0 == 2 || r0 == 4 || r0 == 8) && V % r0 == 0;
    const bool fuse_out = !PLANE && S.nstage >= 2 && (rl == 2 || rl == 4 || rl == 8) && V % rl == 0;
    if (fuse_in) {
      if (r0 == 8) load_first<R, INV, 8, CAP, NTH>(buf, S, d, data, lstride, ls);
      else if (r0 == 4) load_first<R, INV, 4, CAP, NTH>(buf, S, d, data, lstride, ls);
      else load_first<R, INV, 2, CAP, NTH>(buf, S, d, data, lstride, ls);
    } else {
      // load: addresses, then data, then LDS (branch-free; absent elements
      // read element 0 of the tile and are dropped)
      long long a[V];
      int sl[V];
      bool ok[V];
#pragma unroll
      for (int k = 0; k < V; k++) {
        int e = threadIdx.x + k * NTH;
        int ec = min(e, tot - 1);
        int l, x;
        if (d.row) {
          l = fdiv(ec, S.m_n);
          x = ec - l * N;
        } else {
          x = fdiv(ec, S.m_t);
          l = ec - x * T;
        }
        ok[k] = e < tot && l < nl;
        sl[k] = e < tot ? lds_index(l, x, ls) : -1;
        a[k] = ok[k] ? (long long)l * lstride + (long long)x * d.ps : 0;
      }
      cx<R> v[V];
#pragma unroll
      for (int k = 0; k < V; k++) v[k] = gload(data, a[k]);
#pragma unroll
      for (int k = 0; k < V; k++)
        if (sl[k] >= 0) buf[sl[k]] = ok[k] ? v[k] : cx<R>{0, 0};
    }
    __syncthreads();
    const int s0 = fuse_in ? 1 : 0, s1 = fuse_out ? S.nstage - 1 : S.nstage;
    auto run_stages = [&](auto tw) {
      // roots of the direct-sum stages: after the schedule's N stage
      // entries, p per such stage in stage order (a fused first stage is a
      // power of two, so none is skipped)
      int roff = N;
      for (int s = s0; s < s1; s++) {
        switch (S.radix[s]) {
          case 2: stage_lds<R, INV, 2, CAP, NTH, CT>(buf, S.st[s], T, ls, tw); break;
          case 3: stage_lds<R, INV, 3, CAP, NTH, CT>(buf, S.st[s], T, ls, tw); break;
          case 4: stage_lds<R, INV, 4, CAP, NTH, CT>(buf, S.st[s], T, ls, tw); break;
          case 5: if constexpr (PRIMES) stage_lds<R, INV, 5, CAP, NTH, CT>(buf, S.st[s], T, ls, tw); break;
          case 7: if constexpr (PRIMES) stage_lds<R, INV, 7, CAP, NTH, CT>(buf, S.st[s], T, ls, tw); break;
          case 8: stage_lds<R, INV, 8, CAP, NTH, CT>(buf, S.st[s], T, ls, tw); break;
          case 11: if constexpr (PRIMES) stage_lds<R, INV, 11, CAP, NTH, CT>(buf, S.st[s], T, ls, tw); break;
          case 13: if constexpr (PRIMES) stage_lds<R, INV, 13, CAP, NTH, CT>(buf, S.st[s], T, ls, tw); break;
          case 16: stage_lds<R, INV, 16, CAP, NTH, CT>(buf, S.st[s], T, ls, tw); break;
          default:
            if constexpr (ANY) {
              stage_lds_any<R, INV, CAP, NTH, CT>(buf, S.st[s], S.radix[s], T, ls, tw, twall + S.tw_off + roff);
              roff += S.radix[s];
            }
            break;
        }
      }
      if (fuse_out) {
        if (rl == 8) last_store<R, INV, 8, CAP, NTH>(buf, S, d, data, lstride, ls, tw);
        else if (rl == 4) last_store<R, INV, 4, CAP, NTH>(buf, S, d, data, lstride, ls, tw);
        else last_store<R, INV, 2, CAP, NTH>(buf, S, d, data, lstride, ls, tw);
      }
    };
    if constexpr (TWL) run_stages((const lds_twg<R> *)twl);
    else run_stages(tw);
    if constexpr (PLANE) {
      // the columns: S2.T = N lines of S2.N = T points, point x of column l
      // at x * ls + pad(l) (stage_lds ends on a barrier)
      const BatchSched &S2 = sched[d.sched2];
      auto run_cols = [&](auto tw2) {
        typedef decltype(tw2) TW2;
        for (int s = 0; s < S2.nstage; s++) {
          switch (S2.radix[s]) {
            case 2: stage_lds<R, INV, 2, CAP, NTH, CT, TW2, true>(buf, S2.st[s], S2.T, ls, tw2, S2.m_t); break;
            case 3: stage_lds<R, INV, 3, CAP, NTH, CT, TW2, true>(buf, S2.st[s], S2.T, ls, tw2, S2.m_t); break;
            case 4: stage_lds<R, INV, 4, CAP, NTH, CT, TW2, true>(buf, S2.st[s], S2.T, ls, tw2, S2.m_t); break;
            case 5: if constexpr (PRIMES) stage_lds<R, INV, 5, CAP, NTH, CT, TW2, true>(buf, S2.st[s], S2.T, ls, tw2, S2.m_t); break;
            case 7: if constexpr (PRIMES) stage_lds<R, INV, 7, CAP, NTH, CT, TW2, true>(buf, S2.st[s], S2.T, ls, tw2, S2.m_t); break;
            case 8: stage_lds<R, INV, 8, CAP, NTH, CT, TW2, true>(buf, S2.st[s], S2.T, ls, tw2, S2.m_t); break;
            case 11: if constexpr (PRIMES) stage_lds<R, INV, 11, CAP, NTH, CT, TW2, true>(buf, S2.st[s], S2.T, ls, tw2, S2.m_t); break;
            case 13: if constexpr (PRIMES) stage_lds<R, INV, 13, CAP, NTH, CT, TW2, true>(buf, S2.st[s], S2.T, ls, tw2, S2.m_t); break;
            case 16: stage_lds<R, INV, 16, CAP, NTH, CT, TW2, true>(buf, S2.st[s], S2.T, ls, tw2, S2.m_t); break;
            default: break;
          }
        }
      };
      if constexpr (TWL) run_cols((const lds_twg<R> *)twl + N);
      else run_cols(twall + S2.tw_off);
    }
    if (!fuse_out) {
      // store: same element walk as the load, straight from LDS
#pragma unroll
      for (int k = 0; k < V; k++) {
        int e = threadIdx.x + k * NTH;
        int ec = min(e, tot - 1);
        int l, x;
        if (d.row) {
          l = fdiv(ec, S.m_n);
          x = ec - l * N;
        } else {
          x = fdiv(ec, S.m_t);
          l = ec - x * T;
        }
        if (e < tot && l < nl) {
          const cx<R> w = buf[lds_index(l, x, ls)];
          const long long at = (long long)l * lstride + (long long)x * d.ps;
          gstore(data, at, w);
        }
      }
    }
    __syncthreads();  // the next tile reuses the LDS tile
  }
}

template <typename R, bool INV, int CAP, bool PRIMES, bool TWL, bool PLANE = false, bool ANY = false>
hipError_t launch_typed(const BatchLaunch &p, hipStream_t s) {
  constexpr int NTH = BNTH;  // every form, the large planes at 16 points per thread
  auto k = batch_pass_kernel<R, INV, CAP, NTH, PRIMES, TWL, PLANE, ANY>;
  // once per kernel (a thread-safe static)
  static const hipError_t attr =
      hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
  if (attr != hipSuccess) return attr;
  hipLaunchKernelGGL(k, dim3((unsigned)p.grid), dim3(NTH), p.lds, s, p);
  return hipGetLastError();
}

}  // namespace

int batch_tile_capacity(int dbl) { return dbl ? BCAP_D : BCAP_F; }

int batch_plane_capacity(int dbl) { return dbl ? BCAP_D : PCAP_F; }

size_t batch_plane_lds_bytes(int dbl, int T, int N) {
  if ((long long)T * N > batch_plane_capacity(dbl) || !batch_tw_in_lds(N) || !batch_tw_in_lds(T)) return 0;
  const size_t b = batch_tw_at(T, N, dbl ? 16 : 8) + (size_t)(N + T) * 16;
  return b <= 160 * 1024 ? b : 0;
}

bool batch_tw_in_lds(int N) { return N <= BTW_LDS_N; }

size_t batch_lds_bytes(int dbl, int T, int N) {
  return batch_tw_at(T, N, dbl ? 16 : 8) + (batch_tw_in_lds(N) ? (size_t)N * 16 : 0);
}

// Workgroups resident at once: by LDS (160 KiB per CU) and by registers
// (512 VGPRs per SIMD lane, 4 SIMDs per CU, at most 8 waves per SIMD).
long long batch_resident_groups(int dbl, int primes, size_t lds) {
  const void *k;
  if (dbl) k = primes ? (const void *)batch_pass_kernel<double, false, BCAP_D, BNTH, true, true, false, false>
                      : (const void *)batch_pass_kernel<double, false, BCAP_D, BNTH, false, true, false, false>;
  else k = primes ? (const void *)batch_pass_kernel<float, false, BCAP_F, BNTH, true, true, false, false>
                  : (const void *)batch_pass_kernel<float, false, BCAP_F, BNTH, false, true, false, false>;
  hipFuncAttributes fa;
  int vgprs = (hipFuncGetAttributes(&fa, k) == hipSuccess && fa.numRegs > 0) ? fa.numRegs : 128;
  int dev = 0, cus = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (cus <= 0) cus = 256;
  int waves_per_simd = std::min(8, 512 / (((vgprs + 7) / 8) * 8));
  int by_regs = (waves_per_simd * 4) / (BNTH / 64);
  int by_lds = (int)((160 * 1024) / std::max<size_t>(lds, 1));
  return (long long)cus * std::max(1, std::min(by_regs, by_lds));
}

// Large-plane kernel: at most two workgroups per CU (128 VGPRs at 512
// threads), fewer if LDS says so.
long long batch_plane_resident_groups(size_t lds) {
  int dev = 0, cus = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (cus <= 0) cus = 256;
  return (long long)cus * std::max(1, std::min(2, (int)((160 * 1024) / std::max<size_t>(lds, 1))));
}

template <typename R, int CAP, bool TWL, bool PLANE = false>
hipError_t launch_forms(const BatchLaunch &p, hipStream_t s) {
  if constexpr (!PLANE)
    if (p.any)
      return p.inverse ? launch_typed<R, true, CAP, true, TWL, false, true>(p, s)
                       : launch_typed<R, false, CAP, true, TWL, false, true>(p, s);
  if (p.primes)
    return p.inverse ? launch_typed<R, true, CAP, true, TWL, PLANE>(p, s) : launch_typed<R, false, CAP, true, TWL, PLANE>(p, s);
  return p.inverse ? launch_typed<R, true, CAP, false, TWL, PLANE>(p, s) : launch_typed<R, false, CAP, false, TWL, PLANE>(p, s);
}

int launch_batch_pass(const BatchLaunch &p, void *stream) {
  if (p.ntiles <= 0 || p.grid <= 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  hipError_t e;
  // plane tiles: twiddles always in LDS (both axes at most BTW_LDS_N
  // points); plane 2: planes within the line tile, plane 1: larger ones
  // (complex float only)
  if (p.plane == 2) {
    e = p.dbl ? launch_forms<double, BCAP_D, true, true>(p, s) : launch_forms<float, BCAP_F, true, true>(p, s);
    return (int)e;
  }
  if (p.plane == 1) {
    if (p.dbl) return (int)hipErrorInvalidValue;
    return (int)launch_forms<float, PCAP_F, true, true>(p, s);
  }
  if (p.dbl) e = p.long_lines ? launch_forms<double, BCAP_D, false>(p, s) : launch_forms<double, BCAP_D, true>(p, s);
  else e = p.long_lines ? launch_forms<float, BCAP_F, false>(p, s) : launch_forms<float, BCAP_F, true>(p, s);
  return (int)e;
}

}  // namespace dm
