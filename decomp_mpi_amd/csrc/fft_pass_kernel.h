// K2: batched 1D FFT pass over one axis of a (virtually) distributed array:
// the kernel template and its launchers.
//
// Replaces the FFTW guru stages of the split FFT (split_fft.c:426-546) and the
// twiddle pass (dimscale.c:7-49).  A workgroup loads a tile of T lines x N
// points into LDS through per-axis offset tables (so any decomposition, rank
// slot layout or factor-reversed order is read and written directly), runs
// mixed-radix Stockham stages (radix 2,3,4,5,7,8,11,13,16) in LDS, and stores
// the tile through the output tables.  The walk over memory follows the
// contiguous direction of each side: along the line (row mode) or across the
// T lines of the tile (column mode), so that global accesses coalesce.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <type_traits>

#include "fft_device.h"
#include "kernels.h"

namespace dm {

namespace fftpass {

using namespace fftdev;

struct Pass {
  FftLaunch p;
  StageInfo st[16];
  unsigned long long m_n, m_t;  // magics for N and T
  int ls;                        // LDS line stride (complex elements)
  int fuse_in;                   // first stage runs on the loaded registers
  int fuse_out;                  // last stage stores straight from registers
  int xcd_pair;                  // tiles sharing 128-byte rows, run back to back on one XCD
  int tw_lds;                    // twiddle table copied to LDS
};

template <typename R, bool INV, int RADIX, int CAP, int NTH, bool CT, typename TW>
__device__ __forceinline__ void stage(cx<R> *buf, const Pass &P, int s, TW tw) {
  stage_lds<R, INV, RADIX, CAP, NTH, CT>(buf, P.st[s], P.p.T, P.ls, tw);
}
template <typename R, bool INV, int CAP, int NTH, bool CT, typename TW, typename RT>
__device__ __forceinline__ void stage_any(cx<R> *buf, const Pass &P, int s, TW tw, RT roots) {
  stage_lds_any<R, INV, CAP, NTH, CT>(buf, P.st[s], P.p.radix[s], P.p.T, P.ls, tw, roots);
}

// Per-line address data of one side of a tile (in LDS).
//   uniform:     base[l] = sum over axes other than t of abs[d][g_d]
//   non-uniform: rbase[l] = owner contribution of those axes, offs[l][k] =
//                their local offsets (k over the axes other than t)
// Pointers into LDS are typed as such (address space 3) so that reads of
// the line data are LDS instructions; through a generic pointer they would
// be flat loads, queued in order with the global data loads.
typedef __attribute__((address_space(3))) i64 lds_i64;
typedef __attribute__((address_space(3))) int lds_int;
template <typename R>
using lds_cx = __attribute__((address_space(3))) cx<R>;
template <typename R, bool CT>
using lds_tw = __attribute__((address_space(3))) typename tw_local<R, CT>::type;

template <bool UNIFORM>
struct LineAddr;

template <>
struct LineAddr<true> {
  lds_i64 *base;
  __host__ __device__ static size_t bytes(int T, int) { return (size_t)T * sizeof(i64); }
  __device__ void init(char *mem, int, int) { base = (lds_i64 *)(mem); }
  __device__ void set_line(const FftLaunch &p, const i64 *const *abs, const int *g, int l) {
    i64 a = 0;
    for (int d = 0; d < p.ndim; d++)
      if (d != p.t) a += abs[d][g[d]];
    base[l] = a;
  }
  __device__ __forceinline__ i64 at(const FftLaunch &, const FftSide &s, const i64 *absT, int l, int x) const {
    if (s.aff) return base[l] + (i64)x * s.aff;  // affine along t: no table load
    if (s.jump) return base[l] + (i64)(x >> s.rsh) * s.jump + (i64)(x & s.rmask) * s.rstride;
    return base[l] + absT[x];
  }
  // the same with the side's form fixed at compile time (no branch)
  template <bool AFF>
  __device__ __forceinline__ i64 at_form(const FftSide &s, const i64 *absT, int l, int x) const {
    if constexpr (AFF) return base[l] + (i64)x * s.aff;
    else return base[l] + absT[x];
  }
};

template <>
struct LineAddr<false> {
  lds_int *rbase;
  lds_int *offs;
  int nd;
  __host__ __device__ static size_t bytes(int T, int ndim) { return (size_t)T * sizeof(int) * (1 + ndim); }
  __device__ void init(char *mem, int T, int ndim) {
    rbase = (lds_int *)(mem);
    offs = rbase + T;
    nd = ndim;
  }
  __device__ void set_line(const FftLaunch &p, const FftSide &s, const int *g, int l) {
    int r = 0;
    for (int d = 0; d < p.ndim; d++) {
      if (d == p.t) continue;
      r += s.rank[d][g[d]];
      offs[l * nd + d] = s.off[d][g[d]];
    }
    rbase[l] = r;
  }
  __device__ __forceinline__ i64 at(const FftLaunch &p, const FftSide &s, const i64 *, int l, int x) const {
    int R = rbase[l] + s.rank[p.t][x];
    const i64 *m = s.strides + (i64)R * p.ndim;
    i64 a = (i64)R * s.slot + m[p.t] * s.off[p.t][x];
    for (int d = 0; d < p.ndim; d++)
      if (d != p.t) a += m[d] * offs[l * nd + d];
    return a;
  }
};

// Tile t -> (tile along u, batch coordinates); fills the per-line address
// data of both sides for that tile (threads l < T).
template <bool UIN, bool UOUT>
__device__ __forceinline__ int setup_lines(const FftLaunch &p, long long tile, LineAddr<UIN> &lin,
                                           LineAddr<UOUT> &lout, int T, int NTHR) {
  int tu = (int)(tile % p.tiles_u);
  const int u0 = tu * T;
  const int mk = p.merge_k;
  const int nlines = mk > 1 ? T : (p.u >= 0 ? min(T, p.L[p.u] - u0) : 1);
  if constexpr (UIN && UOUT) {
    if (p.line_in_tab != nullptr) {  // precomputed line bases: one load each
      for (int l = threadIdx.x; l < T; l += NTHR) {
        lin.base[l] = p.line_in_tab[tile * T + l];
        lout.base[l] = p.line_out_tab[tile * T + l];
      }
      return nlines;
    }
  }
  long long bid = tile / p.tiles_u;
  int gb[DM_MAX_DIMS];
  for (int d = 0; d < DM_MAX_DIMS; d++) gb[d] = 0;
  for (int k = p.nbatch - 1; k >= 0; k--) {
    int d = p.batch_axis[k];
    const bool merged = mk > 1 && d == p.merge_v;
    const int len = merged ? p.L[d] / mk : p.L[d];
    const int c = (int)(bid % len);
    gb[d] = merged ? c * mk : c;
    bid /= len;
  }
  for (int l = threadIdx.x; l < T; l += NTHR) {
    int g[DM_MAX_DIMS];
    for (int d = 0; d < p.ndim; d++) g[d] = gb[d];
    g[p.t] = 0;
    if (mk > 1) {
      // all L[u] lines for each of the tile's merge_k values of axis merge_v
      const int lu = p.L[p.u];
      const int cu = l % lu;
      g[p.u] = p.ord_u ? p.ord_u[cu] : cu;
      g[p.merge_v] += l / lu;
    } else if (p.u >= 0) {
      g[p.u] = (l < nlines) ? (p.ord_u ? p.ord_u[u0 + l] : u0 + l) : 0;
    }
    if constexpr (UIN) lin.set_line(p, p.off_in, g, l);
    else lin.set_line(p, p.sin, g, l);
    if constexpr (UOUT) lout.set_line(p, p.off_out, g, l);
    else lout.set_line(p, p.sout, g, l);
  }
  return nlines;
}

// (line, point) of tile element e for a row walk (along the line) or a
// column walk (across the tile's lines); e must be below T * N.
__device__ __forceinline__ void element(const Pass &P, int row, const int *ord, int e, int &l, int &x) {
  if (row) {
    l = fdiv(e, P.m_n);
    int i = e - l * P.p.N;
    x = ord ? ord[i] : i;
  } else {
    x = fdiv(e, P.m_t);
    l = e - x * P.p.T;
  }
}

// Butterfly q of this thread in a stage with nr butterflies per line:
// column walks put consecutive threads on consecutive lines, row walks on
// consecutive butterflies of a line, so each input/output index r of the
// butterfly is one coalesced access across the wave.
__device__ __forceinline__ void butterfly_of(int bi, int row, int nr, unsigned long long m_nr, int T,
                                             unsigned long long m_t, int &l, int &j) {
  if (row) {
    l = fdiv(bi, m_nr);
    j = bi - l * nr;
  } else {
    j = fdiv(bi, m_t);
    l = bi - j * T;
  }
}

// First stage on the loaded values: each thread loads whole butterflies
// (points j + nr*r of line l), transforms them in registers (no twiddle in
// the first stage) and writes the outputs to their Stockham places in LDS.
template <typename R, bool INV, int RADIX, int CAP, int NTH, bool UIN>
__device__ __forceinline__ void load_first_stage(cx<R> *buf, const Pass &P, const LineAddr<UIN> &lin,
                                                 const cx<R> *in, const i64 *inT, int nl) {
  constexpr int K = (CAP / NTH) / RADIX;
  const FftLaunch &p = P.p;
  const StageInfo &si = P.st[0];
  const int nr = si.nr, nb = p.T * nr;
  i64 a[K][RADIX];
  int lk[K], jk[K];
  bool ok[K];
#pragma unroll
  for (int q = 0; q < K; q++) {
    int bi = threadIdx.x + q * NTH;
    int bc = min(bi, nb - 1);
    butterfly_of(bc, p.in_row, nr, si.m_nr, p.T, P.m_t, lk[q], jk[q]);
    ok[q] = bi < nb && lk[q] < nl;
    int ls = ok[q] ? lk[q] : 0;
#pragma unroll
    for (int r = 0; r < RADIX; r++) a[q][r] = lin.at(p, p.sin, inT, ls, jk[q] + r * nr);
  }
  cx<R> v[K][RADIX];
  gather<K * RADIX>(&v[0][0], in, &a[0][0], p.stream_in);
#pragma unroll
  for (int q = 0; q < K; q++) {
    if (!ok[q])
#pragma unroll
      for (int r = 0; r < RADIX; r++) v[q][r] = {0, 0};
    dft<R, INV, RADIX>(v[q]);
    if (threadIdx.x + q * NTH < nb) {
      int base = lk[q] * P.ls, idxd = jk[q] * RADIX;
#pragma unroll
      for (int r = 0; r < RADIX; r++) buf[base + lds_pad(idxd + r)] = v[q][r];
    }
  }
}

// Last stage straight to memory: read butterfly inputs from LDS, twiddle,
// transform, and store outputs j + nr*r of line l.
// AFF (uniform output sides, compensated twiddles): the output is affine
// along the line, chosen once per launch, so that no branches sit between
// the twiddle products and the stores (with them the compensated form
// spills at 64 VGPRs).
template <typename R, bool INV, int RADIX, int CAP, int NTH, bool UOUT, bool CT, bool AFF, typename TW>
__device__ __forceinline__ void last_stage_store(const cx<R> *buf, const Pass &P, const LineAddr<UOUT> &lout,
                                                 cx<R> *out, const i64 *outT, TW tw, int nl) {
  constexpr int K = (CAP / NTH) / RADIX;
  const FftLaunch &p = P.p;
  const StageInfo &si = P.st[p.nstage - 1];
  const int nr = si.nr, nb = p.T * nr;
  cx<R> v[K][RADIX];
  i64 a[K][RADIX];
  bool ok[K];
#pragma unroll
  for (int q = 0; q < K; q++) {
    int bi = threadIdx.x + q * NTH;
    int bc = min(bi, nb - 1);
    int l, j;
    butterfly_of(bc, p.out_row, nr, si.m_nr, p.T, P.m_t, l, j);
    ok[q] = bi < nb && l < nl;
    int base = l * P.ls;
#pragma unroll
    for (int r = 0; r < RADIX; r++) v[q][r] = buf[base + lds_pad(j + r * nr)];
    // in the last stage ns = nr, so j % ns = j (stage-ordered table)
#pragma unroll
    for (int r = 1; r < RADIX; r++) v[q][r] = twmul<INV, CT>(v[q][r], tw, (nr - 1) + (r - 1) * nr + j);
    dft<R, INV, RADIX>(v[q]);
    int ls = ok[q] ? l : 0;
#pragma unroll
    for (int r = 0; r < RADIX; r++) {
      if constexpr (UOUT && CT) a[q][r] = lout.template at_form<AFF>(p.sout, outT, ls, j + r * nr);
      else a[q][r] = lout.at(p, p.sout, outT, ls, j + r * nr);
    }
  }
  scatter<K * RADIX, RADIX>(out, &a[0][0], &v[0][0], ok, p.stream_out);
}

// One tile per workgroup.  Loads and stores run in three branch-free
// phases -- index lookups and addresses, then all data accesses back to
// back, then the LDS side -- because vector memory operations complete in
// order: a lookup issued after a data load would make the data load's
// latency part of every later address computation.
// Complex float at 8 points per thread is held to 64 VGPRs (8 waves per
// SIMD: two 1024-thread or four 512-thread workgroups per CU).
// CT: compensated twiddle products (complex float only).
// ANY: the schedule has a prime radix above 13 (direct-sum stages); such
// schedules get kernels of their own, so that the stage's registers do not
// weigh on the others.
template <typename R, bool INV, int CAP, int NTH, bool POW2, bool UIN, bool UOUT, bool CT, bool ANY>
__global__ __launch_bounds__(NTH, (sizeof(R) == 4 && CAP / NTH == 8) ? 8 : 1) void fft_pass_kernel(Pass P) {
  extern __shared__ char smem_raw[];
  cx<R> *buf = reinterpret_cast<cx<R> *>(smem_raw);
  const FftLaunch &p = P.p;
  const int T = p.T;
  constexpr int V = CAP / NTH;
  const int tot = T * p.N;

  char *lmem = smem_raw + (size_t)T * P.ls * sizeof(cx<R>);
  LineAddr<UIN> lin;
  LineAddr<UOUT> lout;
  lin.init(lmem, T, p.ndim);
  lout.init(lmem + LineAddr<UIN>::bytes(T, p.ndim), T, p.ndim);
  // Workgroups are dispatched round-robin over the 8 XCDs.  When a tile's
  // rows are shorter than a 128-byte line, the xcd_pair tiles that share
  // those lines are given to consecutive workgroups of one XCD, so its L2
  // merges the partial lines instead of each XCD fetching whole lines.
  long long tile = blockIdx.x;
  if (P.xcd_pair > 1) {
    const long long G = 8LL * P.xcd_pair;
    if (tile < (p.nblocks / G) * G) {
      long long r = tile % G;
      tile = tile - r + (r % 8) * P.xcd_pair + r / 8;
    }
  }
  tile += p.tile_base;
  const int nl = setup_lines<UIN, UOUT>(p, tile, lin, lout, T, NTH);
  // twiddle table in LDS after the line data
  size_t tw_at = (size_t)T * P.ls * sizeof(cx<R>) + LineAddr<UIN>::bytes(T, p.ndim) +
                 LineAddr<UOUT>::bytes(T, p.ndim);
  typedef typename tw_local<R, CT>::type TwL;
  typedef typename tw_global<R>::type TwG;
  TwL *twl = reinterpret_cast<TwL *>(smem_raw + ((tw_at + 15) & ~(size_t)15));
  const TwG *twg = reinterpret_cast<const TwG *>(p.twiddle);
  if (P.tw_lds)
    for (int i = threadIdx.x; i < p.N; i += NTH) {
      if constexpr (sizeof(TwL) == sizeof(TwG)) twl[i] = twg[i];
      else twl[i] = TwL{twg[i].x, twg[i].y};
    }
  __syncthreads();

  const i64 *inT = p.off_in[p.t];
  const i64 *outT = p.off_out[p.t];
  const cx<R> *in = reinterpret_cast<const cx<R> *>(p.in);
  cx<R> *out = reinterpret_cast<cx<R> *>(p.out);

  int s0 = 0, s1 = p.nstage;
  if (P.fuse_in) {
    switch (p.radix[0]) {
      case 2: if constexpr ((CAP / NTH) % 2 == 0) load_first_stage<R, INV, 2, CAP, NTH, UIN>(buf, P, lin, in, inT, nl); break;
      case 4: if constexpr ((CAP / NTH) % 4 == 0) load_first_stage<R, INV, 4, CAP, NTH, UIN>(buf, P, lin, in, inT, nl); break;
      case 8: if constexpr ((CAP / NTH) % 8 == 0) load_first_stage<R, INV, 8, CAP, NTH, UIN>(buf, P, lin, in, inT, nl); break;
      case 16: if constexpr ((CAP / NTH) % 16 == 0) load_first_stage<R, INV, 16, CAP, NTH, UIN>(buf, P, lin, in, inT, nl); break;
      default: break;
    }
    s0 = 1;
  } else {
    i64 a[V];
    int sl[V];
    bool ok[V];
#pragma unroll
    for (int k = 0; k < V; k++) {
      int e = threadIdx.x + k * NTH;
      int ec = min(e, tot - 1);
      int l, x;
      element(P, p.in_row, p.ord_in, ec, l, x);
      ok[k] = e < tot && l < nl;
      sl[k] = e < tot ? lds_index(l, x, P.ls) : -1;
      a[k] = lin.at(p, p.sin, inT, ok[k] ? l : 0, x);
    }
    cx<R> v[V];
    gather<V>(v, in, a, p.stream_in);
#pragma unroll
    for (int k = 0; k < V; k++)
      if (sl[k] >= 0) buf[sl[k]] = ok[k] ? v[k] : cx<R>{0, 0};
  }
  __syncthreads();
  if (P.fuse_out) s1 = p.nstage - 1;

  // stages and the fused store, instantiated for LDS or global twiddles
  auto run_stages = [&](auto tw) {
    // roots of the direct-sum stages: after the N stage entries of the
    // device table, p per such stage in stage order (the first stage is
    // never one when s0 = 1: the fused first stage is a power of two)
    int roff = p.N;
    for (int s = s0; s < s1; s++) {
      switch (p.radix[s]) {
        case 2: stage<R, INV, 2, CAP, NTH, CT>(buf, P, s, tw); break;
        case 3: if constexpr (!POW2) stage<R, INV, 3, CAP, NTH, CT>(buf, P, s, tw); break;
        case 4: stage<R, INV, 4, CAP, NTH, CT>(buf, P, s, tw); break;
        case 5: if constexpr (!POW2) stage<R, INV, 5, CAP, NTH, CT>(buf, P, s, tw); break;
        case 7: if constexpr (!POW2) stage<R, INV, 7, CAP, NTH, CT>(buf, P, s, tw); break;
        case 8: stage<R, INV, 8, CAP, NTH, CT>(buf, P, s, tw); break;
        case 11: if constexpr (!POW2) stage<R, INV, 11, CAP, NTH, CT>(buf, P, s, tw); break;
        case 13: if constexpr (!POW2) stage<R, INV, 13, CAP, NTH, CT>(buf, P, s, tw); break;
        case 16: stage<R, INV, 16, CAP, NTH, CT>(buf, P, s, tw); break;
        default:
          if constexpr (ANY) {
            if constexpr (sizeof(R) == 4 && !CT)
              stage_any<R, INV, CAP, NTH, CT>(buf, P, s, tw, tw_hi{reinterpret_cast<const cx<R> *>(twg)} + roff);
            else
              stage_any<R, INV, CAP, NTH, CT>(buf, P, s, tw, twg + roff);
            roff += p.radix[s];
          }
          break;
      }
    }

    if (P.fuse_out) {
      auto store = [&](auto aff) {
        constexpr bool AFF = decltype(aff)::value;
        switch (p.radix[p.nstage - 1]) {
          case 2: if constexpr ((CAP / NTH) % 2 == 0) last_stage_store<R, INV, 2, CAP, NTH, UOUT, CT, AFF>(buf, P, lout, out, outT, tw, nl); break;
          case 4: if constexpr ((CAP / NTH) % 4 == 0) last_stage_store<R, INV, 4, CAP, NTH, UOUT, CT, AFF>(buf, P, lout, out, outT, tw, nl); break;
          case 8: if constexpr ((CAP / NTH) % 8 == 0) last_stage_store<R, INV, 8, CAP, NTH, UOUT, CT, AFF>(buf, P, lout, out, outT, tw, nl); break;
          case 16: if constexpr ((CAP / NTH) % 16 == 0) last_stage_store<R, INV, 16, CAP, NTH, UOUT, CT, AFF>(buf, P, lout, out, outT, tw, nl); break;
          default: break;
        }
      };
      if constexpr (CT && !UOUT) {
        // not built: the host turns the fused store off for this form
      } else if (CT && p.sout.aff) {
        store(std::true_type{});
      } else {
        store(std::false_type{});
      }
    }
  };
  if (P.tw_lds) run_stages((const lds_tw<R, CT> *)twl);
  else if constexpr (sizeof(R) == 4 && !CT) run_stages(tw_hi{reinterpret_cast<const cx<R> *>(twg)});
  else run_stages(twg);
  if (!P.fuse_out) {
    i64 a[V];
    int sl[V];
    bool ok[V];
#pragma unroll
    for (int k = 0; k < V; k++) {
      int e = threadIdx.x + k * NTH;
      int ec = min(e, tot - 1);
      int l, x;
      element(P, p.out_row, p.ord_out, ec, l, x);
      ok[k] = e < tot && l < nl;
      sl[k] = lds_index(l, x, P.ls);
      a[k] = lout.at(p, p.sout, outT, ok[k] ? l : 0, x);
    }
    cx<R> w[V];
#pragma unroll
    for (int k = 0; k < V; k++) w[k] = buf[sl[k]];
    scatter<V, 1>(out, a, w, ok, p.stream_out);
  }
}

inline unsigned long long magic_for(int d) { return (0x100000000ULL / (unsigned long long)d) + 1ULL; }

// Big tiles: 8 lines of 512 complex doubles on 512 threads (two such
// workgroups share a CU), 16 lines of 512 complex floats on 1024 threads.
// For complex double, a 1024-thread tile of 16 lines (256-byte column rows)
// measured slower except on column walks at a 64 KiB or larger line
// stride, where it is used (launch_tiles).
constexpr int CAP_D = 4096;  // complex double points per big tile
constexpr int CAP_F = 8192;  // complex float points per big tile
constexpr int NTH_BIG = 512;
template <typename R, bool INV, int CAP, int NTH, bool POW2, bool UIN, bool UOUT, bool CT, bool ANY>
hipError_t launch_typed(const Pass &P, size_t lds, hipStream_t s) {
  auto k = fft_pass_kernel<R, INV, CAP, NTH, POW2, UIN, UOUT, CT, ANY>;
  // Once per kernel (a thread-safe static): the LDS limit, and the CU count
  // and VGPR use for the residency report.
  struct Once {
    hipError_t attr;
    int cus, vgprs;
  };
  static const Once once = [k] {
    Once o{hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024), 0, 128};
    int dev = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&o.cus, hipDeviceAttributeMultiprocessorCount, dev);
    hipFuncAttributes fa;
    if (hipFuncGetAttributes(&fa, (const void *)k) == hipSuccess && fa.numRegs > 0) o.vgprs = fa.numRegs;
    return o;
  }();
  if (once.attr != hipSuccess) return once.attr;
  const long long grid = P.p.nblocks;  // one tile per workgroup
  static const bool debug = getenv("DM_DEBUG") != nullptr;
  if (debug) {
    // Residency from LDS and VGPR use (512 VGPRs per SIMD lane, at most 8
    // waves per SIMD, 4 SIMDs per CU): the runtime's occupancy query assumes
    // 64 KiB of LDS per CU, and gfx950 has 160 KiB.
    int waves_per_wg = NTH / 64;
    int lds_fit = (int)((160 * 1024) / std::max<size_t>(lds, 1));
    int waves_per_simd = std::min(8, 512 / (((once.vgprs + 7) / 8) * 8));
    int vgpr_fit = (waves_per_simd * 4) / waves_per_wg;
    int per_cu = std::max(1, std::min(lds_fit, vgpr_fit));
    fprintf(stderr, "fft pass N=%d T=%d tiles=%lld grid=%lld per_cu=%d cus=%d lds=%zu\n", P.p.N, P.p.T,
            P.p.nblocks, grid, per_cu, once.cus, lds);
  }
  hipLaunchKernelGGL(k, dim3((unsigned)grid), dim3(NTH), lds, s, P);
  return hipGetLastError();
}

template <typename R, bool INV, int CAP, int NTH, bool POW2, bool CT, bool ANY>
hipError_t launch_layouts(const Pass &P, size_t lds, hipStream_t s) {
  bool ui = P.p.sin.uniform, uo = P.p.sout.uniform;
  if (ui && uo) return launch_typed<R, INV, CAP, NTH, POW2, true, true, CT, ANY>(P, lds, s);
  if (ui) return launch_typed<R, INV, CAP, NTH, POW2, true, false, CT, ANY>(P, lds, s);
  if (uo) return launch_typed<R, INV, CAP, NTH, POW2, false, true, CT, ANY>(P, lds, s);
  return launch_typed<R, INV, CAP, NTH, POW2, false, false, CT, ANY>(P, lds, s);
}

template <typename R, bool INV, int CAP, int NTH, bool POW2, bool ANY>
hipError_t launch_twiddles(const Pass &P, size_t lds, hipStream_t s) {
  if constexpr (sizeof(R) == 4)
    if (P.p.ctw) return launch_layouts<R, INV, CAP, NTH, POW2, true, ANY>(P, lds, s);
  return launch_layouts<R, INV, CAP, NTH, POW2, false, ANY>(P, lds, s);
}

// Power-of-two radix schedules get a kernel without the prime butterflies
// (fewer registers).  Every tile holds 8 points per thread in the load and
// store phases.  Complex double: CAP points on 512 threads, or CAP/2 on 256
// (half the LDS, twice the workgroups per CU), or the wide tile.  Complex
// float: CAP points on 1024 threads or CAP/2 on 512 (at 16 points per
// thread, 512 threads for CAP measured slower).
template <typename R, bool INV, bool POW2, bool ANY>
hipError_t launch_tiles(const Pass &P, size_t lds, hipStream_t s) {
  constexpr int CAP = sizeof(R) == 8 ? CAP_D : CAP_F;
  const bool small = P.p.T * P.p.N <= CAP / 2;
  if constexpr (sizeof(R) == 4) {
    if (small) return launch_twiddles<R, INV, CAP / 2, NTH_BIG, POW2, ANY>(P, lds, s);
    return launch_twiddles<R, INV, CAP, 2 * NTH_BIG, POW2, ANY>(P, lds, s);
  } else {
    // wide complex-double tile: 16 lines (256-byte column rows) on 1024
    // threads, for column walks at a large line stride (for complex float
    // the same 256-byte rows measured slower: 2.66 vs 2.58 ms per 512^3
    // step)
    if (P.p.T * P.p.N > CAP) return launch_twiddles<R, INV, 2 * CAP, 2 * NTH_BIG, POW2, ANY>(P, lds, s);
    if (small) return launch_twiddles<R, INV, CAP / 2, NTH_BIG / 2, POW2, ANY>(P, lds, s);
    return launch_twiddles<R, INV, CAP, NTH_BIG, POW2, ANY>(P, lds, s);
  }
}

// One translation unit per (precision, direction, radix family) holds the
// kernels of that family (fft_inst.hip, built once per combination).
hipError_t launch_d_fwd_p2(const Pass &P, size_t lds, hipStream_t s);
hipError_t launch_d_fwd_mx(const Pass &P, size_t lds, hipStream_t s);
hipError_t launch_d_inv_p2(const Pass &P, size_t lds, hipStream_t s);
hipError_t launch_d_inv_mx(const Pass &P, size_t lds, hipStream_t s);
hipError_t launch_f_fwd_p2(const Pass &P, size_t lds, hipStream_t s);
hipError_t launch_f_fwd_mx(const Pass &P, size_t lds, hipStream_t s);
hipError_t launch_f_inv_p2(const Pass &P, size_t lds, hipStream_t s);
hipError_t launch_f_inv_mx(const Pass &P, size_t lds, hipStream_t s);
hipError_t launch_d_fwd_any(const Pass &P, size_t lds, hipStream_t s);
hipError_t launch_d_inv_any(const Pass &P, size_t lds, hipStream_t s);
hipError_t launch_f_fwd_any(const Pass &P, size_t lds, hipStream_t s);
hipError_t launch_f_inv_any(const Pass &P, size_t lds, hipStream_t s);

}  // namespace fftpass
}  // namespace dm

