// K2: batched 1D FFT pass over one axis of a (virtually) distributed array.
//
// Replaces the FFTW guru stages of the split FFT (split_fft.c:426-546) and the
// twiddle pass (dimscale.c:7-49).  A workgroup loads a tile of T lines x N
// points into LDS through per-axis offset tables (so any decomposition, rank
// slot layout or factor-reversed order is read and written directly), runs
// mixed-radix Stockham stages (radix 2,3,4,5,7,8,11,13,16) in LDS, and stores
// the tile through the output tables.  The walk over memory follows the
// contiguous direction of each side: along the line (row mode) or across the
// T lines of the tile (column mode), so that global accesses coalesce.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace dm {

namespace {

constexpr int NT = 256;  // threads per workgroup

template <typename R>
struct alignas(2 * sizeof(R)) cx {
  R x, y;
};
template <typename R>
__device__ __forceinline__ cx<R> operator+(cx<R> a, cx<R> b) { return {a.x + b.x, a.y + b.y}; }
template <typename R>
__device__ __forceinline__ cx<R> operator-(cx<R> a, cx<R> b) { return {a.x - b.x, a.y - b.y}; }
template <typename R>
__device__ __forceinline__ cx<R> cmul(cx<R> a, cx<R> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// multiply by -i (forward) or +i (inverse)
template <typename R, bool INV>
__device__ __forceinline__ cx<R> rot(cx<R> a) {
  return INV ? cx<R>{-a.y, a.x} : cx<R>{a.y, -a.x};
}

// ---- butterflies: v[k] <- sum_n v[n] exp(-+2 pi i n k / RADIX) --------------

template <typename R, bool INV>
__device__ __forceinline__ void dft2(cx<R> *v) {
  cx<R> a = v[0], b = v[1];
  v[0] = a + b;
  v[1] = a - b;
}

template <typename R, bool INV>
__device__ __forceinline__ void dft4(cx<R> &a0, cx<R> &a1, cx<R> &a2, cx<R> &a3) {
  cx<R> t0 = a0 + a2, t1 = a0 - a2, t2 = a1 + a3, t3 = rot<R, INV>(a1 - a3);
  a0 = t0 + t2;
  a2 = t0 - t2;
  a1 = t1 + t3;
  a3 = t1 - t3;
}

template <typename R, bool INV>
__device__ __forceinline__ void dft8(cx<R> *v) {
  const R c = (R)0.7071067811865476;
  cx<R> e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
  cx<R> o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
  dft4<R, INV>(e0, e1, e2, e3);
  dft4<R, INV>(o0, o1, o2, o3);
  // o_k *= w8^k
  cx<R> w1, w3;
  if (!INV) {
    w1 = {(o1.x + o1.y) * c, (o1.y - o1.x) * c};
    w3 = {(o3.y - o3.x) * c, -(o3.x + o3.y) * c};
  } else {
    w1 = {(o1.x - o1.y) * c, (o1.x + o1.y) * c};
    w3 = {-(o3.x + o3.y) * c, (o3.x - o3.y) * c};
  }
  cx<R> w2 = rot<R, INV>(o2);
  v[0] = e0 + o0;
  v[4] = e0 - o0;
  v[1] = e1 + w1;
  v[5] = e1 - w1;
  v[2] = e2 + w2;
  v[6] = e2 - w2;
  v[3] = e3 + w3;
  v[7] = e3 - w3;
}

template <typename R, bool INV>
__device__ __forceinline__ void dft16(cx<R> *v) {
  // cos(pi k / 8), sin(pi k / 8), k = 0..7
  const R c16[8] = {(R)1.0, (R)0.9238795325112867, (R)0.7071067811865476, (R)0.3826834323650898,
                    (R)0.0, (R)-0.3826834323650898, (R)-0.7071067811865476, (R)-0.9238795325112867};
  const R s16[8] = {(R)0.0, (R)0.3826834323650898, (R)0.7071067811865476, (R)0.9238795325112867,
                    (R)1.0, (R)0.9238795325112867, (R)0.7071067811865476, (R)0.3826834323650898};
  cx<R> e[8], o[8];
#pragma unroll
  for (int k = 0; k < 8; k++) {
    e[k] = v[2 * k];
    o[k] = v[2 * k + 1];
  }
  dft8<R, INV>(e);
  dft8<R, INV>(o);
#pragma unroll
  for (int k = 0; k < 8; k++) {
    cx<R> w = {c16[k], INV ? s16[k] : -s16[k]};
    cx<R> t = cmul(o[k], w);
    v[k] = e[k] + t;
    v[k + 8] = e[k] - t;
  }
}

template <int P> struct Roots;
template <> struct Roots<3> {
  static constexpr double c[3] = {1.0, -0.5, -0.5};
  static constexpr double s[3] = {0.0, 0.8660254037844386, -0.8660254037844386};
};
template <> struct Roots<5> {
  static constexpr double c[5] = {1.0, 0.30901699437494745, -0.8090169943749475, -0.8090169943749475,
                                  0.30901699437494745};
  static constexpr double s[5] = {0.0, 0.9510565162951535, 0.5877852522924731, -0.5877852522924731,
                                  -0.9510565162951535};
};
template <> struct Roots<7> {
  static constexpr double c[7] = {1.0, 0.6234898018587335, -0.2225209339563144, -0.9009688679024191,
                                  -0.9009688679024191, -0.2225209339563144, 0.6234898018587335};
  static constexpr double s[7] = {0.0, 0.7818314824680298, 0.9749279121818236, 0.4338837391175581,
                                  -0.4338837391175581, -0.9749279121818236, -0.7818314824680298};
};
template <> struct Roots<11> {
  static constexpr double c[11] = {1.0, 0.8412535328311812, 0.41541501300188644, -0.14231483827328514,
                                   -0.6548607339452851, -0.9594929736144974, -0.9594929736144974,
                                   -0.6548607339452851, -0.14231483827328514, 0.41541501300188644,
                                   0.8412535328311812};
  static constexpr double s[11] = {0.0, 0.5406408174555976, 0.9096319953545183, 0.9898214418809327,
                                   0.7557495743542583, 0.28173255684142967, -0.28173255684142967,
                                   -0.7557495743542583, -0.9898214418809327, -0.9096319953545183,
                                   -0.5406408174555976};
};
template <> struct Roots<13> {
  static constexpr double c[13] = {1.0, 0.8854560256532099, 0.5680647467311558, 0.12053668025532305,
                                   -0.3546048870425356, -0.7485107481711011, -0.970941817426052,
                                   -0.970941817426052, -0.7485107481711011, -0.3546048870425356,
                                   0.12053668025532305, 0.5680647467311558, 0.8854560256532099};
  static constexpr double s[13] = {0.0, 0.46472317204376856, 0.8229838658936564, 0.992708874098054,
                                   0.9350162426854148, 0.6631226582407952, 0.23931566428755777,
                                   -0.23931566428755777, -0.6631226582407952, -0.9350162426854148,
                                   -0.992708874098054, -0.8229838658936564, -0.46472317204376856};
};

// Odd prime P: pair n with P-n so each output needs (P-1)/2 real
// multiply-adds per part.
template <typename R, bool INV, int P>
__device__ __forceinline__ void dft_prime(cx<R> *v) {
  constexpr int H = (P - 1) / 2;
  cx<R> sum[H + 1], dif[H + 1];
#pragma unroll
  for (int n = 1; n <= H; n++) {
    sum[n] = v[n] + v[P - n];
    dif[n] = v[n] - v[P - n];
  }
  cx<R> a0 = v[0];
  cx<R> x0 = a0;
#pragma unroll
  for (int n = 1; n <= H; n++) x0 = x0 + sum[n];
#pragma unroll
  for (int k = 1; k <= H; k++) {
    cx<R> re = a0;  // a0 + sum_n cos * s_n
    R imx = 0, imy = 0;  // sum_n sin * d_n
#pragma unroll
    for (int n = 1; n <= H; n++) {
      int m = (n * k) % P;
      R c = (R)Roots<P>::c[m], s = (R)Roots<P>::s[m];
      re.x += c * sum[n].x;
      re.y += c * sum[n].y;
      imx += s * dif[n].x;
      imy += s * dif[n].y;
    }
    // forward: X[k] = re - i * (imx + i imy) = (re.x + imy, re.y - imx)
    if (!INV) {
      v[k] = {re.x + imy, re.y - imx};
      v[P - k] = {re.x - imy, re.y + imx};
    } else {
      v[k] = {re.x - imy, re.y + imx};
      v[P - k] = {re.x + imy, re.y - imx};
    }
  }
  v[0] = x0;
}

template <typename R, bool INV, int RADIX>
__device__ __forceinline__ void dft(cx<R> *v) {
  if constexpr (RADIX == 2) dft2<R, INV>(v);
  else if constexpr (RADIX == 4) dft4<R, INV>(v[0], v[1], v[2], v[3]);
  else if constexpr (RADIX == 8) dft8<R, INV>(v);
  else if constexpr (RADIX == 16) dft16<R, INV>(v);
  else dft_prime<R, INV, RADIX>(v);
}

// ---- fast division by runtime constants (numerators below 2^16) -------------

__device__ __forceinline__ int fdiv(int n, unsigned long long magic) {
  return (int)(((unsigned long long)(unsigned)n * magic) >> 32);
}

struct StageInfo {
  unsigned long long m_nr;   // magic for N / radix
  unsigned long long m_ns;   // magic for Ns
  int nr, ns, twstride;
};

struct Pass {
  FftLaunch p;
  StageInfo st[16];
  unsigned long long m_n, m_t;  // magics for N and T
  int ls;                        // LDS line stride (complex elements)
};

__device__ __forceinline__ int lds_index(int line, int x, int ls) { return line * ls + x + (x >> 4); }

template <typename R, bool INV, int RADIX, int CAP>
__device__ __forceinline__ void stage(cx<R> *buf, const Pass &P, int s, const cx<R> *tw) {
  constexpr int K = (CAP / RADIX + NT - 1) / NT;
  const StageInfo &si = P.st[s];
  const int nr = si.nr, ns = si.ns;
  const int nb = P.p.T * nr;
  cx<R> v[K][RADIX];
  int dst[K];
#pragma unroll
  for (int k = 0; k < K; k++) {
    int b = threadIdx.x + k * NT;
    dst[k] = -1;
    if (b < nb) {
      int l = fdiv(b, si.m_nr);
      int j = b - l * nr;
      int jn = fdiv(j, si.m_ns);
      int jj = j - jn * ns;
      int base = l * P.ls;
#pragma unroll
      for (int r = 0; r < RADIX; r++) {
        int x = j + r * nr;
        v[k][r] = buf[base + x + (x >> 4)];
      }
      if (ns > 1) {
#pragma unroll
        for (int r = 1; r < RADIX; r++) {
          cx<R> w = tw[jj * r * si.twstride];
          if (INV) w.y = -w.y;
          v[k][r] = cmul(v[k][r], w);
        }
      }
      dft<R, INV, RADIX>(v[k]);
      // line and Stockham-expanded index of the first output
      int idxd = jn * ns * RADIX + jj;
      dst[k] = (l << 16) | idxd;
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; k++) {
    if (dst[k] >= 0) {
      int l = dst[k] >> 16, idxd = dst[k] & 0xffff;
      int base = l * P.ls;
#pragma unroll
      for (int r = 0; r < RADIX; r++) {
        int x = idxd + r * ns;
        buf[base + x + (x >> 4)] = v[k][r];
      }
    }
  }
  __syncthreads();
}

// Per-line address data of one side of a tile (in LDS).
//   uniform:     base[l] = sum over axes other than t of abs[d][g_d]
//   non-uniform: rbase[l] = owner contribution of those axes, offs[l][k] =
//                their local offsets (k over the axes other than t)
template <bool UNIFORM>
struct LineAddr;

template <>
struct LineAddr<true> {
  i64 *base;
  __host__ __device__ static size_t bytes(int T, int) { return (size_t)T * sizeof(i64); }
  __device__ void init(char *mem, int, int) { base = reinterpret_cast<i64 *>(mem); }
  __device__ void set_line(const FftLaunch &p, const i64 *const *abs, const int *g, int l) {
    i64 a = 0;
    for (int d = 0; d < p.ndim; d++)
      if (d != p.t) a += abs[d][g[d]];
    base[l] = a;
  }
  __device__ __forceinline__ i64 at(const FftLaunch &, const FftSide &, const i64 *absT, int l, int x) const {
    return base[l] + absT[x];
  }
};

template <>
struct LineAddr<false> {
  int *rbase;
  int *offs;
  int nd;
  __host__ __device__ static size_t bytes(int T, int ndim) { return (size_t)T * sizeof(int) * (1 + ndim); }
  __device__ void init(char *mem, int T, int ndim) {
    rbase = reinterpret_cast<int *>(mem);
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

template <typename R, bool INV, int CAP, bool UIN, bool UOUT>
__global__ __launch_bounds__(NT) void fft_pass_kernel(Pass P) {
  extern __shared__ char smem_raw[];
  cx<R> *buf = reinterpret_cast<cx<R> *>(smem_raw);
  const FftLaunch &p = P.p;
  const int N = p.N, T = p.T;

  // block -> (tile along u, batch coordinates)
  long long bid = blockIdx.x;
  int tu = (int)(bid % p.tiles_u);
  bid /= p.tiles_u;
  int gb[DM_MAX_DIMS];
  for (int k = p.nbatch - 1; k >= 0; k--) {
    int d = p.batch_axis[k];
    gb[d] = (int)(bid % p.L[d]);
    bid /= p.L[d];
  }
  const int u0 = tu * T;
  const int nlines = p.u >= 0 ? min(T, p.L[p.u] - u0) : 1;

  char *lmem = smem_raw + (size_t)T * P.ls * sizeof(cx<R>);
  LineAddr<UIN> lin;
  LineAddr<UOUT> lout;
  lin.init(lmem, T, p.ndim);
  lout.init(lmem + LineAddr<UIN>::bytes(T, p.ndim), T, p.ndim);
  for (int l = threadIdx.x; l < T; l += NT) {
    int g[DM_MAX_DIMS];
    for (int d = 0; d < p.ndim; d++) g[d] = (d == p.u || d == p.t) ? 0 : gb[d];
    if (p.u >= 0) g[p.u] = (l < nlines) ? (p.ord_u ? p.ord_u[u0 + l] : u0 + l) : 0;
    if constexpr (UIN) lin.set_line(p, p.off_in, g, l);
    else lin.set_line(p, p.sin, g, l);
    if constexpr (UOUT) lout.set_line(p, p.off_out, g, l);
    else lout.set_line(p, p.sout, g, l);
  }
  __syncthreads();

  const i64 *inT = p.off_in[p.t];
  const i64 *outT = p.off_out[p.t];
  const cx<R> *in = reinterpret_cast<const cx<R> *>(p.in);
  cx<R> *out = reinterpret_cast<cx<R> *>(p.out);
  const cx<R> *tw = reinterpret_cast<const cx<R> *>(p.twiddle);
  const int tot = T * N;

  for (int e = threadIdx.x; e < tot; e += NT) {
    int l, x;
    if (p.in_row) {
      l = fdiv(e, P.m_n);
      int i = e - l * N;
      x = p.ord_in ? p.ord_in[i] : i;
    } else {
      x = fdiv(e, P.m_t);
      l = e - x * T;
    }
    cx<R> v = {0, 0};
    if (l < nlines) v = in[lin.at(p, p.sin, inT, l, x)];
    buf[lds_index(l, x, P.ls)] = v;
  }
  __syncthreads();

  for (int s = 0; s < p.nstage; s++) {
    switch (p.radix[s]) {
      case 2: stage<R, INV, 2, CAP>(buf, P, s, tw); break;
      case 3: stage<R, INV, 3, CAP>(buf, P, s, tw); break;
      case 4: stage<R, INV, 4, CAP>(buf, P, s, tw); break;
      case 5: stage<R, INV, 5, CAP>(buf, P, s, tw); break;
      case 7: stage<R, INV, 7, CAP>(buf, P, s, tw); break;
      case 8: stage<R, INV, 8, CAP>(buf, P, s, tw); break;
      case 11: stage<R, INV, 11, CAP>(buf, P, s, tw); break;
      case 13: stage<R, INV, 13, CAP>(buf, P, s, tw); break;
      case 16: stage<R, INV, 16, CAP>(buf, P, s, tw); break;
      default: break;
    }
  }

  for (int e = threadIdx.x; e < tot; e += NT) {
    int l, x;
    if (p.out_row) {
      l = fdiv(e, P.m_n);
      int i = e - l * N;
      x = p.ord_out ? p.ord_out[i] : i;
    } else {
      x = fdiv(e, P.m_t);
      l = e - x * T;
    }
    if (l < nlines) out[lout.at(p, p.sout, outT, l, x)] = buf[lds_index(l, x, P.ls)];
  }
}

unsigned long long magic_for(int d) { return (0x100000000ULL / (unsigned long long)d) + 1ULL; }

constexpr int CAP_D = 4096;  // complex double points per tile
constexpr int CAP_F = 8192;  // complex float points per tile

template <typename R, bool INV, int CAP, bool UIN, bool UOUT>
void launch_typed(const Pass &P, size_t lds, hipStream_t s) {
  static bool attr_set = false;
  if (!attr_set) {
    (void)hipFuncSetAttribute((const void *)fft_pass_kernel<R, INV, CAP, UIN, UOUT>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    attr_set = true;
  }
  hipLaunchKernelGGL((fft_pass_kernel<R, INV, CAP, UIN, UOUT>), dim3((unsigned)P.p.nblocks), dim3(NT), lds, s, P);
}

template <typename R, bool INV, int CAP>
void launch_layouts(const Pass &P, size_t lds, hipStream_t s) {
  bool ui = P.p.sin.uniform, uo = P.p.sout.uniform;
  if (ui && uo) launch_typed<R, INV, CAP, true, true>(P, lds, s);
  else if (ui) launch_typed<R, INV, CAP, true, false>(P, lds, s);
  else if (uo) launch_typed<R, INV, CAP, false, true>(P, lds, s);
  else launch_typed<R, INV, CAP, false, false>(P, lds, s);
}

}  // namespace

int fft_tile_capacity(int dbl) { return dbl ? CAP_D : CAP_F; }

size_t fft_lds_bytes(const FftLaunch &p) {
  size_t elem = p.dbl ? 16 : 8;
  size_t ls = (size_t)p.N + (p.N >> 4);
  auto side = [&](int uniform) { return uniform ? (size_t)p.T * 8 : (size_t)p.T * 4 * (1 + p.ndim); };
  return (size_t)p.T * ls * elem + side(p.sin.uniform) + side(p.sout.uniform);
}

void launch_fft_pass(const FftLaunch &p, void *stream) {
  Pass P;
  P.p = p;
  P.m_n = magic_for(p.N);
  P.m_t = magic_for(p.T);
  P.ls = p.N + (p.N >> 4);
  int ns = 1;
  for (int s = 0; s < p.nstage; s++) {
    int r = p.radix[s];
    P.st[s].nr = p.N / r;
    P.st[s].ns = ns;
    P.st[s].m_nr = magic_for(p.N / r);
    P.st[s].m_ns = magic_for(ns);
    P.st[s].twstride = p.N / (ns * r);
    ns *= r;
  }
  long long batch = 1;
  for (int k = 0; k < p.nbatch; k++) batch *= p.L[p.batch_axis[k]];
  P.p.nblocks = (long long)p.tiles_u * batch;
  if (P.p.nblocks <= 0) return;
  size_t lds = fft_lds_bytes(p);
  hipStream_t s = (hipStream_t)stream;
  if (p.dbl) {
    if (p.inverse) launch_layouts<double, true, CAP_D>(P, lds, s);
    else launch_layouts<double, false, CAP_D>(P, lds, s);
  } else {
    if (p.inverse) launch_layouts<float, true, CAP_F>(P, lds, s);
    else launch_layouts<float, false, CAP_F>(P, lds, s);
  }
}

}  // namespace dm
