// Device building blocks shared by the FFT pass kernels: complex type,
// radix-2/3/4/5/7/8/11/13/16 butterflies, magic-number division, the padded
// LDS tile layout and one Stockham stage over an LDS tile.
#pragma once
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace dm {
namespace fftdev {

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
// Twiddle table entries in device memory, 16 bytes per root.  Complex
// double: the root itself.  Complex float: a compensated pair, the root w
// (formed in long double, rounded to double) as hi = fl32(w) and lo =
// w - hi rounded to bfloat16 (both parts of lo packed in one word, real
// part in the low half), so the entry holds w to about 2^-33 where a float
// alone holds it to 2^-25.  (hi.re, hi.im, lo) is one vector: a single
// 12-byte load.
typedef float twf __attribute__((ext_vector_type(3)));
template <typename R>
struct tw_global {
  typedef cx<R> type;
};
template <>
struct tw_global<float> {
  typedef twf type;
};
// Entry of a table copied to LDS: the whole entry for compensated products
// (CT), the root's (re, im) otherwise.
template <typename R, bool CT>
struct tw_local {
  typedef cx<R> type;
};
template <>
struct tw_local<float, true> {
  typedef twf type;
};

// The float parts (hi) of a compensated table in device memory, read as
// 8-byte complex floats at a 16-byte stride, for products without the
// compensation.
struct tw_hi {
  const cx<float> *p;
  __device__ __forceinline__ cx<float> operator[](int i) const { return p[2 * i]; }
  __device__ __forceinline__ tw_hi operator+(int k) const { return {p + 2 * k}; }
};

// a * w for table entry tw[i] (the conjugate root for the inverse).  CT
// (complex float only): compensated product, a * lo formed first and added
// inside the fused multiply-adds of a * hi, so that the result carries the
// rounding of the float arithmetic but not the rounding of w to float.
// Without CT only the entry's (re, im) is read.
template <bool INV, bool CT, typename R, typename TW>
__device__ __forceinline__ cx<R> twmul(cx<R> a, TW tw, int i) {
  if constexpr (CT) {
    static_assert(sizeof(R) == 4, "compensated twiddles are for complex float");
    const twf e = tw[i];
    const unsigned lo = __float_as_uint(e.z);
    const float lx = __uint_as_float(lo << 16), ly0 = __uint_as_float(lo & 0xffff0000u);
    const float hx = e.x, hy = INV ? -e.y : e.y, ly = INV ? -ly0 : ly0;
    const float cr = a.x * lx - a.y * ly;
    const float ci = a.x * ly + a.y * lx;
    return {__builtin_fmaf(a.x, hx, __builtin_fmaf(-a.y, hy, cr)), __builtin_fmaf(a.x, hy, __builtin_fmaf(a.y, hx, ci))};
  } else {
    const R wx = tw[i].x, wy = INV ? -tw[i].y : tw[i].y;
    return {a.x * wx - a.y * wy, a.x * wy + a.y * wx};
  }
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


// LDS placement of point x of tile line `line`: one pad element every 16
// points and an odd line stride in elements, so a column walk over
// consecutive lines and strided Stockham writes spread over the banks.
__host__ __device__ __forceinline__ int lds_pad(int x) { return x + (x >> 4); }
__host__ __device__ __forceinline__ int lds_line(int n) { return n + (n >> 4) + 1; }
__device__ __forceinline__ int lds_index(int line, int x, int ls) { return line * ls + lds_pad(x); }

// One Stockham stage over the T lines of an LDS tile (line stride ls):
// butterflies read points j + r * nr of their line, get the stage twiddles
// (stage-ordered table), transform, and write back to their Stockham places.
// TW: pointer to the twiddles, typed by where they live (an LDS-qualified
// pointer reads with LDS instructions; a generic one would use flat loads);
// CT: compensated twiddle products (complex float).
// COL: the lines are the columns of the LDS tile (point x of line l at
// x * ls + pad(l): a plane's second axis), consecutive butterflies on
// consecutive lines; m_t: magic for T.
template <typename R, bool INV, int RADIX, int CAP, int NTH, bool CT, typename TW, bool COL = false>
__device__ __forceinline__ void stage_lds(cx<R> *buf, const StageInfo &si, int T, int ls, TW tw,
                                          unsigned long long m_t = 0) {
  constexpr int K = (CAP / RADIX + NTH - 1) / NTH;
  const int nr = si.nr, ns = si.ns;
  const int nb = T * nr;
  auto at = [&](int l, int x) { return COL ? x * ls + lds_pad(l) : l * ls + lds_pad(x); };
  cx<R> v[K][RADIX];
  int dst[K];
#pragma unroll
  for (int k = 0; k < K; k++) {
    int b = threadIdx.x + k * NTH;
    dst[k] = -1;
    if (b < nb) {
      int l, j;
      if constexpr (COL) {
        j = fdiv(b, m_t);
        l = b - j * T;
      } else {
        l = fdiv(b, si.m_nr);
        j = b - l * nr;
      }
      int jn = fdiv(j, si.m_ns);
      int jj = j - jn * ns;
#pragma unroll
      for (int r = 0; r < RADIX; r++) v[k][r] = buf[at(l, j + r * nr)];
      if (ns > 1) {
        // stage-ordered table: consecutive jj read consecutive entries
        TW tws = tw + (ns - 1) + jj;
#pragma unroll
        for (int r = 1; r < RADIX; r++) v[k][r] = twmul<INV, CT>(v[k][r], tws, (r - 1) * ns);
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
#pragma unroll
      for (int r = 0; r < RADIX; r++) buf[at(l, idxd + r * ns)] = v[k][r];
    }
  }
  __syncthreads();
}

// One Stockham stage of a prime radix p without a kernel of its own (a
// runtime value above 13), as direct sums: each work item is one output q of
// one butterfly, y_q = sum_r (x_r * stage twiddle_r) * w_p^(r q mod p),
// reading the butterfly's p inputs from LDS (lanes of a wave share
// butterflies, so most reads are broadcasts).  O(p) per point; outputs wait
// in registers until every item of the tile has read its inputs.  roots:
// w_p^m for m < p, in the table's twiddle form.  The p products are summed in
// double in either precision, so that a complex-float sum of up to 8191
// terms keeps the error of one product (a float sum would grow with p).
template <typename R, bool INV, int CAP, int NTH, bool CT, typename TW, typename RT>
__device__ __forceinline__ void stage_lds_any(cx<R> *buf, const StageInfo &si, int p, int T, int ls, TW tw,
                                              RT roots) {
  constexpr int K = (CAP + NTH - 1) / NTH;
  const int nr = si.nr, ns = si.ns;
  const int items = T * nr * p;  // T * N <= CAP
  const unsigned long long m_p = 0x100000000ull / (unsigned)p + 1;
  cx<R> y[K];
  int dst[K];
#pragma unroll
  for (int k = 0; k < K; k++) {
    int e = threadIdx.x + k * NTH;
    dst[k] = -1;
    if (e < items) {
      int b = fdiv(e, m_p);
      int q = e - b * p;
      int l = fdiv(b, si.m_nr);
      int j = b - l * nr;
      int jn = fdiv(j, si.m_ns);
      int jj = j - jn * ns;
      int base = l * ls;
      TW tws = tw + (ns - 1) + jj;
      const cx<R> x0 = buf[base + lds_pad(j)];
      double ax = x0.x, ay = x0.y;
      int m = 0;
      for (int r = 1; r < p; r++) {
        cx<R> v = buf[base + lds_pad(j + r * nr)];
        if (ns > 1) v = twmul<INV, CT>(v, tws, (r - 1) * ns);
        m += q;
        if (m >= p) m -= p;
        const cx<R> t = twmul<INV, CT>(v, roots, m);
        ax += (double)t.x;
        ay += (double)t.y;
      }
      y[k] = cx<R>{(R)ax, (R)ay};
      dst[k] = (l << 16) | (jn * ns * p + jj + q * ns);
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; k++)
    if (dst[k] >= 0) buf[(dst[k] >> 16) * ls + lds_pad(dst[k] & 0xffff)] = y[k];
  __syncthreads();
}

// Global data accesses of a pass: N elements at offsets a[k] from base,
// optionally with the non-temporal hint (data streamed once per launch).
// The hint is uniform across the launch, so the branch is taken once.
template <int N, typename R, typename P>
__device__ __forceinline__ void gather(cx<R> *v, P src, const i64 *a, bool nt) {
  typedef R v2 __attribute__((ext_vector_type(2)));
  if (nt) {
#pragma unroll
    for (int k = 0; k < N; k++) {
      const v2 t = __builtin_nontemporal_load(reinterpret_cast<const v2 *>(&src[a[k]]));
      v[k] = {t.x, t.y};
    }
  } else {
#pragma unroll
    for (int k = 0; k < N; k++) v[k] = src[a[k]];
  }
}

// Stores element k when ok[k / PER] holds.
template <int N, int PER, typename R, typename P>
__device__ __forceinline__ void scatter(P dst, const i64 *a, const cx<R> *v, const bool *ok, bool nt) {
  typedef R v2 __attribute__((ext_vector_type(2)));
  if (nt) {
#pragma unroll
    for (int k = 0; k < N; k++)
      if (ok[k / PER]) __builtin_nontemporal_store(v2{v[k].x, v[k].y}, reinterpret_cast<v2 *>(&dst[a[k]]));
  } else {
#pragma unroll
    for (int k = 0; k < N; k++)
      if (ok[k / PER]) dst[a[k]] = v[k];
  }
}

}  // namespace fftdev
}  // namespace dm
