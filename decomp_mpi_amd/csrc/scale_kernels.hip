// K4: multiply every owned slot of every rank's local block by the product
// of per-axis factors, the device form of decomp_scale (dimscale.c:7-67).
//
// Rank r's local block is the box of count[r][d] positions per axis at
// memory strides stride[r][d]; the factor of axis d at local position p is
// table_d[first[r][d] + p] (the rows of rank r's position along that axis).
// The product is formed outer axis to inner in double precision, as the
// reference's recursion forms it; a zero product marks a slot that holds
// no coordinate and leaves it untouched.  The backward transform multiplies
// by the conjugate.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace dm {
namespace {

constexpr int kThreads = 256;
constexpr int kPerThread = 1;  // 2 and 4 measured slower (2.61 and 2.84 against 2.10 ms per 512^3 step)

// IDX: the integer type of a rank's element index (32-bit unless a rank
// holds 2^32 elements or more: a 64-bit division costs several times more).
template <typename R, bool CONJ, typename IDX>
__global__ __launch_bounds__(kThreads) void scale_kernel(ScaleLaunch p) {
  const int r = blockIdx.y;
  const int nd = p.ndim;
  const int *rt = p.rank_tab + (size_t)r * nd * 2;
  const i64 *st = p.strides + (size_t)r * nd;
  IDX total = 1;
  for (int d = 0; d < nd; d++) total *= (IDX)rt[2 * d + 1];
  typedef R v2 __attribute__((ext_vector_type(2)));
  v2 *data = reinterpret_cast<v2 *>(p.data) + (i64)r * p.slot;
  const IDX base = (IDX)blockIdx.x * kThreads * kPerThread + threadIdx.x;
#pragma unroll
  for (int k = 0; k < kPerThread; k++) {
    IDX e = base + (IDX)k * kThreads;
    if (e >= total) break;
    // local coordinates, last axis fastest
    i64 off = 0;
    int loc[DM_MAX_DIMS];
    for (int d = nd - 1; d >= 0; d--) {
      const IDX c = (IDX)rt[2 * d + 1];
      const IDX q = e / c;
      loc[d] = (int)(e - q * c);
      e = q;
      off += (i64)loc[d] * st[d];
    }
    double fr = 1.0, fi = 0.0;
    for (int d = 0; d < nd; d++) {
      const double *t = p.tw[d] + 2 * (size_t)(rt[2 * d] + loc[d]);
      const double tr = t[0], ti = t[1];
      const double nr = fr * tr - fi * ti;
      fi = fr * ti + fi * tr;
      fr = nr;
    }
    if (fr == 0.0 && fi == 0.0) continue;
    if (CONJ) fi = -fi;
    // each element is read once and written once: non-temporal accesses
    const v2 x = __builtin_nontemporal_load(data + off);
    const double xr = (double)x.x, xi = (double)x.y;
    __builtin_nontemporal_store(v2{(R)(fr * xr - fi * xi), (R)(fr * xi + fi * xr)}, data + off);
  }
}

}  // namespace

int launch_scale(const ScaleLaunch &p, void *stream) {
  if (p.nranks <= 0 || p.max_elems <= 0) return 0;
  const long long per_block = (long long)kThreads * kPerThread;
  dim3 grid((unsigned)((p.max_elems + per_block - 1) / per_block), (unsigned)p.nranks);
  hipStream_t s = (hipStream_t)stream;
  if (p.max_elems + per_block >= (1LL << 32)) {
    if (p.dbl) {
      if (p.conj) hipLaunchKernelGGL((scale_kernel<double, true, unsigned long long>), grid, dim3(kThreads), 0, s, p);
      else hipLaunchKernelGGL((scale_kernel<double, false, unsigned long long>), grid, dim3(kThreads), 0, s, p);
    } else {
      if (p.conj) hipLaunchKernelGGL((scale_kernel<float, true, unsigned long long>), grid, dim3(kThreads), 0, s, p);
      else hipLaunchKernelGGL((scale_kernel<float, false, unsigned long long>), grid, dim3(kThreads), 0, s, p);
    }
  } else if (p.dbl) {
    if (p.conj) hipLaunchKernelGGL((scale_kernel<double, true, unsigned>), grid, dim3(kThreads), 0, s, p);
    else hipLaunchKernelGGL((scale_kernel<double, false, unsigned>), grid, dim3(kThreads), 0, s, p);
  } else {
    if (p.conj) hipLaunchKernelGGL((scale_kernel<float, true, unsigned>), grid, dim3(kThreads), 0, s, p);
    else hipLaunchKernelGGL((scale_kernel<float, false, unsigned>), grid, dim3(kThreads), 0, s, p);
  }
  return (int)hipGetLastError();
}

}  // namespace dm
