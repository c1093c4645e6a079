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

constexpr int kThreads = 256;  // most threads per workgroup

// A workgroup takes one chunk of one row (the last axis), or, for rows
// shorter than the workgroup, several whole rows.  The product of a row's
// outer factors comes from the plan's row table (formed on the host, outer
// axis first, as the reference's recursion forms it), so an element takes
// two table loads -- its row's and its last-axis factor -- and two complex
// products.
// IDX: the integer type of a rank's element and chunk indices (32-bit
// unless a rank holds 2^32 elements or more).
template <typename R, bool CONJ, typename IDX>
__global__ __launch_bounds__(kThreads) void scale_kernel(ScaleLaunch p) {
  const int r = blockIdx.y;
  const int nd = p.ndim;
  const int *rt = p.rank_tab + (size_t)r * nd * 2;
  const i64 *st = p.strides + (size_t)r * nd;
  const int inner = rt[2 * (nd - 1) + 1];
  if (inner <= 0) return;  // an empty block
  IDX rows = 1;
  for (int d = 0; d < nd - 1; d++) rows *= (IDX)rt[2 * d + 1];
  const int bd = (int)blockDim.x, tid = (int)threadIdx.x;
  const IDX bx = (IDX)blockIdx.x;
  IDX row;
  int col;
  if (inner >= bd) {  // chunks of one row
    const IDX cpr = ((IDX)inner + (IDX)bd - 1) / (IDX)bd;
    if (bx >= rows * cpr) return;
    row = bx / cpr;
    col = (int)(bx - row * cpr) * bd + tid;
    if (col >= inner) return;
  } else {  // bd / inner whole rows
    const int rpb = bd / inner, sub = tid / inner;
    col = tid - sub * inner;
    row = bx * (IDX)rpb + (IDX)sub;
    if (sub >= rpb || row >= rows) return;
  }
  // the row's outer coordinates
  const IDX row0 = row;
  i64 off = (i64)col * st[nd - 1];
  for (int d = nd - 2; d >= 0; d--) {
    const IDX c = (IDX)rt[2 * d + 1];
    const IDX q = row / c;
    off += (i64)(row - q * c) * st[d];
    row = q;
  }
  // the element's load goes out before the factors' loads return (every
  // slot of the box is inside the rank's slot; a zero product only skips
  // the store), so its latency overlaps theirs
  typedef R v2 __attribute__((ext_vector_type(2)));
  v2 *data = reinterpret_cast<v2 *>(p.data) + (i64)r * p.slot + off;
  const v2 x = __builtin_nontemporal_load(data);
  // the outer axes' product from the plan's row table, then the last axis
  const double *rp = p.rows + 2 * ((size_t)p.row_off[r] + (size_t)row0);
  double fr = rp[0], fi = rp[1];
  {
    const double *t = p.tw[nd - 1] + 2 * (size_t)(rt[2 * (nd - 1)] + col);
    const double tr = t[0], ti = t[1];
    const double nr = fr * tr - fi * ti;
    fi = fr * ti + fi * tr;
    fr = nr;
  }
  if (fr == 0.0 && fi == 0.0) return;
  if (CONJ) fi = -fi;
  // each element is read once and written once: non-temporal accesses
  const double xr = (double)x.x, xi = (double)x.y;
  __builtin_nontemporal_store(v2{(R)(fr * xr - fi * xi), (R)(fr * xi + fi * xr)}, data);
}

}  // namespace

int launch_scale(const ScaleLaunch &p, void *stream) {
  if (p.nranks <= 0 || p.max_elems <= 0 || p.max_blocks <= 0) return 0;
  if (p.block < 64 || p.block > kThreads || p.block % 64 != 0 || p.max_blocks >= (1LL << 31))
    return (int)hipErrorInvalidValue;
  const dim3 grid((unsigned)p.max_blocks, (unsigned)p.nranks), block((unsigned)p.block);
  hipStream_t s = (hipStream_t)stream;
  if (p.max_elems + p.block >= (1LL << 32)) {
    if (p.dbl) {
      if (p.conj) hipLaunchKernelGGL((scale_kernel<double, true, unsigned long long>), grid, block, 0, s, p);
      else hipLaunchKernelGGL((scale_kernel<double, false, unsigned long long>), grid, block, 0, s, p);
    } else {
      if (p.conj) hipLaunchKernelGGL((scale_kernel<float, true, unsigned long long>), grid, block, 0, s, p);
      else hipLaunchKernelGGL((scale_kernel<float, false, unsigned long long>), grid, block, 0, s, p);
    }
  } else if (p.dbl) {
    if (p.conj) hipLaunchKernelGGL((scale_kernel<double, true, unsigned>), grid, block, 0, s, p);
    else hipLaunchKernelGGL((scale_kernel<double, false, unsigned>), grid, block, 0, s, p);
  } else {
    if (p.conj) hipLaunchKernelGGL((scale_kernel<float, true, unsigned>), grid, block, 0, s, p);
    else hipLaunchKernelGGL((scale_kernel<float, false, unsigned>), grid, block, 0, s, p);
  }
  return (int)hipGetLastError();
}

}  // namespace dm
