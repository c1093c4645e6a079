// Building FFT pass programs from layouts (shared by the split plan, the
// per-axis FFT and the batched transforms).
#pragma once
#include <string>
#include <vector>

#include "internal.h"
#include "kernels.h"

namespace dm {

// One side (input or output) of a pass: a layout bound to a buffer slot size.
struct Side {
  HostLayout h;
  i64 slot = 0;                         // elements per rank slot of the buffer
  std::vector<std::vector<i64>> abs;    // absolute element offsets per axis
  std::vector<std::vector<int>> order;  // per axis: coordinates by increasing offset
};
Side make_side(const HostLayout &h, i64 slot);

bool choose_radices(int n, std::vector<int> *out, bool any_prime = false);
// Whether the pass kernel transforms lines of n points in this precision
// (n within one tile: radices 2..16, 3, 5, 7, 11, 13 by their own stage
// kernels, larger primes by direct sums); why: the reason if not.
bool axis_supported(int n, int dbl, std::string *why);
// exp(-2 pi i k / n), k < n, interleaved (re, im), accurate to double rounding
std::vector<double> roots(int n);
// Device image of a twiddle table of n roots w (interleaved (re, im)
// doubles) into dst, 16 bytes per root in either precision: complex double
// stores the root; complex float stores the compensated pair hi = fl32(w)
// (two floats) and lo = bf16(w - hi) (real part in the low half of the
// third word; the fourth word is 0).
void twiddle_image(const double *w, size_t n, int dbl, void *dst);

// Host image of the device tables of a program.
struct TableBuilder {
  std::vector<char> host;
  size_t put(const void *src, size_t bytes);
};

// Table offsets of one side's non-uniform addressing data.
struct PendingSide {
  size_t rank[DM_MAX_DIMS];
  size_t off[DM_MAX_DIMS];
  size_t strides;
};

struct PendingPass {
  FftLaunch L;
  size_t in_off[DM_MAX_DIMS];
  size_t out_off[DM_MAX_DIMS];
  PendingSide sin, sout;
  long ord_in = -1, ord_out = -1, ord_u = -1, tw = -1;
  long line_in = -1, line_out = -1;
};

// Plan the pass over axis t from layout I (buffer in) to layout U (buffer
// out); ctw: complex float with compensated twiddle products.  Returns
// false after raise_error when the axis cannot be done.
bool plan_pass(dm_comm_t comm, const Side &I, const Side &U, int t, int dbl, bool inverse, const void *in,
               void *out, TableBuilder &tb, PendingPass *pp, int ctw = 1);
// Upload the tables and resolve table pointers; returns the device block.
void *finalize_passes(dm_comm_t comm, TableBuilder &tb, std::vector<PendingPass> &pp,
                      std::vector<FftLaunch> *out);

// Axis d of h split into sub-axes 2d (y, A[d] points) and 2d+1 (z, L_d /
// A[d] points) over g = y + A[d] * z, when owner and offset separate into
// y and z parts; false (with the reason) otherwise.
bool split_layout(const HostLayout &h, const std::vector<int> &A, HostLayout *out, std::string *why);

// Layout validity for FFT use: uniform strides, every coordinate owned,
// every rank's data inside region_elems.
bool layout_usable(dm_comm_t comm, const char *what, const HostLayout &h, size_t region_elems);

}  // namespace dm
