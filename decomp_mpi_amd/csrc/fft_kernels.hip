// K2 host side: launch parameters of one FFT pass (magics, fused stages, XCD
// pairing, LDS size) and the dispatch to the kernel family that runs it.
// The kernel itself is fft_pass_kernel.h; its instantiations live in
// fft_inst.hip, one translation unit per precision, direction and radix
// family.
//
// Replaces the FFTW guru stages of the split FFT (split_fft.c:426-546) and the
// twiddle pass (dimscale.c:7-49).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>

#include "fft_pass_kernel.h"

namespace dm {
// Points per tile the planner should use.  Column walks need the big tile
// (T * element = 128-byte rows); pure row walks use the half tile (twice the
// workgroups per CU).  DM_FFT_TILE=big|small forces one size.
int fft_tile_capacity(int dbl, int column, int wide) {
  using namespace fftpass;
  static const int force = [] {
    const char *e = getenv("DM_FFT_TILE");
    return e == nullptr ? -1 : (e[0] == 'b' ? 0 : 1);
  }();
  int cap = dbl ? CAP_D : CAP_F;
  if (dbl && column && wide && force < 0) return 2 * cap;
  bool small = force >= 0 ? force == 1 : !column;
  return small ? cap / 2 : cap;
}

// Twiddle table in LDS: 16 bytes per entry for a complex double or a
// compensated complex-float pair, 8 for a plain complex float.  A
// compensated table stays within 8 KiB there, so that two 8192-point
// complex-float tiles still share a CU.
static size_t twiddle_lds_entry(const FftLaunch &p) { return (p.dbl || p.ctw) ? 16 : 8; }
static bool twiddle_in_lds(const FftLaunch &p) {
  return (size_t)p.N * twiddle_lds_entry(p) <= ((!p.dbl && p.ctw) ? 8192u : 16384u);
}

size_t fft_lds_bytes(const FftLaunch &p) {
  using namespace fftpass;
  size_t elem = p.dbl ? 16 : 8;
  size_t ls = (size_t)lds_line(p.N);
  auto side = [&](int uniform) { return uniform ? (size_t)p.T * 8 : (size_t)p.T * 4 * (1 + p.ndim); };
  // tile data + per-line address data of both sides
  size_t b = (size_t)p.T * ls * elem + side(p.sin.uniform) + side(p.sout.uniform);
  b = (b + 15) & ~(size_t)15;
  if (twiddle_in_lds(p)) b += (size_t)p.N * twiddle_lds_entry(p);
  return b;
}

int launch_fft_pass(const FftLaunch &p, void *stream) {
  using namespace fftpass;
  Pass P{};
  P.p = p;
  P.m_n = magic_for(p.N);
  P.m_t = magic_for(p.T);
  P.ls = lds_line(p.N);
  {
    // Fused first/last stages need whole butterflies per thread (points per
    // thread a multiple of the radix) and an unpermuted walk along the line
    // in row mode (a permuted one goes through LDS to stay coalesced).
    const int V = 8;  // points per thread in every tile
    auto pow2 = [](int r) { return r == 2 || r == 4 || r == 8 || r == 16; };
    int r0 = p.radix[0], rl = p.radix[p.nstage > 0 ? p.nstage - 1 : 0];
    P.fuse_in = p.nstage >= 1 && pow2(r0) && V % r0 == 0 && (!p.in_row || p.ord_in == nullptr);
    P.fuse_out = p.nstage >= 2 && pow2(rl) && V % rl == 0 && (!p.out_row || p.ord_out == nullptr);
    static const bool off = getenv("DM_NO_FUSE") != nullptr;
    if (off) P.fuse_in = P.fuse_out = 0;
    // compensated twiddles with per-rank output strides: the fused store's
    // address arithmetic and the twiddle pairs do not fit 64 VGPRs, so the
    // last stage runs in LDS (the kernel has no fused store for this form)
    if (!p.dbl && p.ctw && !p.sout.uniform) P.fuse_out = 0;
    // column walks with rows shorter than 128 bytes: pair the tiles that
    // share lines on one XCD
    int elem = p.dbl ? 16 : 8;
    int row_bytes = p.T * elem;
    bool col = !p.in_row || !p.out_row;
    P.xcd_pair = (col && p.u >= 0 && row_bytes < 128) ? 128 / row_bytes : 1;
    P.tw_lds = twiddle_in_lds(p);
  }
  int ns = 1;
  for (int s = 0; s < p.nstage; s++) {
    int r = p.radix[s];
    P.st[s].nr = p.N / r;
    P.st[s].ns = ns;
    P.st[s].m_nr = magic_for(p.N / r);
    P.st[s].m_ns = magic_for(ns);
    ns *= r;
  }
  const long long all = fft_pass_tiles(p);
  if (p.nblocks <= 0) {  // the whole pass
    P.p.nblocks = all;
    P.p.tile_base = 0;
  }
  if (P.p.nblocks <= 0) return 0;
  if (P.p.tile_base < 0 || P.p.tile_base + P.p.nblocks > all) return (int)hipErrorInvalidValue;
  size_t lds = fft_lds_bytes(p);
  hipStream_t s = (hipStream_t)stream;
  bool pow2 = true, any = false;
  for (int i = 0; i < p.nstage; i++) {
    pow2 = pow2 && (p.radix[i] & (p.radix[i] - 1)) == 0;
    any = any || !has_radix_kernel(p.radix[i]);
  }
  hipError_t e;
  if (p.dbl) {
    if (any) e = p.inverse ? launch_d_inv_any(P, lds, s) : launch_d_fwd_any(P, lds, s);
    else if (p.inverse) e = pow2 ? launch_d_inv_p2(P, lds, s) : launch_d_inv_mx(P, lds, s);
    else e = pow2 ? launch_d_fwd_p2(P, lds, s) : launch_d_fwd_mx(P, lds, s);
  } else {
    if (any) e = p.inverse ? launch_f_inv_any(P, lds, s) : launch_f_fwd_any(P, lds, s);
    else if (p.inverse) e = pow2 ? launch_f_inv_p2(P, lds, s) : launch_f_inv_mx(P, lds, s);
    else e = pow2 ? launch_f_fwd_p2(P, lds, s) : launch_f_fwd_mx(P, lds, s);
  }
  return (int)e;
}

}  // namespace dm
