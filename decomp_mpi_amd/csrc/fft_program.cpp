// Pass programs: tile axes, walk modes, radix schedules and device tables.
#include "fft_program.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#include "logger.h"

namespace dm {

Side make_side(const HostLayout &h, i64 slot) {
  Side s;
  s.h = h;
  s.slot = slot;
  s.abs.resize(h.ndim);
  s.order.resize(h.ndim);
  for (int d = 0; d < h.ndim; d++) {
    s.abs[d].resize(h.L[d]);
    for (int g = 0; g < h.L[d]; g++) s.abs[d][g] = (i64)h.rank_part[d][g] * slot + h.m[d] * h.off[d][g];
    s.order[d].resize(h.L[d]);
    std::iota(s.order[d].begin(), s.order[d].end(), 0);
    const auto &a = s.abs[d];
    std::stable_sort(s.order[d].begin(), s.order[d].end(), [&](int x, int y) { return a[x] < a[y]; });
  }
  return s;
}

// Odd prime factors one stage each; powers of two in radix-8 stages with one
// 16/4/2 stage for the remainder.  any_prime: primes above 13 become stages
// of the direct-sum kernel, last in the schedule (the pass kernel's
// general-radix stage); without it such a length is refused.
bool choose_radices(int n, std::vector<int> *out, bool any_prime) {
  out->clear();
  if (n < 1) return false;
  std::vector<int> odd, big;
  const int primes[] = {13, 11, 7, 5, 3};
  for (int p : primes)
    while (n % p == 0) {
      odd.push_back(p);
      n /= p;
    }
  int k = 0;
  while (n % 2 == 0) {
    n /= 2;
    k++;
  }
  if (any_prime)
    for (int p = 17; n > 1 && p <= n; p += 2)
      while (n % p == 0) {
        big.push_back(p);
        n /= p;
      }
  if (n != 1) return false;
  std::vector<int> pw;
  while (k >= 3) {
    pw.push_back(8);
    k -= 3;
  }
  if (k == 1) {
    if (!pw.empty()) pw.back() = 16;
    else pw.push_back(2);
  } else if (k == 2) {
    pw.push_back(4);
  }
  out->insert(out->end(), pw.begin(), pw.end());
  out->insert(out->end(), odd.begin(), odd.end());
  out->insert(out->end(), big.begin(), big.end());
  return true;
}

std::vector<double> roots(int n) {
  std::vector<double> w(2 * (size_t)n);
  const long double tau = 6.283185307179586476925286766559005768L;
  const long double h = 0.70710678118654752440084436210484903L;
  const long double tc[8] = {1, h, 0, -h, -1, -h, 0, h};
  const long double ts[8] = {0, h, 1, h, 0, -h, -1, -h};
  for (int k = 0; k < n; k++) {
    long double c, s;
    if ((8LL * k) % n == 0) {  // multiples of pi/4 exactly
      c = tc[8LL * k / n];
      s = ts[8LL * k / n];
    } else {
      long double a = tau * (long double)k / (long double)n;
      c = cosl(a);
      s = sinl(a);
    }
    w[2 * (size_t)k] = (double)c;
    w[2 * (size_t)k + 1] = (double)(-s);
  }
  return w;
}

void twiddle_image(const double *w, size_t n, int dbl, void *dst) {
  if (dbl) {
    memcpy(dst, w, n * 2 * sizeof(double));
    return;
  }
  // bfloat16 of x: the high half of its float bits, rounded to nearest even
  auto bf16 = [](double x) -> uint32_t {
    float f = (float)x;
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return u >> 16;
  };
  uint32_t *e = (uint32_t *)dst;
  for (size_t i = 0; i < n; i++) {
    const double re = w[2 * i], im = w[2 * i + 1];
    const float hr = (float)re, hi = (float)im;
    memcpy(&e[4 * i], &hr, 4);
    memcpy(&e[4 * i + 1], &hi, 4);
    e[4 * i + 2] = bf16(re - (double)hr) | (bf16(im - (double)hi) << 16);
    e[4 * i + 3] = 0;
  }
}

bool axis_supported(int n, int dbl, std::string *why) {
  char msg[160];
  std::vector<int> rad;
  const int cap = fft_tile_capacity(dbl, 1);
  if (n > cap) {
    snprintf(msg, sizeof msg, "axis length %d exceeds the %d-point tile of the %s pass", n, cap,
             dbl ? "complex-double" : "complex-float");
  } else if (!choose_radices(n, &rad, true) || (int)rad.size() > 16) {
    snprintf(msg, sizeof msg, "axis length %d has more than 16 prime factors", n);
  } else {
    return true;
  }
  if (why) *why = msg;
  return false;
}

size_t TableBuilder::put(const void *src, size_t bytes) {
  size_t at = host.size();
  host.resize(at + ((bytes + 255) & ~(size_t)255), 0);
  memcpy(&host[at], src, bytes);
  return at;
}

static bool is_identity(const std::vector<int> &o) {
  for (size_t i = 0; i < o.size(); i++)
    if (o[i] != (int)i) return false;
  return true;
}

bool plan_pass(dm_comm_t comm, const Side &I, const Side &U, int t, int dbl, bool inverse, const void *in,
               void *out, TableBuilder &tb, PendingPass *pp, int ctw) {
  int nd = I.h.ndim;
  FftLaunch &L = pp->L;
  memset(&L, 0, sizeof L);
  L.ndim = nd;
  for (int d = 0; d < nd; d++) L.L[d] = I.h.L[d];
  L.t = t;
  L.N = I.h.L[t];
  int cap = fft_tile_capacity(dbl, 1);
  if (L.N > cap) {
    raise_error(comm, "FFT: axis %d of length %d exceeds the %d-point pass tile", t, L.N, cap);
    return false;
  }
  std::vector<int> rad;
  if (!choose_radices(L.N, &rad, true) || (int)rad.size() > 16) {
    raise_error(comm, "FFT: axis length %d has more than 16 prime factors", L.N);
    return false;
  }
  L.nstage = (int)rad.size();
  for (int s = 0; s < L.nstage; s++) L.radix[s] = rad[s];
  auto fastest_other = [&](const Side &s) {
    int best = -1;
    for (int d = 0; d < nd; d++) {
      if (d == t || s.h.L[d] <= 1) continue;
      if (best < 0 || s.h.m[d] < s.h.m[best] || (s.h.m[d] == s.h.m[best] && d > best)) best = d;
    }
    return best;
  };
  bool in_row = (I.h.fastest == t) || fastest_other(I) < 0;
  bool out_row = (U.h.fastest == t) || fastest_other(U) < 0;
  int u = out_row ? fastest_other(I) : fastest_other(U);
  L.u = u;
  L.in_row = in_row ? 1 : 0;
  L.out_row = out_row ? 1 : 0;
  if (u >= 0) {
    // column walk at a line stride of 64 KiB or more: wide rows pay for the
    // lower occupancy of the wide tile -- for lines of 256 points or more
    // (the phase order's 16- and 32-point sub-axis passes ran slower in it:
    // FFT_2 6.15 -> 5.54 ms per 512^3 step without)
    const size_t es = dbl ? 16 : 8;
    bool wide = L.N >= 256 && ((!in_row && (size_t)I.h.m[t] * es >= (64u << 10)) ||
                               (!out_row && (size_t)U.h.m[t] * es >= (64u << 10)));
    static const bool nowide = getenv("DM_NO_WIDE") != nullptr;
    int tcap = fft_tile_capacity(dbl, !(in_row && out_row), wide && !nowide);
    L.T = std::max(1, std::min(std::min(tcap / L.N, I.h.L[u]), 512));
    L.tiles_u = (I.h.L[u] + L.T - 1) / L.T;
  } else {
    L.T = 1;
    L.tiles_u = 1;
  }
  // Non-temporal data accesses measured faster on every pass (512^3: up to
  // 10%) except the stores of a pass whose output is a column walk at a line
  // stride of 64 KiB or more: there the hint on the loads alone is best
  // (complex float: 2.58 -> 2.50 ms per 512^3 step against the hint on both).
  {
    // (complex double: no hint at all there, 4.63 ms against 4.85 with the
    // hint on its loads)
    const size_t es = dbl ? 16 : 8;
    static const bool nont = getenv("DM_NO_NONTEMPORAL") != nullptr;
    const bool far = !out_row && (size_t)U.h.m[t] * es >= (64u << 10);
    L.stream_in = (nont || (far && dbl)) ? 0 : 1;
    L.stream_out = (nont || far) ? 0 : 1;
  }
  L.nbatch = 0;
  for (int d = 0; d < nd; d++)
    if (d != t && d != u) L.batch_axis[L.nbatch++] = d;
  // A row walk on both sides whose u axis holds less than half a tile of
  // lines: each tile takes all L[u] lines for k consecutive values of the
  // fastest batch axis (k dividing its length), so tiles stay full (the
  // sub-axis passes of the phase order: 16- or 32-point lines, 16 or 32 of
  // them along u).
  L.merge_v = -1;
  L.merge_k = 1;
  if (u >= 0 && L.nbatch > 0 && in_row && out_row) {
    const int per = std::min(fft_tile_capacity(dbl, 0) / L.N, 512);
    const int lu = I.h.L[u];
    const int v = L.batch_axis[L.nbatch - 1];
    if (2 * lu <= per) {
      for (int k = per / lu; k >= 2; k--)
        if (I.h.L[v] % k == 0) {
          L.merge_v = v;
          L.merge_k = k;
          L.T = lu * k;
          L.tiles_u = 1;
          break;
        }
    }
  }
  const long long ntiles = fft_pass_tiles(L);
  long long batch = 1;
  for (int k = 0; k < L.nbatch; k++) batch *= I.h.L[L.batch_axis[k]];
  if (ntiles >= (1LL << 31)) {
    raise_error(comm, "FFT: %lld workgroups exceed the launch limit", ntiles);
    return false;
  }
  L.inverse = inverse ? 1 : 0;
  L.dbl = dbl;
  L.ctw = (!dbl && ctw) ? 1 : 0;
  for (int d = 0; d < nd; d++) {
    pp->in_off[d] = tb.put(I.abs[d].data(), I.abs[d].size() * sizeof(i64));
    pp->out_off[d] = tb.put(U.abs[d].data(), U.abs[d].size() * sizeof(i64));
  }
  // rank-dependent strides: owner, local offset and per-rank stride tables
  auto side = [&](const Side &s, FftSide &fs, PendingSide &ps) {
    fs.uniform = s.h.uniform ? 1 : 0;
    fs.proc_used = s.h.proc_used;
    fs.slot = s.slot;
    // offsets along t affine (x * m_t: axis t local to every rank, in
    // order), or in runs of R = 2^k points at a jump J per run (axis t in
    // equal blocks over the ranks: (x >> k) * J + (x & (R - 1)) * m_t)
    fs.aff = 0;
    fs.jump = 0;
    if (s.h.uniform && s.h.m[t] > 0) {
      bool affine = true;
      for (int x = 0; x < s.h.L[t] && affine; x++) affine = s.h.rank_part[t][x] == 0 && s.h.off[t][x] == x;
      static const bool noaff = getenv("DM_NO_AFFINE") != nullptr;
      if (affine && !noaff) fs.aff = s.h.m[t];
      const std::vector<i64> &a = s.abs[t];
      const i64 m = s.h.m[t];
      const int n = s.h.L[t];
      if (!affine && !noaff && a[0] == 0) {
        int R = 1;
        while (R < n && a[R] == a[R - 1] + m) R++;
        if (R < n && (R & (R - 1)) == 0) {
          int k = 0;
          while ((1 << k) < R) k++;
          const i64 J = a[R];
          bool ok = J != 0;
          for (int x = 0; x < n && ok; x++) ok = a[x] == (i64)(x >> k) * J + (i64)(x & (R - 1)) * m;
          if (ok) {
            fs.jump = J;
            fs.rstride = m;
            fs.rsh = k;
            fs.rmask = R - 1;
          }
        }
      }
    }
    if (s.h.uniform) return;
    for (int d = 0; d < nd; d++) {
      ps.rank[d] = tb.put(s.h.rank_part[d].data(), s.h.L[d] * sizeof(int));
      ps.off[d] = tb.put(s.h.off[d].data(), s.h.L[d] * sizeof(int));
    }
    ps.strides = tb.put(s.h.strides.data(), s.h.strides.size() * sizeof(i64));
  };
  side(I, L.sin, pp->sin);
  side(U, L.sout, pp->sout);
  if (in_row && !is_identity(I.order[t])) pp->ord_in = (long)tb.put(I.order[t].data(), L.N * sizeof(int));
  if (out_row && !is_identity(U.order[t])) pp->ord_out = (long)tb.put(U.order[t].data(), L.N * sizeof(int));
  if (u >= 0) {
    // tile lines follow the address order of the column-walked side
    const Side &cs = out_row ? I : U;
    if (!is_identity(cs.order[u])) pp->ord_u = (long)tb.put(cs.order[u].data(), I.h.L[u] * sizeof(int));
  }
  // Per-tile line bases for uniform sides: one coalesced load per tile in
  // the kernel instead of a chain of lookups (bounded to 64 MiB of tables).
  {
    long long nlines = (long long)L.tiles_u * batch * L.T;
    if (I.h.uniform && U.h.uniform && L.merge_k == 1 && nlines * 16 <= (64LL << 20)) {
      std::vector<i64> bi(nlines), bo(nlines);
      const std::vector<int> *ou = nullptr;
      if (pp->ord_u >= 0) ou = &((out_row ? I : U).order[u]);
      std::vector<int> g(nd, 0);
      for (long long tile = 0; tile < (long long)L.tiles_u * batch; tile++) {
        int tu = (int)(tile % L.tiles_u);
        long long bid = tile / L.tiles_u;
        for (int k = L.nbatch - 1; k >= 0; k--) {
          int d = L.batch_axis[k];
          g[d] = (int)(bid % I.h.L[d]);
          bid /= I.h.L[d];
        }
        i64 base_i = 0, base_o = 0;
        for (int k = 0; k < L.nbatch; k++) {
          int d = L.batch_axis[k];
          base_i += I.abs[d][g[d]];
          base_o += U.abs[d][g[d]];
        }
        for (int l = 0; l < L.T; l++) {
          i64 ai = base_i, ao = base_o;
          if (u >= 0) {
            int cu = tu * L.T + l;
            if (cu >= I.h.L[u]) cu = 0;
            else if (ou) cu = (*ou)[cu];
            ai += I.abs[u][cu];
            ao += U.abs[u][cu];
          }
          bi[tile * L.T + l] = ai;
          bo[tile * L.T + l] = ao;
        }
      }
      pp->line_in = (long)tb.put(bi.data(), bi.size() * sizeof(i64));
      pp->line_out = (long)tb.put(bo.data(), bo.size() * sizeof(i64));
    }
  }
  // Twiddles in stage order: stage s (Ns = product of the earlier radices,
  // radix R) holds w^(jj * r * N / (Ns * R)) for r = 1..R-1, jj < Ns at
  // index (Ns - 1) + (r - 1) * Ns + jj, so the lanes of a wave (consecutive
  // jj) read consecutive entries.  The stages telescope to N - 1 entries.
  std::vector<double> root = roots(L.N);
  std::vector<double> w(2 * (size_t)L.N, 0.0);
  {
    int ns = 1;
    for (int st = 0; st < L.nstage; st++) {
      int R = L.radix[st];
      int stride = L.N / (ns * R);
      for (int r = 1; r < R; r++)
        for (int jj = 0; jj < ns; jj++) {
          size_t at = (size_t)(ns - 1) + (size_t)(r - 1) * ns + jj;
          size_t k = (size_t)jj * r * stride;
          w[2 * at] = root[2 * k];
          w[2 * at + 1] = root[2 * k + 1];
        }
      ns *= R;
    }
    // after the N stage entries: w_p^m = w^(m N / p), m < p, for each stage
    // of a radix without a kernel of its own, in stage order
    for (int st = 0; st < L.nstage; st++) {
      int R = L.radix[st];
      if (has_radix_kernel(R)) continue;
      for (int m = 0; m < R; m++) {
        size_t k = (size_t)m * (L.N / R);
        w.push_back(root[2 * k]);
        w.push_back(root[2 * k + 1]);
      }
    }
  }
  const size_t ntw = w.size() / 2;
  std::vector<char> img(ntw * 16);
  twiddle_image(w.data(), ntw, dbl, img.data());
  pp->tw = (long)tb.put(img.data(), img.size());
  L.in = in;
  L.out = out;
  auto form = [](const FftSide &f) {
    if (f.aff) return std::string(" affine");
    if (f.jump) return " affine in runs of " + std::to_string(f.rmask + 1);
    return std::string("");
  };
  logger(comm, "FFT pass: axis %d (N %d, %d stages) lines along %d, T %d, %lld workgroups, in %s stride %lld%s, out %s stride %lld%s%s%s\n",
         t, L.N, L.nstage, u, L.T, ntiles, in_row ? "row" : "column", (long long)I.h.m[t],
         form(L.sin).c_str(), out_row ? "row" : "column", (long long)U.h.m[t], form(L.sout).c_str(),
         L.stream_in ? ", non-temporal loads" : "", L.stream_out ? ", non-temporal stores" : "");
  return true;
}

void *finalize_passes(dm_comm_t comm, TableBuilder &tb, std::vector<PendingPass> &pp,
                      std::vector<FftLaunch> *out) {
  void *mem = device_alloc(std::max<size_t>(tb.host.size(), 256));
  if (mem == nullptr) {
    raise_error(comm, "FFT: device allocation of %zu bytes of tables failed", tb.host.size());
    return nullptr;
  }
  if (!tb.host.empty()) h2d(mem, tb.host.data(), tb.host.size());
  char *b = (char *)mem;
  out->clear();
  for (auto &p : pp) {
    FftLaunch L = p.L;
    for (int d = 0; d < L.ndim; d++) {
      L.off_in[d] = (const i64 *)(b + p.in_off[d]);
      L.off_out[d] = (const i64 *)(b + p.out_off[d]);
    }
    auto side = [&](FftSide &fs, const PendingSide &ps) {
      if (fs.uniform) return;
      for (int d = 0; d < L.ndim; d++) {
        fs.rank[d] = (const int *)(b + ps.rank[d]);
        fs.off[d] = (const int *)(b + ps.off[d]);
      }
      fs.strides = (const i64 *)(b + ps.strides);
    };
    side(L.sin, p.sin);
    side(L.sout, p.sout);
    L.ord_in = p.ord_in >= 0 ? (const int *)(b + p.ord_in) : nullptr;
    L.ord_out = p.ord_out >= 0 ? (const int *)(b + p.ord_out) : nullptr;
    L.ord_u = p.ord_u >= 0 ? (const int *)(b + p.ord_u) : nullptr;
    L.line_in_tab = p.line_in >= 0 ? (const i64 *)(b + p.line_in) : nullptr;
    L.line_out_tab = p.line_out >= 0 ? (const i64 *)(b + p.line_out) : nullptr;
    L.twiddle = b + p.tw;
    out->push_back(L);
  }
  return mem;
}

bool layout_usable(dm_comm_t comm, const char *what, const HostLayout &h, size_t region_elems) {
  if (!h.offsets_ok) {
    raise_error(comm, "FFT: the %s decomposition has a negative offset or stride", what);
    return false;
  }
  for (int d = 0; d < h.ndim; d++)
    for (int g = 0; g < h.L[d]; g++)
      if (h.rank_part[d][g] >= DM_NO_OWNER) {
        raise_error(comm, "FFT: coordinate %d of axis %d of the %s decomposition has no owner", g, d, what);
        return false;
      }
  for (int r = 0; r < h.proc_used; r++)
    if ((size_t)h.slot_extent[r] > region_elems) {
      raise_error(comm, "FFT: rank %d needs %lld elements of %s, the buffer region holds %zu", r,
                  (long long)h.slot_extent[r], what, region_elems);
      return false;
    }
  return true;
}

}  // namespace dm

namespace dm {

// Axis d of h as two sub-axes over g = y + A_d * z: sub-axis 2d is y (A_d
// points), sub-axis 2d+1 is z (L_d / A_d points).  The owner and the local
// offset must be sums of a y part and a z part (the split FFT's A and B
// decompositions are: Interleave(Block(A, ga), B) gives rank(y) and
// z + B * off(y); BlockWithFactor(A * B, A, gb) gives rank(z) and
// y + A * (z % (bs / A))).  Each part's offsets are divided by their common
// step, which goes into the sub-axis' memory stride.
bool split_layout(const HostLayout &h, const std::vector<int> &A, HostLayout *out, std::string *why) {
  const int nd = h.ndim;
  if (2 * nd > DM_MAX_DIMS) {
    *why = "too many axes to split";
    return false;
  }
  HostLayout s;
  s.ndim = 2 * nd;
  s.proc_used = h.proc_used;
  s.L.assign(2 * nd, 1);
  s.rank_part.assign(2 * nd, std::vector<int>());
  s.off.assign(2 * nd, std::vector<int>());
  s.strides.assign((size_t)h.proc_used * 2 * nd, 0);
  s.uniform = h.uniform;
  s.slot_extent = h.slot_extent;
  s.offsets_ok = h.offsets_ok;
  for (int d = 0; d < nd; d++) {
    const int L = h.L[d], a = A[d];
    if (a < 1 || L % a != 0) {
      *why = "factor does not divide the axis";
      return false;
    }
    const int b = L / a;
    const auto &rp = h.rank_part[d];
    const auto &of = h.off[d];
    std::vector<int> ry(a), rz(b), oy(a), oz(b);
    for (int z = 0; z < b; z++) {
      rz[z] = rp[a * z];
      oz[z] = of[a * z];
    }
    for (int y = 0; y < a; y++) {
      ry[y] = rp[y] - rp[0];
      oy[y] = of[y] - of[0];
    }
    for (int z = 0; z < b; z++)
      for (int y = 0; y < a; y++) {
        const int g = y + a * z;
        if (rp[g] >= DM_NO_OWNER || rp[g] != ry[y] + rz[z] || of[g] != oy[y] + oz[z]) {
          *why = "the decomposition is not separable along the split";
          return false;
        }
      }
    auto step_of = [](const std::vector<int> &o) {
      int st = 0;
      for (int v : o) st = std::gcd(st, v < 0 ? -v : v);
      return st > 0 ? st : 1;
    };
    const int sy = step_of(oy), sz = step_of(oz);
    for (auto &v : oy) v /= sy;
    for (auto &v : oz) v /= sz;
    s.L[2 * d] = a;
    s.L[2 * d + 1] = b;
    s.rank_part[2 * d] = ry;
    s.rank_part[2 * d + 1] = rz;
    s.off[2 * d] = oy;
    s.off[2 * d + 1] = oz;
    for (int r = 0; r < h.proc_used; r++) {
      const i64 m = h.strides[(size_t)r * nd + d];
      s.strides[(size_t)r * 2 * nd + 2 * d] = m * sy;
      s.strides[(size_t)r * 2 * nd + 2 * d + 1] = m * sz;
    }
  }
  s.m.assign(s.strides.begin(), s.strides.begin() + 2 * nd);
  // fastest sub-axis: smallest stride among those longer than one; ties ->
  // the later sub-axis
  int best = 2 * nd - 1;
  i64 bs = -1;
  for (int i = 2 * nd - 1; i >= 0; i--) {
    if (s.L[i] <= 1) continue;
    if (bs < 0 || s.m[i] < bs) {
      bs = s.m[i];
      best = i;
    }
  }
  s.fastest = best;
  *out = s;
  return true;
}

}  // namespace dm
