/*
 * oracle.c -- CPU restatement of the reference's hot path, for tests only.
 *
 * TEST INFRASTRUCTURE.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load this file's library, and only as the checker (or as
 * the timed CPU baseline).  The product library never calls it.
 *
 * Each function restates one piece of the reference, cited by file:line
 * (paths relative to the reference tree):
 *   descriptors        block_decomp.c:6-13,48  local_decomp.c:6-11
 *                      single_decomp.c:6-12   rotate_decomp.c:7-14,28-33
 *                      interleave_decomp.c:6-20  factor_reverse_decomp.c:8-34
 *   axis/decomp        reshape.c:47-89 (positions, local counts),
 *                      reshape.c:174-241 (proc_used, packed strides),
 *                      reshape.c:317-324 (extent)
 *   reshape            reshape.c:104-136 (index lists), reshape.c:330-350
 *                      (sortOrder), reshape.c:351-449 + typelist.c:74-119
 *                      (nested message types), reshape.c:461-492 (message
 *                      pairing: the k-th element sent is the k-th received)
 *   twiddles           split_fft.c:348-420
 *   twiddle multiply   dimscale.c:7-67
 *   split FFT          split_fft.c:36-77 (A/B decomps), 426-546 (stage
 *                      transforms as FFTW guru dims), 622-682 (forward);
 *                      backward: the exact inverse of forward
 *   processor grid     processor_grid.c:24-97
 *
 * The stage DFTs stand in for FFTW's guru interface (a third-party library
 * absent from this image): iterative radix-2 for powers of two, direct
 * summation otherwise, both in double precision with twiddles computed in
 * long double.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define OR_BLOCK 1
#define OR_LOCAL 2
#define OR_SINGLE 3
#define OR_ROTATE 4
#define OR_INTERLEAVE 5
#define OR_FREV 6

#define OR_MAXD 8

/* A descriptor is a node list; node i = spec[5i .. 5i+4] =
 * {type, p0, p1, p2, nested node index or -1}; node 0 is the root. */

static const int *node(const int *spec, int i) { return spec + 5 * i; }

int or_coord_size(const int *spec, int i) {
  const int *n = node(spec, i);
  switch (n[0]) {
    case OR_BLOCK: case OR_LOCAL: case OR_SINGLE: return n[1];
    case OR_ROTATE: case OR_FREV: return or_coord_size(spec, n[4]);
    case OR_INTERLEAVE: return or_coord_size(spec, n[4]) * n[1];
  }
  return -1;
}

int or_proc_size(const int *spec, int i) {
  const int *n = node(spec, i);
  switch (n[0]) {
    case OR_BLOCK: return n[3];
    case OR_LOCAL: return 1;
    case OR_SINGLE: return n[3];
    default: return or_proc_size(spec, n[4]);
  }
}

/* block_decomp.c:48 */
static int block_size(const int *n) {
  int len = n[1], factor = n[2], procs = n[3];
  return factor * (((len / factor) + procs - 1) / procs);
}

/* rank (want_rank=1) or local offset of global coordinate g */
static int eval(const int *spec, int i, int g, int want_rank) {
  const int *n = node(spec, i);
  switch (n[0]) {
    case OR_BLOCK: {
      int bs = block_size(n);
      return want_rank ? g / bs : g % bs;
    }
    case OR_LOCAL: return want_rank ? 0 : g;
    case OR_SINGLE: return want_rank ? n[2] : g;
    case OR_ROTATE: {
      int len = or_coord_size(spec, i);
      return eval(spec, n[4], (g + len + n[1]) % len, want_rank);
    }
    case OR_INTERLEAVE: {
      int A = or_coord_size(spec, n[4]), B = n[1];
      int y = g % A, z = g / A;
      if (want_rank) return eval(spec, n[4], y, 1);
      return z + B * eval(spec, n[4], y, 0);
    }
    case OR_FREV: {
      int A = n[1], B = n[2];
      int k = g / A + B * (g % A);
      return eval(spec, n[4], k, want_rank);
    }
  }
  return -1;
}

int or_rank(const int *spec, int g) { return eval(spec, 0, g, 1); }
int or_offset(const int *spec, int g) { return eval(spec, 0, g, 0); }

/* ---------------------------------------------------------------- decomps */

typedef struct {
  int ndim, P;
  const int *spec[OR_MAXD];
  int L[OR_MAXD], procs[OR_MAXD], rs[OR_MAXD];
  int proc_used;
  long *stride; /* [P*ndim] */
  int *rank_of[OR_MAXD], *off_of[OR_MAXD];
  int *max_local[OR_MAXD]; /* per position */
} Dec;

static int pos_of(const Dec *d, int r, int i) { return (r / d->rs[i]) % d->procs[i]; }

/* stride_mode 1: packed (reshape.c:228-236), else explicit strides[ndim]
 * applied on every rank. */
static void dec_init(Dec *d, int P, int ndim, const int *specs, const int *spec_off, const int *rs,
                     int packed, const long *explicit_strides) {
  memset(d, 0, sizeof *d);
  d->ndim = ndim;
  d->P = P;
  int max = 0;
  for (int i = 0; i < ndim; i++) {
    d->spec[i] = specs + spec_off[i];
    d->L[i] = or_coord_size(d->spec[i], 0);
    d->procs[i] = or_proc_size(d->spec[i], 0);
    d->rs[i] = rs[i];
    max += rs[i] * (d->procs[i] - 1);
    d->rank_of[i] = malloc(sizeof(int) * d->L[i]);
    d->off_of[i] = malloc(sizeof(int) * d->L[i]);
    d->max_local[i] = calloc(d->procs[i], sizeof(int));
    for (int g = 0; g < d->L[i]; g++) {
      int r = or_rank(d->spec[i], g), o = or_offset(d->spec[i], g);
      d->rank_of[i][g] = r;
      d->off_of[i][g] = o;
      if (r >= 0 && r < d->procs[i] && o > d->max_local[i][r]) d->max_local[i][r] = o;
    }
  }
  d->proc_used = max + 1;
  d->stride = malloc(sizeof(long) * P * ndim);
  for (int r = 0; r < P; r++) {
    long m = 1;
    for (int i = ndim - 1; i >= 0; i--) {
      if (packed) {
        d->stride[r * ndim + i] = m;
        m *= d->max_local[i][pos_of(d, r, i)] + 1;
      } else {
        d->stride[r * ndim + i] = explicit_strides[i];
      }
    }
  }
}

static void dec_free(Dec *d) {
  for (int i = 0; i < d->ndim; i++) {
    free(d->rank_of[i]);
    free(d->off_of[i]);
    free(d->max_local[i]);
  }
  free(d->stride);
}

static long dec_extent(const Dec *d, int r) {
  long e = 1;
  for (int i = 0; i < d->ndim; i++) e += d->stride[r * d->ndim + i] * d->max_local[i][pos_of(d, r, i)];
  return e;
}

/* Per-rank strides, proc_used and extents for Python. */
int or_decomp_info(int P, int ndim, const int *specs, const int *spec_off, const int *rs, int packed,
                   const long *explicit_strides, long *strides_out, long *extent_out) {
  Dec d;
  dec_init(&d, P, ndim, specs, spec_off, rs, packed, explicit_strides);
  memcpy(strides_out, d.stride, sizeof(long) * P * ndim);
  for (int r = 0; r < P; r++) extent_out[r] = dec_extent(&d, r);
  int pu = d.proc_used;
  dec_free(&d);
  return pu;
}

/* ---------------------------------------------------------------- reshape */

/* reshape.c:330-350 with the calling rank's source strides */
static void sort_order(const Dec *src, int me, int *order) {
  int n = src->ndim;
  for (int i = 0; i < n; i++) order[i] = n - 1 - i;
  for (int i = 0; i < n - 1; i++)
    for (int j = i + 1; j < n; j++)
      if (src->stride[me * n + order[j]] < src->stride[me * n + order[i]]) {
        int t = order[j];
        order[j] = order[i];
        order[i] = t;
      }
}

/* Element offsets of the message between `mine` rank me and `remote`
 * rank peer, in the order the nested datatype lists them: the last
 * expanded axis outermost, coordinates ascending on every axis.  Returns
 * the count, or -1 when the message does not exist on this side. */
static long message(const Dec *mine, int me, const Dec *remote, int peer, const int *order, long **out) {
  int n = mine->ndim;
  if (me >= mine->proc_used || peer >= remote->proc_used) return -1;
  int *lists[OR_MAXD], cnt[OR_MAXD];
  long total = 1;
  for (int a = 0; a < n; a++) {
    int i = order[a];
    int mypos = pos_of(mine, me, i), rpos = pos_of(remote, peer, i);
    lists[a] = malloc(sizeof(int) * (mine->L[i] + 1));
    cnt[a] = 0;
    for (int g = 0; g < mine->L[i]; g++)
      if (mine->rank_of[i][g] == mypos && remote->rank_of[i][g] == rpos) lists[a][cnt[a]++] = g;
    total *= cnt[a];
  }
  /* the peer's rank must be what the expansion produces: sum of positions
   * times remote rank strides (typelist.c:93) */
  long nr = 0;
  for (int i = 0; i < n; i++) nr += (long)pos_of(remote, peer, i) * remote->rs[i];
  if (total == 0 || nr != peer) {
    for (int a = 0; a < n; a++) free(lists[a]);
    return -1;
  }
  long *o = malloc(sizeof(long) * total);
  int idx[OR_MAXD] = {0};
  for (long k = 0; k < total; k++) {
    long off = 0;
    for (int a = 0; a < n; a++) {
      int i = order[a];
      off += mine->stride[me * n + i] * mine->off_of[i][lists[a][idx[a]]];
    }
    o[k] = off;
    /* order[0] fastest */
    for (int a = 0; a < n; a++) {
      if (++idx[a] < cnt[a]) break;
      idx[a] = 0;
    }
  }
  for (int a = 0; a < n; a++) free(lists[a]);
  *out = o;
  return total;
}

/* The message of rank me (of `mine`) to or from rank peer (of `remote`)
 * as the reference's nested datatype lists it (typelist.c:74-119 with the
 * axis order of reshape.c:394): element offsets into out (at most max),
 * in the order the sender or receiver sorts by its own view of the source
 * strides.  send: mine = src, remote = dst; else mine = dst, remote = src.
 * Returns the element count, or -1 when there is no such message. */
long or_message(int P, int ndim, const int *specs_s, const int *off_s, const int *rs_s, int packed_s,
                const long *expl_s, const int *specs_d, const int *off_d, const int *rs_d, int packed_d,
                const long *expl_d, int me, int peer, int send, long *out, long max) {
  Dec S, D;
  dec_init(&S, P, ndim, specs_s, off_s, rs_s, packed_s, expl_s);
  dec_init(&D, P, ndim, specs_d, off_d, rs_d, packed_d, expl_d);
  int order[OR_MAXD];
  sort_order(&S, me, order);
  long *o = NULL;
  long n = send ? message(&S, me, &D, peer, order, &o) : message(&D, me, &S, peer, order, &o);
  for (long k = 0; k < n && k < max; k++) out[k] = o[k];
  free(o);
  dec_free(&S);
  dec_free(&D);
  return n;
}

/* runReshapePlan restated over all ranks: every (S, D) message pair is
 * matched element by element.  Returns 0, or a negative code when a send
 * has no matching receive or the counts differ (where MPI would hang or
 * report truncation). */
int or_reshape(int P, int ndim, const int *specs_s, const int *off_s, const int *rs_s, int packed_s,
               const long *expl_s, const int *specs_d, const int *off_d, const int *rs_d, int packed_d,
               const long *expl_d, int elem, const char *src, long src_slot, char *dst, long dst_slot) {
  Dec S, D;
  dec_init(&S, P, ndim, specs_s, off_s, rs_s, packed_s, expl_s);
  dec_init(&D, P, ndim, specs_d, off_d, rs_d, packed_d, expl_d);
  int rc = 0;
  for (int s = 0; s < S.proc_used && rc == 0; s++) {
    int order_s[OR_MAXD];
    sort_order(&S, s, order_s);
    for (int d = 0; d < D.proc_used && rc == 0; d++) {
      int order_d[OR_MAXD];
      sort_order(&S, d, order_d); /* the receiver sorts by its own view of src */
      long *so = NULL, *ro = NULL;
      long ns = message(&S, s, &D, d, order_s, &so);
      long nrv = message(&D, d, &S, s, order_d, &ro);
      if (ns < 0 && nrv < 0) continue;
      if (ns != nrv) {
        rc = -1;
      } else {
        for (long k = 0; k < ns; k++)
          memcpy(dst + d * dst_slot + ro[k] * elem, src + s * src_slot + so[k] * elem, elem);
      }
      free(so);
      free(ro);
    }
  }
  dec_free(&S);
  dec_free(&D);
  return rc;
}

/* ---------------------------------------------------------------- twiddles */

/* split_fft.c:348-420 for rank r of the A decomposition; table[i] has
 * max_local+1 entries (re, im pairs), unused slots 0. */
static void twiddle_tables(const Dec *A, int r, const int *Af, const int *L, double **tab) {
  for (int i = 0; i < A->ndim; i++) {
    int pos = pos_of(A, r, i);
    int ml = A->max_local[i][pos];
    tab[i] = calloc(2 * (ml + 1), sizeof(double));
    for (int g = 0; g < L[i]; g++) {
      if (A->rank_of[i][g] != pos) continue;
      int local = A->off_of[i][g];
      int y = g % Af[i], l = g / Af[i];
      double angle = -(2.0 * M_PI) * ((double)(y * l) / (double)L[i]);
      double re = cos(angle), im = sin(angle);
      int rem = (y * l) % L[i];
      if (rem == 0) re = 1.0, im = 0.0;
      else if (4 * rem == L[i]) re = 0.0, im = -1.0;
      else if (2 * rem == L[i]) re = -1.0, im = 0.0;
      else if (4 * rem == 3 * L[i]) re = -0.0, im = 1.0;
      tab[i][2 * local] = re;
      tab[i][2 * local + 1] = im;
    }
  }
}

int or_twiddle_table(int P, int ndim, const int *specs, const int *spec_off, const int *rs, int rank,
                     const int *Af, const int *L, int axis, double *out) {
  Dec A;
  dec_init(&A, P, ndim, specs, spec_off, rs, 1, NULL);
  double *tab[OR_MAXD];
  twiddle_tables(&A, rank, Af, L, tab);
  int ml = A.max_local[axis][pos_of(&A, rank, axis)];
  memcpy(out, tab[axis], sizeof(double) * 2 * (ml + 1));
  for (int i = 0; i < ndim; i++) free(tab[i]);
  dec_free(&A);
  return ml + 1;
}

/* dimscale.c:7-49 */
static void dim_scale(const Dec *d, int r, int dim, double complex factor, double **tab, double complex *data,
                      long offset, int sign) {
  int n = d->ndim;
  int count = d->max_local[dim][pos_of(d, r, dim)] + 1;
  long ms = d->stride[r * n + dim];
  for (int pos = 0; pos < count; pos++) {
    long o = offset + pos * ms;
    double complex f = factor;
    if (tab != NULL && tab[dim] != NULL) f = factor * (tab[dim][2 * pos] + I * tab[dim][2 * pos + 1]);
    if (f == 0.0) continue;
    if (dim == n - 1) {
      if (sign == 1) f = conj(f);
      data[o] = f * data[o];
    } else {
      dim_scale(d, r, dim + 1, f, tab, data, o, sign);
    }
  }
}

int or_decomp_scale(int P, int ndim, const int *specs, const int *spec_off, const int *rs, int rank,
                    const int *Af, const int *L, double *data, int sign) {
  Dec A;
  dec_init(&A, P, ndim, specs, spec_off, rs, 1, NULL);
  double *tab[OR_MAXD];
  twiddle_tables(&A, rank, Af, L, tab);
  dim_scale(&A, rank, 0, 1.0, tab, (double complex *)data, 0, sign);
  for (int i = 0; i < ndim; i++) free(tab[i]);
  dec_free(&A);
  return 0;
}

/* ---------------------------------------------------------------- stage DFTs */

/* exp(sign 2 pi i k / n) for k < n, computed in long double; the last
 * table is kept since consecutive calls share n. */
static const double complex *root_table(int n, int sign) {
  static double complex *tab = NULL;
  static int tn = 0, ts = 0;
  if (tab != NULL && tn == n && ts == sign) return tab;
  free(tab);
  tab = malloc(sizeof(double complex) * n);
  for (int k = 0; k < n; k++) {
    long double a = (long double)sign * 2.0L * 3.141592653589793238462643383279503L * k / n;
    tab[k] = (double)cosl(a) + I * (double)sinl(a);
  }
  tn = n;
  ts = sign;
  return tab;
}

static void dft1(double complex *x, int n, int sign, double complex *tmp) {
  if (n <= 1) return;
  const double complex *w = root_table(n, sign);
  if ((n & (n - 1)) == 0) {
    /* iterative radix-2, decimation in time */
    for (int i = 1, j = 0; i < n; i++) {
      int bit = n >> 1;
      for (; j & bit; bit >>= 1) j ^= bit;
      j ^= bit;
      if (i < j) {
        double complex t = x[i];
        x[i] = x[j];
        x[j] = t;
      }
    }
    for (int len = 2; len <= n; len <<= 1) {
      for (int k = 0; k < len / 2; k++) {
        double complex wk = w[(long)k * (n / len)];
        for (int i = 0; i < n; i += len) {
          double complex u = x[i + k], v = x[i + k + len / 2] * wk;
          x[i + k] = u + v;
          x[i + k + len / 2] = u - v;
        }
      }
    }
    return;
  }
  for (int k = 0; k < n; k++) {
    double complex s = 0;
    for (int j = 0; j < n; j++) {
      s += x[j] * w[((long)j * k) % n];
    }
    tmp[k] = s;
  }
  memcpy(x, tmp, sizeof(double complex) * n);
}

/* FFTW guru semantics: transform dims (n, stride) and loop dims (n, stride),
 * identical input and output strides, out-of-place in -> out. */
static void guru(int nt, const int *tn, const long *ts, int nl, const int *ln, const long *ls,
                 const double complex *in, double complex *out, int sign) {
  long tot = 1;
  for (int a = 0; a < nt; a++) tot *= tn[a];
  double complex *buf = malloc(sizeof(double complex) * tot);
  int maxn = 1;
  for (int a = 0; a < nt; a++) maxn = tn[a] > maxn ? tn[a] : maxn;
  double complex *line = malloc(sizeof(double complex) * maxn), *tmp = malloc(sizeof(double complex) * maxn);
  long nloop = 1;
  for (int a = 0; a < nl; a++) nloop *= ln[a];
  for (long li = 0; li < nloop; li++) {
    long base = 0, rem = li;
    for (int a = nl - 1; a >= 0; a--) {
      base += (rem % ln[a]) * ls[a];
      rem /= ln[a];
    }
    /* gather the sub-array, row-major over transform dims */
    for (long k = 0; k < tot; k++) {
      long off = base, r2 = k;
      for (int a = nt - 1; a >= 0; a--) {
        off += (r2 % tn[a]) * ts[a];
        r2 /= tn[a];
      }
      buf[k] = in[off];
    }
    /* 1D transforms along every transform dim */
    long inner = 1;
    for (int a = nt - 1; a >= 0; a--) {
      long outer = tot / (inner * tn[a]);
      for (long o = 0; o < outer; o++)
        for (long i = 0; i < inner; i++) {
          long b0 = o * inner * tn[a] + i;
          for (int j = 0; j < tn[a]; j++) line[j] = buf[b0 + j * inner];
          dft1(line, tn[a], sign, tmp);
          for (int j = 0; j < tn[a]; j++) buf[b0 + j * inner] = line[j];
        }
      inner *= tn[a];
    }
    for (long k = 0; k < tot; k++) {
      long off = base, r2 = k;
      for (int a = nt - 1; a >= 0; a--) {
        off += (r2 % tn[a]) * ts[a];
        r2 /= tn[a];
      }
      out[off] = buf[k];
    }
  }
  free(buf);
  free(line);
  free(tmp);
}

/* split_fft.c:426-458 (stage 1, factors F = B over decomp X = A) and
 * 491-519 (stage 2, F = A over X = B) for rank r. */
static void stage(const Dec *X, int r, const int *F, const double complex *in, double complex *out, int sign) {
  int n = X->ndim, nt = 0, nl = 0;
  int tn[OR_MAXD], ln[OR_MAXD];
  long ts[OR_MAXD], ls[OR_MAXD];
  for (int i = 0; i < n; i++) {
    long m = X->stride[r * n + i];
    if (F[i] > 1) {
      tn[nt] = F[i];
      ts[nt] = m;
      nt++;
    }
    int nloc = 0, pos = pos_of(X, r, i);
    for (int g = 0; g < X->L[i]; g++) nloc += X->rank_of[i][g] == pos;
    int count = nloc / F[i];
    if (count > 1) {
      ln[nl] = count;
      ls[nl] = m * F[i];
      nl++;
    }
  }
  if (nt == 0) return;
  if (nl == 0) {
    ln[0] = 1;
    ls[0] = 0;
    nl = 1;
  }
  guru(nt, tn, ts, nl, ln, ls, in, out, sign);
}

/* processor_grid.c:24-97 */
static void search_grid(int *best_len, int *best_load, double *best_asym, int ndim, int mydim, int *grid,
                        const int *tasks, int max_proc, int proc, int load, int mn, int mx) {
  int newproc = proc;
  for (grid[mydim] = 1; grid[mydim] <= max_proc && grid[mydim] <= tasks[mydim] && newproc <= max_proc;
       grid[mydim]++) {
    newproc = proc * grid[mydim];
    if (newproc <= max_proc) {
      int lo = tasks[mydim] / grid[mydim], hi = (tasks[mydim] + grid[mydim] - 1) / grid[mydim];
      int newload = load * hi;
      if (mn < lo) lo = mn;
      if (mx > hi) hi = mx;
      if (mydim < ndim - 1) {
        search_grid(best_len, best_load, best_asym, ndim, mydim + 1, grid, tasks, max_proc, newproc, newload, lo,
                    hi);
      } else {
        double ratio = (double)hi / (double)lo;
        if (newload <= *best_load || *best_load == 0) {
          if (*best_load == 0 || newload < *best_load || ratio < *best_asym) {
            *best_load = newload;
            *best_asym = ratio;
            for (int i = 0; i < ndim; i++) best_len[i] = grid[i];
          }
        }
      }
    }
  }
}

void or_processor_grid(int P, int ndim, const int *tasks, int *grid_out) {
  int load = 1, mx = 1, grid[OR_MAXD];
  double asym = 10000.0;
  for (int i = 0; i < ndim; i++) {
    grid_out[i] = 1;
    load *= tasks[i];
    if (tasks[i] > mx) mx = tasks[i];
  }
  search_grid(grid_out, &load, &asym, ndim, 0, grid, tasks, P, 1, 1, mx, 0);
}

/* ---------------------------------------------------------------- split FFT */

/* Spec node lists for the A and B decompositions of one axis
 * (split_fft.c:47-50 and 70). */
static void a_spec(int *s, int Ai, int Bi, int ga) {
  int t[10] = {OR_INTERLEAVE, Bi, 0, 0, 1, OR_BLOCK, Ai, 1, ga, -1};
  memcpy(s, t, sizeof t);
}
static void b_spec(int *s, int Ai, int Bi, int gb) {
  int t[5] = {OR_BLOCK, Ai * Bi, Ai, gb, -1};
  memcpy(s, t, sizeof t);
}

/* One forward or backward split transform over P ranks.
 *   start: specs/offs/rs/packed/expl as or_reshape; dat: P slots of
 *   dat_slot complex doubles; work: P slots of work_slot complex doubles,
 *   extent elements each half (extent from or_split_extent).
 * Forward: dat (start) -> work[0:extent) (B layout).  Backward: the exact
 * inverse, work -> dat.  Returns 0, or negative on a pairing error. */
int or_split_fft(int P, int ndim, const int *specs_s, const int *off_s, const int *rs_s, int packed_s,
                 const long *expl_s, const int *Af, const int *Bf, const int *ga, const int *gb, long extent,
                 double *dat, long dat_slot, double *work, long work_slot, int direction) {
  int L[OR_MAXD], use_first = 1, use_second = 1;
  int aspec[OR_MAXD * 10], bspec[OR_MAXD * 5], aoff[OR_MAXD], boff[OR_MAXD], ars[OR_MAXD], brs[OR_MAXD];
  for (int i = 0; i < ndim; i++) {
    L[i] = Af[i] * Bf[i];
    use_first = use_first && Af[i] == 1;
    use_second = use_second && Bf[i] == 1;
    a_spec(aspec + 10 * i, Af[i], Bf[i], ga[i]);
    aoff[i] = 10 * i;
    b_spec(bspec + 5 * i, Af[i], Bf[i], gb[i]);
    boff[i] = 5 * i;
  }
  if (use_second) return -3; /* the reference skips A->B here; not restated */
  /* packed rank strides: last axis fastest (reshape.c:228-235) */
  int r = 1, rb = 1;
  for (int i = ndim - 1; i >= 0; i--) {
    ars[i] = r;
    r *= ga[i];
    brs[i] = rb;
    rb *= gb[i];
  }
  Dec A, B;
  dec_init(&A, P, ndim, aspec, aoff, ars, 1, NULL);
  if (!use_first) dec_init(&B, P, ndim, bspec, boff, brs, 1, NULL);
  long es = 16;
  long wbytes = work_slot * es, dbytes = dat_slot * es;
  char *w0 = (char *)work, *w1 = (char *)(work + 2 * extent);
  int rc = 0;
  if (direction == -1) {
    rc = or_reshape(P, ndim, specs_s, off_s, rs_s, packed_s, expl_s, aspec, aoff, ars, 1, NULL, 16, (char *)dat,
                    dbytes, w1, wbytes);
    for (int q = 0; q < A.proc_used && rc == 0; q++)
      stage(&A, q, Bf, (double complex *)(w1 + q * wbytes), (double complex *)(w0 + q * wbytes), -1);
    if (!use_first && rc == 0) {
      for (int q = 0; q < A.proc_used; q++)
        or_decomp_scale(P, ndim, aspec, aoff, ars, q, Af, L, (double *)(w0 + q * wbytes), -1);
      rc = or_reshape(P, ndim, aspec, aoff, ars, 1, NULL, bspec, boff, brs, 1, NULL, 16, w0, wbytes, w1, wbytes);
      for (int q = 0; q < B.proc_used && rc == 0; q++)
        stage(&B, q, Af, (double complex *)(w1 + q * wbytes), (double complex *)(w0 + q * wbytes), -1);
    }
  } else {
    if (!use_first) {
      for (int q = 0; q < B.proc_used; q++)
        stage(&B, q, Af, (double complex *)(w0 + q * wbytes), (double complex *)(w1 + q * wbytes), +1);
      rc = or_reshape(P, ndim, bspec, boff, brs, 1, NULL, aspec, aoff, ars, 1, NULL, 16, w1, wbytes, w0, wbytes);
      for (int q = 0; q < A.proc_used && rc == 0; q++)
        or_decomp_scale(P, ndim, aspec, aoff, ars, q, Af, L, (double *)(w0 + q * wbytes), +1);
    }
    for (int q = 0; q < A.proc_used && rc == 0; q++)
      stage(&A, q, Bf, (double complex *)(w0 + q * wbytes), (double complex *)(w1 + q * wbytes), +1);
    if (rc == 0)
      rc = or_reshape(P, ndim, aspec, aoff, ars, 1, NULL, specs_s, off_s, rs_s, packed_s, expl_s, 16, w1, wbytes,
                      (char *)dat, dbytes);
  }
  dec_free(&A);
  if (!use_first) dec_free(&B);
  return rc;
}

/* extent = max over ranks of the A and B extents (split_fft.c:286-303) */
long or_split_extent(int P, int ndim, const int *Af, const int *Bf, const int *ga, const int *gb) {
  int aspec[OR_MAXD * 10] = {0}, bspec[OR_MAXD * 5] = {0}, aoff[OR_MAXD] = {0}, boff[OR_MAXD] = {0};
  int ars[OR_MAXD] = {0}, brs[OR_MAXD] = {0};
  int use_first = 1;
  int r = 1, rb = 1;
  for (int i = ndim - 1; i >= 0; i--) {
    a_spec(aspec + 10 * i, Af[i], Bf[i], ga[i]);
    aoff[i] = 10 * i;
    b_spec(bspec + 5 * i, Af[i], Bf[i], gb[i]);
    boff[i] = 5 * i;
    ars[i] = r;
    r *= ga[i];
    brs[i] = rb;
    rb *= gb[i];
    use_first = use_first && Af[i] == 1;
  }
  Dec A;
  dec_init(&A, P, ndim, aspec, aoff, ars, 1, NULL);
  long ext = 0;
  for (int q = 0; q < P; q++) ext = dec_extent(&A, q) > ext ? dec_extent(&A, q) : ext;
  dec_free(&A);
  if (!use_first) {
    Dec B;
    dec_init(&B, P, ndim, bspec, boff, brs, 1, NULL);
    for (int q = 0; q < P; q++) ext = dec_extent(&B, q) > ext ? dec_extent(&B, q) : ext;
    dec_free(&B);
  }
  return ext;
}
