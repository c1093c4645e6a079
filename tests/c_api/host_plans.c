/* Host-side planning over many shapes, virtual rank counts and element
 * sizes, written for a build of the C layer with the address and undefined
 * behaviour sanitizers: processor grids, packed block / pencil / slab
 * decompositions, reshape plans between every pair of them in both
 * directions (a pair whose fastest axis is whole on one side and split on
 * the other included), split FFT plans in both precisions from the slab
 * and the sqrt-factor planners, one refused plan in error mode 1, and every
 * object freed again.  Registered host memory stands in for the device
 * buffers, which planning never reads: nothing here needs a GPU.  Last,
 * the 512^3 shapes on 32 ranks (the benchmark's; byte offsets above 2^31)
 * over address space reserved without access rights, so that any read of
 * the buffers while planning would fault.
 * Prints the number of plans made and "OK" when every check held. */
#define _DEFAULT_SOURCE
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/mman.h>

#include "block_decomp.h"
#include "dm_batch.h"
#include "dm_runtime.h"
#include "local_decomp.h"
#include "processor_grid.h"
#include "reshape.h"
#include "split_fft.h"

static int failures = 0, plans = 0;

#define CHECK(cond, ...)                  \
  do {                                    \
    if (!(cond)) {                        \
      failures++;                         \
      fprintf(stderr, "FAILED: " __VA_ARGS__); \
      fprintf(stderr, "\n");              \
    }                                     \
  } while (0)

/* axis `whole` local, the other two blocked over a 2D processor grid */
static Decomp pencil(MPI_Comm world, const int L[3], int whole) {
  int b = (whole + 1) % 3, c = (whole + 2) % 3;
  int lo = b < c ? b : c, hi = b < c ? c : b;
  int tasks[2] = {L[lo], L[hi]};
  int *g = makeProcessorGrid(world, 2, tasks);
  DimDecomp dd[3];
  dd[whole] = makeLocalDecomp(L[whole]);
  dd[lo] = makeBlockDecomp(L[lo], g[0]);
  dd[hi] = makeBlockDecomp(L[hi], g[1]);
  free(g);
  return makePackedDecomp(world, 3, dd);
}

/* the first axis blocked over every rank, the others local */
static Decomp slab(MPI_Comm world, const int L[3], int axis, int P) {
  DimDecomp dd[3];
  for (int i = 0; i < 3; i++) dd[i] = (i == axis) ? makeBlockDecomp(L[i], P) : makeLocalDecomp(L[i]);
  return makePackedDecomp(world, 3, dd);
}

static Decomp packed_block(MPI_Comm world, const int L[3]) {
  int *g = makeProcessorGrid(world, 3, (int *)L);
  DimDecomp dd[3];
  for (int i = 0; i < 3; i++) dd[i] = makeBlockDecomp(L[i], g[i]);
  free(g);
  return makePackedDecomp(world, 3, dd);
}

static void reshape_pair(MPI_Comm world, MPI_Datatype t, int es, Decomp a, Decomp b, const char *what) {
  ReshapePlan f = makeReshapePlan(world, t, a, b);
  ReshapePlan r = makeReshapePlan(world, t, b, a);
  CHECK(f != NULL && r != NULL, "reshape %s (element %d B)", what, es);
  if (f != NULL) CHECK(checkReshapePlanBounds(f, es, extent(a), extent(b)) == 0, "bounds %s", what);
  if (r != NULL) CHECK(checkReshapePlanBounds(r, es, extent(b), extent(a)) == 0, "bounds back %s", what);
  plans += (f != NULL) + (r != NULL);
  freeReshapePlan(f);
  freeReshapePlan(r);
}

static void *reserve(size_t bytes) {
  void *p = mmap(NULL, bytes, PROT_NONE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0);
  return p == MAP_FAILED ? NULL : p;
}

static void split_plans(MPI_Comm world, Decomp start, int P, int untouched) {
  int per = extent(start);
  int length = untouched ? 2 * per : 4 * per + 4096;
  size_t es = sizeof(fftw_complex);
  void *dat = untouched ? reserve((size_t)P * per * es) : calloc((size_t)P * per, es);
  void *work = untouched ? reserve((size_t)P * length * es) : calloc((size_t)P * length, es);
  CHECK(dat != NULL && work != NULL, "buffers");
  if (dat == NULL || work == NULL) return;
  CHECK(dm_register(world, dat, (size_t)per * es) == 0, "register dat");
  CHECK(dm_register(world, work, (size_t)length * es) == 0, "register work");
  for (int s = 0; s < 2; s++) {
    SplitPlan d = makeSplitPlan(s, start, (fftw_complex *)dat, length, (fftw_complex *)work);
    SplitPlan f = makeSplitPlanf(s, start, (fftwf_complex *)dat, length, (fftwf_complex *)work);
    CHECK(d != NULL && f != NULL, "split plan slab=%d P=%d", s, P);
    if (d != NULL) {
      CHECK(d->extent <= length, "split extent %d > %d", d->extent, length);
      CHECK(getTransformed(d) != NULL || d->A_to_B == NULL, "transformed decomposition");
      CHECK(dm_split_set_twiddles(d, DM_TWIDDLES_PLAIN) == 0, "twiddle mode on FP64 (no effect)");
    }
    if (f != NULL) {
      CHECK(dm_split_set_twiddles(f, DM_TWIDDLES_PLAIN) == 0, "plain twiddles on FP32");
      CHECK(dm_split_twiddles(f) == DM_TWIDDLES_PLAIN, "twiddle mode read back");
    }
    plans += (d != NULL) + (f != NULL);
    freeSplitPlan(d);
    freeSplitPlan(f);
  }
  dm_unregister(dat);
  dm_unregister(work);
  if (untouched) {
    munmap(dat, (size_t)P * per * es);
    munmap(work, (size_t)P * length * es);
  } else {
    free(dat);
    free(work);
  }
}

int main(void) {
  static const int shapes[][3] = {{64, 64, 64}, {96, 80, 112}, {48, 40, 36}, {30, 42, 70}, {128, 16, 8}};
  static const int ranks[] = {1, 2, 3, 4, 6, 8};
  MPI_Datatype c48;
  MPI_Type_contiguous(3, MPI_C_DOUBLE_COMPLEX, &c48);
  MPI_Type_commit(&c48);
  MPI_Datatype types[] = {MPI_FLOAT, MPI_DOUBLE, MPI_C_DOUBLE_COMPLEX, c48};
  const int sizes[] = {4, 8, 16, 48};

  for (size_t si = 0; si < sizeof shapes / sizeof shapes[0]; si++) {
    const int *L = shapes[si];
    for (size_t ri = 0; ri < sizeof ranks / sizeof ranks[0]; ri++) {
      int P = ranks[ri];
      MPI_Comm world = dm_world_create(P);
      Decomp start = packed_block(world, L);
      Decomp pen[3], sl[3];
      for (int a = 0; a < 3; a++) {
        pen[a] = pencil(world, L, a);
        sl[a] = slab(world, L, a, P);
      }
      for (int t = 0; t < 4; t++) {
        char what[64];
        for (int a = 0; a < 3; a++) {
          snprintf(what, sizeof what, "start<->pencil%d P=%d", a, P);
          reshape_pair(world, types[t], sizes[t], start, pen[a], what);
          for (int b = a + 1; b < 3; b++) {
            snprintf(what, sizeof what, "pencil%d<->pencil%d P=%d", a, b, P);
            reshape_pair(world, types[t], sizes[t], pen[a], pen[b], what);
            /* slab over the first axis against slab over the last: the
             * last (fastest) axis is whole on one side, split on the other */
            snprintf(what, sizeof what, "slab%d<->slab%d P=%d", a, b, P);
            reshape_pair(world, types[t], sizes[t], sl[a], sl[b], what);
          }
        }
      }
      split_plans(world, start, P, 0);

      /* a reshape between arrays of different shapes is refused */
      int M[3] = {L[0] + 1, L[1], L[2]};
      Decomp other = packed_block(world, M);
      dm_set_error_mode(1);
      ReshapePlan bad = makeReshapePlan(world, MPI_DOUBLE, start, other);
      CHECK(bad == NULL && dm_last_error() != NULL && strstr(dm_last_error(), "not compatible") != NULL,
            "incompatible reshape refused");
      dm_clear_error();
      dm_set_error_mode(0);
      freeDecomp(other);

      for (int a = 0; a < 3; a++) {
        freeDecomp(pen[a]);
        freeDecomp(sl[a]);
      }
      freeDecomp(start);
      dm_world_free(world);
    }
  }
  /* the benchmark's 512^3 on 32 ranks: the pencil start, the packed block
   * start, a z-pencil, split plans in both precisions, 16-byte reshapes */
  {
    const int L[3] = {512, 512, 512}, P = 32;
    MPI_Comm world = dm_world_create(P);
    Decomp start = packed_block(world, L);
    Decomp pen0 = pencil(world, L, 0), pen2 = pencil(world, L, 2);
    reshape_pair(world, MPI_C_DOUBLE_COMPLEX, 16, start, pen2, "512^3 start<->pencil2");
    reshape_pair(world, MPI_C_DOUBLE_COMPLEX, 16, pen0, pen2, "512^3 pencil0<->pencil2");
    split_plans(world, pen0, P, 1);
    split_plans(world, start, P, 1);
    freeDecomp(start);
    freeDecomp(pen0);
    freeDecomp(pen2);
    dm_world_free(world);
  }
  /* batched plans (host side only: validation, schedules, byte counts) */
  {
    static const int shapes[] = {17, 34, 16, 19, 32, 23, 64, 96, 128, 4096, 1, 1, 2, 16, 289};
    for (int dbl = 0; dbl < 2; dbl++) {
      DmBatch b = dm_batch_plan(5, 3, shapes, dbl, -1);
      CHECK(b != NULL, "batch plan precision %d", dbl);
      if (b == NULL) continue;
      const double alg = dm_batch_bytes(b);
      CHECK(dm_batch_bytes_moved(b) == alg, "line tiles move the algorithmic bytes");
      CHECK(dm_batch_set_schedule(b, DM_BATCH_TILE_PLANES, DM_BATCH_RADIX_SMALL_FIRST) == 0, "plane schedule");
      CHECK(dm_batch_bytes_moved(b) < alg, "planes move fewer bytes");
      dm_batch_free(b);
      plans++;
    }
    static const int too_long[] = {8192, 2, 2};
    dm_set_error_mode(1);
    CHECK(dm_batch_plan(1, 3, too_long, 0, -1) == NULL, "an axis longer than the batched tile is refused");
    dm_clear_error();
    dm_set_error_mode(0);
  }
  MPI_Type_free(&c48);
  printf("plans=%d\n%s\n", plans, failures ? "FAILED" : "OK");
  return failures != 0;
}
