/* The reference's split_plan_test.c driver logic (split_plan_test.c:240-330),
 * written again for this library's SPMD layer: one host thread per rank.
 * Each rank builds the packed block start decomposition on the processor
 * grid makeProcessorGrid chooses, allocates its own arrays, plans with
 * makeSplitPlan, fills its local elements with their global row-major
 * index (split_plan_test.c:60-110), writes the array to a file in global
 * order, runs the forward transform, prints the Timers line the reference
 * appends to split_timings_<size>l.txt, writes the transform through
 * getTransformed's decomposition, runs the backward transform and checks
 * that its elements came back scaled by size^3.
 *
 * The round-trip check has a tolerance: the reference compares a floating-
 * point round trip with exact equality (split_plan_test.c:141), which no
 * FFT meets in general.
 *
 * Usage: split_plan_test_c size npe [slab] [phase] [outdir]
 *   slab:  0 or 1 (makeSplitPlan's slab argument)
 *   phase: 1 runs the plan in the reference's phase order (every timer)
 *   outdir: where the initial and transformed files go (none: no files) */
#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "block_decomp.h"
#include "dm_runtime.h"
#include "dm_spmd.h"
#include "processor_grid.h"
#include "reshape.h"
#include "split_fft.h"
#include "timer.h"

#define FUDGE 4

static int size, npe, slab, phase;
static const char *outdir;
static long long total_errors;

/* every local (k, j, i) of d holds i + size (j + size k) at getOffset */
static void set_array(Decomp d, fftw_complex *host, long len) {
  for (long p = 0; p < len; p++) host[p] = -99.0 - 33.0 * I;
  for (int k = 0; k < size; k++) {
    if (!isLocal(d->dims[0], k)) continue;
    for (int j = 0; j < size; j++) {
      if (!isLocal(d->dims[1], j)) continue;
      for (int i = 0; i < size; i++) {
        if (!isLocal(d->dims[2], i)) continue;
        int coord[3] = {k, j, i};
        int pos = getOffset(d, coord);
        if (pos >= 0 && pos < len) host[pos] = (double)(i + size * (j + size * k)) + 0.0 * I;
      }
    }
  }
}

/* errors among the local elements of d: norm * value against the index */
static long long check_array(Decomp d, const fftw_complex *host, long len, double norm, double tol) {
  long long errors = 0;
  for (int k = 0; k < size; k++) {
    if (!isLocal(d->dims[0], k)) continue;
    for (int j = 0; j < size; j++) {
      if (!isLocal(d->dims[1], j)) continue;
      for (int i = 0; i < size; i++) {
        if (!isLocal(d->dims[2], i)) continue;
        int coord[3] = {k, j, i};
        int pos = getOffset(d, coord);
        double want = (double)(i + size * (j + size * k));
        if (pos < 0 || pos >= len) {
          errors++;
          continue;
        }
        fftw_complex got = norm * host[pos];
        if (cabs(got - want) > tol * (1.0 + fabs(want))) {
          if (errors < 5)
            fprintf(stderr, "Incorrect array value [%d,%d,%d] got (%g,%g) expected (%g,%g)\n", k, j, i,
                    creal(got), cimag(got), want, 0.0);
          errors++;
        }
      }
    }
  }
  return errors;
}

static void body(MPI_Comm me, void *arg) {
  (void)arg;
  int rank;
  MPI_Comm_rank(me, &rank);
  int L[3] = {size, size, size};
  long len = 2L * size * size * size / npe;
  fftw_complex *in = (fftw_complex *)dm_spmd_malloc(me, sizeof(fftw_complex) * (size_t)len);
  fftw_complex *out = (fftw_complex *)dm_spmd_malloc(me, sizeof(fftw_complex) * (size_t)len * 2 * FUDGE);
  fftw_complex *host = (fftw_complex *)malloc(sizeof(fftw_complex) * (size_t)len);
  int *grid = makeProcessorGrid(me, 3, L);
  if (rank == 0) printf("initial grid=[%d,%d,%d]\n", grid[0], grid[1], grid[2]);
  DimDecomp dims[3];
  for (int i = 0; i < 3; i++) dims[i] = makeBlockDecomp(size, grid[i]);
  Decomp start = makePackedDecomp(me, 3, dims);
  SplitPlan plan = makeSplitPlan(slab, start, in, (int)(2 * len * FUDGE), out);
  if (phase) dm_split_set_phase_order(plan, 1);
  if (rank == 0) printf("A=(%d,%d,%d)\n", plan->A_factors[0], plan->A_factors[1], plan->A_factors[2]);

  set_array(start, host, len);
  dm_spmd_put(me, in, 0, host, sizeof(fftw_complex) * (size_t)len);
  dm_spmd_get(me, host, in, 0, sizeof(fftw_complex) * (size_t)len);
  long long errors = check_array(start, host, len, 1.0, 0.0);
  char name[512];
  if (outdir) {
    snprintf(name, sizeof name, "%s/initial_split_%dl_%dp.dat", outdir, size, npe);
    if (dm_file_write(name, MPI_C_DOUBLE_COMPLEX, start, in) != 0) errors++;
  }
  dm_spmd_barrier(me);
  double t0 = timer();
  runSplitPlan(plan, FFTW_FORWARD);
  double t1 = timer();
  Decomp end = getTransformed(plan);
  if (rank == 0) {
    printf("elapsed time=%e\n", t1 - t0);
    /* the fields and order of split_timings_<size>l.txt */
    printf("split_timings %d %d %e %e %e %e %e %e\n", size, npe, plan->times[ALL], plan->times[RESHAPE_1],
           plan->times[FFT_1], plan->times[TWIDDLE], plan->times[RESHAPE_2], plan->times[FFT_2]);
  }
  if (outdir && end != NULL) {
    snprintf(name, sizeof name, "%s/parallel_split_%dl_%dp.dat", outdir, size, npe);
    if (dm_file_write(name, MPI_C_DOUBLE_COMPLEX, end, out) != 0) errors++;
  }
  runSplitPlan(plan, FFTW_BACKWARD);
  dm_spmd_get(me, host, in, 0, sizeof(fftw_complex) * (size_t)len);
  errors += check_array(start, host, len, 1.0 / ((double)size * size * size), 1e-9);
  long long sum = dm_spmd_allreduce_sum(me, errors);
  if (rank == 0) total_errors = sum;
  freeSplitPlan(plan);
  freeDecomp(start);
  free(grid);
  free(host);
  dm_spmd_free(me, in);
  dm_spmd_free(me, out);
}

int main(int argc, char **argv) {
  size = argc > 1 ? atoi(argv[1]) : 16;
  npe = argc > 2 ? atoi(argv[2]) : 4;
  slab = argc > 3 ? atoi(argv[3]) : 0;
  phase = argc > 4 ? atoi(argv[4]) : 0;
  outdir = argc > 5 ? argv[5] : NULL;
  MPI_Comm world = dm_world_create(npe);
  dm_spmd_run(world, body, NULL);
  dm_world_free(world);
  printf("size=%d npe=%d errors=%lld %s\n", size, npe, total_errors, total_errors ? "FAILED" : "OK");
  return total_errors != 0;
}
