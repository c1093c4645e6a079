/* The reference's plan_compare.c driver logic (plan_compare.c:270-391),
 * written again for this library's SPMD layer: one host thread per rank.
 * Each rank builds the packed block start decomposition, a slab plan
 * (makeParallelPlan with B = (1, size, size)) and an autotuned plan
 * (autotuneParallelPlan, collective), times both forward transforms over
 * TRIAL trials of REPEAT runs with the largest time over the ranks kept
 * (MPI_Allreduce MAX), and rank 0 prints the line the reference appends to
 * compare_results_<size>.txt: both plans' A factors and the proc_size of
 * every axis of their A and B decompositions (struct fields read directly,
 * as plan_compare.c:373-388 reads them), worst and best time per run.
 * Then the autotuned plan transforms an array of global row-major indices
 * and writes the result through getTransformed's decomposition.
 *
 * Usage: plan_compare_c size npe [outdir] */
#include <complex.h>
#include <stdio.h>
#include <stdlib.h>

#include "block_decomp.h"
#include "dm_runtime.h"
#include "dm_spmd.h"
#include "processor_grid.h"
#include "reshape.h"
#include "split_fft.h"
#include "timer.h"

#define FUDGE 4
#define TRIAL 3
#define REPEAT 4

static int size, npe;
static const char *outdir;
static int failed;

static void time_plan(MPI_Comm me, SplitPlan p, double *worst, double *best) {
  *worst = 0.0;
  *best = 1e9;
  for (int trial = 0; trial < TRIAL; trial++) {
    dm_spmd_barrier(me);
    double t0 = timer();
    for (int i = 0; i < REPEAT; i++) runSplitPlan(p, FFTW_FORWARD);
    double el = dm_spmd_allreduce_max(me, timer() - t0);
    if (el > *worst) *worst = el;
    if (el < *best) *best = el;
  }
}

static void body(MPI_Comm me, void *arg) {
  (void)arg;
  int rank;
  MPI_Comm_rank(me, &rank);
  int L[3] = {size, size, size};
  long len = 2L * size * size * size / npe;
  fftw_complex *in = (fftw_complex *)dm_spmd_malloc(me, sizeof(fftw_complex) * (size_t)len);
  fftw_complex *out = (fftw_complex *)dm_spmd_malloc(me, sizeof(fftw_complex) * (size_t)len * 2 * FUDGE);
  int *grid = makeProcessorGrid(me, 3, L);
  DimDecomp dims[3];
  for (int i = 0; i < 3; i++) dims[i] = makeBlockDecomp(size, grid[i]);
  Decomp start = makePackedDecomp(me, 3, dims);
  int slab_B[3] = {1, size, size};
  SplitPlan slab_plan = makeParallelPlan(start, slab_B, in, (int)(2 * len * FUDGE), out);
  SplitPlan auto_plan = autotuneParallelPlan(start, in, (int)(2 * len * FUDGE), out);
  if (slab_plan == NULL || auto_plan == NULL) {
    if (rank == 0) failed = 1;
    return;
  }
  if (rank == 0)
    printf("Auto plan is: A[0]=%d A[1]=%d A[2]=%d\n", auto_plan->A_factors[0], auto_plan->A_factors[1],
           auto_plan->A_factors[2]);
  double worst_slab, best_slab, worst_auto, best_auto;
  time_plan(me, slab_plan, &worst_slab, &best_slab);
  time_plan(me, auto_plan, &worst_auto, &best_auto);
  if (rank == 0) {
    /* the fields and order of compare_results_<size>.txt */
    printf("compare npe %d initial-grid %d %d %d slab-A %d %d %d slab-grid-a %d %d %d slab-grid-b %d %d %d "
           "slab time %.2e %.2e auto-A %d %d %d auto-grid-A %d %d %d auto-grid-B %d %d %d auto-time %.2e %.2e\n",
           npe, grid[0], grid[1], grid[2], slab_plan->A_factors[0], slab_plan->A_factors[1],
           slab_plan->A_factors[2], slab_plan->A->dims[0]->decomp->proc_size,
           slab_plan->A->dims[1]->decomp->proc_size, slab_plan->A->dims[2]->decomp->proc_size,
           slab_plan->B->dims[0]->decomp->proc_size, slab_plan->B->dims[1]->decomp->proc_size,
           slab_plan->B->dims[2]->decomp->proc_size, worst_slab / REPEAT, best_slab / REPEAT,
           auto_plan->A_factors[0], auto_plan->A_factors[1], auto_plan->A_factors[2],
           auto_plan->A->dims[0]->decomp->proc_size, auto_plan->A->dims[1]->decomp->proc_size,
           auto_plan->A->dims[2]->decomp->proc_size, auto_plan->B ? auto_plan->B->dims[0]->decomp->proc_size : 0,
           auto_plan->B ? auto_plan->B->dims[1]->decomp->proc_size : 0,
           auto_plan->B ? auto_plan->B->dims[2]->decomp->proc_size : 0, worst_auto / REPEAT, best_auto / REPEAT);
  }
  if (outdir) {
    /* the autotuned plan's transform of the global row-major indices */
    fftw_complex *host = (fftw_complex *)calloc((size_t)len, sizeof(fftw_complex));
    for (int k = 0; k < size; k++) {
      if (!isLocal(start->dims[0], k)) continue;
      for (int j = 0; j < size; j++) {
        if (!isLocal(start->dims[1], j)) continue;
        for (int i = 0; i < size; i++) {
          if (!isLocal(start->dims[2], i)) continue;
          int coord[3] = {k, j, i};
          host[getOffset(start, coord)] = (double)(i + size * (j + size * k));
        }
      }
    }
    dm_spmd_put(me, in, 0, host, sizeof(fftw_complex) * (size_t)len);
    free(host);
    runSplitPlan(auto_plan, FFTW_FORWARD);
    char name[512];
    snprintf(name, sizeof name, "%s/parallel_split_%d-%d-%d_%dl_%dp.dat", outdir, auto_plan->A_factors[0],
             auto_plan->A_factors[1], auto_plan->A_factors[2], size, npe);
    if (dm_file_write(name, MPI_C_DOUBLE_COMPLEX, getTransformed(auto_plan), out) != 0 && rank == 0) failed = 1;
  }
  freeSplitPlan(slab_plan);
  freeSplitPlan(auto_plan);
  freeDecomp(start);
  free(grid);
  dm_spmd_free(me, in);
  dm_spmd_free(me, out);
}

int main(int argc, char **argv) {
  size = argc > 1 ? atoi(argv[1]) : 16;
  npe = argc > 2 ? atoi(argv[2]) : 4;
  outdir = argc > 3 ? argv[3] : NULL;
  MPI_Comm world = dm_world_create(npe);
  dm_spmd_run(world, body, NULL);
  dm_world_free(world);
  printf("size=%d npe=%d %s\n", size, npe, failed ? "FAILED" : "OK");
  return failed;
}
