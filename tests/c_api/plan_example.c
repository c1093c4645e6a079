/* A C caller of the library's reference-compatible API, written the way
 * the reference's drivers are (test.c:92-180, split_plan_test.c:240-330):
 * build decompositions, plan a reshape and a split FFT, read the plan's
 * fields.  Plans only: nothing here needs a GPU.  Compiled by the CPU test
 * suite twice, with and without <complex.h> ahead of the headers, to check
 * that the headers are valid C in both fftw_complex forms. */
#ifdef WITH_COMPLEX_H
#include <complex.h>
#endif
#include <stdio.h>
#include <stdlib.h>

#include "block_decomp.h"
#include "dm_runtime.h"
#include "local_decomp.h"
#include "processor_grid.h"
#include "reshape.h"
#include "rotate_decomp.h"
#include "single_decomp.h"
#include "split_fft.h"

#define SIZE 10

int main(void) {
  int npe = 4, failures = 0;
  MPI_Comm world = dm_world_create(npe);

  /* test.c:118-150: a 10x10 array on rank 0 -> rotated local/block */
  AxisDecomp o[2], n[2];
  o[0] = makeAxisDecomp(world, makeSingleDecomp(SIZE, 0, 1), 1, 1);
  o[1] = makeAxisDecomp(world, makeSingleDecomp(SIZE, 0, npe), SIZE, 1);
  Decomp orig = makeDecomp(2, o);
  /* makeRotatedDecomp keeps a clone of what it wraps (rotate_decomp.c:47) */
  DimDecomp inner0 = makeLocalDecomp(SIZE), inner1 = makeBlockDecomp(SIZE, npe);
  n[0] = makeAxisDecomp(world, makeRotatedDecomp(2, inner0), 1, 1);
  n[1] = makeAxisDecomp(world, makeRotatedDecomp(1, inner1), SIZE, 1);
  freeDimDecomp(inner0);
  freeDimDecomp(inner1);
  Decomp dest = makeDecomp(2, n);
  ReshapePlan plan = makeReshapePlan(world, MPI_INT, orig, dest);
  if (plan == NULL) failures++;
  if (checkReshapePlanBounds(plan, sizeof(int), SIZE * SIZE, SIZE * ((SIZE + npe - 1) / npe)) != 0) failures++;
  int rank = myRank(dest);
  printf("reshape plan ok, myRank=%d extent=%d\n", rank, extent(dest));

  /* split_plan_test.c: processor grid, packed start, default split plan */
  int L[3] = {64, 64, 64};
  int *grid = makeProcessorGrid(world, 3, L);
  DimDecomp dd[3];
  for (int i = 0; i < 3; i++) dd[i] = makeBlockDecomp(L[i], grid[i]);
  Decomp start = makePackedDecomp(world, 3, dd);
  int per = extent(start);
  int length = 2 * per + 1024;
  /* registered memory stands in for device buffers: planning never reads it */
  void *dat = calloc((size_t)npe * per, sizeof(fftw_complex));
  void *work = calloc((size_t)npe * length, sizeof(fftw_complex));
  dm_register(world, dat, (size_t)per * sizeof(fftw_complex));
  dm_register(world, work, (size_t)length * sizeof(fftw_complex));
  SplitPlan sp = makeSplitPlan(0, start, (fftw_complex *)dat, length, (fftw_complex *)work);
  if (sp == NULL || getTransformed(sp) == NULL) failures++;
  printf("grid=(%d,%d,%d) A=(%d,%d,%d) extent=%d timers=%d\n", grid[0], grid[1], grid[2], sp->A_factors[0],
         sp->A_factors[1], sp->A_factors[2], sp->extent, NUM_TIMERS);
  if (sp->A_factors[0] != 8 || grid[0] != 1 || grid[1] != 2 || grid[2] != 2) failures++;
  freeSplitPlan(sp);
  dm_unregister(dat);
  dm_unregister(work);
  free(dat);
  free(work);
  free(grid);
  freeReshapePlan(plan);
  /* each Decomp owns its axes and their DimDecomps */
  freeDecomp(start);
  freeDecomp(orig);
  freeDecomp(dest);
  dm_world_free(world);
  printf("%s\n", failures ? "FAILED" : "OK");
  return failures != 0;
}
