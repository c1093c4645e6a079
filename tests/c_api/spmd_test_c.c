/* The reference's test.c driver logic (test.c:92-231), written again for
 * this library's SPMD layer: one host thread per rank, each building its
 * own decompositions with its own view, each passing its own local arrays
 * to the collective runReshapePlan, each checking its own elements with
 * isLocal/getOffset, errors summed over the ranks.  Usage: spmd_test_c npe. */
#include <stdio.h>
#include <stdlib.h>

#include "block_decomp.h"
#include "dm_runtime.h"
#include "dm_spmd.h"
#include "local_decomp.h"
#include "reshape.h"
#include "rotate_decomp.h"
#include "single_decomp.h"

#define SIZE 10

static int npe;
static long long total_errors;

/* test.c:16-45: every local (j, i) of d holds j + SIZE * i at getOffset */
static int check_slab(MPI_Comm me, Decomp d, const int *host) {
  int errors = 0;
  for (int i = 0; i < SIZE; i++) {
    if (!isLocal(d->dims[1], i)) continue;
    for (int j = 0; j < SIZE; j++) {
      if (!isLocal(d->dims[0], j)) continue;
      int coord[2] = {j, i};
      int pos = getOffset(d, coord);
      if (host[pos] != j + SIZE * i) errors++;
    }
  }
  (void)me;
  return errors;
}

static void body(MPI_Comm me, void *arg) {
  (void)arg;
  int rank;
  MPI_Comm_rank(me, &rank);
  int local = SIZE * ((SIZE + npe - 1) / npe);

  AxisDecomp oa[2] = {makeAxisDecomp(me, makeSingleDecomp(SIZE, 0, 1), 1, 1),
                      makeAxisDecomp(me, makeSingleDecomp(SIZE, 0, npe), SIZE, 1)};
  Decomp orig = makeDecomp(2, oa);
  AxisDecomp na[2] = {makeAxisDecomp(me, makeRotatedDecomp(2, makeLocalDecomp(SIZE)), 1, 1),
                      makeAxisDecomp(me, makeRotatedDecomp(1, makeBlockDecomp(SIZE, npe)), SIZE, 1)};
  Decomp fresh = makeDecomp(2, na);
  long long errors = 0;
  if (myRank(orig) != rank || myRank(fresh) != rank) errors++;

  /* local arrays: the global 10x10 on rank 0 only, destinations everywhere */
  int *src = (int *)dm_spmd_malloc(me, SIZE * SIZE * sizeof(int));
  int *dst = (int *)dm_spmd_malloc(me, (size_t)local * sizeof(int));
  int *d2 = (int *)dm_spmd_malloc(me, (size_t)local * sizeof(int));
  int *host = (int *)malloc(SIZE * SIZE * sizeof(int));
  if (rank == 0) {
    for (int i = 0; i < SIZE * SIZE; i++) host[i] = i;
    dm_spmd_put(me, src, 0, host, SIZE * SIZE * sizeof(int));
  }
  for (int i = 0; i < local; i++) host[i] = -rank;
  dm_spmd_put(me, dst, 0, host, (size_t)local * sizeof(int));
  dm_spmd_put(me, d2, 0, host, (size_t)local * sizeof(int));

  ReshapePlan plan = makeReshapePlan(me, MPI_INT, orig, fresh);
  runReshapePlan(me, plan, src, dst);
  dm_spmd_get(me, host, dst, 0, (size_t)local * sizeof(int));
  errors += check_slab(me, fresh, host);

  if (npe > 1 && npe % 2 == 0) {
    AxisDecomp da[2] = {makeAxisDecomp(me, makeRotatedDecomp(3, makeBlockDecomp(SIZE, 2)), 1, 1),
                        makeAxisDecomp(me, makeRotatedDecomp(3, makeBlockDecomp(SIZE, npe / 2)), SIZE / 2, 2)};
    Decomp dd = makeDecomp(2, da);
    ReshapePlan p2a = makeReshapePlan(me, MPI_INT, orig, dd);
    runReshapePlan(me, p2a, src, d2);
    dm_spmd_get(me, host, d2, 0, (size_t)local * sizeof(int));
    errors += check_slab(me, dd, host);
    for (int i = 0; i < local; i++) host[i] = -rank * 1000;
    dm_spmd_put(me, d2, 0, host, (size_t)local * sizeof(int));
    ReshapePlan p2b = makeReshapePlan(me, MPI_INT, fresh, dd);
    runReshapePlan(me, p2b, dst, d2);
    dm_spmd_get(me, host, d2, 0, (size_t)local * sizeof(int));
    errors += check_slab(me, dd, host);
    freeReshapePlan(p2a);
    freeReshapePlan(p2b);
    freeDecomp(dd);
  }
  long long sum = dm_spmd_allreduce_sum(me, errors);
  if (rank == 0) total_errors = sum;
  freeReshapePlan(plan);
  free(host);
  dm_spmd_free(me, src);
  dm_spmd_free(me, dst);
  dm_spmd_free(me, d2);
  freeDecomp(orig);
  freeDecomp(fresh);
}

int main(int argc, char **argv) {
  npe = argc > 1 ? atoi(argv[1]) : 4;
  MPI_Comm world = dm_world_create(npe);
  dm_spmd_run(world, body, NULL);
  dm_world_free(world);
  printf("npe=%d errors=%lld %s\n", npe, total_errors, total_errors ? "FAILED" : "OK");
  return total_errors != 0;
}
