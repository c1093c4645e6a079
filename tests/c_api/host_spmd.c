/* The SPMD layer's host-side synchronisation, written for a build of the C
 * layer with the thread sanitizer: P threads (one per virtual rank) call
 * timer() (whose first call from any thread sets the base) and
 * the all-reduces and the barrier many times in a row, each builds and
 * frees its own decompositions and reshape plans between them, and in one
 * session a collective reshape whose ranks pass different destination
 * offsets is refused on every rank (the argument check runs before any
 * device work, so nothing here needs a GPU).  Prints "OK" when every rank
 * saw the expected sums, maxima and errors. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "block_decomp.h"
#include "dm_runtime.h"
#include "dm_spmd.h"
#include "local_decomp.h"
#include "reshape.h"
#include "timer.h"

#define P 6
#define ROUNDS 2000

static int bad[P];
static char errors[P][256];
static char slots[P][4096];

static void body(MPI_Comm me, void *arg) {
  (void)arg;
  int r;
  MPI_Comm_rank(me, &r);
  const double t0 = timer();
  for (int i = 0; i < ROUNDS; i++) {
    long long s = dm_spmd_allreduce_sum(me, r + i);
    if (s != (long long)P * (P - 1) / 2 + (long long)P * i) bad[r]++;
    double m = dm_spmd_allreduce_max(me, (double)(r * 10 + i));
    if (m != (double)((P - 1) * 10 + i)) bad[r]++;
    if (i % 16 == 0) {
      DimDecomp a[2] = {makeBlockDecomp(24 + i % 5, P), makeLocalDecomp(7)};
      DimDecomp b[2] = {makeLocalDecomp(24 + i % 5), makeBlockDecomp(7, P)};
      Decomp da = makePackedDecomp(me, 2, a);
      Decomp db = makePackedDecomp(me, 2, b);
      ReshapePlan plan = makeReshapePlan(me, MPI_DOUBLE, da, db);
      if (plan == NULL || checkReshapePlanBounds(plan, 8, extent(da), extent(db)) != 0) bad[r]++;
      freeReshapePlan(plan);
      freeDecomp(da);
      freeDecomp(db);
    }
    dm_spmd_barrier(me);
  }
  if (!(timer() >= t0 && t0 >= 0.0)) bad[r]++;
  /* rank 1 passes its destination 16 bytes further in than the others */
  DimDecomp dd[1] = {makeBlockDecomp(12, P)};
  Decomp d = makePackedDecomp(me, 1, dd);
  ReshapePlan plan = makeReshapePlan(me, MPI_INT, d, d);
  char *mine = slots[r];
  dm_clear_error();
  runReshapePlan(me, plan, mine, mine + 2048 + (r == 1 ? 16 : 0));
  const char *e = dm_last_error();
  snprintf(errors[r], sizeof errors[r], "%s", e ? e : "");
  freeReshapePlan(plan);
  freeDecomp(d);
}

int main(void) {
  MPI_Comm world = dm_world_create(P);
  int failures = 0;
  if (dm_register(world, slots, sizeof slots[0]) != 0) failures++;
  dm_set_error_mode(1);
  if (dm_spmd_run(world, body, NULL) != 0) failures++;
  dm_set_error_mode(0);
  for (int r = 0; r < P; r++) {
    if (bad[r]) failures++;
    if (strstr(errors[r], "the ranks' arguments differ") == NULL) failures++;
  }
  dm_unregister(slots);
  dm_world_free(world);
  printf("%s\n", failures ? "FAILED" : "OK");
  return failures != 0;
}
