/*
 * dm_spmd.h -- run a driver written the MPI way: one host thread per rank.
 *
 * No counterpart in the reference (which is MPI).  dm_spmd_run starts P
 * threads, thread r calling body(view of rank r, arg).  Inside a session
 * the reference API is used as under MPI: every rank builds its own
 * decompositions and plans with its view, and the collective calls
 * (runReshapePlan, runSplitPlan, autotuneParallelPlan, dm_file_write,
 * dm_file_read) are made by every rank: they rendezvous, rank 0's call runs
 * the one launch that moves every rank's share, and all ranks return when
 * it has finished.  Each rank passes pointers into its own slot of a
 * distributed buffer (dm_spmd_malloc), as it would pass its own local array
 * under MPI; the ranks' arguments are checked to agree (the same plan or
 * decomposition, their own slots of the same buffers at the same offset),
 * and a call whose arguments differ raises an error on every rank.
 */
#ifndef DM_SPMD_H
#define DM_SPMD_H
#include <stddef.h>
#include "dm_mpi.h"
#ifdef __cplusplus
extern "C" {
#endif

/** Runs body on world's P ranks, one thread each; returns when all return. */
int dm_spmd_run(MPI_Comm world, void (*body)(MPI_Comm me, void *arg), void *arg);
/** All ranks of the session wait for each other (MPI_Barrier). */
void dm_spmd_barrier(MPI_Comm me);
/** Sum of value over the ranks, returned to every rank (MPI_Allreduce). */
long long dm_spmd_allreduce_sum(MPI_Comm me, long long value);
/** Largest value over the ranks, returned to every rank (MPI_Allreduce
 *  with MPI_MAX, as the reference's timing loops use it). */
double dm_spmd_allreduce_max(MPI_Comm me, double value);
/** Collective: a distributed buffer of bytes per rank; each rank gets the
 *  pointer to its own slot. */
void *dm_spmd_malloc(MPI_Comm me, size_t bytes_per_rank);
/** Collective: frees the buffer (each rank passes its own slot pointer). */
void dm_spmd_free(MPI_Comm me, void *mine);
/** Copies between host memory and this rank's own slot (offset bytes in). */
int dm_spmd_put(MPI_Comm me, void *mine, size_t offset, const void *host, size_t bytes);
int dm_spmd_get(MPI_Comm me, void *host, const void *mine, size_t offset, size_t bytes);

#ifdef __cplusplus
}
#endif
#endif
