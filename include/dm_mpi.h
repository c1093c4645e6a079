/*
 * dm_mpi.h -- the virtual-rank world that stands in for an MPI communicator.
 *
 * The reference API carries MPI_Comm, MPI_Datatype and MPI_Aint in its
 * signatures (reshape.h:11,105,272,293 and split_fft.h:11 in the reference).
 * On one MI355X the processor grid becomes P virtual ranks whose local arrays
 * are slots of one HBM buffer, so these types name objects of this library:
 *
 *   MPI_Comm      a view of a virtual world of P ranks, seen from one rank.
 *                 Collective calls (makeReshapePlan, runReshapePlan,
 *                 makeSplitPlan, runSplitPlan, ...) act for all P ranks at
 *                 once: call them once, from one host thread, with any view
 *                 of the world.  Rank-local queries (myRank, isLocal,
 *                 getOffset, ...) answer for the view's rank.
 *   MPI_Datatype  an element type: only its extent (bytes) matters to the
 *                 data movement, as in the reference where every message is
 *                 a nested type over one base type.
 *
 * The MPI_* spellings below are macros over dm_* symbols so that a program
 * can be linked against this library and a real MPI at the same time.
 * Define DM_NO_MPI_NAMES before including this header to get only the dm_*
 * names.
 */
#ifndef DM_MPI_H
#define DM_MPI_H

#include <stddef.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dm_comm *dm_comm_t;
typedef struct dm_datatype *dm_datatype_t;
typedef ptrdiff_t dm_aint_t;

#define DM_SUCCESS 0
#define DM_ERR_ARG 1

/* ---- virtual world ---------------------------------------------------- */

/** Create a world of nranks virtual ranks; returns its view from rank 0. */
dm_comm_t dm_world_create(int nranks);
/** The same world seen from virtual rank `rank` (owned by the world). */
dm_comm_t dm_comm_view(dm_comm_t comm, int rank);
/** Free a world and every view of it (pass any view). */
void dm_world_free(dm_comm_t comm);
int dm_comm_rank(dm_comm_t comm, int *rank);
int dm_comm_size(dm_comm_t comm, int *size);
/** Stop the process after an error: message on stderr, then abort(). */
void dm_abort(dm_comm_t comm, int code);

/* ---- element types ---------------------------------------------------- */

extern struct dm_datatype dm_type_byte, dm_type_char, dm_type_int,
    dm_type_long, dm_type_long_long, dm_type_float, dm_type_double,
    dm_type_c_float_complex, dm_type_c_double_complex;

int dm_type_contiguous(int count, dm_datatype_t old, dm_datatype_t *result);
int dm_type_dup(dm_datatype_t old, dm_datatype_t *result);
int dm_type_commit(dm_datatype_t *type);
int dm_type_free(dm_datatype_t *type);
int dm_type_get_extent(dm_datatype_t type, dm_aint_t *lb, dm_aint_t *extent);
int dm_type_size(dm_datatype_t type, int *size);

#ifndef DM_NO_MPI_NAMES
typedef dm_comm_t MPI_Comm;
typedef dm_datatype_t MPI_Datatype;
typedef dm_aint_t MPI_Aint;
#define MPI_SUCCESS DM_SUCCESS
#define MPI_BYTE (&dm_type_byte)
#define MPI_CHAR (&dm_type_char)
#define MPI_INT (&dm_type_int)
#define MPI_LONG (&dm_type_long)
#define MPI_LONG_LONG (&dm_type_long_long)
#define MPI_INT64_T (&dm_type_long_long)
#define MPI_FLOAT (&dm_type_float)
#define MPI_DOUBLE (&dm_type_double)
#define MPI_C_FLOAT_COMPLEX (&dm_type_c_float_complex)
#define MPI_C_DOUBLE_COMPLEX (&dm_type_c_double_complex)
#define MPI_Comm_rank dm_comm_rank
#define MPI_Comm_size dm_comm_size
#define MPI_Abort dm_abort
#define MPI_Type_contiguous dm_type_contiguous
#define MPI_Type_dup dm_type_dup
#define MPI_Type_commit dm_type_commit
#define MPI_Type_free dm_type_free
#define MPI_Type_get_extent dm_type_get_extent
#define MPI_Type_size dm_type_size
#endif

#ifdef __cplusplus
}
#endif
#endif /* DM_MPI_H */
