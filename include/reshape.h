/*
 * reshape.h -- decomposition descriptors and decomposition-driven reshape.
 *
 * Drop-in for the reference's reshape.h (reshape.h:1-302).  Names, argument
 * meaning, struct field names and error behaviour are the reference's; the
 * implementation is this library's: a reshape plan compiles to a device
 * gather program over per-axis lookup tables and runs as HIP kernels on one
 * MI355X, where the reference posts MPI_Isend/Irecv of nested datatypes.
 *
 * Structs carry one extra trailing field (`ext` / `prog`) that holds the
 * per-virtual-rank tables of the global view; callers never touch it.
 */
#ifndef RESHAPE
#define RESHAPE
#include <stdio.h>
#include "dm_mpi.h"
#include "index.h"

#ifdef __cplusplus
extern "C" {
#endif

/** Print "<rank>: Error in reshape:<msg>" on stderr and abort the process
 *  (reference reshape.c:12-21 calls MPI_Abort(comm,555)).  Returns only
 *  after dm_set_error_mode(1), which records the message instead. */
void raise_error(MPI_Comm comm, const char *msg, ...);

typedef struct dim_decomp *DimDecomp;

/** global coordinate -> position along the processor axis (reshape.h:21) */
typedef int (*rank_mapper)(DimDecomp, int);
/** global coordinate -> local index in units of the axis memory stride
 *  (reshape.h:26) */
typedef int (*offset_mapper)(DimDecomp, int);
typedef void (*print_decomp)(FILE *, DimDecomp);
typedef DimDecomp (*clone_decomp)(DimDecomp);

/** One axis' decomposition strategy (reshape.h:45-54).  User strategies may
 *  fill the function table themselves; allocate with calloc. */
struct dim_decomp {
  rank_mapper rank_offset;
  offset_mapper type_offset;
  void (*destroy)(DimDecomp);
  print_decomp printer;
  clone_decomp cloner;
  char *type;
  int coord_size;
  int proc_size;
};
void freeDimDecomp(DimDecomp);
DimDecomp cloneDimDecomp(DimDecomp);
int sizeDimDecomp(DimDecomp d);
void describeDimDecomp(FILE *out, DimDecomp d);

typedef struct axis_decomp *AxisDecomp;
struct dm_axis_ext;
/** A DimDecomp placed in memory and on the rank grid (reshape.h:75-86).
 *  Fields describe the comm view's rank; `ext` holds every position's. */
struct axis_decomp {
  DimDecomp decomp;
  MPI_Comm comm;
  int memory_stride;
  int rank_stride;
  int my_rank_in_dimension;
  int n_local;
  int max_local;
  Index index;
  struct dm_axis_ext *ext;
};

Index *makeIndexList(AxisDecomp local, DimDecomp remote);
/** reshape.c:137-148: runs of the global coordinates local to this rank. */
Index makeGlobalIndex(AxisDecomp local);
/** reshape.c:158-172: the view of one axis -- src placed at g * extent *
 *  stride bytes for every local coordinate g, ascending; returns
 *  stride * coord_size. */
int expandIOType(MPI_Datatype *result, MPI_Aint extent, MPI_Datatype src,
                 AxisDecomp mine, int stride);
AxisDecomp makeAxisDecomp(MPI_Comm comm, DimDecomp d, int stride,
                          int rank_stride);
void freeAxisDecomp(AxisDecomp);
void describeAxisDecomp(FILE *out, AxisDecomp d);
int isLocal(AxisDecomp d, int global);

typedef struct decomp *Decomp;
struct dm_decomp_ext;
/** Multi-axis decomposition (reshape.h:134-142). */
struct decomp {
  int ndim;
  MPI_Comm comm;
  int proc_used;
  AxisDecomp *dims;
  struct dm_decomp_ext *ext;
};

Decomp makeDecomp(int ndim, AxisDecomp dims[]);
Decomp makePackedDecomp(MPI_Comm comm, int ndim, DimDecomp dims[]);
void freeDecomp(Decomp d);
void describeDeomp(FILE *out, Decomp d);
int extent(Decomp d);
int compatible(Decomp a, Decomp b);
#define COMPATIBLE 1
#define NOT_COMPATIBLE 0
/** Declared by the reference (reshape.h:177) but defined nowhere there.
 *  Here: compatible, and every global coordinate has the same owner rank
 *  and local offset on every virtual rank in both. */
int equivalent(Decomp a, Decomp b);
int myRank(Decomp);
int isCoordLocal(Decomp d, int dim, int pos);
int areCoordsLocal(Decomp d, int *pos);
int getOffset(Decomp d, int *pos);

/* ---- message lists (reshape.h:206-251) -------------------------------- */
struct type_entry {
  int rank;
  MPI_Datatype type;
};
struct type_entry_list {
  int size;
  int count;
  struct type_entry *list;
};
typedef struct type_entry_list *TypeEntryList;
#define INITIAL_TYPELIST_SIZE 16
TypeEntryList makeTypeEntryList(void);
void commitTypeEntryList(TypeEntryList);
void freeTypeEntryList(TypeEntryList);
void addTypeEntry(TypeEntryList list, int rank, MPI_Datatype type);
TypeEntryList expandList(int max_proc, MPI_Aint extent, TypeEntryList list,
                         AxisDecomp mine, AxisDecomp remote);
void printTypeEntryList(FILE *out, TypeEntryList list);

struct dm_reshape_program;
/** A compiled reshape (reshape.h:255-262).  sends/recvs list the view
 *  rank's messages (peer rank and a datatype with the message's element
 *  count and true extent); `prog` is the device gather program. */
struct reshape_plan {
  MPI_Comm comm;
  Decomp src;
  Decomp dest;
  int *order;
  TypeEntryList sends;
  TypeEntryList recvs;
  struct dm_reshape_program *prog;
};
typedef struct reshape_plan *ReshapePlan;

/** Build a plan moving elements of type `base` from the src to the dest
 *  decomposition (reshape.h:272).  Aliases src and dest: keep them alive. */
ReshapePlan makeReshapePlan(MPI_Comm comm, MPI_Datatype base, Decomp src,
                            Decomp dest);
void freeReshapePlan(ReshapePlan plan);
/** Move the data (reshape.h:285).  src and dest are pointers into
 *  distributed device buffers (dm_malloc / dm_register): rank r's local
 *  array is at the same offset in slot r.  Runs every rank's share in one
 *  launch; returns when the data has landed (unless the world is in
 *  asynchronous mode, see dm_set_async). */
void runReshapePlan(MPI_Comm comm, ReshapePlan plan, void *src, void *dest);
/** 0 when the messages stay inside arrays of the given lengths, else 1
 *  after a message on stderr (reshape.h:293).  Inside a dm_spmd_run
 *  session the calling rank's own messages are checked, as under MPI;
 *  outside one, every rank's against the same lengths. */
int checkReshapePlanBounds(ReshapePlan plan, int type_size, int src_len,
                           int dest_len);
/** reshape.c:494-521: the calling rank's file view of d over the global
 *  row-major file (last axis fastest): its elements' byte offsets, in the
 *  order of its loops over its coordinates (first axis outermost), in a
 *  datatype whose extent is the whole file (0 on ranks d does not use).
 *  dm_type_view reads the offsets; dm_file_write / dm_file_read move the
 *  whole distributed array to and from such a file. */
void makeIOType(MPI_Datatype *result, MPI_Datatype src, Decomp d);
/** No counterpart in the reference (its view is an MPI datatype): number of
 *  base elements of type t's view, their byte offsets written to
 *  offsets[0..max). */
long long dm_type_view(MPI_Datatype t, long long *offsets, long long max);
/** No counterpart in the reference, whose drivers write through MPI-IO with
 *  makeIOType's views (split_plan_test.c:160-200).  Collective: the whole
 *  distributed array buf (decomposition d, elements of type base) to or
 *  from the file at path as raw native-endian elements in global row-major
 *  order (the reference's "native" representation).  0 on success. */
int dm_file_write(const char *path, MPI_Datatype base, Decomp d, const void *buf);
int dm_file_read(const char *path, MPI_Datatype base, Decomp d, void *buf);

#ifdef __cplusplus
}
#endif
#endif
