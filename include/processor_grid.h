/* processor_grid.h -- processor grid search (reference processor_grid.c:1-98;
 * its header is not shipped, fields follow their use there). */
#ifndef DM_PROCESSOR_GRID_H
#define DM_PROCESSOR_GRID_H
#include "dm_mpi.h"
#ifdef __cplusplus
extern "C" {
#endif
struct grid_result {
  int ndim;
  int *length;
  int load;
  int procs;
  double asym;
};
typedef struct grid_result *GridResult;
/** Grid over at most comm-size ranks minimising the largest per-rank task
 *  count prod(ceil(tasks/grid)), ties broken by a smaller max/min block
 *  ratio (processor_grid.c:24-97).  Returns a malloc'd array of ndim. */
int *makeProcessorGrid(MPI_Comm comm, int ndim, int tasks[]);
#ifdef __cplusplus
}
#endif
#endif
