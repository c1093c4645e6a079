/* logger.h -- debug logging (reference logger.c:6-27; its header is not
 * shipped).  Prints only when DM_DEBUG is set in the environment. */
#ifndef DM_LOGGER_H
#define DM_LOGGER_H
#include "dm_mpi.h"
#ifdef __cplusplus
extern "C" {
#endif
void logger(MPI_Comm comm, const char *msg, ...);
void fragment_logger(const char *msg, ...);
#ifdef __cplusplus
}
#endif
#endif
