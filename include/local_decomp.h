/* local_decomp.h -- an axis held whole by every rank (reference
 * local_decomp.h:1-18, local_decomp.c:6-28): rank 0, offset = g. */
#ifndef DM_LOCAL_DECOMP_H
#define DM_LOCAL_DECOMP_H
#include "reshape.h"
#ifdef __cplusplus
extern "C" {
#endif
struct local_decomp {
  struct dim_decomp parent;
};
typedef struct local_decomp *LocalDecomp;
int local_rank_offset(DimDecomp d, int global);
int local_type_offset(DimDecomp d, int global);
DimDecomp local_clone(DimDecomp d);
DimDecomp makeLocalDecomp(int len);
#ifdef __cplusplus
}
#endif
#endif
