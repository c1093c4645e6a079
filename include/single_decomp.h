/* single_decomp.h -- all of an axis on one position of the rank axis
 * (reference single_decomp.h:1-24, single_decomp.c:6-30). */
#ifndef DM_SINGLE_DECOMP_H
#define DM_SINGLE_DECOMP_H
#include "reshape.h"
#ifdef __cplusplus
extern "C" {
#endif
struct single_decomp {
  struct dim_decomp parent;
  int rank;
};
typedef struct single_decomp *SingleDecomp;
int single_rank_offset(DimDecomp d, int global);
int single_type_offset(DimDecomp d, int global);
DimDecomp single_clone(DimDecomp d);
/** coord_size coordinates, all at position `rank` of `size` positions. */
DimDecomp makeSingleDecomp(int coord_size, int rank, int size);
#ifdef __cplusplus
}
#endif
#endif
