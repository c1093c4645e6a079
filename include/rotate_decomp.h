/* rotate_decomp.h -- cyclic coordinate shift decorator (reference
 * rotate_decomp.h:1-14, rotate_decomp.c:7-50): g' = (g + len + rotate) % len,
 * then the nested decomposition.  The nested one is cloned. */
#ifndef DM_ROTATE_DECOMP_H
#define DM_ROTATE_DECOMP_H
#include "reshape.h"
#ifdef __cplusplus
extern "C" {
#endif
struct rotate_decomp {
  struct dim_decomp parent;
  int rotate;
  DimDecomp nested;
};
typedef struct rotate_decomp *RotateDecomp;
int rotate_rank_offset(DimDecomp d, int global);
int rotate_type_offset(DimDecomp d, int global);
DimDecomp clone_rotate(DimDecomp d);
DimDecomp makeRotatedDecomp(int rotate, DimDecomp nested);
#ifdef __cplusplus
}
#endif
#endif
