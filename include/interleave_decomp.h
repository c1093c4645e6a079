/* interleave_decomp.h -- interleaved sequences (reference
 * interleave_decomp.h:1-30, interleave_decomp.c:6-56): with A the nested
 * coord_size, g = y + A*z; rank = nested.rank(y),
 * offset = z + B * nested.offset(y).  The nested one is cloned. */
#ifndef DM_INTERLEAVE_DECOMP_H
#define DM_INTERLEAVE_DECOMP_H
#include "reshape.h"
#ifdef __cplusplus
extern "C" {
#endif
struct interleave_decomp {
  struct dim_decomp parent;
  DimDecomp nested;
  int A;
  int B;
};
typedef struct interleave_decomp *InterleaveDecomp;
int interleave_rank_offset(DimDecomp d, int global);
int interleave_type_offset(DimDecomp d, int global);
void print_interleave(FILE *out, DimDecomp d);
DimDecomp clone_interleave(DimDecomp d);
DimDecomp makeInterleaveDecomp(DimDecomp nested, int B);
#ifdef __cplusplus
}
#endif
#endif
