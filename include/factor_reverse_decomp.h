/* factor_reverse_decomp.h -- swap the significance of two factors of the
 * coordinate (reference factor_reverse_decomp.h:1-24,
 * factor_reverse_decomp.c:8-58): k = g / A + B * (g % A), then nested.
 * NULL when A * B differs from the nested coord_size. */
#ifndef DM_FACTOR_REVERSE_DECOMP_H
#define DM_FACTOR_REVERSE_DECOMP_H
#include "reshape.h"
#ifdef __cplusplus
extern "C" {
#endif
struct factor_reverse_decomp {
  struct dim_decomp parent;
  int A;
  int B;
  DimDecomp nested;
};
typedef struct factor_reverse_decomp *FactorReverseDecomp;
int factor_reverse_rank_offset(DimDecomp d, int global);
int factor_reverse_type_offset(DimDecomp d, int global);
DimDecomp clone_factor_reverse(DimDecomp d);
DimDecomp makeFactorReverseDecomp(int A, int B, DimDecomp nested);
/** AxisDecomp with the same comm, strides and rank stride as orig
 *  (factor_reverse_decomp.c:56-58). */
AxisDecomp makeFactorReverseAxisDecomp(int A, int B, AxisDecomp orig);
#ifdef __cplusplus
}
#endif
#endif
