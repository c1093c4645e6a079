/* block_decomp.h -- contiguous blocks (reference block_decomp.h:1-28,
 * block_decomp.c:6-51): rank = g / blocksize, offset = g % blocksize,
 * blocksize = factor * ceil((len / factor) / procs). */
#ifndef DM_BLOCK_DECOMP_H
#define DM_BLOCK_DECOMP_H
#include "reshape.h"
#ifdef __cplusplus
extern "C" {
#endif
struct block_decomp {
  struct dim_decomp parent;
  int blocksize;
};
typedef struct block_decomp *BlockDecomp;
int block_rank_offset(DimDecomp d, int global);
int block_type_offset(DimDecomp d, int global);
/** Block decomposition of len coordinates over proc ranks
 *  (block_decomp.c:23-25). */
DimDecomp makeBlockDecomp(int len, int proc);
/** Block size a multiple of factor; NULL when factor does not divide len
 *  (block_decomp.c:33-51). */
DimDecomp makeBlockDecompWithFactor(int len, int factor, int procs);
#ifdef __cplusplus
}
#endif
#endif
