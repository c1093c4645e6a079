// Index run-lists and the six DimDecomp strategies.
//
// Formulas follow the reference descriptors (cited per function); the code is
// this library's.  Every strategy is allocated with calloc so that fields a
// caller's extension does not set stay zero, as reshape.h:37-44 asks.
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "internal.h"
#include "block_decomp.h"
#include "local_decomp.h"
#include "single_decomp.h"
#include "rotate_decomp.h"
#include "interleave_decomp.h"
#include "factor_reverse_decomp.h"
#include "logger.h"

// ---------------------------------------------------------------- Index
// index.c:7-59: runs of consecutive offsets.

extern "C" Index makeIndex(void) {
  Index ix = (Index)calloc(1, sizeof(struct index));
  ix->size = INDEX_INITIAL_SIZE;
  ix->offsets = (int *)calloc(ix->size, sizeof(int));
  ix->counts = (int *)calloc(ix->size, sizeof(int));
  return ix;
}

extern "C" void freeIndex(Index ix) {
  if (ix == nullptr) return;
  free(ix->offsets);
  free(ix->counts);
  free(ix);
}

extern "C" void addIndex(Index ix, int offset) {
  if (ix->count > 0) {
    int last = ix->count - 1;
    if (ix->offsets[last] + ix->counts[last] == offset) {
      ix->counts[last]++;
      return;
    }
  }
  if (ix->count == ix->size) {
    ix->size *= 2;
    ix->offsets = (int *)realloc(ix->offsets, ix->size * sizeof(int));
    ix->counts = (int *)realloc(ix->counts, ix->size * sizeof(int));
  }
  ix->offsets[ix->count] = offset;
  ix->counts[ix->count] = 1;
  ix->count++;
}

extern "C" void freeIndexList(Index *list) {
  if (list == nullptr) return;
  for (Index *p = list; *p != nullptr; p++) freeIndex(*p);
  free(list);
}

extern "C" void printIndex(FILE *out, Index ix) {
  for (int r = 0; r < ix->count; r++)
    for (int k = 0; k < ix->counts[r]; k++) fprintf(out, "%d ", ix->offsets[r] + k);
}

extern "C" void logIndex(Index ix) {
  for (int r = 0; r < ix->count; r++)
    for (int k = 0; k < ix->counts[r]; k++) fragment_logger("%d ", ix->offsets[r] + k);
}

// ---------------------------------------------------------------- base
// reshape.c:22-45

extern "C" void freeDimDecomp(DimDecomp d) {
  if (d == nullptr) return;
  if (d->destroy != nullptr) d->destroy(d);
  free(d);
}

extern "C" DimDecomp cloneDimDecomp(DimDecomp d) {
  if (d == nullptr || d->cloner == nullptr) return nullptr;
  return d->cloner(d);
}

extern "C" int sizeDimDecomp(DimDecomp d) { return d->coord_size; }

extern "C" void describeDimDecomp(FILE *out, DimDecomp d) {
  fprintf(out, "DimDecomp(%s) coords=%d procs=%d", d->type, d->coord_size, d->proc_size);
  if (d->printer != nullptr) d->printer(out, d);
}

template <typename T>
static T *copy_of(DimDecomp d) {
  T *r = (T *)calloc(1, sizeof(T));
  memcpy(r, d, sizeof(T));
  return r;
}

// ---------------------------------------------------------------- block
// block_decomp.c:6-51

extern "C" int block_rank_offset(DimDecomp d, int global) {
  return global / ((BlockDecomp)d)->blocksize;
}
extern "C" int block_type_offset(DimDecomp d, int global) {
  return global % ((BlockDecomp)d)->blocksize;
}
static void print_block(FILE *out, DimDecomp d) {
  fprintf(out, " blocksize=%d", ((BlockDecomp)d)->blocksize);
}
static DimDecomp clone_block(DimDecomp d) { return (DimDecomp)copy_of<struct block_decomp>(d); }

extern "C" DimDecomp makeBlockDecompWithFactor(int len, int factor, int procs) {
  if (factor < 1 || procs < 1 || len % factor != 0) return nullptr;
  BlockDecomp b = (BlockDecomp)calloc(1, sizeof(struct block_decomp));
  b->parent.coord_size = len;
  b->parent.proc_size = procs;
  b->parent.rank_offset = block_rank_offset;
  b->parent.type_offset = block_type_offset;
  b->parent.printer = print_block;
  b->parent.cloner = clone_block;
  b->parent.type = (char *)"block";
  int units = len / factor;                     // blocks are whole multiples of factor
  b->blocksize = factor * ((units + procs - 1) / procs);
  if (b->blocksize < 1) b->blocksize = factor;  // len == 0: keep divisions defined
  return (DimDecomp)b;
}

extern "C" DimDecomp makeBlockDecomp(int len, int procs) {
  return makeBlockDecompWithFactor(len, 1, procs);
}

// ---------------------------------------------------------------- local
// local_decomp.c:6-28

extern "C" int local_rank_offset(DimDecomp, int) { return 0; }
extern "C" int local_type_offset(DimDecomp, int global) { return global; }
extern "C" DimDecomp local_clone(DimDecomp d) { return (DimDecomp)copy_of<struct local_decomp>(d); }

extern "C" DimDecomp makeLocalDecomp(int len) {
  LocalDecomp l = (LocalDecomp)calloc(1, sizeof(struct local_decomp));
  l->parent.coord_size = len;
  l->parent.proc_size = 1;
  l->parent.rank_offset = local_rank_offset;
  l->parent.type_offset = local_type_offset;
  l->parent.cloner = local_clone;
  l->parent.type = (char *)"local";
  return (DimDecomp)l;
}

// ---------------------------------------------------------------- single
// single_decomp.c:6-30

extern "C" int single_rank_offset(DimDecomp d, int) { return ((SingleDecomp)d)->rank; }
extern "C" int single_type_offset(DimDecomp, int global) { return global; }
extern "C" DimDecomp single_clone(DimDecomp d) { return (DimDecomp)copy_of<struct single_decomp>(d); }

extern "C" DimDecomp makeSingleDecomp(int coord_size, int rank, int size) {
  SingleDecomp s = (SingleDecomp)calloc(1, sizeof(struct single_decomp));
  s->parent.coord_size = coord_size;
  s->parent.proc_size = size;
  s->parent.rank_offset = single_rank_offset;
  s->parent.type_offset = single_type_offset;
  s->parent.cloner = single_clone;
  s->parent.type = (char *)"single";
  s->rank = rank;
  return (DimDecomp)s;
}

// ---------------------------------------------------------------- rotate
// rotate_decomp.c:7-50: cyclic shift of the coordinate before the nested map.

static inline int rotated(RotateDecomp r, int g) {
  int len = r->parent.coord_size;
  return (g + len + r->rotate) % len;
}
extern "C" int rotate_rank_offset(DimDecomp d, int global) {
  RotateDecomp r = (RotateDecomp)d;
  return r->nested->rank_offset(r->nested, rotated(r, global));
}
extern "C" int rotate_type_offset(DimDecomp d, int global) {
  RotateDecomp r = (RotateDecomp)d;
  return r->nested->type_offset(r->nested, rotated(r, global));
}
static void rotate_destroy(DimDecomp d) { freeDimDecomp(((RotateDecomp)d)->nested); }
extern "C" DimDecomp clone_rotate(DimDecomp d) {
  RotateDecomp r = copy_of<struct rotate_decomp>(d);
  r->nested = cloneDimDecomp(((RotateDecomp)d)->nested);
  return (DimDecomp)r;
}

extern "C" DimDecomp makeRotatedDecomp(int rotate, DimDecomp nested) {
  if (nested == nullptr) return nullptr;
  RotateDecomp r = (RotateDecomp)calloc(1, sizeof(struct rotate_decomp));
  r->parent.coord_size = nested->coord_size;
  r->parent.proc_size = nested->proc_size;
  r->parent.rank_offset = rotate_rank_offset;
  r->parent.type_offset = rotate_type_offset;
  r->parent.destroy = rotate_destroy;
  r->parent.cloner = clone_rotate;
  r->parent.type = (char *)"rotate";
  r->rotate = rotate;
  r->nested = cloneDimDecomp(nested);
  return (DimDecomp)r;
}

// ---------------------------------------------------------------- interleave
// interleave_decomp.c:6-56: g = y + A z, A = nested coord_size.

extern "C" int interleave_rank_offset(DimDecomp d, int global) {
  InterleaveDecomp v = (InterleaveDecomp)d;
  return v->nested->rank_offset(v->nested, global % v->A);
}
extern "C" int interleave_type_offset(DimDecomp d, int global) {
  InterleaveDecomp v = (InterleaveDecomp)d;
  int y = global % v->A, z = global / v->A;
  return z + v->B * v->nested->type_offset(v->nested, y);
}
static void interleave_destroy(DimDecomp d) { freeDimDecomp(((InterleaveDecomp)d)->nested); }
extern "C" void print_interleave(FILE *out, DimDecomp d) {
  InterleaveDecomp v = (InterleaveDecomp)d;
  fprintf(out, " A=%d B=%d nested:", v->A, v->B);
  describeDimDecomp(out, v->nested);
}
extern "C" DimDecomp clone_interleave(DimDecomp d) {
  InterleaveDecomp v = copy_of<struct interleave_decomp>(d);
  v->nested = cloneDimDecomp(((InterleaveDecomp)d)->nested);
  return (DimDecomp)v;
}

extern "C" DimDecomp makeInterleaveDecomp(DimDecomp nested, int B) {
  if (nested == nullptr) return nullptr;
  InterleaveDecomp v = (InterleaveDecomp)calloc(1, sizeof(struct interleave_decomp));
  v->A = nested->coord_size;
  v->B = B;
  v->parent.coord_size = v->A * B;
  v->parent.proc_size = nested->proc_size;
  v->parent.rank_offset = interleave_rank_offset;
  v->parent.type_offset = interleave_type_offset;
  v->parent.destroy = interleave_destroy;
  v->parent.printer = print_interleave;
  v->parent.cloner = clone_interleave;
  v->parent.type = (char *)"interleave";
  v->nested = cloneDimDecomp(nested);
  return (DimDecomp)v;
}

// ---------------------------------------------------------------- factor reverse
// factor_reverse_decomp.c:8-58: k = g / A + B * (g % A).

static inline int reversed(FactorReverseDecomp f, int g) { return g / f->A + f->B * (g % f->A); }
extern "C" int factor_reverse_rank_offset(DimDecomp d, int global) {
  FactorReverseDecomp f = (FactorReverseDecomp)d;
  return f->nested->rank_offset(f->nested, reversed(f, global));
}
extern "C" int factor_reverse_type_offset(DimDecomp d, int global) {
  FactorReverseDecomp f = (FactorReverseDecomp)d;
  return f->nested->type_offset(f->nested, reversed(f, global));
}
static void factor_reverse_destroy(DimDecomp d) { freeDimDecomp(((FactorReverseDecomp)d)->nested); }
extern "C" DimDecomp clone_factor_reverse(DimDecomp d) {
  FactorReverseDecomp f = copy_of<struct factor_reverse_decomp>(d);
  f->nested = cloneDimDecomp(((FactorReverseDecomp)d)->nested);
  return (DimDecomp)f;
}

extern "C" DimDecomp makeFactorReverseDecomp(int A, int B, DimDecomp nested) {
  if (nested == nullptr) return nullptr;
  if (A * B != nested->coord_size) {
    fprintf(stderr, "FactorReverse: factors do not equal size\n");
    return nullptr;
  }
  FactorReverseDecomp f = (FactorReverseDecomp)calloc(1, sizeof(struct factor_reverse_decomp));
  f->parent.coord_size = nested->coord_size;
  f->parent.proc_size = nested->proc_size;
  f->parent.rank_offset = factor_reverse_rank_offset;
  f->parent.type_offset = factor_reverse_type_offset;
  f->parent.destroy = factor_reverse_destroy;
  f->parent.cloner = clone_factor_reverse;
  f->parent.type = (char *)"factor_reverse";
  f->A = A;
  f->B = B;
  f->nested = cloneDimDecomp(nested);
  return (DimDecomp)f;
}
