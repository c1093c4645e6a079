/*
 * index.h -- run-length lists of local offsets.
 * Same interface as the reference's index.h:1-25 (struct index, makeIndex,
 * addIndex, freeIndex, freeIndexList, printIndex, logIndex).
 */
#ifndef INDEX_H
#define INDEX_H
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

/** Offsets stored as (offset, count) runs; a new offset that extends the
 *  last run by one is merged into it. */
struct index {
  int size;     /* capacity of offsets/counts */
  int count;    /* number of runs */
  int *offsets; /* first offset of each run */
  int *counts;  /* length of each run */
};
typedef struct index *Index;

#define INDEX_INITIAL_SIZE 16

Index makeIndex(void);
void freeIndex(Index);
void addIndex(Index i, int offset);
/** Free a NULL-terminated array of Index objects and the array itself. */
void freeIndexList(Index *list);
void printIndex(FILE *out, Index index);
void logIndex(Index index);

#ifdef __cplusplus
}
#endif
#endif
