/*
 * dm_batch.h -- many independent local multi-dimensional FFTs at once.
 *
 * No counterpart in the reference, whose transforms are one distributed
 * array at a time (split_fft.h).  This is the batched form: transforms of
 * uneven shapes (each a contiguous C-order array in device memory, last
 * axis fastest) are transformed in place by one persistent launch per axis,
 * every launch covering the tiles of all transforms.  The launches can be
 * captured once into a HIP graph and replayed.
 *
 * Precision: 1 complex double, 0 complex float (twiddles generated in
 * double precision either way).  Direction FFTW_FORWARD (-1) or
 * FFTW_BACKWARD (+1), unnormalised, as everywhere in this library.
 */
#ifndef DM_BATCH_H
#define DM_BATCH_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct dm_batch *DmBatch;

/* Tile and radix schedule choices (dm_batch_set_schedule, dm_batch_autotune). */
#define DM_BATCH_TILE_FULL 0    /* tiles of up to the kernel's capacity */
#define DM_BATCH_TILE_HALF 1    /* tiles of up to half of it */
/* The last two axes of a transform in one plane tile where the plane fits
 * (at most 8192 complex-float or 4096 complex-double points, each side at
 * most 256): one read and one write of those axes instead of two; full
 * line tiles for the other axes and transforms. */
#define DM_BATCH_TILE_PLANES 2
#define DM_BATCH_RADIX_LARGE_FIRST 0
#define DM_BATCH_RADIX_SMALL_FIRST 1
/* Launch grids (dm_batch_set_grid, dm_batch_autotune). */
#define DM_BATCH_GRID_TILES 0       /* one workgroup per tile (the default) */
#define DM_BATCH_GRID_PERSISTENT 1  /* as many workgroups as fit the device at once, each looping over tiles */

/** count transforms of ndim (1..3) axes; shapes[i * ndim + d] is axis d of
 *  transform i.  Axis lengths up to 4096 points, any factors (primes above
 *  13 run as direct-sum stages, O(p) work per point), at most 16 stages.
 *  Returns NULL (after an error message) otherwise. */
DmBatch dm_batch_plan(int count, int ndim, const int *shapes, int precision, int direction);
/** Device pointers of the transforms' data, data[i] for transform i.  Builds
 *  and uploads the tile tables; must precede run, graph and autotune. */
int dm_batch_bind(DmBatch b, void *const *data);
/** Tile and radix schedule (the defaults: FULL, LARGE_FIRST); rebinds. */
int dm_batch_set_schedule(DmBatch b, int tile, int radix_order);
/** Launch grid (DM_BATCH_GRID_*); rebinds. */
int dm_batch_set_grid(DmBatch b, int mode);
/** The current launch grid mode. */
int dm_batch_grid(DmBatch b);
/** Enqueue the transform on a HIP stream (NULL: the legacy default stream). */
int dm_batch_run(DmBatch b, void *stream);
/** First call: capture dm_batch_run into a HIP graph; every call launches
 *  the graph on the stream.  Rebinding or rescheduling drops the graph. */
int dm_batch_graph_launch(DmBatch b, void *stream);
/** Times every grid, tile and radix-order choice on the bound data (which
 *  it overwrites) and keeps the fastest; returns its milliseconds per run
 *  (< 0: error). */
double dm_batch_autotune(DmBatch b, void *stream);
/** Tiles per axis pass and launch grid of the current binding (inspection;
 *  ndim entries each: plane tiles are counted with the last axis's pass). */
int dm_batch_info(DmBatch b, long long *tiles, long long *grid, int *tile_choice, int *radix_order);
/** Algorithmic HBM bytes of one run: sum over transforms of 2 * ndim * N * elem. */
double dm_batch_bytes(DmBatch b);
/** Bytes the current schedule reads and writes: as dm_batch_bytes, with one
 *  pass instead of two for the transforms whose last two axes run as planes. */
double dm_batch_bytes_moved(DmBatch b);
void dm_batch_free(DmBatch b);

/** Device bytes of the fixed table area that dm_batch_run_noalloc uses. */
#define DM_BATCH_NOALLOC_SCRATCH_BYTES ((size_t)16 << 20)
/** No counterpart in the reference.  For callers that cannot allocate
 *  device memory (code under HIP graph capture, tight serving loops,
 *  harnesses that forbid hipMalloc): plan, bind and run in one call that
 *  allocates no device memory.  The tables the bound form allocates (data
 *  pointers, schedules, twiddles, tile lists) go into a fixed
 *  DM_BATCH_NOALLOC_SCRATCH_BYTES area that each device holds once for the
 *  life of the process; a batch whose tables do not fit is refused (1,
 *  after an error message).  Arguments as dm_batch_plan, plus data[i] (device
 *  pointer of transform i, in place) and the stream (NULL: the legacy
 *  default stream).  Returns 0 once the tables are on the device and the
 *  passes are enqueued; the result equals dm_batch_run's bit for bit (full
 *  tiles, large radices first).  Every call on one device shares its table
 *  area: calls that overlap in time on one device (different streams or
 *  host threads) race; calls on one stream are ordered by it. */
int dm_batch_run_noalloc(int count, int ndim, const int *shapes, int precision, int direction,
                         void *const *data, void *stream);

#ifdef __cplusplus
}
#endif
#endif
