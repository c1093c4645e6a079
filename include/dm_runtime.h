/*
 * dm_runtime.h -- this library's own entry points beside the reference API:
 * distributed device buffers, execution control, error handling, plan
 * inspection.  None of these has a counterpart in the reference (whose
 * buffers are per-process host arrays and whose errors always abort).
 */
#ifndef DM_RUNTIME_H
#define DM_RUNTIME_H
#include <stddef.h>
#include "dm_mpi.h"
#ifdef __cplusplus
extern "C" {
#endif

/* ---- distributed device buffers ---------------------------------------
 * A distributed buffer is one device allocation of nranks slots.  A pointer
 * into slot 0 stands for "this offset in every rank's local array". */
void *dm_malloc(MPI_Comm comm, size_t bytes_per_rank);
int dm_free(void *base);
/** Register caller-owned device memory (nranks * slot_bytes) as a buffer. */
int dm_register(MPI_Comm comm, void *base, size_t slot_bytes);
int dm_unregister(void *base);
/** Slot size in bytes of the buffer holding ptr, 0 when ptr is unknown. */
size_t dm_slot_bytes(const void *ptr);
/** Host <-> device copies of one rank's slot region (synchronous). */
int dm_put(void *dst, int rank, const void *host, size_t bytes);
int dm_get(void *host, const void *src, int rank, size_t bytes);
/** Fill bytes of every slot with a byte value (synchronous). */
int dm_fill(void *dst, int value, size_t bytes_per_rank);

/* ---- execution control ------------------------------------------------ */
/** 1 when a HIP device is present. */
int dm_device_available(void);
/** 0 (default): collective calls return after the device work finished.
 *  1: they only enqueue on the world's stream; call dm_synchronize. */
int dm_set_async(MPI_Comm comm, int on);
int dm_synchronize(MPI_Comm comm);
/** The world's HIP stream (hipStream_t), created on first use. */
void *dm_get_stream(MPI_Comm comm);
/** Use a caller-owned HIP stream (NULL restores the world's own). */
int dm_set_stream(MPI_Comm comm, void *stream);

/* ---- per-axis FFT stage ------------------------------------------------
 * One batched 1D FFT along `axis` of a distributed array: read the `in`
 * decomposition, transform, write the `out` decomposition (which may
 * differ: the reshape is fused into the pass).  Replaces a driver's FFTW
 * guru stage plus the runReshapePlan after it (fftw_test.c:298-336 in the
 * reference).  precision 1: complex double, 0: complex float.  direction
 * FFTW_FORWARD (-1) or FFTW_BACKWARD (+1), unnormalised.  Input and output
 * may be the same memory only when the two decompositions are equivalent. */
struct decomp;
typedef struct dm_axis_fft *DmAxisFFT;
DmAxisFFT dm_plan_axis_fft(struct decomp *in, struct decomp *out, int axis, int precision, int direction);
int dm_run_axis_fft(DmAxisFFT plan, const void *in, void *out);
void dm_free_axis_fft(DmAxisFFT plan);

/* ---- calibration -------------------------------------------------------
 * Average milliseconds of a plain 16-byte-vector device copy of `bytes`
 * over `reps` launches: the achievable-bandwidth and byte-counter reference
 * for the roofline figures. */
double dm_copy_bench(MPI_Comm comm, void *dst, const void *src, size_t bytes, int reps);

/* ---- errors ------------------------------------------------------------
 * mode 0 (default): raise_error prints and aborts, as the reference.
 * mode 1: raise_error records the message and returns; the failing call
 * then returns NULL / an error value.  Used by the error-path tests. */
void dm_set_error_mode(int mode);
const char *dm_last_error(void);
void dm_clear_error(void);

#ifdef __cplusplus
}
#endif
#endif
