/*
 * dm_serve.h -- a request loop for mixed-size FFTs with continuous batching.
 *
 * No counterpart in the reference.  Clients submit independent transforms
 * (1..3 axes, complex float or double, either direction) with host input
 * and output buffers.  The server admits queued requests in arrival order
 * while their device memory fits an HBM budget (device memory comes from a
 * pool, so the steady state allocates nothing), groups admitted requests of
 * the same precision and direction into batches, and runs each batch on one
 * of several HIP streams: input copy, one batched launch per axis
 * (dm_batch.h kernels), output copy.  A request that arrives while others
 * are on the device joins the next batch that a free stream takes, so work
 * of different batches overlaps.  Plans stay on the device: the tables of a
 * (shape, precision, direction) are built and uploaded by its first request
 * and reused by the later ones.  The transform of a request does not depend
 * on what it is batched with.
 *
 * One host thread drives a server (submit / poll / wait); servers are not
 * thread-safe.  Host buffers should be pinned (dm_host_alloc) for copies to
 * overlap with compute.
 */
#ifndef DM_SERVE_H
#define DM_SERVE_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct dm_server *DmServer;

/** budget: device bytes the requests' buffers may occupy at once; streams:
 *  concurrent batches (1..16); max_batch: requests per batch (1..16). */
DmServer dm_server_create(size_t budget, int streams, int max_batch);
/** Queue a request: shape[ndim] (C order), precision 1 complex double / 0
 *  complex float, direction FFTW_FORWARD or FFTW_BACKWARD (unnormalised).
 *  in / out: host buffers of prod(shape) elements (may be the same).
 *  Returns the request id (>= 0) or -1 (bad arguments, an unsupported
 *  shape, or larger than the whole budget; the reason is in dm_last_error,
 *  whatever the error mode). */
long long dm_server_submit(DmServer s, int ndim, const int *shape, int precision, int direction,
                           const void *in, void *out);
/** Make progress without blocking: retire finished batches, admit, launch.
 *  Returns the number of requests not yet finished. */
int dm_server_poll(DmServer s);
/** Block until request id has finished (its output is in the host buffer). */
int dm_server_wait(DmServer s, long long id);
/** Block until every submitted request has finished. */
int dm_server_drain(DmServer s);
/** Seconds (steady clock, server-relative) at which request id was
 *  submitted, launched on the device and seen finished; 0 until reached. */
int dm_server_times(DmServer s, long long id, double *submitted, double *launched, double *finished);
/** Where a finished request's latency went, in milliseconds: ms[0] queue
 *  (submitted until its batch started to be enqueued), ms[2] input copies,
 *  ms[3] kernels, ms[4] output copies (ms[2..4] are its batch's, from events
 *  on the batch's stream), ms[5] the rest until a poll saw it finished (the
 *  stream starting late behind other streams' work, and poll spacing);
 *  ms[0] + ms[2] + ms[3] + ms[4] + ms[5] is the latency.  ms[1] is the host
 *  time spent enqueueing the batch, which overlaps ms[2..4].  Returns 1
 *  until the request has finished. */
int dm_server_breakdown(DmServer s, long long id, double *ms);
/** Counters: batches launched, largest batch, peak device bytes in use,
 *  device bytes held by the pool. */
int dm_server_stats(DmServer s, long long *batches, long long *max_batch_seen, size_t *peak_bytes,
                    size_t *pool_bytes);
/** Device-resident plans: shapes cached now, plans built so far, device
 *  bytes of the cached tables. */
int dm_server_plan_stats(DmServer s, long long *cached, long long *builds, size_t *bytes);
void dm_server_destroy(DmServer s);

/** Pinned host memory for request buffers. */
void *dm_host_alloc(size_t bytes);
void dm_host_free(void *p);

#ifdef __cplusplus
}
#endif
#endif
