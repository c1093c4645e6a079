/*
 * split_fft.h -- two-stage ("split", four-step) parallel 3D FFT.
 *
 * Drop-in for the reference's split_fft.h:1-65.  Same names, arguments,
 * struct fields, timer and flag enums.  Execution differs by design: the
 * plan compiles to a device pass program (batched radix FFT passes whose
 * loads and stores go through per-axis layout tables, so the start->A and
 * A->B reshapes and the twiddle multiply are fused into the FFT passes) and
 * runs on one MI355X.  dat / workspace are pointers into distributed device
 * buffers (dm_malloc / dm_register; see dm_mpi.h), length counts elements
 * per virtual rank as in the reference.
 *
 * Result placement (forward): work[0:extent) in the B decomposition.  Along
 * each axis d, frequency k sits at B-position p = k / B_d + A_d * (k % B_d);
 * getTransformed() returns the decomposition that says so.  Backward runs
 * the exact inverse (unnormalised): B layout in work -> start layout in dat.
 */
#ifndef SPLIT_FFT
#define SPLIT_FFT
#include "reshape.h"
#include "dm_fftw.h"

#ifdef __cplusplus
extern "C" {
#endif

struct dm_split_ext;
struct split_fft_plan {
  MPI_Comm comm;
  int *A_factors;
  Decomp A;
  Decomp B;
  Decomp B_flipped;   /* output decomposition, see getTransformed */
  fftw_complex *src;  /* fftwf_complex data for an FP32 plan */
  fftw_complex *work;
  int extent;         /* elements to hold either an A or a B decomposition */
  fftw_plan plan_a;
  fftw_plan plan_b;
  fftw_plan plan_a_back;
  fftw_plan plan_b_back;
  ReshapePlan start_to_A;
  ReshapePlan A_to_start;
  _Complex double **twiddle; /* per axis, indexed by local A offset */
  ReshapePlan A_to_B;
  ReshapePlan B_to_A;
  double *times;      /* seconds per phase, accumulated per run */
  struct dm_split_ext *ext;
};

/* timer fields (split_fft.h:32-40) */
enum Timers { RESHAPE_1 = 0, FFT_1, TWIDDLE, RESHAPE_2, FFT_2, ALL, NUM_TIMERS };

/* plan search flags (split_fft.h:42-47); RUN_PHASE_ORDER is this library's:
 * the plan runs in the reference's phase order (see dm_split_set_phase_order) */
enum SearchFlags {
  NO_FLAGS = 0x00,
  PREFER_LARGE_FFT = 0x01,
  PREFER_STAGE1 = 0x02,
  FORCE_USE_FIRST = 0x04,
  RUN_PHASE_ORDER = 0x100
};
typedef struct split_fft_plan *SplitPlan;

/** Called once per local slot of d with its global coordinate
 *  (split_fft.h:50-51, split_fft.c:862-889).  dat is host memory. */
typedef void (*data_visitor)(int ndim, int *global_position, _Complex double *dat);
void visitData(Decomp d, data_visitor vis, _Complex double *dat);

/** split_fft.c:249-569 */
SplitPlan buildParallelPlan(int flags, int use_first, int use_second, Decomp start,
                            int ndim, int A[], int B[], int L[], int a_grid[],
                            int b_grid[], fftw_complex *dat, int length,
                            fftw_complex *workspace);
/** split_fft.c:747-788 */
SplitPlan proposeParallelPlan(int flags, Decomp start, int *initial_factor,
                              fftw_complex *dat, int length, fftw_complex *workspace);
/** Model-guided replacement of the reference's exhaustive search
 *  (split_fft.c:78-247): candidate factor/grid tuples are ranked by a
 *  traffic model of this library's pass program and the best few are
 *  timed on the device. */
SplitPlan autotuneParallelPlan(Decomp start, fftw_complex *dat, int length,
                               fftw_complex *workspace);
/** split_fft.c:573-614: slab (B = L except B_last = 1) or A = the largest
 *  divisor of L not above sqrt(L). */
SplitPlan makeSplitPlan(int slab, Decomp start, fftw_complex *dat, int length,
                        fftw_complex *workspace);
/** split_fft.c:570-572 */
SplitPlan makeParallelPlan(Decomp start, int *initial_factor, fftw_complex *dat,
                           int length, fftw_complex *workspace);

/* FP32 form: same planners, single-precision butterflies, FP64-generated
 * twiddles.  No counterpart in the reference (complex-double only). */
SplitPlan buildParallelPlanf(int flags, int use_first, int use_second, Decomp start,
                             int ndim, int A[], int B[], int L[], int a_grid[],
                             int b_grid[], fftwf_complex *dat, int length,
                             fftwf_complex *workspace);
SplitPlan proposeParallelPlanf(int flags, Decomp start, int *initial_factor,
                               fftwf_complex *dat, int length, fftwf_complex *workspace);
SplitPlan autotuneParallelPlanf(Decomp start, fftwf_complex *dat, int length,
                                fftwf_complex *workspace);
SplitPlan makeSplitPlanf(int slab, Decomp start, fftwf_complex *dat, int length,
                         fftwf_complex *workspace);
SplitPlan makeParallelPlanf(Decomp start, int *initial_factor, fftwf_complex *dat,
                            int length, fftwf_complex *workspace);

/** No counterpart in the reference.  on = 1: run the plan in the
 *  reference's phase order, each phase its own device work -- RESHAPE_1
 *  (start -> A), FFT_1 (the B_d-point DFTs), TWIDDLE (the plan's twiddle
 *  tables), RESHAPE_2 (A -> B), FFT_2 (the A_d-point DFTs) -- with every
 *  phase timed into plan->times.  on = 0 (the default): the fused program,
 *  one pass per axis.  Both give the same result placement.  A plan with an
 *  axis longer than one pass tile (4096 complex-double or 8192
 *  complex-float points) whose A_d and B_d each fit one runs in phase order
 *  only (on = 0 is refused).  Returns 0 or an error code. */
int dm_split_set_phase_order(SplitPlan plan, int on);
/** 1 when the plan runs in phase order. */
int dm_split_phase_order(SplitPlan plan);
/** No counterpart in the reference (its transforms are complex double).
 *  Twiddle products of a complex-float plan: DM_TWIDDLES_COMPENSATED (the
 *  default) multiplies by each root as a float part plus a bfloat16
 *  correction, formed from the FP64 root and added inside the fused
 *  multiply-adds of the product, so the root is held to about 2^-33;
 *  DM_TWIDDLES_PLAIN uses the float part alone (about 2^-25).  No effect on
 *  complex-double plans.  Returns 0 or an error code. */
enum { DM_TWIDDLES_COMPENSATED = 0, DM_TWIDDLES_PLAIN = 1 };
int dm_split_set_twiddles(SplitPlan plan, int mode);
/** The plan's twiddle mode, or -1 for a NULL plan. */
int dm_split_twiddles(SplitPlan plan);
/** No counterpart in the reference.  The forward result goes to
 *  work[0:extent) in the caller's decomposition `out` (compatible with the
 *  start, every coordinate owned, each rank's block within extent elements)
 *  instead of the B layout, and the backward reads it from there;
 *  getTransformed then returns `out`, which the plan aliases.  For example
 *  FFTW-MPI's FFTW_MPI_TRANSPOSED_OUT layout of an n0 x n1 x n2 transform
 *  over P ranks -- axis 1 in blocks over the ranks, local order [k1][k0][k2]
 *  -- is makeDecomp of makeAxisDecomp(Local(n0), n2, 1),
 *  makeAxisDecomp(Block(n1, P), n0 * n2, 1), makeAxisDecomp(Local(n2), 1, 1)
 *  (fftw_test.c:277-281).  Fused plans only.  Returns 0 or an error code. */
int dm_split_set_output(SplitPlan plan, Decomp out);

/** The decomposition of the forward result: frequency -> rank and offset
 *  in work (split_fft.c:615-621; NULL for single-stage plans). */
Decomp getTransformed(SplitPlan plan);
/** direction FFTW_FORWARD or FFTW_BACKWARD (split_fft.c:622-744). */
void runSplitPlan(SplitPlan plan, int direction);
void freeSplitPlan(SplitPlan plan);
/** per axis Interleave(Block(A[i], grid[i]), B[i]), packed (split_fft.c:36-61) */
Decomp makeInterleavedDecomp(int ndim, MPI_Comm comm, int lengths[], int counts[],
                             int grid[]);
/** per axis BlockWithFactor(A[i]*B[i], A[i], grid[i]), packed (split_fft.c:63-77) */
Decomp makeBlockedDecomp(int ndim, MPI_Comm comm, int A[], int B[], int grid[]);
/** binary gcd (split_fft.c:890-926) */
unsigned int gcd(unsigned int u, unsigned int v);
/** twiddle multiply over a local block (dimscale.c:53-67): host memory. */
void decomp_scale(Decomp decomp, _Complex double **factors, _Complex double *data,
                  int sign);

#ifdef __cplusplus
}
#endif
#endif
