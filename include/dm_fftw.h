/*
 * dm_fftw.h -- the FFTW names that appear in the split-FFT API.
 *
 * The reference's split_fft.h takes fftw_complex buffers, stores fftw_plan
 * handles in its plan struct and uses FFTW_FORWARD / FFTW_BACKWARD as the
 * direction of runSplitPlan (split_fft.h:16-22,52-59, split_fft.c:628,683).
 * FFTW itself is not used: a plan handle here names a compiled device pass
 * program of this library.  fftw_complex is 16 bytes (re, im), as in FFTW:
 * C99 `double _Complex` when <complex.h> came first, else double[2].
 * fftwf_complex is the 8-byte single-precision element of the FP32 form.
 */
#ifndef DM_FFTW_H
#define DM_FFTW_H

#if !defined(FFTW3_H)
#if defined(_Complex_I) && defined(complex) && defined(I) && !defined(__cplusplus)
typedef double _Complex fftw_complex;
typedef float _Complex fftwf_complex;
#else
typedef double fftw_complex[2];
typedef float fftwf_complex[2];
#endif
#define FFTW_FORWARD (-1)
#define FFTW_BACKWARD (+1)
/** Opaque device pass program (fills the fftw_plan fields of a SplitPlan). */
typedef struct dm_fft_program *fftw_plan;
#endif

#endif /* DM_FFTW_H */
