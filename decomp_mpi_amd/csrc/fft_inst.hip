// Instantiations of the FFT pass kernel for one precision, direction and
// radix family (powers of two only, mixed radices up to 13, or mixed radices
// with direct-sum stages for larger primes).  The build compiles
// this file once per combination with DM_INST_* set; the defaults give the
// complex-double forward power-of-two family.
#include "fft_pass_kernel.h"

#ifndef DM_INST_NAME
#define DM_INST_NAME launch_d_fwd_p2
#define DM_INST_R double
#define DM_INST_INV false
#define DM_INST_POW2 true
#endif
#ifndef DM_INST_ANY
#define DM_INST_ANY false
#endif

namespace dm {
namespace fftpass {
hipError_t DM_INST_NAME(const Pass &P, size_t lds, hipStream_t s) {
  return launch_tiles<DM_INST_R, DM_INST_INV, DM_INST_POW2, DM_INST_ANY>(P, lds, s);
}
}  // namespace fftpass
}  // namespace dm
