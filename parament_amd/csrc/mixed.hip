// Mixed-precision Chebyshev slice kernel (bf16 three-way split on the matrix cores,
// FP64 first-order terms).  Not yet enabled: the host falls back to the FP64 path.
#include "kernels.hpp"

namespace pa {

bool mixed_supported(int) { return false; }

hipError_t launch_cheb_mixed(const MixedArgs&, hipStream_t) { return hipErrorNotSupported; }

}  // namespace pa
