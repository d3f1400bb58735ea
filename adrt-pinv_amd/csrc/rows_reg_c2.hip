// Register row kernel instantiations: compute type float, compensated=true, Chebyshev / Gram levels 2..n.
#include "rows_reg.hpp"

namespace adrt {

int rows_reg_launch_c2(const RegArgs& RA, int loge, hipStream_t s) { return rreg::launch_t<float, true, 2>(RA, loge, s); }

}  // namespace adrt
