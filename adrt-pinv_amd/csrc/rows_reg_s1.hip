// Register row kernel instantiations: compute type float, compensated=false, Chebyshev / Gram levels 1..n.
#include "rows_reg.hpp"

namespace adrt {

int rows_reg_launch_s1(const RegArgs& RA, int loge, hipStream_t s) { return rreg::launch_t<float, false, 1>(RA, loge, s); }

}  // namespace adrt
