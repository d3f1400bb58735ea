// Register row kernel instantiations: compute type double, compensated=false, Chebyshev / Gram levels 2..n.
#include "rows_reg.hpp"

namespace adrt {

int rows_reg_launch_d2(const RegArgs& RA, int loge, hipStream_t s) { return rreg::launch_t<double, false, 2>(RA, loge, s); }

}  // namespace adrt
