// Register row kernel instantiations: compute type float, compensated=true, levels 2..n, complex128 rows in and out
// with the explicit level 1 in double.
#include "rows_reg.hpp"

namespace adrt {

int rows_reg_launch_x2(const RegArgs& RA, int loge, hipStream_t s) {
    return rreg::launch_t<float, true, 2, true>(RA, loge, s);
}

}  // namespace adrt
