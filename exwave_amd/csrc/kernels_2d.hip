// kernels_2d.hip -- instantiates every 2D kernel (degrees 1..MAX_DEGREE,
// FP64/FP32 operator arithmetic, Cartesian/general geometry).
#include "launchers.hpp"

namespace exw {
const DimKernels &kernels_2d() {
  static const DimKernels table = [] {
    DimKernels t{};
    fill_dim_kernels<2, MAX_DEGREE>(t);
    return t;
  }();
  return table;
}
}  // namespace exw
