// kernels_3d.hip -- instantiates every 3D kernel (degrees 1..MAX_DEGREE_3D,
// FP64/FP32 operator arithmetic, Cartesian/general geometry).
#include "launchers.hpp"

namespace exw {
const DimKernels &kernels_3d() {
  static const DimKernels table = [] {
    DimKernels t{};
    fill_dim_kernels<3, MAX_DEGREE_3D>(t);
    return t;
  }();
  return table;
}
}  // namespace exw
