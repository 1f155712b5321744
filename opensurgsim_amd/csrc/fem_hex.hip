// Hexahedral (Fem3DElementCube) path.
#include "fem_hex.hpp"

#include <stdexcept>

namespace femgpu {

void hex_setup(HexData&, int64_t, const double*, int64_t, const int32_t*, const double*, std::vector<double>&,
               std::vector<int32_t>&, std::vector<int32_t>&, hipStream_t) {
    throw std::runtime_error("hexahedral elements are not available in this build");
}

void hex_step(HexData&, const fem_options&, const double*, double*, double*, const double*, const double*,
              const uint8_t*, double*, CgScalars*, CgScalars*, int*, hipStream_t, int&, fem_stats&, int) {
    throw std::runtime_error("hexahedral elements are not available in this build");
}

void hex_eval_elements(HexData&, int, int, const double*, const double*, double*, double*, double*, hipStream_t) {
    throw std::runtime_error("hexahedral elements are not available in this build");
}

}  // namespace femgpu
