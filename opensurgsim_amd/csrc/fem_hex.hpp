// Hexahedral (Fem3DElementCube) path: element-by-element operator (Physics/Fem3DElementCube.cpp,
// Physics/FemElement.cpp:168-210).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <vector>

#include "../../include/fem_gpu.h"
#include "fem_kernels.hpp"

namespace femgpu {

struct HexData {
    int64_t num_nodes = 0;
    int64_t num_elements = 0;
    bool ready = false;
};

// Element initialisation (volume/J asserts, lumped masses) and the block pattern of the would-be
// assembled matrix (for the pattern query). Throws std::runtime_error on invalid elements.
void hex_setup(HexData& h, int64_t num_nodes, const double* x0, int64_t num_elements, const int32_t* conn,
               const double* mat, std::vector<double>& mass_per_node, std::vector<int32_t>& rowptr_b,
               std::vector<int32_t>& colidx_b, hipStream_t s);

void hex_step(HexData& h, const fem_options& opt, const double* x0, double* x, double* v, const double* fg,
              const double* fext, const uint8_t* bcmask, double* partial, CgScalars* sc, CgScalars* h_sc,
              int* invalid, hipStream_t s, int& predicted_iters, fem_stats& stats, int max_iter);

void hex_eval_elements(HexData& h, int first, int count, const double* x0, const double* x, double* R, double* K,
                       double* f, hipStream_t s);

}  // namespace femgpu
