// Launchers of the adjoint kernels (fem_adjoint.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace femgpu {

struct AdjointArgs {
    int num_elements;
    const int4* conn;
    const double* x0;      // rest positions
    const double* x;       // x_n (state the step started from)
    const double* v;       // v_n
    const double* v_next;  // v_{n+1} (dv = v_{n+1} - v_n)
    const double* mu;      // solution of S mu = lv + dt lx
    const uint8_t* bcmask;
    const double4* mat;    // (lambda, mu, rho, E) per element
    double dt;
    double* gx_rec;        // [E][12] d phi_e / d x_e
    double* gv_rec;        // [E][12] -dt K_e^T mu~_e
    double* grad_E;        // [E] accumulated dL/dE_e
};

void launch_adjoint_rhs(int n, const double* lx, const double* lv, double dt, double* gv, hipStream_t s);
void launch_adjoint_elements(int kind, const AdjointArgs& a, hipStream_t s);
void launch_adjoint_gather(int num_nodes, const int* node_ptr, const int* node_list, const double* gx_rec,
                           const double* gv_rec, const double* lx, const double* gv, double* lx_out, double* lv_out,
                           hipStream_t s);

}  // namespace femgpu
