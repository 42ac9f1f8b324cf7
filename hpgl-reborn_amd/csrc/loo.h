// loo.h -- leave-one-out kriging of scattered points and its adjoint (loo.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <string>

namespace hpglx {

struct LooArgs {
    int n;               // points
    int k;               // neighbours per point
    int np;              // parameter sets
    int model;           // 0 spherical, 1 exponential, 2 gaussian
    const float* xyz;    // n x 3 coordinates
    const float* val;    // n values
    const int* nbr;      // n x k neighbour indices
    const double* prm;   // np x 5: sill, nugget, rx, ry, rz
    double* est;         // np x n leave-one-out estimates
    double* state;       // forward -> backward state (loo_state_doubles), or null
    int* status;         // np x n: 0 kriged, 2 singular (estimate = neighbour mean), or null
    const double* gest;  // backward: np x n gradient of the loss w.r.t. est (null = zero)
    double* jac;         // backward: np x n x 5 d est / d params, or null
    double* part;        // backward: partial sums (loo_part_doubles)
};

size_t loo_state_doubles(int n, int k, int np);
size_t loo_part_doubles(int n, int k, int np);
int loo_forward(const LooArgs& a, hipStream_t st, std::string& err);
// grad (np x 5) = sum_i gest[p, i] * d est[p, i] / d params[p]
int loo_backward(const LooArgs& a, double* grad, hipStream_t st, std::string& err);
int knn(const float* xyz, int n, int k, int* nbr, hipStream_t st, std::string& err);

}  // namespace hpglx
