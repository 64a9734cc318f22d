// Launchers of the setup, epilogue and layout kernels (internal).
#pragma once
#include <hip/hip_runtime.h>

#include "device_math.hpp"

namespace ausm {

hipError_t launch_grid_interp(double* gx, double* gy, int M, int N, int P, hipStream_t s);
hipError_t launch_grid_jacobi(const double* gx, const double* gy, double* ox, double* oy, int M,
                              int N, int P, hipStream_t s);
hipError_t launch_grid_sor(double* gx, double* gy, int M, int N, int P, int color, double omega,
                           hipStream_t s);
hipError_t launch_f64_to_f32(const double* a, float* b, size_t n, hipStream_t s);
hipError_t launch_metrics(const float* vx, const float* vy, int M, int N, int P, float* area,
                          float* flen, hipStream_t s);
int wave_speed_blocks(int M, int N);
hipError_t launch_wave_speed(const float* q, const float* vx, const float* vy, const DevParams* par,
                             int M, int N, int P, int PS, float* ws, float* part, hipStream_t s);
hipError_t launch_reduce_max(const float* part, int n, float* out, hipStream_t s);
hipError_t launch_reduce_sum_f64(const double* part, int n, double* out, hipStream_t s);
hipError_t launch_forces(const float* q, const float* vx, const float* vy, const DevParams* par,
                         int M, int PS, double ref_len, double* seg, double* out, double* cp,
                         hipStream_t s);
hipError_t launch_aos_to_soa(const float* aos, float* q, int M, int N, int P, int PS,
                             hipStream_t s);
hipError_t launch_soa_to_aos(const float* q, const float* vx, const float* vy,
                             const DevParams* par, int M, int N, int P, int PS, float* aos,
                             hipStream_t s);
hipError_t launch_planes_to_aos(const float* q, int M, int N, int P, int PS, float* aos,
                                hipStream_t s);
hipError_t launch_fill_state(float* q, const DevParams* par, int M, int N, int P, int PS,
                             int freestream, hipStream_t s);

}  // namespace ausm
