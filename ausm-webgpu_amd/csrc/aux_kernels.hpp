// Launchers of the setup, epilogue and layout kernels (internal).
#pragma once
#include <hip/hip_runtime.h>

#include "device_math.hpp"

namespace ausm {

hipError_t launch_grid_interp(double* gx, double* gy, int M, int N, int P, hipStream_t s);
hipError_t launch_grid_jacobi(const double* gx, const double* gy, double* ox, double* oy, int M,
                              int N, int P, hipStream_t s);
// grid_block_sweeps() Jacobi sweeps (or red-black half-sweeps, colours
// alternating from color0) in one launch, bit-for-bit as many
// launch_grid_jacobi (launch_grid_sor) calls
int grid_block_sweeps();
hipError_t launch_grid_jacobi_block(const double* gx, const double* gy, double* ox, double* oy,
                                    int M, int N, int P, hipStream_t s);
hipError_t launch_grid_sor_block(const double* gx, const double* gy, double* ox, double* oy, int M,
                                 int N, int P, int color0, double omega, hipStream_t s);
// one red-black half-sweep of colour `color`, from (gx, gy) into (ox, oy)
hipError_t launch_grid_sor(const double* gx, const double* gy, double* ox, double* oy, int M,
                           int N, int P, int color, double omega, hipStream_t s);
hipError_t launch_f64_to_f32(const double* a, float* b, size_t n, hipStream_t s);
hipError_t launch_metrics(const float* vx, const float* vy, int M, int N, int P, float* area,
                          float* flen, hipStream_t s);
int wave_speed_blocks(int M, int N);
// R = float (the solver state) or double (the all-FP64 solve)
template <typename R>
hipError_t launch_wave_speed(const R* q, const R* vx, const R* vy, const DevParamsT<R>* par, int M,
                             int N, int P, int PS, R* ws, R* part, hipStream_t s);
template <typename R>
hipError_t launch_reduce_max(const R* part, int n, R* out, hipStream_t s);
hipError_t launch_reduce_sum_f64(const double* part, int n, double* out, hipStream_t s);
template <typename R>
hipError_t launch_forces(const R* q, const R* vx, const R* vy, const DevParamsT<R>* par, int M,
                         int PS, double ref_len, double* seg, double* out, double* cp,
                         hipStream_t s);
template <typename R>
hipError_t launch_aos_to_soa(const R* aos, R* q, int M, int N, int P, int PS, hipStream_t s);
template <typename R>
hipError_t launch_soa_to_aos(const R* q, const R* vx, const R* vy, const DevParamsT<R>* par, int M,
                             int N, int P, int PS, R* aos, hipStream_t s);
template <typename R>
hipError_t launch_planes_to_aos(const R* q, int M, int N, int P, int PS, R* aos, hipStream_t s);
template <typename R>
hipError_t launch_fill_state(R* q, const DevParamsT<R>* par, int M, int N, int P, int PS,
                             int freestream, hipStream_t s);
template <typename A, typename B>
hipError_t launch_convert_planes(const A* a, B* b, int M, int N, int P, int PS, hipStream_t s);
// derived display field of every interior cell (N x M floats) and its
// free-stream value
template <typename R>
hipError_t launch_derived_field(const R* q, const DevParamsT<R>* par, int M, int N, int P, int PS,
                                int mode, float* out, float* fs_out, hipStream_t s);
// compensated pair hi + lo (exact in FP64) <-> FP64 state planes
hipError_t launch_comp_to_f64(const float* hi, const float* lo, double* q, int M, int N, int P,
                              int PS, hipStream_t s);
hipError_t launch_f64_to_comp(const double* q, float* hi, float* lo, int M, int N, int P, int PS,
                              hipStream_t s);

// recip_rn(d) against the IEEE 1/d for every FP32 divisor in [2^-126, 2^126)
// and sqrt_rn(x) against the IEEE square root for x = +0 and every FP32 x in
// [2^-96, 2^128); adds the number of mismatches to *bad (which the caller
// clears)
hipError_t launch_arithmetic_check(unsigned long long* bad, hipStream_t s);

}  // namespace ausm
