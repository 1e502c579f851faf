// Launchers of the simulation kernels in sim.hip (internal).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fft_lds.hpp"

namespace cmbl {

template <typename T>
hipError_t launch_grf(const double* cov, const double* mul, const cplx<T>* add, cplx<T>* out,
                      int N, int C, long ntrials, uint64_t seed, uint64_t stream0,
                      hipStream_t s);

}  // namespace cmbl
