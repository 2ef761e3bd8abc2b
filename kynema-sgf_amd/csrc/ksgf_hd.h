// Function qualifier for code that runs in the HIP kernels and, compiled as
// plain host C++, in the CPU checks of those kernels.
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define KSGF_HD __host__ __device__ __forceinline__
#else
#include <cmath>
#define KSGF_HD inline
#endif
