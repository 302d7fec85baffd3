// Shared definitions for the gfx950 kernels of the TACS assembly / block-CSR /
// Krylov / adjoint path.  Wavefront size is 64 (CDNA).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define TACS_WAVE 64

// Kernel-side status codes written to a device status word.
enum {
  TACS_HIP_OK = 0,
  TACS_HIP_ERR_ROW_TOO_LONG = 1001,
  TACS_HIP_ERR_TOO_MANY_ELEMS = 1002,
  TACS_HIP_ERR_MISSING_ENTRY = 1003,
  TACS_HIP_ERR_ARG = 1004,
  TACS_HIP_ERR_SINGULAR = 1005,
};

#define TACS_HIP_CHECK(expr)                      \
  do {                                            \
    hipError_t _e = (expr);                       \
    if (_e != hipSuccess) return (int)_e;         \
  } while (0)

// Sum across the 64 lanes of a wavefront; every lane receives the result.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, TACS_WAVE);
  return v;
}
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, TACS_WAVE);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, TACS_WAVE));
  return v;
}

// SolidMaterial (packed C[21], density, yield stress) is declared in the C ABI.
#include "../../include/tacs_hip.h"

// Number of launches' worth of workgroups for grid-stride kernels.
static inline int tacs_grid_for(long work, int per_block, int cap) {
  long g = (work + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (int)g;
}
