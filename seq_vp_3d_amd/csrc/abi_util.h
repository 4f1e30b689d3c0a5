// Helpers shared by the C-ABI entry points.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/seqvp.h"

// Status of the launches just enqueued by this host thread: SEQVP_ELAUNCH if any failed.
static inline int launch_status() {
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? SEQVP_OK : SEQVP_ELAUNCH;
}

// dst = src, or dst = 0 when src is NULL (n doubles).
static __global__ void __launch_bounds__(256) copy_or_zero_kernel(const double* __restrict__ src, double* __restrict__ dst,
                                                           int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) dst[i] = src ? src[i] : 0.0;
}

// Per-module workspace sizes (defined in each translation unit).
size_t seqvp_q1_workspace_bytes(int op, int64_t nx, int64_t ny, int64_t nz);
size_t seqvp_dst_workspace_bytes(int64_t mx, int64_t my, int64_t mz, int64_t ncomp);
size_t seqvp_vec_workspace_bytes(int64_t n);
