// FP64 issue-rate probes for the roofline: what one MI355X sustains on register-only streams of
// v_mfma_f64_16x16x4f64 and of FP64 vector FMAs, beside the published 78.6 TFLOP/s of both pipes.
//
// Every lane keeps NACC independent accumulators, so no instruction waits for the previous one's result, and
// writes their sum once at the end (out holds blocks x 256 doubles). The host times a launch with events and
// divides the operations by the time: MFMA blocks x 4 waves x iters x NACC x 2048 flop, VALU blocks x 256 lanes
// x iters x NACC x 2 flop.
#include "abi_util.h"

namespace seqvp {

typedef double d4 __attribute__((ext_vector_type(4)));
constexpr int CEIL_NACC = 8;

__global__ void __launch_bounds__(256) mfma_f64_rate_kernel(double* __restrict__ out, int64_t iters) {
  d4 acc[CEIL_NACC];
#pragma unroll
  for (int j = 0; j < CEIL_NACC; ++j) acc[j] = d4{0.0, 0.0, 0.0, 0.0};
  const double a = 1.0 + threadIdx.x * 1e-9, b = 1.0 - threadIdx.x * 1e-9;
  for (int64_t i = 0; i < iters; ++i) {
#pragma unroll
    for (int j = 0; j < CEIL_NACC; ++j) acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[j], 0, 0, 0);
  }
  double s = 0.0;
#pragma unroll
  for (int j = 0; j < CEIL_NACC; ++j) s += (acc[j][0] + acc[j][1]) + (acc[j][2] + acc[j][3]);
  out[(int64_t)blockIdx.x * 256 + threadIdx.x] = s;
}

__global__ void __launch_bounds__(256) valu_f64_rate_kernel(double* __restrict__ out, int64_t iters) {
  double x[CEIL_NACC];
#pragma unroll
  for (int j = 0; j < CEIL_NACC; ++j) x[j] = 1.0 + (threadIdx.x + j) * 1e-9;
  const double m = 0.999999999, c = 1e-9;
  for (int64_t i = 0; i < iters; ++i) {
#pragma unroll
    for (int j = 0; j < CEIL_NACC; ++j) x[j] = __builtin_fma(x[j], m, c);
  }
  double s = 0.0;
#pragma unroll
  for (int j = 0; j < CEIL_NACC; ++j) s += x[j];
  out[(int64_t)blockIdx.x * 256 + threadIdx.x] = s;
}

}  // namespace seqvp

using namespace seqvp;

extern "C" int seqvp_fp64_ceiling_probe(int kind, int64_t blocks, int64_t iters, double* out, void* stream) {
  if (!out || blocks < 1 || blocks > 65536 || iters < 1 || (kind != 0 && kind != 1)) return SEQVP_EINVAL;
  if (kind == 0)
    hipLaunchKernelGGL(mfma_f64_rate_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, out, iters);
  else
    hipLaunchKernelGGL(valu_f64_rate_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, out, iters);
  return launch_status();
}
