// FP32 atomic line-request rate: the ceiling the backward kernels' scatters
// are measured against.  Each wave instruction adds 1.0 to 64 floats that
// fall in `lines_per_instr` distinct 64-byte lines (64 / lines_per_instr
// consecutive lanes share a line, on consecutive floats), the lines spread
// over the target by a multiplicative hash, with the same hardware atomic
// (global_atomic_add_f32, no return) as the scatters.
#include "ff_common.h"

namespace ff {
namespace {

__global__ __launch_bounds__(256) void atomic_rate_kernel(float* __restrict__ dst, int64_t n_lines, int64_t n_instr,
                                                          int lanes_per_line) {
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  const int lane = threadIdx.x & 63;
  const int grp = lane / lanes_per_line, sub = lane - grp * lanes_per_line;
  for (int64_t w = wave; w < n_instr; w += nwaves) {
    const uint64_t h = ((uint64_t)(w * 64 + grp) * 0x9E3779B97F4A7C15ull) >> 20;
    const int64_t line = (int64_t)(h % (uint64_t)n_lines);
    unsafeAtomicAdd(dst + line * 16 + sub, 1.f);
  }
}

}  // namespace
}  // namespace ff

extern "C" {

int ff_atomic_rate_probe(float* dst, int64_t n_lines, int64_t n_instr, int32_t lines_per_instr, void* stream) {
  if (!dst || n_lines <= 0 || n_instr < 0) return FF_ERR_ARG;
  if (lines_per_instr != 4 && lines_per_instr != 16 && lines_per_instr != 64) return FF_ERR_ARG;
  if (n_instr == 0) return FF_OK;
  const int64_t blocks = (int64_t)ff::kWave * 1024 * 4 / 256;  // 4 waves per SIMD on 256 CUs, grid-stride
  hipLaunchKernelGGL(ff::atomic_rate_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, dst, n_lines,
                     n_instr, 64 / lines_per_instr);
  return hipGetLastError() == hipSuccess ? FF_OK : FF_ERR_LAUNCH;
}

}  // extern "C"
