// Fused multi-tensor Adam step (torch.optim.Adam semantics, no weight decay,
// no amsgrad), with an optional BF16 shadow copy of each parameter for the
// kernels that gather from BF16 tables.
//
// Reference: train_per_scene.py:132 (Adam, betas (0.9, 0.99)), the notebooks'
// optimisers (scripts/2D_regression.ipynb:177, scripts/sdf_regression.ipynb:356).
// Per element, in torch's order:
//   m = m + (1-b1) * (g - m)           (lerp_ with weight 1-b1 < 0.5)
//   v = v * b2 + (1-b2) * g * g        (mul_ then addcmul_)
//   p = p - step_size * m / (sqrt(v) / sqrt(bc2) + eps)
#include "ff_common.h"

namespace ff {

constexpr int kAdamMax = 32;
// Elements per block: 2048 balances the grid (measured over 1,023 NeRF steps,
// 12.8 M parameters: 63.3 us per step against 64.4 at 1024, 66.3 at 4096,
// 69.8 at 8192).
constexpr int kAdamChunk = 2048;

struct AdamBatch {
  float* param[kAdamMax];
  float* grad[kAdamMax];
  float* m[kAdamMax];
  float* v[kAdamMax];
  uint16_t* shadow[kAdamMax];
  int64_t n[kAdamMax];
  float step_size[kAdamMax];  // lr / bias_correction1
  int32_t reps[kAdamMax];
  int64_t rstride[kAdamMax];
  int chunk_start[kAdamMax + 1];
  int count;
};

__global__ __launch_bounds__(256) void adam_kernel(AdamBatch b, float beta1, float beta2, float eps,
                                                   float bc2_sqrt, int zero_grad) {
  const int blk = blockIdx.x;
  int t = 0;
  while (t + 1 < b.count && b.chunk_start[t + 1] <= blk) ++t;
  const int64_t base = (int64_t)(blk - b.chunk_start[t]) * kAdamChunk;
  const int64_t n = b.n[t];
  float* __restrict__ P = b.param[t];
  float* __restrict__ G = b.grad[t];
  float* __restrict__ M = b.m[t];
  float* __restrict__ V = b.v[t];
  uint16_t* __restrict__ S = b.shadow[t];
  const float ss = b.step_size[t];
  const int R = b.reps[t];
  const int64_t rs = b.rstride[t];
  const float w1 = 1.f - beta1, w2 = 1.f - beta2;
  auto update = [&](float g, float& m, float& v, float& p) {
    m = m + w1 * (g - m);
    v = v * beta2;
    v = v + w2 * g * g;
    const float denom = sqrtf(v) / bc2_sqrt + eps;
    p = p + (-ss) * m / denom;
  };
  // one gradient copy, whole chunk in range, 16-byte aligned rows: four
  // elements per lane per pass, the loads of a pass issued together
  const bool vec = R == 1 && base + kAdamChunk <= n &&
                   ((reinterpret_cast<uintptr_t>(P) | reinterpret_cast<uintptr_t>(G) |
                     reinterpret_cast<uintptr_t>(M) | reinterpret_cast<uintptr_t>(V)) & 15) == 0 &&
                   (!S || (reinterpret_cast<uintptr_t>(S) & 7) == 0);
  if (vec) {
    for (int k = 4 * threadIdx.x; k < kAdamChunk; k += 4 * blockDim.x) {
      const int64_t i = base + k;
      const float4 g4 = *reinterpret_cast<const float4*>(G + i);
      float4 m4 = *reinterpret_cast<const float4*>(M + i);
      float4 v4 = *reinterpret_cast<const float4*>(V + i);
      float4 p4 = *reinterpret_cast<const float4*>(P + i);
      update(g4.x, m4.x, v4.x, p4.x);
      update(g4.y, m4.y, v4.y, p4.y);
      update(g4.z, m4.z, v4.z, p4.z);
      update(g4.w, m4.w, v4.w, p4.w);
      *reinterpret_cast<float4*>(M + i) = m4;
      *reinterpret_cast<float4*>(V + i) = v4;
      *reinterpret_cast<float4*>(P + i) = p4;
      if (S) {
        const uint32_t lo = (uint32_t)f32_to_bf16(p4.x) | ((uint32_t)f32_to_bf16(p4.y) << 16);
        const uint32_t hi = (uint32_t)f32_to_bf16(p4.z) | ((uint32_t)f32_to_bf16(p4.w) << 16);
        *reinterpret_cast<uint2*>(S + i) = make_uint2(lo, hi);
      }
      if (zero_grad) *reinterpret_cast<float4*>(G + i) = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    return;
  }
  for (int k = threadIdx.x; k < kAdamChunk; k += blockDim.x) {
    const int64_t i = base + k;
    if (i >= n) break;
    float g = G[i];
    for (int r = 1; r < R; ++r) g += G[i + r * rs];
    float m = M[i], v = V[i], p = P[i];
    update(g, m, v, p);
    M[i] = m;
    V[i] = v;
    P[i] = p;
    if (S) S[i] = f32_to_bf16(p);
    if (zero_grad)
      for (int r = 0; r < R; ++r) G[i + r * rs] = 0.f;
  }
}

__global__ __launch_bounds__(256) void reduce_replicas_kernel(float* __restrict__ a, int R, int64_t n,
                                                              int64_t stride) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = a[i];
  for (int r = 1; r < R; ++r) {
    s += a[i + r * stride];
    a[i + r * stride] = 0.f;
  }
  a[i] = s;
}

__global__ __launch_bounds__(256) void cast_bf16_kernel(const float* __restrict__ src,
                                                        uint16_t* __restrict__ dst, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = f32_to_bf16(src[i]);
}

}  // namespace ff

using namespace ff;

extern "C" {

int ff_adam_multi(const ff_adam_tensor* tensors, int count, float beta1, float beta2, float eps,
                  int64_t step, int zero_grad, void* stream) {
  if (!tensors || count < 0 || step < 1) return FF_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  const double bc1 = 1.0 - std::pow((double)beta1, (double)step);
  const double bc2 = 1.0 - std::pow((double)beta2, (double)step);
  const float bc2_sqrt = (float)std::sqrt(bc2);
  int done = 0;
  while (done < count) {
    AdamBatch b{};
    int chunks = 0;
    int k = 0;
    while (done + k < count && k < kAdamMax) {
      const ff_adam_tensor& t = tensors[done + k];
      if (t.n < 0 || (t.n > 0 && (!t.param || !t.grad || !t.exp_avg || !t.exp_avg_sq)))
        return FF_ERR_ARG;
      const int64_t c = (t.n + kAdamChunk - 1) / kAdamChunk;
      if ((int64_t)chunks + c > (int64_t)0x7fffffff) break;
      b.param[k] = t.param; b.grad[k] = t.grad; b.m[k] = t.exp_avg; b.v[k] = t.exp_avg_sq;
      b.shadow[k] = t.shadow_bf16; b.n[k] = t.n;
      b.step_size[k] = (float)((double)t.lr / bc1);
      b.reps[k] = t.grad_replicas > 1 ? t.grad_replicas : 1;
      b.rstride[k] = t.grad_stride;
      if (b.reps[k] > 1 && t.grad_stride < t.n) return FF_ERR_ARG;
      b.chunk_start[k] = chunks;
      chunks += (int)c;
      ++k;
    }
    if (k == 0) return FF_ERR_ARG;
    b.chunk_start[k] = chunks;
    b.count = k;
    if (chunks > 0) {
      hipLaunchKernelGGL(adam_kernel, dim3(chunks), dim3(256), 0, s, b, beta1, beta2, eps, bc2_sqrt,
                         zero_grad);
      if (hipGetLastError() != hipSuccess) return FF_ERR_LAUNCH;
    }
    done += k;
  }
  return FF_OK;
}

int ff_adam_step(float* param, float* grad, float* exp_avg, float* exp_avg_sq, uint16_t* shadow_bf16,
                 int64_t n, float lr, float beta1, float beta2, float eps, int64_t step,
                 int zero_grad, void* stream) {
  ff_adam_tensor t{param, grad, exp_avg, exp_avg_sq, shadow_bf16, n, lr, 1, 0};
  return ff_adam_multi(&t, 1, beta1, beta2, eps, step, zero_grad, stream);
}

int ff_reduce_replicas(float* arena, int32_t replicas, int64_t n, int64_t stride, void* stream) {
  if (!arena || replicas < 1 || n < 0 || (replicas > 1 && stride < n)) return FF_ERR_ARG;
  if (replicas == 1 || n == 0) return FF_OK;
  hipLaunchKernelGGL(reduce_replicas_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, arena, replicas, n, stride);
  return hipGetLastError() == hipSuccess ? FF_OK : FF_ERR_LAUNCH;
}

int ff_cast_bf16(const float* src, uint16_t* dst, int64_t n, void* stream) {
  if (n < 0 || (n > 0 && (!src || !dst))) return FF_ERR_ARG;
  if (n == 0) return FF_OK;
  hipLaunchKernelGGL(cast_bf16_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, src, dst, n);
  return hipGetLastError() == hipSuccess ? FF_OK : FF_ERR_LAUNCH;
}

}  // extern "C"
