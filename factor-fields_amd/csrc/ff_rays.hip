// Camera rays and training-batch gathers on the device.
//
// ff_camera_rays: get_ray_directions / get_ray_directions_blender
// (dataLoader/ray_utils.py:27-66), the blender loader's normalisation
// (dataLoader/blender.py:54-55) and get_rays (ray_utils.py:69-90) fused, one
// thread per ray, in the reference's FP32 operation order: (x + 0.5 - cx) /
// fx, the norm as sqrt(x^2 + y^2 + z^2), the rotation as the three-term sum
// d0 R[i][0] + d1 R[i][1] + d2 R[i][2].
//
// ff_gather_rays: allrays[ray_idx], allrgbs[ray_idx] (train_per_scene.py:
// 151-152), one thread per output float: consecutive threads write
// consecutive floats and read the 24 (12) bytes of one source row together.
#include <cmath>

#include "ff_common.h"

namespace ff {
namespace {

__global__ __launch_bounds__(256) void camera_rays_kernel(ff_cameras c, const int64_t* __restrict__ view,
                                                          const int64_t* __restrict__ pix, int64_t n,
                                                          float* __restrict__ rays) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int64_t hw = (int64_t)c.H * c.W;
  int64_t v, p;
  if (view) {
    v = view[t];
    p = pix[t];
  } else {
    v = t / hw;
    p = t - v * hw;
  }
  float* o = rays + t * 6;
  if (v < 0 || v >= c.n_cams || p < 0 || p >= hw) {
#pragma unroll
    for (int k = 0; k < 6; ++k) o[k] = __builtin_nanf("");
    return;
  }
  const float x = (float)(p % c.W) + 0.5f;
  const float y = (float)(p / c.W) + 0.5f;
  float d[3] = {(x - c.cx) / c.fx, (y - c.cy) / c.fy, 1.f};
  if (c.blender) {
    d[1] = -d[1];
    d[2] = -1.f;
  }
  if (c.normalize) {
    const float nrm = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
#pragma unroll
    for (int k = 0; k < 3; ++k) d[k] = d[k] / nrm;
  }
  const float* M = c.c2w + v * 12;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    o[i] = M[i * 4 + 3];
    o[3 + i] = d[0] * M[i * 4 + 0] + d[1] * M[i * 4 + 1] + d[2] * M[i * 4 + 2];
  }
}

__global__ __launch_bounds__(256) void gather_rays_kernel(const float* __restrict__ rays,
                                                          const float* __restrict__ rgbs, int64_t n_total,
                                                          const int64_t* __restrict__ idx, int64_t B,
                                                          float* __restrict__ out_rays,
                                                          float* __restrict__ out_rgbs) {
  const int per = out_rgbs ? 9 : 6;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B * per) return;
  const int64_t b = t / per;
  const int k = (int)(t - b * per);
  const int64_t j = idx[b];
  const bool ok = j >= 0 && j < n_total;
  if (k < 6)
    out_rays[b * 6 + k] = ok ? rays[j * 6 + k] : __builtin_nanf("");
  else
    out_rgbs[b * 3 + k - 6] = ok ? rgbs[j * 3 + k - 6] : __builtin_nanf("");
}

int rays_status() { return hipGetLastError() == hipSuccess ? FF_OK : FF_ERR_LAUNCH; }

}  // namespace
}  // namespace ff

using namespace ff;

extern "C" {

int ff_camera_rays(const ff_cameras* cams, const int64_t* view, const int64_t* pix, int64_t n, float* rays,
                   void* stream) {
  if (!cams || n < 0 || !cams->c2w || cams->n_cams <= 0 || cams->H <= 0 || cams->W <= 0) return FF_ERR_ARG;
  if (!(cams->fx != 0.f) || !(cams->fy != 0.f)) return FF_ERR_ARG;
  if (n == 0) return FF_OK;  // (an empty pair list may come with NULL pointers)
  if ((view == nullptr) != (pix == nullptr)) return FF_ERR_ARG;
  if (!view && n != (int64_t)cams->n_cams * cams->H * cams->W) return FF_ERR_ARG;
  if (!rays) return FF_ERR_ARG;
  hipLaunchKernelGGL(camera_rays_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                     *cams, view, pix, n, rays);
  return rays_status();
}

int ff_gather_rays(const float* rays, const float* rgbs, int64_t n_total, const int64_t* idx, int64_t B,
                   float* out_rays, float* out_rgbs, void* stream) {
  if (B < 0 || n_total < 0) return FF_ERR_ARG;
  if (B == 0) return FF_OK;  // (an empty batch may come with NULL pointers)
  if ((rgbs == nullptr) != (out_rgbs == nullptr)) return FF_ERR_ARG;
  if (!rays || !idx || !out_rays) return FF_ERR_ARG;
  const int64_t total = B * (out_rgbs ? 9 : 6);
  hipLaunchKernelGGL(gather_rays_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, rays, rgbs, n_total, idx, B, out_rays, out_rgbs);
  return rays_status();
}

}  // extern "C"
