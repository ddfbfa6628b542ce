// Ray sampling with occupancy skipping and alpha compositing (gfx950, wave64).
//
// Sampler (models/FactorFields.py:586-602 sample_point, :91-110 AlphaGridMask,
// :859-867 masking): one wavefront per ray, lanes take 64 consecutive
// candidates at a time; validity is a ballot, and surviving samples are
// written ray-major at (ray offset + popcount of lower lanes), which is the
// order of the reference's boolean indexing.  Counting and writing are two
// passes over the same arithmetic; the per-ray counts are turned into
// offsets by a one-block scan, so the total never travels to the host.
//
// Compositing (models/FactorFields.py:82-88 raw2alpha, :878-896): one
// wavefront per ray; each round of 64 samples does a multiplicative prefix
// scan of (1 - alpha + 1e-10) across lanes (shuffles, no LDS) and carries the
// running transmittance to the next round.  The backward pass recomputes the
// forward scan, then walks the rounds in reverse with a suffix scan of the
// affine maps x -> e_i*alpha_i + (1-alpha_i+1e-10)*x, which yields
// dL/dalpha_i = T_i * (e_i - R_i) without dividing by (1 - alpha_i).
//
// Arithmetic that decides the masks is written operation by operation in the
// reference's order (built with -ffp-contract=off).
#include "ff_common.h"

namespace ff {

struct RaySetup {
  float o[3], d[3];
  float tmin;
};

__device__ __forceinline__ void ray_setup(const ff_ray_sampler& s, const float* __restrict__ rays,
                                          int64_t b, RaySetup& r) {
  float tmin = -INFINITY;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    r.o[a] = rays[b * 6 + a];
    r.d[a] = rays[b * 6 + 3 + a];
    const float vec = r.d[a] == 0.f ? 1e-6f : r.d[a];
    const float ra = (s.aabb_max[a] - r.o[a]) / vec;
    const float rb = (s.aabb_min[a] - r.o[a]) / vec;
    tmin = fmaxf(tmin, fminf(ra, rb));
  }
  r.tmin = fminf(fmaxf(tmin, 0.05f), 1e3f);
}

// t of candidate i: t_min + stepSize * (i [+ u]).
__device__ __forceinline__ float cand_t(const ff_ray_sampler& s, const RaySetup& r, int i, float u,
                                        bool jit) {
  float rng = (float)i;
  if (jit) rng = rng + u;
  const float step = s.step_size * rng;
  return r.tmin + step;
}

// Occupancy test: trilinear grid_sample (align_corners=True, zeros) of the
// 0/1 volume, > 0.5.  The sum runs over in-bounds corners in the reference's
// order with unfused arithmetic so the mask is bit-exact.
__device__ __forceinline__ bool occ_test(const ff_ray_sampler& s, const float* p) {
  float pos[3];
  const int R[3] = {s.occ_reso[0], s.occ_reso[1], s.occ_reso[2]};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float c = (p[a] - s.occ_origin[a]) * s.occ_inv[a] - 1.f;
    pos[a] = unnorm_ac(c, R[a]);
  }
  Corners<3> cr;
  linear_corners<3>(pos, R, cr);
  float acc = 0.f;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const float v = s.occ[cr.tex[k]] ? cr.w[k] : 0.f;
    acc = acc + v;
  }
  return acc > 0.5f;
}

__device__ __forceinline__ bool cand_valid(const ff_ray_sampler& s, const RaySetup& r, float t,
                                           float* p) {
  bool inner = true;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float dt = r.d[a] * t;
    p[a] = r.o[a] + dt;
    inner = inner && !(s.aabb_min[a] > p[a] || p[a] > s.aabb_max[a]);
  }
  if (!inner) return false;
  if (s.occ) return occ_test(s, p);
  return true;
}

// pass 0: count valid samples per ray into ray_off[b+1]; pass 1: write them.
__global__ __launch_bounds__(256) void sample_rays_kernel(ff_ray_sampler s, const float* __restrict__ rays,
                                                          const float* __restrict__ jitter, int64_t B,
                                                          int pass, ff_samples out) {
  const int lane = threadIdx.x & 63;
  const int64_t b = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (b >= B) return;
  RaySetup r;
  ray_setup(s, rays, b, r);
  const bool jit = jitter != nullptr;
  const float u = jit ? jitter[b] : 0.f;
  const int Ns = s.n_samples;
  int base = pass == 1 ? out.ray_off[b] : 0;
  int cnt = 0;
  const uint64_t lower = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (int i0 = 0; i0 < Ns; i0 += 64) {
    const int i = i0 + lane;
    float p[3];
    float t = 0.f;
    bool valid = false;
    if (i < Ns) {
      t = cand_t(s, r, i, u, jit);
      valid = cand_valid(s, r, t, p);
    }
    const uint64_t mask = __ballot(valid);
    if (pass == 1 && valid) {
      const int64_t pos = (int64_t)base + cnt + __popcll(mask & lower);
      const float tn = (i + 1 < Ns) ? cand_t(s, r, i + 1, u, jit) : t;
      const float dist = (i + 1 < Ns) ? (tn - t) : 0.f;
      out.xyz[pos * 3 + 0] = p[0];
      out.xyz[pos * 3 + 1] = p[1];
      out.xyz[pos * 3 + 2] = p[2];
      out.z[pos] = t;
      out.dist[pos] = dist * s.distance_scale;
      out.ray[pos] = (int32_t)b;
      out.idx[pos] = i;
    }
    cnt += __popcll(mask);
  }
  if (pass == 0 && lane == 0) out.ray_off[b + 1] = cnt;
}

// In-place inclusive scan of off[1..B] (one block of 1024), off[0] = 0.
__global__ __launch_bounds__(1024) void scan_offsets_kernel(int32_t* off, int64_t B) {
  __shared__ int32_t part[1024];
  const int t = threadIdx.x;
  const int64_t per = (B + 1023) / 1024;
  const int64_t lo = t * per, hi = lo + per < B ? lo + per : B;
  int32_t s = 0;
  for (int64_t i = lo; i < hi; ++i) s += off[i + 1];
  part[t] = s;
  __syncthreads();
  for (int d = 1; d < 1024; d <<= 1) {
    int32_t v = t >= d ? part[t - d] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int32_t run = t ? part[t - 1] : 0;
  for (int64_t i = lo; i < hi; ++i) {
    run += off[i + 1];
    off[i + 1] = run;
  }
  if (t == 0) off[0] = 0;
}

// ------------------------------------------------------------ compositing
__device__ __forceinline__ float shfl_up_f(float v, int d) { return __shfl_up(v, d, 64); }
__device__ __forceinline__ float shfl_down_f(float v, int d) { return __shfl_down(v, d, 64); }

// Inclusive multiplicative scan across the wave.
__device__ __forceinline__ float wave_prod_scan(float v, int lane) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const float t = shfl_up_f(v, d);
    if (lane >= d) v = v * t;
  }
  return v;
}

__global__ __launch_bounds__(256) void composite_fwd_kernel(ff_samples s, int64_t B,
                                                            const float* __restrict__ sigma,
                                                            const float* __restrict__ rgb, int white_bg,
                                                            float thres, float* __restrict__ rgb_map,
                                                            float* __restrict__ rgb_raw,
                                                            float* __restrict__ depth,
                                                            float* __restrict__ acc_out,
                                                            float* __restrict__ weights) {
  const int lane = threadIdx.x & 63;
  const int64_t b = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (b >= B) return;
  const int64_t s0 = s.ray_off[b];
  const int cnt = s.ray_off[b + 1] - (int)s0;
  float carry = 1.f, acc = 0.f, c0 = 0.f, c1 = 0.f, c2 = 0.f, dep = 0.f;
  for (int r0 = 0; r0 < cnt; r0 += 64) {
    const int j = r0 + lane;
    const bool act = j < cnt;
    const int64_t k = s0 + j;
    const float sg = act ? sigma[k] : 0.f;
    const float dd = act ? s.dist[k] : 0.f;
    const float alpha = 1.f - expf(-sg * dd);
    const float f = act ? (1.f - alpha) + 1e-10f : 1.f;
    const float incl = wave_prod_scan(f, lane);
    float excl = shfl_up_f(incl, 1);
    if (lane == 0) excl = 1.f;
    const float T = carry * excl;
    const float w = act ? alpha * T : 0.f;
    if (act) {
      const bool app = w > thres;
      acc += w;
      if (app) {
        c0 += w * rgb[k * 3 + 0];
        c1 += w * rgb[k * 3 + 1];
        c2 += w * rgb[k * 3 + 2];
      }
      dep += w * s.z[k];
      if (weights) weights[k] = w;
    }
    carry = carry * __shfl(incl, 63, 64);
  }
  acc = wave_sum(acc);
  c0 = wave_sum(c0);
  c1 = wave_sum(c1);
  c2 = wave_sum(c2);
  dep = wave_sum(dep);
  if (lane == 0) {
    const float bg = white_bg ? (1.f - acc) : 0.f;
    const float r[3] = {c0 + bg, c1 + bg, c2 + bg};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      if (rgb_raw) rgb_raw[b * 3 + c] = r[c];
      rgb_map[b * 3 + c] = fminf(fmaxf(r[c], 0.f), 1.f);
    }
    if (depth) depth[b] = dep;
    if (acc_out) acc_out[b] = acc;
  }
}

// App samples (models/FactorFields.py:878-885): weight > thres, with the
// arithmetic of composite_fwd_kernel so both kernels agree on every sample.
// pass 0: count per ray into off[b + 1]; pass 1: write the sample indices in
// ray-major order, and each live sample's position (or -1) into app_pos.
__global__ __launch_bounds__(256) void select_app_kernel(ff_samples s, int64_t B, const float* __restrict__ sigma,
                                                         float thres, int pass, int32_t* __restrict__ off,
                                                         int32_t* __restrict__ rows,
                                                         int32_t* __restrict__ app_pos) {
  const int lane = threadIdx.x & 63;
  const int64_t b = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (b >= B) return;
  const int64_t s0 = s.ray_off[b];
  const int cnt = s.ray_off[b + 1] - (int)s0;
  const uint64_t lower = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  int base = pass == 1 ? off[b] : 0;
  int napp = 0;
  float carry = 1.f;
  for (int r0 = 0; r0 < cnt; r0 += 64) {
    const int j = r0 + lane;
    const bool act = j < cnt;
    const int64_t k = s0 + j;
    const float sg = act ? sigma[k] : 0.f;
    const float dd = act ? s.dist[k] : 0.f;
    const float alpha = 1.f - expf(-sg * dd);
    const float f = act ? (1.f - alpha) + 1e-10f : 1.f;
    const float incl = wave_prod_scan(f, lane);
    float excl = shfl_up_f(incl, 1);
    if (lane == 0) excl = 1.f;
    const float T = carry * excl;
    const float w = act ? alpha * T : 0.f;
    const bool app = act && w > thres;
    const uint64_t mask = __ballot(app);
    if (pass == 1 && act) {
      const int pos = base + napp + __popcll(mask & lower);
      if (app) rows[pos] = (int32_t)k;
      if (app_pos) app_pos[k] = app ? pos : -1;
    }
    napp += __popcll(mask);
    carry = carry * __shfl(incl, 63, 64);
  }
  if (pass == 0 && lane == 0) off[b + 1] = napp;
}

constexpr int kMaxRounds = 32;  // n_samples <= 2048

__global__ __launch_bounds__(256) void composite_bwd_kernel(
    ff_samples s, int64_t B, const float* __restrict__ sigma, const float* __restrict__ rgb,
    int white_bg, float thres, const float* __restrict__ rgb_raw, const float* __restrict__ d_rgb_map,
    const float* __restrict__ target, float loss_scale, float* __restrict__ loss_sum,
    float* __restrict__ d_sample) {
  const int lane = threadIdx.x & 63;
  const int wib = threadIdx.x >> 6;
  const int64_t b = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  // the ray's sample range, loaded beside the colour gradient's inputs (not
  // after the loss barrier: one memory round trip fewer per ray)
  const int64_t s0 = b < B ? s.ray_off[b] : 0;
  const int cnt = b < B ? (int)(s.ray_off[b + 1] - s0) : 0;
  // dL/d rgb_map, then through the clamp
  float g[3];
  float lsum = 0.f;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const float raw = b < B ? rgb_raw[b * 3 + c] : 0.f;
    float gc;
    if (target) {
      const float mapped = fminf(fmaxf(raw, 0.f), 1.f);
      const float e = b < B ? mapped - target[b * 3 + c] : 0.f;
      gc = (2.f * e) * (loss_scale / (float)(3.0 * (double)B));
      lsum += e * e;
    } else {
      gc = b < B ? d_rgb_map[b * 3 + c] : 0.f;
    }
    g[c] = (raw >= 0.f && raw <= 1.f) ? gc : 0.f;
  }
  // loss: one atomic per block (the rays' sums added in LDS), not one per ray
  if (target && loss_sum) {
    __shared__ float part[4];
    if (lane == 0) part[wib] = lsum;
    __syncthreads();
    if (threadIdx.x == 0) unsafeAtomicAdd(loss_sum, (part[0] + part[1]) + (part[2] + part[3]));
  }
  if (b >= B) return;
  const float gbg = white_bg ? (g[0] + g[1]) + g[2] : 0.f;
  const int rounds = (cnt + 63) / 64;
  if (rounds > kMaxRounds) return;  // the sampler never produces more than 2048 per ray
  // forward pass: transmittance at the start of each round; lane r keeps round r's
  // (a ray of one round starts at transmittance 1: nothing to compute)
  float carry = 1.f, my_carry = 1.f;
  for (int r = 0; rounds > 1 && r < rounds; ++r) {
    const int j = r * 64 + lane;
    const bool act = j < cnt;
    const int64_t k = s0 + j;
    const float alpha = act ? 1.f - expf(-sigma[k] * s.dist[k]) : 0.f;
    const float f = act ? (1.f - alpha) + 1e-10f : 1.f;
    const float incl = wave_prod_scan(f, lane);
    if (lane == r) my_carry = carry;
    carry = carry * __shfl(incl, 63, 64);
  }
  // reverse pass
  float carryS = 0.f;
  for (int r = rounds - 1; r >= 0; --r) {
    const int j = r * 64 + lane;
    const bool act = j < cnt;
    const int64_t k = s0 + j;
    const float sg = act ? sigma[k] : 0.f;
    const float dd = act ? s.dist[k] : 0.f;
    const float ex = expf(-sg * dd);
    const float alpha = 1.f - ex;
    const float f = act ? (1.f - alpha) + 1e-10f : 1.f;
    const float incl = wave_prod_scan(f, lane);
    float excl = shfl_up_f(incl, 1);
    if (lane == 0) excl = 1.f;
    const float T = __shfl(my_carry, r, 64) * excl;
    const float w = act ? alpha * T : 0.f;
    const bool app = act && w > thres;
    float c[3] = {0.f, 0.f, 0.f};
    if (act) {
      c[0] = rgb[k * 3 + 0];
      c[1] = rgb[k * 3 + 1];
      c[2] = rgb[k * 3 + 2];
    }
    const float e = act ? ((app ? (g[0] * c[0] + g[1] * c[1]) + g[2] * c[2] : 0.f) - gbg) : 0.f;
    // suffix scan of x -> a + f x, with a = e * alpha
    float A = f, Bv = act ? e * alpha : 0.f;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const float A2 = shfl_down_f(A, d);
      const float B2 = shfl_down_f(Bv, d);
      if (lane + d < 64) {
        Bv = Bv + A * B2;
        A = A * A2;
      }
    }
    const float S = Bv + A * carryS;
    float R = shfl_down_f(S, 1);
    if (lane == 63) R = carryS;
    if (act) {
      const float dalpha = T * (e - R);
      const float dsig = dalpha * (dd * ex);
      float4 out;
      out.x = dsig;
      out.y = app ? g[0] * w : 0.f;
      out.z = app ? g[1] * w : 0.f;
      out.w = app ? g[2] * w : 0.f;
      reinterpret_cast<float4*>(d_sample)[k] = out;
    }
    carryS = __shfl(S, 0, 64);
  }
}

// alpha/times clamped to [0,1], 3x3x3 max (out-of-range neighbours ignored), >= thres.
__global__ __launch_bounds__(256) void occupancy_pool_kernel(const float* __restrict__ alpha, int D, int H,
                                                             int W, float times, float thres,
                                                             uint8_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = (int64_t)D * H * W;
  if (i >= n) return;
  const int x = (int)(i % W), y = (int)((i / W) % H), z = (int)(i / ((int64_t)W * H));
  float m = -INFINITY;
  for (int dz = -1; dz <= 1; ++dz)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        const int zz = z + dz, yy = y + dy, xx = x + dx;
        if (zz < 0 || zz >= D || yy < 0 || yy >= H || xx < 0 || xx >= W) continue;
        float v = alpha[((int64_t)zz * H + yy) * W + xx] / times;
        v = fminf(fmaxf(v, 0.f), 1.f);
        m = fmaxf(m, v);
      }
  out[i] = m >= thres ? 1 : 0;
}

// Zero a small int32 range with a kernel (not a memset node, so that a caller
// may capture the entry point in a HIP graph).
__global__ void zero_i32_kernel(int32_t* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

int launch_zero_i32(int32_t* p, int n, hipStream_t s) {
  if (n <= 0) return FF_OK;
  hipLaunchKernelGGL(zero_i32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n);
  return hipGetLastError() == hipSuccess ? FF_OK : FF_ERR_LAUNCH;
}

int launch_scan_offsets(int32_t* off, int64_t B, hipStream_t s) {
  hipLaunchKernelGGL(scan_offsets_kernel, dim3(1), dim3(1024), 0, s, off, B);
  return hipGetLastError() == hipSuccess ? FF_OK : FF_ERR_LAUNCH;
}

}  // namespace ff

using namespace ff;

extern "C" {

int ff_sample_rays(const ff_ray_sampler* sampler, const float* rays, const float* jitter,
                   int64_t n_rays, const ff_samples* out, void* stream) {
  if (!sampler || !out || n_rays < 0 || (n_rays > 0 && !rays)) return FF_ERR_ARG;
  if (!out->ray_off) return FF_ERR_ARG;
  const ff_ray_sampler& s = *sampler;
  if (s.n_samples < 1 || s.n_samples > 64 * kMaxRounds) return FF_ERR_ARG;
  if (!out->xyz || !out->z || !out->dist || !out->ray || !out->idx) return FF_ERR_ARG;
  if (out->cap < n_rays * (int64_t)s.n_samples) return FF_ERR_ARG;
  if (n_rays * (int64_t)s.n_samples > 0x7fffffffLL) return FF_ERR_ARG;
  if (s.occ)
    for (int a = 0; a < 3; ++a)
      if (s.occ_reso[a] < 2) return FF_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  if (n_rays == 0) return launch_zero_i32(out->ray_off, 1, st);
  const unsigned blocks = (unsigned)((n_rays * 64 + 255) / 256);
  hipLaunchKernelGGL(sample_rays_kernel, dim3(blocks), dim3(256), 0, st, s, rays, jitter, n_rays, 0,
                     *out);
  hipLaunchKernelGGL(scan_offsets_kernel, dim3(1), dim3(1024), 0, st, out->ray_off, n_rays);
  hipLaunchKernelGGL(sample_rays_kernel, dim3(blocks), dim3(256), 0, st, s, rays, jitter, n_rays, 1,
                     *out);
  return hipGetLastError() == hipSuccess ? FF_OK : FF_ERR_LAUNCH;
}

int ff_composite_fwd(const ff_samples* s, int64_t n_rays, const float* sigma, const float* rgb,
                     int white_bg, float weight_thres, float* rgb_map, float* rgb_raw, float* depth,
                     float* acc, float* weights, void* stream) {
  if (!s || !s->ray_off || !s->dist || !s->z || n_rays < 0) return FF_ERR_ARG;
  if (n_rays > 0 && (!sigma || !rgb || !rgb_map)) return FF_ERR_ARG;
  if (n_rays == 0) return FF_OK;
  const unsigned blocks = (unsigned)((n_rays * 64 + 255) / 256);
  hipLaunchKernelGGL(composite_fwd_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, *s,
                     n_rays, sigma, rgb, white_bg, weight_thres, rgb_map, rgb_raw, depth, acc, weights);
  return hipGetLastError() == hipSuccess ? FF_OK : FF_ERR_LAUNCH;
}

int ff_composite_bwd(const ff_samples* s, int64_t n_rays, const float* sigma, const float* rgb,
                     int white_bg, float weight_thres, const float* rgb_raw, const float* d_rgb_map,
                     const float* target, float loss_scale, float* loss_sum, float* d_sample,
                     void* stream) {
  if (!s || !s->ray_off || !s->dist || n_rays < 0) return FF_ERR_ARG;
  if (n_rays > 0 && (!sigma || !rgb || !rgb_raw || !d_sample || (!d_rgb_map && !target)))
    return FF_ERR_ARG;
  if (n_rays == 0) return FF_OK;
  const unsigned blocks = (unsigned)((n_rays * 64 + 255) / 256);
  hipLaunchKernelGGL(composite_bwd_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, *s, n_rays,
                     sigma, rgb, white_bg, weight_thres, rgb_raw, d_rgb_map, target, loss_scale,
                     loss_sum, d_sample);
  return hipGetLastError() == hipSuccess ? FF_OK : FF_ERR_LAUNCH;
}

int ff_select_app(const ff_samples* s, int64_t n_rays, const float* sigma, float weight_thres, int32_t* app_off,
                  int32_t* rows, int32_t* app_pos, void* stream) {
  if (!s || !s->ray_off || !s->dist || n_rays < 0 || !app_off) return FF_ERR_ARG;
  if (n_rays > 0 && (!sigma || !rows)) return FF_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  if (n_rays == 0) return launch_zero_i32(app_off, 1, st);
  const unsigned blocks = (unsigned)((n_rays * 64 + 255) / 256);
  hipLaunchKernelGGL(select_app_kernel, dim3(blocks), dim3(256), 0, st, *s, n_rays, sigma, weight_thres, 0, app_off,
                     rows, app_pos);
  hipLaunchKernelGGL(scan_offsets_kernel, dim3(1), dim3(1024), 0, st, app_off, n_rays);
  hipLaunchKernelGGL(select_app_kernel, dim3(blocks), dim3(256), 0, st, *s, n_rays, sigma, weight_thres, 1, app_off,
                     rows, app_pos);
  return hipGetLastError() == hipSuccess ? FF_OK : FF_ERR_LAUNCH;
}

int ff_occupancy_pool(const float* alpha, int32_t D, int32_t H, int32_t W, float times, float thres,
                      uint8_t* out, void* stream) {
  if (!alpha || !out || D < 1 || H < 1 || W < 1 || !(times > 0.f)) return FF_ERR_ARG;
  const int64_t n = (int64_t)D * H * W;
  hipLaunchKernelGGL(occupancy_pool_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, alpha, D, H, W, times, thres, out);
  return hipGetLastError() == hipSuccess ? FF_OK : FF_ERR_LAUNCH;
}

}  // extern "C"
