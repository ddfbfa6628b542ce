// Learned proposal sampler (gfx950, wave64): one wavefront per ray, one lane
// per proposal sample (64 per ray).
//
// Forward: per lane, trilinear sample of the 8-channel proposal grid, MLP
// 8 -> 32 -> 1 (weights through the scalar cache), softplus density; the
// proposal weights come from a multiplicative prefix scan across the lanes
// (as in compositing, models/FactorFields.py:82-88); sample_pdf
// (dataLoader/ray_utils.py:132-174) is an additive prefix scan for the CDF
// and, for each fine sample, a binary search over the lane-held CDF with
// lane shuffles.  Fine samples are written ray-major like the occupancy
// sampler's output, so the field and compositing kernels consume either.
//
// Backward: recomputes the proposal forward, evaluates the interlevel bound
// loss against the field's fine weights, and back-propagates through the
// proposal compositing (reverse affine scan), the MLP and the grid.
#include "ff_common.h"
#include "ff_field_impl.h"

namespace ff {

int launch_scan_offsets(int32_t* off, int64_t B, hipStream_t s);  // ff_render.hip
int launch_zero_i32(int32_t* p, int n, hipStream_t s);             // ff_render.hip

typedef const __attribute__((address_space(4))) float* pcfp;
constexpr int kPH = 32;  // proposal MLP hidden width
constexpr int kPC = 8;   // proposal grid channels

struct PropRay {
  float o[3], d[3];
  float tn, tf;
  bool hit;
};

__device__ __forceinline__ void prop_ray(const ff_proposal_desc& p, const float* __restrict__ rays, int64_t b,
                                         PropRay& r) {
  float tn = -INFINITY, tf = INFINITY;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    r.o[a] = rays[b * 6 + a];
    r.d[a] = rays[b * 6 + 3 + a];
    const float vec = r.d[a] == 0.f ? 1e-6f : r.d[a];
    const float ta = (p.aabb_max[a] - r.o[a]) / vec;
    const float tb = (p.aabb_min[a] - r.o[a]) / vec;
    tn = fmaxf(tn, fminf(ta, tb));
    tf = fminf(tf, fmaxf(ta, tb));
  }
  r.tn = fmaxf(tn, p.near);
  r.tf = tf;
  r.hit = tf > r.tn;
}

__device__ __forceinline__ float prop_softplus(float x) { return x > 20.f ? x : log1pf(expf(x)); }
__device__ __forceinline__ float prop_softplus_grad(float x) {
  if (x > 20.f) return 1.f;
  const float z = expf(x);
  return z / (z + 1.f);
}

// Proposal density at x; keeps the features, hidden activations and corners for
// the backward pass.
template <bool BF>
__device__ __forceinline__ float prop_density(const ff_proposal_desc& p, const float* x, float* f, float* h,
                                              float& pre, Corners<3>& cr) {
  float pos[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float inv = 2.f / (p.aabb_max[a] - p.aabb_min[a]);
    pos[a] = unnorm_border((x[a] - p.aabb_min[a]) * inv - 1.f, p.reso[a]);
  }
  linear_corners<3>(pos, p.reso, cr);
#pragma unroll
  for (int c = 0; c < kPC; ++c) f[c] = 0.f;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    float v[kPC];
    load_ch<kPC, BF>(p.grid, (int64_t)cr.tex[k] * kPC, v);
#pragma unroll
    for (int c = 0; c < kPC; ++c) f[c] = fmaf(v[c], cr.w[k], f[c]);
  }
  const pcfp w0 = (pcfp)p.w0;
  const pcfp b0 = (pcfp)p.b0;
  const pcfp w1 = (pcfp)p.w1;
  float o = 0.f;
#pragma unroll
  for (int i = 0; i < kPH; ++i) {
    float a = b0[i];
#pragma unroll
    for (int c = 0; c < kPC; ++c) a = fmaf(w0[i * kPC + c], f[c], a);
    h[i] = fmaxf(a, 0.f);
    o = fmaf(w1[i], h[i], o);
  }
  pre = o + p.density_shift;
  return prop_softplus(pre);
}

__device__ __forceinline__ float wave_incl_scan_mul(float v, int lane) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const float t = __shfl_up(v, d, 64);
    if (lane >= d) v = v * t;
  }
  return v;
}

__device__ __forceinline__ float wave_incl_scan_add(float v, int lane) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const float t = __shfl_up(v, d, 64);
    if (lane >= d) v = v + t;
  }
  return v;
}

// Everything the forward computes for one ray (lane j = proposal sample j).
struct PropState {
  float s, s_next;     // bin edges of lane j
  float t, x[3];       // proposal sample
  float sigma, pre, alpha, f, T, w, delta, ex;
  float feat[kPC], h[kPH];
  Corners<3> cr;
};

template <bool BF>
__device__ __forceinline__ void prop_forward(const ff_proposal_desc& p, const PropRay& r,
                                             const float* __restrict__ prop_u, int64_t b, int lane,
                                             PropState& st) {
  const float span = r.tf - r.tn;
  st.s = r.tn + span * ((float)lane / 64.f);
  st.s_next = lane == 63 ? r.tf : r.tn + span * ((float)(lane + 1) / 64.f);
  const float u = prop_u ? prop_u[b * 64 + lane] : 0.5f;
  st.t = st.s + u * (st.s_next - st.s);
#pragma unroll
  for (int a = 0; a < 3; ++a) st.x[a] = r.o[a] + r.d[a] * st.t;
  st.sigma = prop_density<BF>(p, st.x, st.feat, st.h, st.pre, st.cr);
  st.delta = (st.s_next - st.s) * p.distance_scale;
  st.ex = expf(-st.sigma * st.delta);
  st.alpha = 1.f - st.ex;
  st.f = (1.f - st.alpha) + 1e-10f;
  const float incl = wave_incl_scan_mul(st.f, lane);
  float excl = __shfl_up(incl, 1, 64);
  if (lane == 0) excl = 1.f;
  st.T = excl;
  st.w = st.alpha * st.T;
}

// Bin edge i (0..64) of the ray: lanes hold edges 0..63, edge 64 is t_far.
__device__ __forceinline__ float edge_at(float s_lane, float tf, int i) {
  const float v = __shfl(s_lane, i < 64 ? i : 63, 64);
  return i < 64 ? v : tf;
}

template <bool BF>
__global__ __launch_bounds__(256) void proposal_sample_kernel(ff_proposal_desc p, const float* __restrict__ rays,
                                                              const float* __restrict__ prop_u,
                                                              const float* __restrict__ fine_u, int64_t B,
                                                              int pass, ff_samples out) {
  const int lane = threadIdx.x & 63;
  const int64_t b = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (b >= B) return;
  PropRay r;
  prop_ray(p, rays, b, r);
  const int nf = p.n_fine;
  if (pass == 0) {
    if (lane == 0) out.ray_off[b + 1] = r.hit ? nf : 0;
    return;
  }
  if (!r.hit) return;
  PropState st;
  prop_forward<BF>(p, r, prop_u, b, lane, st);
  // sample_pdf: weights + 1e-5, normalised, inclusive CDF (lane j holds cdf[j+1])
  const float wp = st.w + 1e-5f;
  const float total = wave_sum(wp);
  const float cdf = wave_incl_scan_add(wp / total, lane);
  float u;
  if (fine_u) u = ((float)lane + fine_u[b * nf + (lane < nf ? lane : 0)]) / (float)nf;
  else u = nf > 1 ? (float)lane / (float)(nf - 1) : 0.f;
  // inds = searchsorted(cdf, u, right=True) = #{i in 0..64 : cdf[i] <= u}, cdf[0] = 0
  int lo = 0, hi = 65;
#pragma unroll
  for (int it = 0; it < 7; ++it) {
    const int mid = (lo + hi) >> 1;
    const float cm = __shfl(cdf, mid > 0 ? mid - 1 : 0, 64);
    const float v = mid == 0 ? 0.f : cm;
    const bool le = (lo < hi) && v <= u;
    lo = le ? mid + 1 : lo;
    hi = (lo < hi && !le) ? mid : hi;
  }
  const int inds = lo;
  const int below = inds - 1 > 0 ? inds - 1 : 0;
  const int above = inds < 64 ? inds : 64;
  const float c0v = __shfl(cdf, below > 0 ? below - 1 : 0, 64);
  const float c1v = __shfl(cdf, above > 0 ? above - 1 : 0, 64);
  const float cdf0 = below == 0 ? 0.f : c0v;
  const float cdf1 = above == 0 ? 0.f : c1v;
  const float bin0 = edge_at(st.s, r.tf, below);
  const float bin1 = edge_at(st.s, r.tf, above);
  float denom = cdf1 - cdf0;
  denom = denom < 1e-5f ? 1.f : denom;
  const float tt = (u - cdf0) / denom;
  const float tk = bin0 + tt * (bin1 - bin0);
  const float tnext = __shfl_down(tk, 1, 64);
  if (lane < nf) {
    const int64_t pos = (int64_t)out.ray_off[b] + lane;
#pragma unroll
    for (int a = 0; a < 3; ++a) out.xyz[pos * 3 + a] = r.o[a] + r.d[a] * tk;
    out.z[pos] = tk;
    out.dist[pos] = (lane + 1 < nf ? tnext - tk : 0.f) * p.distance_scale;
    out.ray[pos] = (int32_t)b;
    out.idx[pos] = lane;
  }
}

constexpr int kPropParams = kPH * kPC + kPH + kPH;  // dW0, db0, dW1 (partial row layout)
// Per-wave LDS: MLP transpose H [64][33] | F [64][9] | dout [64], reused as
// the scatter staging area afterwards.
constexpr int kPropHS = 33, kPropFS = 9;
constexpr int kPropLds = 64 * kPropHS + 64 * kPropFS + 64;
constexpr int kPropLdsWords = kPropLds > kStageWords ? kPropLds : kStageWords;

// One 64-thread block = one wavefront; each wave walks rays b = wave,
// wave + nwaves, ...  MLP weight gradients: lane L owns hidden unit i = L/2
// and half h = L%2 of the input channels (dW0[i][4h..4h+3]) plus db0[i]
// (h = 0) or dW1[i] (h = 1), accumulated over the wave's samples from an LDS
// transpose; each wave writes one partial row, reduced afterwards in a fixed
// order.  Grid gradients go through the run-merging coalesced scatter.
template <bool BF>
__global__ __launch_bounds__(64) void proposal_bwd_kernel(ff_proposal_desc p, const float* __restrict__ rays,
                                                          const float* __restrict__ prop_u, int64_t B,
                                                          ff_samples fine, const float* __restrict__ weights,
                                                          float loss_scale, ff_proposal_grads g,
                                                          float* __restrict__ loss_sum, float* __restrict__ ws) {
  __shared__ float lds[kPropLdsWords];
  float* Hs = lds;
  float* Fs = lds + 64 * kPropHS;
  float* Ds = Fs + 64 * kPropFS;
  const int lane = threadIdx.x;
  const int64_t wave = blockIdx.x;
  const int64_t nwaves = gridDim.x;
  float* grid_g = g.grid;
  if (g.replicas > 1) grid_g += (int64_t)(xcc_id() % g.replicas) * g.replica_stride;
  const int hi = lane >> 1, half = lane & 1;
  const float w1i = p.w1[hi];
  float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f, a4 = 0.f;
  float lsum = 0.f;
  const float scale = loss_scale / (float)B;
  for (int64_t b = wave; b < B; b += nwaves) {
    const int cnt = fine.ray_off[b + 1] - fine.ray_off[b];
    if (cnt < 2) continue;
    PropRay r;
    prop_ray(p, rays, b, r);
    PropState st;
    prop_forward<BF>(p, r, prop_u, b, lane, st);
    const int64_t f0 = fine.ray_off[b];
    const float tf_k = lane < cnt ? fine.z[f0 + lane] : 0.f;
    const float wf_k = lane < cnt ? weights[f0 + lane] : 0.f;
    const float wincl = wave_incl_scan_add(st.w, lane);
    const float wexcl = wincl - st.w;
    float dw = 0.f;
    for (int k = 0; k + 1 < cnt; ++k) {
      const float t0 = __shfl(tf_k, k, 64);
      const float t1 = __shfl(tf_k, k + 1, 64);
      const float wf = __shfl(wf_k, k, 64);
      const uint64_t m_lo = __ballot(lane >= 1 && st.s <= t0);
      int j_lo = __popcll(m_lo) + (r.tf <= t0 ? 1 : 0);
      int j_hi = __popcll(__ballot(st.s < t1)) - 1;
      j_lo = j_lo > 63 ? 63 : j_lo;
      j_hi = j_hi < 0 ? 0 : (j_hi > 63 ? 63 : j_hi);
      if (j_hi < j_lo) j_hi = j_lo;
      const float bound = __shfl(wincl, j_hi, 64) - __shfl(wexcl, j_lo, 64);
      const float gap = fmaxf(wf - bound, 0.f);
      const float inv = 1.f / (wf + 1e-7f);
      lsum += (lane == 0) ? gap * gap * inv : 0.f;
      const float gk = -2.f * gap * inv * scale;
      if (lane >= j_lo && lane <= j_hi) dw += gk;
    }
    // proposal compositing backward: reverse affine scan
    float A = st.f, Bv = dw * st.alpha;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const float A2 = __shfl_down(A, d, 64);
      const float B2 = __shfl_down(Bv, d, 64);
      if (lane + d < 64) {
        Bv = Bv + A * B2;
        A = A * A2;
      }
    }
    float R = __shfl_down(Bv, 1, 64);
    if (lane == 63) R = 0.f;
    const float dalpha = st.T * (dw - R);
    const float dsig = dalpha * (st.delta * st.ex);
    const float dout = dsig * prop_softplus_grad(st.pre);
    // MLP backward: d feat per lane (sample); weight gradients via the LDS transpose
    const pcfp w0 = (pcfp)p.w0;
    const pcfp w1 = (pcfp)p.w1;
    float df[kPC];
#pragma unroll
    for (int c = 0; c < kPC; ++c) df[c] = 0.f;
#pragma unroll
    for (int i = 0; i < kPH; ++i) {
      const float dh = st.h[i] > 0.f ? w1[i] * dout : 0.f;
#pragma unroll
      for (int c = 0; c < kPC; ++c) df[c] = fmaf(w0[i * kPC + c], dh, df[c]);
      Hs[lane * kPropHS + i] = st.h[i];
    }
#pragma unroll
    for (int c = 0; c < kPC; ++c) Fs[lane * kPropFS + c] = st.feat[c];
    Ds[lane] = dout;
    __syncthreads();
#pragma unroll 4
    for (int j = 0; j < 64; ++j) {
      const float hv = Hs[j * kPropHS + hi];
      const float d = Ds[j];
      const float dh = hv > 0.f ? w1i * d : 0.f;
      const float* f = Fs + j * kPropFS + half * 4;
      a0 = fmaf(dh, f[0], a0);
      a1 = fmaf(dh, f[1], a1);
      a2 = fmaf(dh, f[2], a2);
      a3 = fmaf(dh, f[3], a3);
      a4 += half ? d * hv : dh;
    }
    __syncthreads();
    // grid scatter (all 64 lanes hold real proposal samples of a hit ray)
    wave_scatter<8, kPC, true>(lds, grid_g, st.cr.tex, st.cr.w, df, true);
  }
  float* row = ws + wave * kPropParams;
  const int base = hi * kPC + half * 4;
  row[base + 0] = a0;
  row[base + 1] = a1;
  row[base + 2] = a2;
  row[base + 3] = a3;
  row[half ? kPH * kPC + kPH + hi : kPH * kPC + hi] = a4;
  lsum = wave_sum(lsum);
  if (lane == 0 && loss_sum) unsafeAtomicAdd(loss_sum, lsum * loss_scale / (float)B);
}

// Sample offsets of the proposal's rays in one launch: n_fine for a ray that
// crosses the box, 0 otherwise (the same test as the sampling pass), then
// the exclusive prefix sum over rays, in one block.  Replaces a counting
// pass of one wave per ray and a separate scan for batches up to
// kPropOffsetsMaxRays rays.
constexpr int64_t kPropOffsetsMaxRays = 1 << 16;

__global__ __launch_bounds__(1024) void proposal_offsets_kernel(ff_proposal_desc p, const float* __restrict__ rays,
                                                                int64_t B, int32_t* __restrict__ off) {
  static_assert(kPropOffsetsMaxRays <= 64 * 1024, "one 64-bit hit mask per thread");
  __shared__ int32_t wsum[1024 / kWave];
  const int t = threadIdx.x, lane = t & (kWave - 1), wv = t / kWave;
  const int64_t per = (B + 1023) / 1024;
  const int64_t lo = t * per;
  const int nf = p.n_fine;
  uint64_t hits = 0;
  int32_t s = 0;
  for (int64_t j = 0; j < per; ++j) {
    if (lo + j >= B) break;
    PropRay r;
    prop_ray(p, rays, lo + j, r);
    if (r.hit) {
      hits |= 1ull << j;
      s += nf;
    }
  }
  // block-wide inclusive scan: in each wave, then over the wave totals
  int32_t incl = s;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const int32_t v = __shfl_up(incl, o);
    if (lane >= o) incl += v;
  }
  if (lane == kWave - 1) wsum[wv] = incl;
  __syncthreads();
  if (wv == 0) {
    int32_t x = lane < 1024 / kWave ? wsum[lane] : 0;
#pragma unroll
    for (int o = 1; o < 1024 / kWave; o <<= 1) {
      const int32_t v = __shfl_up(x, o);
      if (lane >= o) x += v;
    }
    if (lane < 1024 / kWave) wsum[lane] = x;
  }
  __syncthreads();
  int32_t run = incl - s + (wv ? wsum[wv - 1] : 0);
  for (int64_t j = 0; j < per; ++j) {
    if (lo + j >= B) break;
    run += ((hits >> j) & 1) ? nf : 0;
    off[lo + j + 1] = run;
  }
  if (t == 0) off[0] = 0;
}

// sum((a - b)^2) over n floats added to out[0], one block, fixed order.
__global__ __launch_bounds__(256) void sq_err_sum_kernel(const float* __restrict__ a, const float* __restrict__ b,
                                                         int64_t n, float* __restrict__ out) {
  __shared__ float part[256 / kWave];
  float s = 0.f;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    const float d = a[i] - b[i];
    s = fmaf(d, d, s);
  }
  s = wave_sum(s);
  if ((threadIdx.x & (kWave - 1)) == 0) part[threadIdx.x / kWave] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    float tot = 0.f;
    for (int w = 0; w < 256 / kWave; ++w) tot += part[w];
    out[0] += tot;
  }
}

int64_t prop_bwd_waves(int64_t n_rays) {
  int cu = 256, dev = 0;
  if (hipGetDevice(&dev) == hipSuccess) hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev);
  const int64_t cap = (int64_t)cu * 16;
  return n_rays < cap ? n_rays : cap;
}

int prop_check(const ff_proposal_desc* p) {
  if (!p || !p->grid || !p->w0 || !p->b0 || !p->w1) return FF_ERR_ARG;
  for (int a = 0; a < 3; ++a)
    if (p->reso[a] < 1 || !(p->aabb_max[a] > p->aabb_min[a])) return FF_ERR_ARG;
  if (p->n_fine < 2 || p->n_fine > 64) return FF_ERR_ARG;
  if (p->table_dtype != FF_F32 && p->table_dtype != FF_BF16) return FF_ERR_ARG;
  return FF_OK;
}

}  // namespace ff

using namespace ff;

extern "C" {

int ff_proposal_sample(const ff_proposal_desc* p, const float* rays, const float* prop_u,
                       const float* fine_u, int64_t n_rays, const ff_samples* out, void* stream) {
  int st = prop_check(p);
  if (st) return st;
  if (!out || !out->ray_off || n_rays < 0 || (n_rays > 0 && !rays)) return FF_ERR_ARG;
  if (!out->xyz || !out->z || !out->dist || !out->ray || !out->idx) return FF_ERR_ARG;
  if (out->cap < n_rays * (int64_t)p->n_fine) return FF_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  if (n_rays == 0) return launch_zero_i32(out->ray_off, 1, s);
  const unsigned blocks = (unsigned)((n_rays * 64 + 255) / 256);
  const ff_proposal_desc pd = *p;
  auto go = [&](auto BFi) {
    constexpr bool BF = decltype(BFi)::value;
    if (n_rays <= kPropOffsetsMaxRays) {
      hipLaunchKernelGGL(proposal_offsets_kernel, dim3(1), dim3(1024), 0, s, pd, rays, n_rays, out->ray_off);
    } else {
      hipLaunchKernelGGL((proposal_sample_kernel<BF>), dim3(blocks), dim3(256), 0, s, pd, rays, prop_u, fine_u,
                         n_rays, 0, *out);
      int r = launch_scan_offsets(out->ray_off, n_rays, s);
      if (r) return r;
    }
    hipLaunchKernelGGL((proposal_sample_kernel<BF>), dim3(blocks), dim3(256), 0, s, pd, rays, prop_u, fine_u,
                       n_rays, 1, *out);
    return hipGetLastError() == hipSuccess ? (int)FF_OK : (int)FF_ERR_LAUNCH;
  };
  return p->table_dtype == FF_BF16 ? go(std::true_type{}) : go(std::false_type{});
}

size_t ff_proposal_bwd_workspace_bytes(int64_t n_rays) {
  if (n_rays <= 0) return 0;
  return (size_t)(prop_bwd_waves(n_rays) + kReduceSplit) * kPropParams * sizeof(float);
}

int ff_proposal_bwd(const ff_proposal_desc* p, const float* rays, const float* prop_u, int64_t n_rays,
                    const ff_samples* fine, const float* weights, float loss_scale,
                    const ff_proposal_grads* grads, float* loss_sum, void* ws, size_t ws_bytes, void* stream) {
  int st = prop_check(p);
  if (st) return st;
  if (!fine || !fine->ray_off || !fine->z || !weights || !grads || !grads->grid || !grads->w0 || !grads->b0 ||
      !grads->w1 || n_rays < 0 || (n_rays > 0 && !rays))
    return FF_ERR_ARG;
  if (n_rays == 0) return FF_OK;
  if (!ws || ws_bytes < ff_proposal_bwd_workspace_bytes(n_rays)) return FF_ERR_WORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  const int64_t waves = prop_bwd_waves(n_rays);
  const ff_proposal_desc pd = *p;
  const ff_proposal_grads gd = *grads;
  const ff_samples fs = *fine;
  float* wsf = (float*)ws;
  if (p->table_dtype == FF_BF16)
    hipLaunchKernelGGL((proposal_bwd_kernel<true>), dim3((unsigned)waves), dim3(64), 0, s, pd, rays, prop_u,
                       n_rays, fs, weights, loss_scale, gd, loss_sum, wsf);
  else
    hipLaunchKernelGGL((proposal_bwd_kernel<false>), dim3((unsigned)waves), dim3(64), 0, s, pd, rays, prop_u,
                       n_rays, fs, weights, loss_scale, gd, loss_sum, wsf);
  if (hipGetLastError() != hipSuccess) return FF_ERR_LAUNCH;
  ff_grads mg{};
  mg.w0 = grads->w0;
  mg.b0 = grads->b0;
  mg.w1 = grads->w1;
  return launch_reduce_partials(wsf, (int)waves, 2, kPC, kPH, 1, mg, s);
}

int ff_sq_err_sum(const float* a, const float* b, int64_t n, float* out, void* stream) {
  if (n < 0 || !out) return FF_ERR_ARG;
  if (n == 0) return FF_OK;
  if (!a || !b) return FF_ERR_ARG;
  hipLaunchKernelGGL(sq_err_sum_kernel, dim3(1), dim3(256), 0, (hipStream_t)stream, a, b, n, out);
  return hipGetLastError() == hipSuccess ? FF_OK : FF_ERR_LAUNCH;
}

}  // extern "C"
