// NeRF field over a packed sample list: get_coding + linear_mat + density
// activation + colour head in one kernel, and its fused backward.
//
// Reference: FactorFields.forward models/FactorFields.py:874-885 (get_coding,
// linear_mat, basis2density :639-643).  Two heads (template HEAD):
//   kHeadSH: sigmoid(eval_sh(2, feat[1:28].view(3, 9), viewdir))
//            (models/sh.py:34-64), decoder 18 -> 64 -> 28;
//   kHeadFeat: the features feat[1:32] of the reference's own colour module
//            (MLPRender_Fea, evaluated afterwards on the app samples by
//            ff_color_mlp_fwd), decoder 18 -> 64 -> 32; the backward takes
//            their gradients from ff_color_mlp_bwd through app_pos.
// The sample count is read on the device (ff_samples.ray_off[n_rays]); the
// grid is persistent and every wave strides over 64-sample tiles.  With BF16
// tables the decoder runs on the matrix cores (BF16 operands, FP32
// accumulation, ff_mfma.h); the FP32 form keeps the FP32 vector decoder.
#include <mutex>
#include <type_traits>

#include "ff_field_impl.h"
#include "ff_mfma_f32.h"

namespace ff {

constexpr float kC0 = 0.28209479177387814f;
constexpr float kC1 = 0.4886025119029199f;
constexpr float kC2_0 = 1.0925484305920792f;
constexpr float kC2_1 = -1.0925484305920792f;
constexpr float kC2_2 = 0.31539156525252005f;
constexpr float kC2_3 = -1.0925484305920792f;
constexpr float kC2_4 = 0.5462742152960396f;

// eval_sh(2, sh, dir) with the reference's grouping; sh = o[1 + 9c .. 9c + 9].
__device__ __forceinline__ float sh2_eval(const float* sh, float x, float y, float z) {
  float r = kC0 * sh[0];
  r = r - kC1 * y * sh[1];
  r = r + kC1 * z * sh[2];
  r = r - kC1 * x * sh[3];
  const float xx = x * x, yy = y * y, zz = z * z;
  const float xy = x * y, yz = y * z, xz = x * z;
  r = r + kC2_0 * xy * sh[4];
  r = r + kC2_1 * yz * sh[5];
  r = r + kC2_2 * (2.f * zz - xx - yy) * sh[6];
  r = r + kC2_3 * xz * sh[7];
  r = r + kC2_4 * (xx - yy) * sh[8];
  return r;
}

// d(eval_sh)/d(sh_k) = Y_k(dir).
__device__ __forceinline__ void sh2_basis(float x, float y, float z, float* Y) {
  const float xx = x * x, yy = y * y, zz = z * z;
  Y[0] = kC0;
  Y[1] = -(kC1 * y);
  Y[2] = kC1 * z;
  Y[3] = -(kC1 * x);
  Y[4] = kC2_0 * (x * y);
  Y[5] = kC2_1 * (y * z);
  Y[6] = kC2_2 * (2.f * zz - xx - yy);
  Y[7] = kC2_3 * (x * z);
  Y[8] = kC2_4 * (xx - yy);
}

__device__ __forceinline__ float sigmoidf(float x) { return 1.f / (1.f + expf(-x)); }

// torch.nn.functional.softplus (beta 1, threshold 20) and its derivative.
__device__ __forceinline__ float softplusf(float x) { return x > 20.f ? x : log1pf(expf(x)); }
__device__ __forceinline__ float softplus_grad(float x) {
  if (x > 20.f) return 1.f;
  const float z = expf(x);
  return z / (z + 1.f);
}

constexpr int kHeadSH = 0, kHeadFeat = 1;
template <int HEAD> constexpr int head_out() { return HEAD == kHeadSH ? 28 : 32; }

// d out for one sample: softplus/ReLU density and sigmoid(SH) colour heads;
// with the feature head, the colour part is row app_pos[i] of d_feat (zero
// for samples the colour module did not see).
template <int HEAD>
__device__ __forceinline__ void nerf_dout(const ff_render_head& hd, const float* o, float4 gs, float dx, float dy,
                                          float dz, int64_t i, float* dout) {
  const float x0 = o[0] + hd.density_shift;
  dout[0] = gs.x * (hd.density_act == 0 ? softplus_grad(x0) : (x0 > 0.f ? 1.f : 0.f));
  if constexpr (HEAD == kHeadFeat) {
    const int32_t pos = hd.app_pos[i];
    const float* g = hd.d_feat + (int64_t)(pos >= 0 ? pos : 0) * hd.feat_stride;
#pragma unroll
    for (int c = 0; c < 31; ++c) dout[1 + c] = pos >= 0 ? g[c] : 0.f;
    return;
  }
  float Y[9];
  sh2_basis(dx, dy, dz, Y);
  const float gc[3] = {gs.y, gs.z, gs.w};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const float y = sigmoidf(sh2_eval(o + 1 + 9 * c, dx, dy, dz));
    const float dpre = gc[c] * (1.f - y) * y;
#pragma unroll
    for (int k = 0; k < 9; ++k) dout[1 + 9 * c + k] = dpre * Y[k];
  }
}

// Per-sample coding rows ([2C, cap], channel-major: coalesced across lanes).
template <int C>
__device__ __forceinline__ void store_coding(const ff_samples& s, int64_t i, const float* bas, const float* cof) {
#pragma unroll
  for (int k = 0; k < C; ++k) {
    s.coding[k * s.cap + i] = bas[k];
    s.coding[(C + k) * s.cap + i] = cof[k];
  }
}

// The sample's basis and coefficient values: the rows the forward stored, or
// a fresh gather when the samples carry none.  Only the BF16-table kernels
// keep rows: there, reading them saves 14 % of the backward; the FP32 kernels
// gather again, which measured faster than reading the rows in them.
template <int CK, bool BF>
__device__ __forceinline__ void coding_of(const ff_field_desc& fd, const ff_samples& s, int64_t i, const float* xv,
                                          float* bas, float* cof) {
  constexpr int C = Spec18::C;
  if (s.coding) {
#pragma unroll
    for (int k = 0; k < C; ++k) {
      bas[k] = s.coding[k * s.cap + i];
      cof[k] = s.coding[(C + k) * s.cap + i];
    }
  } else {
    coding_fwd<3, CK, BF, Spec18>(fd, xv, bas, cof);
  }
}

// The head's outputs of sample i: SH colour into rgb, or the feature row.
template <int HEAD>
__device__ __forceinline__ void head_out_store(const ff_render_head& hd, const float* o, const float* __restrict__ rays,
                                               const ff_samples& s, int64_t i, float* __restrict__ rgb) {
  if constexpr (HEAD == kHeadFeat) {
    float* f = hd.feat + i * hd.feat_stride;
#pragma unroll
    for (int c = 0; c < 31; ++c) f[c] = o[1 + c];
  } else {
    const int64_t rb = s.ray[i];
    const float dx = rays[rb * 6 + 3], dy = rays[rb * 6 + 4], dz = rays[rb * 6 + 5];
#pragma unroll
    for (int c = 0; c < 3; ++c) rgb[i * 3 + c] = sigmoidf(sh2_eval(o + 1 + 9 * c, dx, dy, dz));
  }
}

template <int CK, bool BF, int HEAD>
__global__ __launch_bounds__(64) void nerf_field_bwd_kernel(ff_field_desc fd, ff_mlp_desc m,
                                                            ff_render_head hd, ff_grads g,
                                                            const float* __restrict__ rays,
                                                            ff_samples s, int64_t n_rays,
                                                            const float* __restrict__ d_sample,
                                                            float* __restrict__ ws) {
  constexpr int C = Spec18::C;
  constexpr int O = head_out<HEAD>();
  using Acc = MlpAcc<C, 64, O, 2>;
  static_assert(kWave * 65 >= kStageWords, "LDS tile must also hold the scatter staging area");
  __shared__ float lds[kWave * 65];
  Acc acc;
  acc.zero();
  const int lane = threadIdx.x;
  const int64_t n = s.ray_off[n_rays];
  const int64_t ntiles = (n + kWave - 1) / kWave;
  const ff_grads gl = local_grads(g);
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int64_t i = tile * kWave + lane;
    const bool active = i < n;
    const int64_t ic = active ? i : n - 1;
    float xv[3] = {s.xyz[ic * 3 + 0], s.xyz[ic * 3 + 1], s.xyz[ic * 3 + 2]};
    const int64_t rb = s.ray[ic];
    const float dx = rays[rb * 6 + 3], dy = rays[rb * 6 + 4], dz = rays[rb * 6 + 5];
    const float4 gs = reinterpret_cast<const float4*>(d_sample)[ic];
    float bas[C], cof[C], f[C], df[C];
    coding_fwd<3, CK, BF, Spec18>(fd, xv, bas, cof);
    combine<C, CK>(fd, bas, cof, f);
    mlp_train_tile_fn<C, 64, O, 2>(
        m, acc, lds, f, active,
        [&](const float* o, float* dout) { nerf_dout<HEAD>(hd, o, gs, dx, dy, dz, ic, dout); },
        df);
    float db[C], dc[C];
    combine_bwd<C, CK>(fd, bas, cof, df, db, dc);
    coding_bwd_wave<3, CK, BF, Spec18>(fd, gl, xv, db, dc, active, lds);
  }
  mlp_acc_store<C, 64, O, 2>(acc, ws + (int64_t)blockIdx.x * Acc::kPartial);
}

// FP32 tables with stored coding rows, backward in two kernels.  The fused
// kernel above holds the gathers, the FP32 decoder and the scatter at once:
// 256 VGPRs + 170 AGPRs, one wave per SIMD, and each phase waits on the one
// before.  Split:
//   nerf_decoder_bwd_kernel: the rows the forward stored -> decoder forward
//     and backward on the FP32 matrix cores -> the rows' gradients (d basis,
//     d coefficient), written over the rows; decoder weight-gradient partials
//     in the fused kernel's row layout;
//   nerf_scatter_kernel: corners and weights from the positions, the row
//     gradients scattered with the run-merging loop; no decoder registers, so
//     several waves per SIMD keep the atomic unit busy.
// (The BF16 form measured no gain from the same split: 0.339 against
// 0.325 ms per 196 K-sample batch, so it keeps the fused kernel.)
template <int HEAD>
__global__ __launch_bounds__(64) void nerf_decoder_bwd_kernel(ff_field_desc fd, ff_mlp_desc m, ff_render_head hd,
                                                              const float* __restrict__ rays, ff_samples s,
                                                              int64_t n_rays, const float* __restrict__ d_sample,
                                                              float* __restrict__ ws) {
  constexpr int C = Spec18::C;
  constexpr int O = head_out<HEAD>();
  __shared__ __attribute__((aligned(16))) float lds[kF32LdsWords];
  MfmaDecoderF32<O> W;
  W.load(m);
  MfmaDecoderF32Bwd<O> WB;
  WB.load(m);
  MfmaDecoderAcc<O> acc;
  acc.zero();
  const int lane = threadIdx.x;
  const int64_t n = s.ray_off[n_rays];
  const int64_t ntiles = (n + kWave - 1) / kWave;
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int64_t i = tile * kWave + lane;
    const bool active = i < n;
    const int64_t ic = active ? i : n - 1;
    float bas[C], cof[C], f[C], df[C];
#pragma unroll
    for (int k = 0; k < C; ++k) {
      bas[k] = s.coding[k * s.cap + ic];
      cof[k] = s.coding[(C + k) * s.cap + ic];
    }
    const int64_t rb = s.ray[ic];
    const float dx = rays[rb * 6 + 3], dy = rays[rb * 6 + 4], dz = rays[rb * 6 + 5];
    const float4 gs = reinterpret_cast<const float4*>(d_sample)[ic];
    combine<C, CK_GRID3>(fd, bas, cof, f);
    mfma_f32_decoder_train(
        W, WB, f,
        [&](const float* o, float* dout) {
          nerf_dout<HEAD>(hd, o, gs, dx, dy, dz, ic, dout);
#pragma unroll
          for (int k = 0; k < 32; ++k) dout[k] = (active && k < O) ? dout[k] : 0.f;
        },
        df, acc, lds);
    float db[C], dc[C];
    combine_bwd<C, CK_GRID3>(fd, bas, cof, df, db, dc);
    if (active) {
#pragma unroll
      for (int k = 0; k < C; ++k) {
        s.coding[k * s.cap + i] = db[k];
        s.coding[(C + k) * s.cap + i] = dc[k];
      }
    }
  }
  acc.store(ws + (int64_t)blockIdx.x * (kMfH * kMfIn + kMfH + O * kMfH));
}

__global__ __launch_bounds__(64) void nerf_scatter_kernel(ff_field_desc fd, ff_grads g, ff_samples s,
                                                          int64_t n_rays) {
  constexpr int C = Spec18::C;
  __shared__ float lds[kStageWords];
  const int lane = threadIdx.x;
  const int64_t n = s.ray_off[n_rays];
  const int64_t ntiles = (n + kWave - 1) / kWave;
  const ff_grads gl = local_grads(g);
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int64_t i = tile * kWave + lane;
    const bool active = i < n;
    const int64_t ic = active ? i : n - 1;
    float xv[3] = {s.xyz[ic * 3 + 0], s.xyz[ic * 3 + 1], s.xyz[ic * 3 + 2]};
    float db[C], dc[C];
#pragma unroll
    for (int k = 0; k < C; ++k) {
      db[k] = s.coding[k * s.cap + ic];
      dc[k] = s.coding[(C + k) * s.cap + ic];
    }
    coding_bwd_wave<3, CK_GRID3, false, Spec18, true>(fd, gl, xv, db, dc, active, lds);
  }
}

// Forward with the matrix-core decoder (BF16 tables: BF16 operands; FP32
// tables: FP32 operands): one wave per 64-sample tile.  BF16: bounded to
// three waves per SIMD (168 VGPRs, a few spills): the gathers are
// latency-bound, and the third wave hides more of them than the spills cost
// (measured 70 -> 54 us per 196 K samples; four waves spill too much).  FP32:
// two waves per SIMD (207-231 VGPRs, no spills; three spilled 130-360 B).
template <int CK, bool BF, int HEAD>
__global__ __launch_bounds__(256, BF ? 3 : 2) void nerf_field_fwd_mfma_kernel(ff_field_desc fd, ff_mlp_desc m,
                                                                  ff_render_head hd,
                                                                  const float* __restrict__ rays,
                                                                  ff_samples s, int64_t n_rays,
                                                                  float* __restrict__ sigma,
                                                                  float* __restrict__ rgb) {
  constexpr int C = Spec18::C;
  constexpr int O = head_out<HEAD>();
  const int lane = threadIdx.x & 63;
  const int64_t n = s.ray_off[n_rays];
  const int64_t ntiles = (n + kWave - 1) / kWave;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  if (wave >= ntiles) return;  // the grid is sized for the capacity, not the live count
  std::conditional_t<BF, MfmaDecoder<O>, MfmaDecoderF32<O>> W;
  W.load(m);
  for (int64_t tile = wave; tile < ntiles; tile += nwaves) {
    const int64_t i = tile * kWave + lane;
    const bool active = i < n;
    const int64_t ic = active ? i : n - 1;
    float xv[3] = {s.xyz[ic * 3 + 0], s.xyz[ic * 3 + 1], s.xyz[ic * 3 + 2]};
    float bas[C], cof[C], f[C], o[32];
    coding_fwd<3, CK, BF, Spec18>(fd, xv, bas, cof);
    if (s.coding && active) store_coding<C>(s, i, bas, cof);
    combine<C, CK>(fd, bas, cof, f);
    if constexpr (BF)
      mfma_decoder_fwd(W, f, o);
    else
      mfma_f32_decoder_fwd(W, f, o);
    if (active) {
      const float x0 = o[0] + hd.density_shift;
      sigma[i] = hd.density_act == 0 ? softplusf(x0) : fmaxf(x0, 0.f);
      head_out_store<HEAD>(hd, o, rays, s, i, rgb);
    }
  }
}

// BF16-table backward: gathers recomputed, matrix-core decoder forward and
// backward, run-merging table scatter; decoder weight gradients accumulate in
// the matrix-core accumulators across the wave's tiles.
template <int CK, int HEAD>
__global__ __launch_bounds__(64) void nerf_field_bwd_mfma_kernel(ff_field_desc fd, ff_mlp_desc m,
                                                                 ff_render_head hd, ff_grads g,
                                                                 const float* __restrict__ rays,
                                                                 ff_samples s, int64_t n_rays,
                                                                 const float* __restrict__ d_sample,
                                                                 float* __restrict__ ws) {
  constexpr int C = Spec18::C;
  // LDS: decoder operands (later the scatter staging area) | gathered values
  constexpr int kDec = kMfLdsWords > kStageWords ? kMfLdsWords : kStageWords;
  constexpr int kValRow = 2 * C + 1;
  __shared__ __attribute__((aligned(16))) uint32_t lds[kDec + kWave * kValRow];
  float* vals = reinterpret_cast<float*>(lds + kDec) + threadIdx.x * kValRow;
  constexpr int O = head_out<HEAD>();
  MfmaDecoder<O> W;
  W.load(m);
  MfmaDecoderBwd<O> WB;
  WB.load(m);
  MfmaDecoderAcc<O> acc;
  acc.zero();
  const int lane = threadIdx.x;
  const int64_t n = s.ray_off[n_rays];
  const int64_t ntiles = (n + kWave - 1) / kWave;
  const ff_grads gl = local_grads(g);
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int64_t i = tile * kWave + lane;
    const bool active = i < n;
    const int64_t ic = active ? i : n - 1;
    float xv[3] = {s.xyz[ic * 3 + 0], s.xyz[ic * 3 + 1], s.xyz[ic * 3 + 2]};
    float dout[32], df[C];
    uint32_t msk[2][2];
    {
      float bas[C], cof[C], f[C], o[32];
      coding_of<CK, true>(fd, s, ic, xv, bas, cof);
      combine<C, CK>(fd, bas, cof, f);
#pragma unroll
      for (int k = 0; k < C; ++k) {
        vals[k] = bas[k];
        vals[C + k] = cof[k];
      }
      mfma_decoder_fwd_train(W, f, o, msk, lds);
      const int64_t rb = s.ray[ic];
      const float dx = rays[rb * 6 + 3], dy = rays[rb * 6 + 4], dz = rays[rb * 6 + 5];
      const float4 gs = reinterpret_cast<const float4*>(d_sample)[ic];
      nerf_dout<HEAD>(hd, o, gs, dx, dy, dz, ic, dout);
#pragma unroll
      for (int k = 0; k < 32; ++k) dout[k] = (active && k < O) ? dout[k] : 0.f;
    }
    mfma_decoder_bwd(WB, msk, dout, df, acc, lds);
    float bas[C], cof[C], db[C], dc[C];
#pragma unroll
    for (int k = 0; k < C; ++k) {
      bas[k] = vals[k];
      cof[k] = vals[C + k];
    }
    combine_bwd<C, CK>(fd, bas, cof, df, db, dc);
    coding_bwd_wave<3, CK, true, Spec18, true>(fd, gl, xv, db, dc, active, reinterpret_cast<float*>(lds));
  }
  acc.store(ws + (int64_t)blockIdx.x * (kMfH * kMfIn + kMfH + O * kMfH));
}

// compute_alpha (models/FactorFields.py:710-727) accumulated into alpha.
template <int CK, bool BF, int HEAD>
__global__ __launch_bounds__(256) void nerf_alpha_kernel(ff_field_desc fd, ff_mlp_desc m, ff_render_head hd,
                                                         ff_ray_sampler occ, const float* __restrict__ xyz,
                                                         int64_t n, float length,
                                                         float* __restrict__ alpha) {
  constexpr int C = Spec18::C;
  constexpr int O = head_out<HEAD>();
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float xv[3] = {xyz[i * 3 + 0], xyz[i * 3 + 1], xyz[i * 3 + 2]};
  if (occ.occ) {
    float pos[3];
    const int R[3] = {occ.occ_reso[0], occ.occ_reso[1], occ.occ_reso[2]};
#pragma unroll
    for (int a = 0; a < 3; ++a) pos[a] = unnorm_ac((xv[a] - occ.occ_origin[a]) * occ.occ_inv[a] - 1.f, R[a]);
    Corners<3> cr;
    linear_corners<3>(pos, R, cr);
    float acc = 0.f;
#pragma unroll
    for (int k = 0; k < 8; ++k) acc = acc + (occ.occ[cr.tex[k]] ? cr.w[k] : 0.f);
    if (!(acc > 0.f)) return;   // sigma = 0 -> alpha += 0
  }
  float bas[C], cof[C], f[C], h[64], o[O];
  coding_fwd<3, CK, BF, Spec18>(fd, xv, bas, cof);
  combine<C, CK>(fd, bas, cof, f);
  mlp_fwd<C, 64, O, 2>(m, f, h, o);
  const float x0 = o[0] + hd.density_shift;
  const float sg = hd.density_act == 0 ? softplusf(x0) : fmaxf(x0, 0.f);
  alpha[i] += 1.f - expf(-sg * length);
}

// The SH head on its own (SHRender): rgb = sigmoid(eval_sh(2, sh, dir)) per
// channel, and its backward d sh_k = d rgb * y (1 - y) * Y_k(dir), the same
// arithmetic as inside the render kernels.
__global__ __launch_bounds__(256) void sh_rgb_kernel(const float* __restrict__ dirs, int64_t dir_stride,
                                                     const float* __restrict__ sh, int64_t sh_stride, int64_t n,
                                                     const float* __restrict__ d_rgb, float* __restrict__ rgb,
                                                     float* __restrict__ d_sh) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float x = dirs[i * dir_stride], y = dirs[i * dir_stride + 1], z = dirs[i * dir_stride + 2];
  const float* o = sh + i * sh_stride;
  float Y[9];
  if (d_rgb) sh2_basis(x, y, z, Y);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const float v = sigmoidf(sh2_eval(o + 9 * c, x, y, z));
    if (rgb) rgb[i * 3 + c] = v;
    if (d_rgb) {
      const float dpre = d_rgb[i * 3 + c] * (1.f - v) * v;
#pragma unroll
      for (int k = 0; k < 9; ++k) d_sh[i * 27 + 9 * c + k] = dpre * Y[k];
    }
  }
}

namespace {

int rf_status() { return hipGetLastError() == hipSuccess ? FF_OK : FF_ERR_LAUNCH; }

int cu_count_rf() {
  int dev = 0, cu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 256;
  if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cu <= 0)
    return 256;
  return cu;
}

bool nerf_supported(const ff_field_desc& d, const ff_mlp_desc& m) {
  static const int dims[6] = {4, 4, 4, 2, 2, 2};
  if (d.in_dim != 3 || d.x_dim != 3 || d.n_channels != 18 || d.basis_kind != FF_BASIS_GRID) return false;
  if (d.n_levels != 6) return false;
  for (int l = 0; l < 6; ++l)
    if (d.basis_dims[l] != dims[l]) return false;
  if (!(d.coeff_kind == FF_COEFF_GRID && d.coeff_ndim == 3) && d.coeff_kind != FF_COEFF_VM) return false;
  if (d.coeff_kind == FF_COEFF_VM && d.table_dtype != FF_F32) return false;
  if (m.n_layers != 2 || m.in_dim != 18 || m.hidden != 64 || (m.out_dim != 28 && m.out_dim != 32)) return false;
  if (!m.w0 || !m.b0 || !m.w1) return false;
  for (int l = 0; l < 6; ++l)
    if (!d.basis[l] || d.basis_reso[l] < 2) return false;
  const int nc = d.coeff_kind == FF_COEFF_VM ? 6 : 1;
  for (int k = 0; k < nc; ++k)
    if (!d.coeff[k]) return false;
  return true;
}

// The head matches the decoder (SH: 28 outputs, features: 32) and carries
// the buffers its pass needs.
bool head_ok(const ff_render_head& h, const ff_mlp_desc& m, bool bwd) {
  if (h.color == FF_COLOR_SH) return m.out_dim == 28;
  if (h.color != FF_COLOR_FEATURES || m.out_dim != 32 || h.feat_stride < 31) return false;
  return bwd ? (h.d_feat && h.app_pos) : (h.feat != nullptr);
}

template <int CK, bool BF, class F>
int with_head(int head, F&& f) {
  if (head == kHeadFeat)
    return f(std::integral_constant<int, CK>{}, std::integral_constant<bool, BF>{},
             std::integral_constant<int, kHeadFeat>{});
  return f(std::integral_constant<int, CK>{}, std::integral_constant<bool, BF>{},
           std::integral_constant<int, kHeadSH>{});
}

// f(CK, BF, HEAD) for the field's coefficient kind, table type and head.
template <class F>
int with_nerf(const ff_field_desc& d, int head, F&& f) {
  if (d.coeff_kind == FF_COEFF_VM) return with_head<CK_VM, false>(head, f);
  if (d.table_dtype == FF_BF16) return with_head<CK_GRID3, true>(head, f);
  return with_head<CK_GRID3, false>(head, f);
}

unsigned rf_blocks(int64_t cap, int per, int64_t per_cu) {
  const int64_t need = (cap + per - 1) / per;
  const int64_t lim = (int64_t)cu_count_rf() * per_cu;
  return (unsigned)(need < lim ? (need > 0 ? need : 1) : lim);
}

constexpr int nerf_partial(int O) { return 64 * 18 + 64 + O * 64; }
// Backward kernels: one wave per SIMD is resident (registers, LDS), so four
// 64-thread blocks per CU; each block writes one decoder-gradient partial row.
constexpr int kBwdBlocksPerCu = 4;
// The scatter kernel of the split FP32 backward: 64-thread blocks, eight per
// CU (two waves per SIMD).
constexpr int kScatterBlocksPerCu = 8;

}  // namespace
}  // namespace ff

using namespace ff;

extern "C" {

int ff_render_field_fwd(const ff_field_desc* desc, const ff_mlp_desc* mlp, const ff_render_head* head,
                        const float* rays, int64_t n_rays, const ff_samples* s, float* sigma,
                        float* rgb, void* stream) {
  if (!desc || !mlp || !head || !s || !s->ray_off || !s->xyz || !s->ray || !rays || !sigma) return FF_ERR_ARG;
  if (n_rays < 0 || (head->color == FF_COLOR_SH && !rgb)) return FF_ERR_ARG;
  if (!nerf_supported(*desc, *mlp)) return FF_ERR_UNSUPPORTED;
  if (!head_ok(*head, *mlp, false)) return FF_ERR_ARG;
  if (s->cap <= 0 || n_rays == 0) return FF_OK;
  hipStream_t st = (hipStream_t)stream;
  const ff_field_desc d = *desc;
  const ff_mlp_desc m = *mlp;
  const ff_render_head h = *head;
  const ff_samples sv = *s;
  const unsigned blocks = rf_blocks(s->cap, 256, 8);
  return with_nerf(d, h.color == FF_COLOR_FEATURES ? kHeadFeat : kHeadSH, [&](auto CKi, auto BFi, auto Hi) {
    constexpr int CK = decltype(CKi)::value;
    constexpr bool BF = decltype(BFi)::value;
    constexpr int HEAD = decltype(Hi)::value;
    hipLaunchKernelGGL((nerf_field_fwd_mfma_kernel<CK, BF, HEAD>), dim3(blocks), dim3(256), 0, st, d, m, h, rays,
                       sv, n_rays, sigma, rgb);
    return rf_status();
  });
}

int ff_sh_rgb(const float* dirs, int64_t dir_stride, const float* sh, int64_t sh_stride, int64_t n, float* rgb,
              void* stream) {
  if (n < 0 || dir_stride < 3 || sh_stride < 27) return FF_ERR_ARG;
  if (n == 0) return FF_OK;
  if (!dirs || !sh || !rgb) return FF_ERR_ARG;
  hipLaunchKernelGGL(sh_rgb_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, dirs,
                     dir_stride, sh, sh_stride, n, (const float*)nullptr, rgb, (float*)nullptr);
  return rf_status();
}

int ff_sh_rgb_bwd(const float* dirs, int64_t dir_stride, const float* sh, int64_t sh_stride, int64_t n,
                  const float* d_rgb, float* d_sh, void* stream) {
  if (n < 0 || dir_stride < 3 || sh_stride < 27) return FF_ERR_ARG;
  if (n == 0) return FF_OK;
  if (!dirs || !sh || !d_rgb || !d_sh) return FF_ERR_ARG;
  hipLaunchKernelGGL(sh_rgb_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, dirs,
                     dir_stride, sh, sh_stride, n, d_rgb, (float*)nullptr, d_sh);
  return rf_status();
}

int ff_render_alpha(const ff_field_desc* desc, const ff_mlp_desc* mlp, const ff_render_head* head,
                    const ff_ray_sampler* sampler, const float* xyz, int64_t n, float length,
                    float* alpha_accum, void* stream) {
  if (!desc || !mlp || !head || n < 0 || (n > 0 && (!xyz || !alpha_accum))) return FF_ERR_ARG;
  if (!nerf_supported(*desc, *mlp)) return FF_ERR_UNSUPPORTED;
  if (n == 0) return FF_OK;
  ff_ray_sampler occ{};
  if (sampler && sampler->occ) {
    occ = *sampler;
    for (int a = 0; a < 3; ++a)
      if (occ.occ_reso[a] < 2) return FF_ERR_ARG;
  }
  hipStream_t st = (hipStream_t)stream;
  const ff_field_desc d = *desc;
  const ff_mlp_desc m = *mlp;
  const ff_render_head h = *head;
  // only the density is evaluated: the head selects the decoder width
  return with_nerf(d, m.out_dim == 32 ? kHeadFeat : kHeadSH, [&](auto CKi, auto BFi, auto Hi) {
    constexpr int CK = decltype(CKi)::value;
    constexpr bool BF = decltype(BFi)::value;
    constexpr int HEAD = decltype(Hi)::value;
    hipLaunchKernelGGL((nerf_alpha_kernel<CK, BF, HEAD>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, d,
                       m, h, occ, xyz, n, length, alpha_accum);
    return rf_status();
  });
}

size_t ff_render_field_bwd_workspace_bytes(const ff_field_desc* desc, const ff_mlp_desc* mlp,
                                           int64_t cap) {
  if (!desc || !mlp || cap <= 0 || !nerf_supported(*desc, *mlp)) return 0;
  return ((size_t)rf_blocks(cap, 64, kBwdBlocksPerCu) + kReduceSplit) * nerf_partial(mlp->out_dim) * 4;
}

}  // extern "C"

namespace ff {
namespace {

int render_field_bwd_impl(const ff_field_desc* desc, const ff_mlp_desc* mlp, const ff_render_head* head,
                          const float* rays, int64_t n_rays, const ff_samples* s, const float* d_sample,
                          const ff_grads* grads, void* ws, size_t ws_bytes, void* stream) {
  if (!desc || !mlp || !head || !s || !s->ray_off || !s->xyz || !s->ray || !rays || !d_sample || !grads)
    return FF_ERR_ARG;
  if (n_rays < 0) return FF_ERR_ARG;
  if (!nerf_supported(*desc, *mlp)) return FF_ERR_UNSUPPORTED;
  if (!head_ok(*head, *mlp, true)) return FF_ERR_ARG;
  for (int l = 0; l < 6; ++l)
    if (!grads->basis[l]) return FF_ERR_ARG;
  const int nc = desc->coeff_kind == FF_COEFF_VM ? 6 : 1;
  for (int k = 0; k < nc; ++k)
    if (!grads->coeff[k]) return FF_ERR_ARG;
  if (!grads->w0 || !grads->b0 || !grads->w1) return FF_ERR_ARG;
  if (s->cap <= 0 || n_rays == 0) return FF_OK;
  const unsigned blocks = rf_blocks(s->cap, 64, kBwdBlocksPerCu);
  if (!ws || ws_bytes < ((size_t)blocks + kReduceSplit) * nerf_partial(mlp->out_dim) * 4) return FF_ERR_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  const ff_field_desc d = *desc;
  const ff_mlp_desc m = *mlp;
  const ff_render_head h = *head;
  const ff_samples sv = *s;
  const ff_grads g = *grads;
  int r = with_nerf(d, h.color == FF_COLOR_FEATURES ? kHeadFeat : kHeadSH, [&](auto CKi, auto BFi, auto Hi) {
    constexpr int CK = decltype(CKi)::value;
    constexpr bool BF = decltype(BFi)::value;
    constexpr int HEAD = decltype(Hi)::value;
    if constexpr (BF) {
      hipLaunchKernelGGL((nerf_field_bwd_mfma_kernel<CK, HEAD>), dim3(blocks), dim3(64), 0, st, d, m, h, g, rays,
                         sv, n_rays, d_sample, (float*)ws);
    } else if (CK == CK_GRID3 && sv.coding) {
      // stored rows: decoder kernel, then the scatter kernel (reads the rows'
      // gradients the decoder kernel wrote over them)
      hipLaunchKernelGGL((nerf_decoder_bwd_kernel<HEAD>), dim3(blocks), dim3(64), 0, st, d, m, h, rays, sv, n_rays,
                         d_sample, (float*)ws);
      hipLaunchKernelGGL(nerf_scatter_kernel, dim3(rf_blocks(sv.cap, 64, kScatterBlocksPerCu)), dim3(64), 0, st, d,
                         g, sv, n_rays);
    } else {
      hipLaunchKernelGGL((nerf_field_bwd_kernel<CK, BF, HEAD>), dim3(blocks), dim3(64), 0, st, d, m, h, g, rays, sv,
                         n_rays, d_sample, (float*)ws);
    }
    return rf_status();
  });
  if (r) return r;
  return launch_reduce_partials((const float*)ws, (int)blocks, 2, 18, 64, m.out_dim, g, st);
}

// Scratch of ff_render_field_bwd_noalloc: the decoder-gradient partial rows
// (one per backward block, then the reduction's kReduceSplit rows).  Every row
// is written before it is read, so nothing in it needs clearing between
// calls.  The name is a host-side stand-in: host code uses only the address
// hipGetSymbolAddress returns for the current device.
constexpr size_t kScratchBytes = size_t(16) << 20;
__device__ __align__(256) float scratch[kScratchBytes / sizeof(float)];

constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_addr[kScratchDevices] = {};

// This device's copy of `scratch`, looked up once per device (with that
// device current) and kept.
int scratch_addr(void** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kScratchDevices) return FF_ERR_LAUNCH;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_addr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return FF_ERR_LAUNCH;
    g_scratch_addr[dev] = p;
  }
  *out = g_scratch_addr[dev];
  return FF_OK;
}

}  // namespace
}  // namespace ff

extern "C" {

int ff_render_field_bwd(const ff_field_desc* desc, const ff_mlp_desc* mlp, const ff_render_head* head,
                        const float* rays, int64_t n_rays, const ff_samples* s, const float* d_sample,
                        const ff_grads* grads, void* ws, size_t ws_bytes, void* stream) {
  return render_field_bwd_impl(desc, mlp, head, rays, n_rays, s, d_sample, grads, ws, ws_bytes, stream);
}

size_t ff_render_field_bwd_noalloc_limit_bytes(void) { return kScratchBytes; }

int ff_render_field_bwd_noalloc(const ff_field_desc* desc, const ff_mlp_desc* mlp, const ff_render_head* head,
                                const float* rays, int64_t n_rays, const ff_samples* s, const float* d_sample,
                                const ff_grads* grads, void* stream) {
  if (!desc || !mlp || !s) return FF_ERR_ARG;
  const size_t need = ff_render_field_bwd_workspace_bytes(desc, mlp, s->cap);
  if (need > kScratchBytes) return FF_ERR_SCRATCH_LIMIT;
  void* ws = nullptr;
  const int st = scratch_addr(&ws);
  if (st) return st;
  return render_field_bwd_impl(desc, mlp, head, rays, n_rays, s, d_sample, grads, ws, kScratchBytes, stream);
}

}  // extern "C"
