// Colour module of the reference's NeRF model, MLPRender_Fea
// (models/FactorFields.py:162-203), on the FP32 matrix cores, forward and
// backward; and linear_mat's input dropout (:150-151).
//
// Input of a sample: x = [features (NF), viewdir (3), sin | cos of
// features * 2^f (f < FPE), sin | cos of viewdir * 2^f (f < VPE)]
// (positional_encoding :74-79: channel c, frequency f at c * F + f), then
// rgb = sigmoid(W2 relu(W1 relu(W0 x + b0) + b1)).
//
// One block of four waves takes 32 samples at a time (persistent grid).  The
// samples are the N dimension of v_mfma_f32_32x32x2_f32; wave w owns hidden
// units 32w .. 32w+31 of each layer.  Inputs and activations are shared
// through LDS, unit-major [unit][sample] with a row stride of 33 words: lane
// (n, q) of a wave, n = lane % 32, q = lane / 32, reads unit 2s+q of sample
// n as its B operand at K step s, and unit 32t+n of sample 2s+q as an operand
// of the weight gradients (K = samples); both patterns are conflict-free or
// two-way.  Weights stream from global memory (L2-resident).  The forward
// reads the transposed copies W0^T and W1^T that the entry point writes into
// the workspace; the backward reads W1 and W0 in torch's layout.  Either way
// the 32 lanes of a half-wave read 128 contiguous bytes per operand.
//
// The backward recomputes the forward, then forms dH2 on the vector units,
// dH1 = W1^T dH2 and dX = W0^T dH1 on the matrix cores, and the feature
// gradient through the positional encoding.  Weight gradients: dW0 and dW1 as
// K = 32-sample products into accumulators each wave keeps across its tiles
// (MTI + 4 tiles of 32 x 32 per wave), db0, db1 and dW2 on the vector units.
// Each block writes one partial row; two kernels sum the rows in a fixed
// order, so the weight gradients are deterministic.
#include "ff_common.h"

namespace ff {

typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16c_t;

__device__ __forceinline__ f32x16c_t mfma32f(float a, float b, const f32x16c_t& c) {
  return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
}

constexpr int kCH = 128;   // hidden width
constexpr int kCS = 33;    // LDS row stride (words)
constexpr int kCT = 32;    // samples per tile
constexpr int kCSplit = 32;

template <int NF, int FPE, int VPE>
struct ColorShape {
  static constexpr int KIN = NF + 3 + 2 * FPE * NF + 2 * VPE * 3;   // in_mlpC
  static constexpr int KP = (KIN + 1) / 2 * 2;                        // K of layer 1, even
  static constexpr int MTI = (KIN + 31) / 32;                         // 32-wide input tiles
  static constexpr int XR = MTI * 32;                                 // input rows in LDS
  // offsets of the segments of a partial row
  static constexpr int PW0 = 0;                    // dW0 [128][KIN]
  static constexpr int PB0 = kCH * KIN;            // db0 [128]
  static constexpr int PW1 = PB0 + kCH;            // dW1 [128][128]
  static constexpr int PB1 = PW1 + kCH * kCH;      // db1 [128]
  static constexpr int PW2 = PB1 + kCH;            // dW2 [3][128]
  static constexpr int P = PW2 + 3 * kCH;
  // input blocks
  static constexpr int IV = NF;                    // viewdir
  static constexpr int IFS = NF + 3;               // PE(features) sin
  static constexpr int IFC = IFS + NF * FPE;       // PE(features) cos
  static constexpr int IVS = IFC + NF * FPE;       // PE(viewdir) sin
  static constexpr int IVC = IVS + 3 * VPE;        // PE(viewdir) cos
};

// Input k of one sample (k < KIN).  The products by 2^f are exact, as in the
// reference's positions[..., None] * freq_bands.
template <int NF, int FPE, int VPE>
__device__ __forceinline__ float color_input(int k, const float* __restrict__ fp, const float* __restrict__ dp) {
  using S = ColorShape<NF, FPE, VPE>;
  if (k < S::IV) return fp[k];
  if (k < S::IFS) return dp[k - S::IV];
  if (k < S::IVS) {
    const bool sn = k < S::IFC;
    const int j = sn ? k - S::IFS : k - S::IFC;
    const float v = fp[j / FPE] * (float)(1 << (j % FPE));
    return sn ? sinf(v) : cosf(v);
  }
  const bool sn = k < S::IVC;
  const int j = sn ? k - S::IVS : k - S::IVC;
  const float v = dp[j / VPE] * (float)(1 << (j % VPE));
  return sn ? sinf(v) : cosf(v);
}

// Row j of the evaluated set: its sample index, feature row and direction.
struct ColorRow {
  int64_t r;
  const float* fp;
  const float* dp;
};

__device__ __forceinline__ ColorRow color_row(const ff_color_rows& R, int64_t j) {
  ColorRow c;
  c.r = R.rows ? (int64_t)R.rows[j] : j;
  c.fp = R.feat + c.r * R.feat_stride;
  const int64_t d = R.dir_index ? (int64_t)R.dir_index[c.r] : c.r;
  c.dp = R.dirs + d * R.dir_stride;
  return c;
}

__device__ __forceinline__ int64_t color_count(const ff_color_rows& R) {
  if (!R.count_dev) return R.n_max;
  const int64_t c = (int64_t)*R.count_dev;
  return c < R.n_max ? (c > 0 ? c : 0) : R.n_max;
}

// C-register i of a 32 x 32 tile belongs to row 8(i/4) + 4q + i%4.
__device__ __forceinline__ int crow(int i, int q) { return 8 * (i >> 2) + 4 * q + (i & 3); }

template <int NF, int FPE, int VPE, bool BWD>
__global__ __launch_bounds__(256, 1) void color_mlp_kernel(ff_color_mlp_desc m, ff_color_rows R,
                                                           const float* __restrict__ w0t,
                                                           const float* __restrict__ w1t,
                                                           float* __restrict__ rgb, int64_t rgb_stride,
                                                           const float* __restrict__ d_rgb, int64_t d_rgb_stride,
                                                           float* __restrict__ d_feat, int64_t d_feat_stride,
                                                           float* __restrict__ partial) {
  using S = ColorShape<NF, FPE, VPE>;
  constexpr int KIN = S::KIN, KP = S::KP, MTI = S::MTI, XR = S::XR;
  __shared__ float xT[XR * kCS];                 // inputs, then dX (backward)
  __shared__ float h1T[kCH * kCS];
  __shared__ float h2T[kCH * kCS];
  __shared__ float g2T[BWD ? kCH * kCS : 1];
  __shared__ float g1T[BWD ? kCH * kCS : 1];
  __shared__ float dO[BWD ? 3 * kCT : 1];        // dL/d(pre-sigmoid output) [3][32]

  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, n = lane & 31, q = lane >> 5;
  const int64_t nn = color_count(R);
  const int64_t ntiles = (nn + kCT - 1) / kCT;
  // rows k >= KIN of the input block stay zero (operand padding of dW0 / dX)
  for (int i = tid; i < (XR - KIN) * kCS; i += 256) xT[KIN * kCS + i] = 0.f;

  f32x16c_t a0[MTI], a1[4];
  float dbias = 0.f, dw2[2] = {0.f, 0.f};
  if constexpr (BWD) {
#pragma unroll
    for (int t = 0; t < MTI; ++t)
#pragma unroll
      for (int i = 0; i < 16; ++i) a0[t][i] = 0.f;
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int i = 0; i < 16; ++i) a1[t][i] = 0.f;
  }

  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    // 1. inputs of the 32 samples: thread (sample tid % 32) fills units tid/32 + 8k
    {
      const int sn = tid & 31;
      const int64_t j = tile * kCT + sn;
      if (j < nn) {
        const ColorRow c = color_row(R, j);
        for (int k = tid >> 5; k < KIN; k += 8) xT[k * kCS + sn] = color_input<NF, FPE, VPE>(k, c.fp, c.dp);
      } else {
        for (int k = tid >> 5; k < KIN; k += 8) xT[k * kCS + sn] = 0.f;
      }
    }
    __syncthreads();
    // 2. layer 1, hidden units 32 wave + 0..31
    uint32_t msk1 = 0, msk2 = 0;
    {
      f32x16c_t acc;
#pragma unroll
      for (int i = 0; i < 16; ++i) acc[i] = m.b0[32 * wave + crow(i, q)];
#pragma unroll 8
      for (int s = 0; s < KP / 2; ++s) {
        const int k = 2 * s + q;
        acc = mfma32f(w0t[k * kCH + 32 * wave + n], xT[k * kCS + n], acc);
      }
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        msk1 |= uint32_t(acc[i] > 0.f) << i;
        h1T[(32 * wave + crow(i, q)) * kCS + n] = fmaxf(acc[i], 0.f);
      }
    }
    __syncthreads();
    // 3. layer 2
    {
      f32x16c_t acc;
#pragma unroll
      for (int i = 0; i < 16; ++i) acc[i] = m.b1[32 * wave + crow(i, q)];
#pragma unroll 8
      for (int s = 0; s < kCH / 2; ++s) {
        const int k = 2 * s + q;
        acc = mfma32f(w1t[k * kCH + 32 * wave + n], h1T[k * kCS + n], acc);
      }
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        msk2 |= uint32_t(acc[i] > 0.f) << i;
        h2T[(32 * wave + crow(i, q)) * kCS + n] = fmaxf(acc[i], 0.f);
      }
    }
    __syncthreads();
    // 4. layer 3 (wave c computes output c) and the sigmoid
    if (wave < 3) {
      const int c = wave;
      float o = 0.f;
#pragma unroll 8
      for (int s = 0; s < kCH / 2; ++s) {
        const int k = 2 * s + q;
        o = fmaf(m.w2[c * kCH + k], h2T[k * kCS + n], o);
      }
      o = o + __shfl_xor(o, 32, kWave);
      const float y = 1.f / (1.f + expf(-o));
      const int64_t j = tile * kCT + n;
      if constexpr (!BWD) {
        if (q == 0 && j < nn) rgb[color_row(R, j).r * rgb_stride + c] = y;
      } else {
        if (q == 0) {
          float g = 0.f;
          if (j < nn) g = d_rgb[color_row(R, j).r * d_rgb_stride + c];
          dO[c * kCT + n] = (g * (1.f - y)) * y;
        }
      }
    }
    if constexpr (BWD) {
      __syncthreads();
      // 5. dH2 = relu'(.) * W2^T dO on the vector units
      {
        const float d0 = dO[n], d1 = dO[kCT + n], d2 = dO[2 * kCT + n];
#pragma unroll
        for (int i = 0; i < 16; ++i) {
          const int u = 32 * wave + crow(i, q);
          float g = m.w2[u] * d0;
          g = fmaf(m.w2[kCH + u], d1, g);
          g = fmaf(m.w2[2 * kCH + u], d2, g);
          g2T[u * kCS + n] = ((msk2 >> i) & 1) ? g : 0.f;
        }
      }
      __syncthreads();
      // 6. dH1 = relu'(.) * W1^T dH2 (W1 [out][in]: A[m = in][k = out] = W1[k][m])
      {
        f32x16c_t acc;
#pragma unroll
        for (int i = 0; i < 16; ++i) acc[i] = 0.f;
#pragma unroll 8
        for (int s = 0; s < kCH / 2; ++s) {
          const int k = 2 * s + q;
          acc = mfma32f(m.w1[k * kCH + 32 * wave + n], g2T[k * kCS + n], acc);
        }
#pragma unroll
        for (int i = 0; i < 16; ++i)
          g1T[(32 * wave + crow(i, q)) * kCS + n] = ((msk1 >> i) & 1) ? acc[i] : 0.f;
      }
      __syncthreads();
      // 7. weight gradients of the tile (K = its 32 samples)
      {
#pragma unroll
        for (int t = 0; t < MTI; ++t) {
          const int qd = wave + 4 * t, mt = qd / MTI, nt = qd % MTI;
#pragma unroll
          for (int s = 0; s < kCT / 2; ++s) {
            const int k = 2 * s + q;
            a0[t] = mfma32f(g1T[(32 * mt + n) * kCS + k], xT[(32 * nt + n) * kCS + k], a0[t]);
          }
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) {
#pragma unroll
          for (int s = 0; s < kCT / 2; ++s) {
            const int k = 2 * s + q;
            a1[t] = mfma32f(g2T[(32 * t + n) * kCS + k], h1T[(32 * wave + n) * kCS + k], a1[t]);
          }
        }
        // db0 (threads 0..127) and db1 (128..255), summed over the tile in order
        {
          const float* src = tid < kCH ? g1T + tid * kCS : g2T + (tid - kCH) * kCS;
          float t = 0.f;
#pragma unroll 8
          for (int sm = 0; sm < kCT; ++sm) t += src[sm];
          dbias += t;
        }
        // dW2 [3][128]: entries tid and tid + 256
#pragma unroll
        for (int jp = 0; jp < 2; ++jp) {
          const int p = tid + 256 * jp;
          if (p < 3 * kCH) {
            const int c = p / kCH, u = p % kCH;
            float t = 0.f;
#pragma unroll 8
            for (int sm = 0; sm < kCT; ++sm) t = fmaf(dO[c * kCT + sm], h2T[u * kCS + sm], t);
            dw2[jp] += t;
          }
        }
      }
      __syncthreads();
      // 8. dX = W0^T dH1 into the input block (W0 [hidden][KIN]: A[m = k][k' = u] = W0[u][k])
#pragma unroll
      for (int tt = 0; tt < 2; ++tt) {
        const int t = wave + 4 * tt;
        if (t < MTI) {
          const int col = 32 * t + n;
          const bool ok = col < KIN;
          const int colc = ok ? col : KIN - 1;
          f32x16c_t acc;
#pragma unroll
          for (int i = 0; i < 16; ++i) acc[i] = 0.f;
#pragma unroll 8
          for (int s = 0; s < kCH / 2; ++s) {
            const int k = 2 * s + q;
            const float a = m.w0[k * KIN + colc];
            acc = mfma32f(ok ? a : 0.f, g1T[k * kCS + n], acc);
          }
#pragma unroll
          for (int i = 0; i < 16; ++i) xT[(32 * t + crow(i, q)) * kCS + n] = acc[i];
        }
      }
      __syncthreads();
      // 9. feature gradient: direct input plus the sin / cos blocks of its encoding
      {
        const int sn = tid & 31;
        const int64_t j = tile * kCT + sn;
        if (j < nn) {
          const ColorRow cr = color_row(R, j);
          for (int c = tid >> 5; c < NF; c += 8) {
            const float x = cr.fp[c];
            float g = xT[c * kCS + sn];
#pragma unroll
            for (int f = 0; f < FPE; ++f) {
              const float sc = (float)(1 << f);
              const float v = x * sc;
              const float gs = xT[(S::IFS + c * FPE + f) * kCS + sn];
              const float gc = xT[(S::IFC + c * FPE + f) * kCS + sn];
              g = g + (gs * cosf(v) - gc * sinf(v)) * sc;
            }
            d_feat[j * d_feat_stride + c] = g;
          }
        }
      }
    }
    __syncthreads();
  }

  if constexpr (BWD) {
    float* row = partial + (int64_t)blockIdx.x * S::P;
#pragma unroll
    for (int t = 0; t < MTI; ++t) {
      const int qd = wave + 4 * t, mt = qd / MTI, nt = qd % MTI;
      const int col = 32 * nt + n;
      if (col < KIN) {
#pragma unroll
        for (int i = 0; i < 16; ++i) row[S::PW0 + (32 * mt + crow(i, q)) * KIN + col] = a0[t][i];
      }
    }
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int i = 0; i < 16; ++i) row[S::PW1 + (32 * t + crow(i, q)) * kCH + 32 * wave + n] = a1[t][i];
    row[tid < kCH ? S::PB0 + tid : S::PB1 + (tid - kCH)] = dbias;
#pragma unroll
    for (int jp = 0; jp < 2; ++jp) {
      const int p = tid + 256 * jp;
      if (p < 3 * kCH) row[S::PW2 + p] = dw2[jp];
    }
  }
}

// W0^T [KP][128] (rows k >= KIN zero) and W1^T [128][128] for the forward's
// coalesced operand reads.
__global__ __launch_bounds__(256) void color_transpose_kernel(const float* __restrict__ w0,
                                                              const float* __restrict__ w1, int KIN, int KP,
                                                              float* __restrict__ w0t, float* __restrict__ w1t) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int n0 = KP * kCH;
  if (i < n0) {
    const int k = i / kCH, u = i % kCH;
    w0t[i] = k < KIN ? w0[u * KIN + k] : 0.f;
  } else if (i < n0 + kCH * kCH) {
    const int e = i - n0, k = e / kCH, u = e % kCH;
    w1t[e] = w1[u * kCH + k];
  }
}

// part2[y][k] = sum of rows y, y + kCSplit, ... (in order).
__global__ __launch_bounds__(256) void color_reduce_stage1(const float* __restrict__ rows, int nrows, int P,
                                                           float* __restrict__ part2) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= P) return;
  float s = 0.f;
  for (int r = blockIdx.y; r < nrows; r += kCSplit) s += rows[(int64_t)r * P + k];
  part2[(int64_t)blockIdx.y * P + k] = s;
}

// g += sum over the kCSplit slices, segment by segment.
__global__ __launch_bounds__(256) void color_reduce_stage2(const float* __restrict__ part2, int P, int pb0,
                                                           int pw1, int pb1, int pw2, ff_color_mlp_grads g) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= P) return;
  float t = 0.f;
#pragma unroll
  for (int y = 0; y < kCSplit; ++y) t += part2[(int64_t)y * P + k];
  if (k < pb0) g.w0[k] += t;
  else if (k < pw1) g.b0[k - pb0] += t;
  else if (k < pb1) g.w1[k - pw1] += t;
  else if (k < pw2) g.b1[k - pb1] += t;
  else g.w2[k - pw2] += t;
}

// ---------------------------------------------------------------- dropout
// Squares: a counter-based generator (B. Widynski, "Squares: A Fast
// Counter-Based RNG", 2020): four rounds of squaring with a key; 32 bits out.
__device__ __forceinline__ uint32_t squares32(uint64_t ctr, uint64_t key) {
  uint64_t x = ctr * key, y = x, z = y + key;
  x = x * x + y;
  x = (x >> 32) | (x << 32);
  x = x * x + z;
  x = (x >> 32) | (x << 32);
  x = x * x + y;
  x = (x >> 32) | (x << 32);
  return (uint32_t)((x * x + z) >> 32);
}

__global__ __launch_bounds__(256) void dropout_kernel(const float* __restrict__ x, int64_t n, float scale,
                                                      float keep_prob, const uint8_t* __restrict__ keep,
                                                      uint64_t key, uint64_t offset, float* __restrict__ out,
                                                      uint8_t* __restrict__ keep_out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  bool k;
  if (keep) {
    k = keep[i] != 0;
  } else {
    const float u = (float)(squares32(offset + (uint64_t)i, key) >> 8) * 0x1p-24f;
    k = u < keep_prob;
  }
  out[i] = x[i] * (k ? scale : 0.f);
  if (keep_out) keep_out[i] = k ? 1 : 0;
}

namespace {

int color_status() { return hipGetLastError() == hipSuccess ? FF_OK : FF_ERR_LAUNCH; }

int color_cus() {
  int dev = 0, cu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 256;
  if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cu <= 0) return 256;
  return cu;
}

// Instantiated shapes: in_feat 31, fea_pe 2, hidden 128, 3 layers, view_pe 6 or 2.
template <class F>
int with_color(const ff_color_mlp_desc& m, F&& f) {
  if (m.in_feat != 31 || m.fea_pe != 2 || m.hidden != kCH || m.n_layers != 3) return FF_ERR_UNSUPPORTED;
  if (m.view_pe == 6) return f(std::integral_constant<int, 6>{});
  if (m.view_pe == 2) return f(std::integral_constant<int, 2>{});
  return FF_ERR_UNSUPPORTED;
}

int kin_of(const ff_color_mlp_desc& m) { return m.in_feat + 3 + 2 * m.fea_pe * m.in_feat + 6 * m.view_pe; }

unsigned color_blocks(int64_t n_max) {
  const int64_t tiles = (n_max + kCT - 1) / kCT;
  const int64_t lim = color_cus();
  return (unsigned)(tiles < lim ? (tiles > 0 ? tiles : 1) : lim);
}

// workspace: W0^T [KP][128] | W1^T [128][128] | (backward) partial rows | reduction slices
size_t color_ws_floats(const ff_color_mlp_desc& m, int64_t n_max, bool bwd) {
  const int KIN = kin_of(m), KP = (KIN + 1) / 2 * 2;
  const int P = kCH * KIN + kCH + kCH * kCH + kCH + 3 * kCH;
  size_t f = (size_t)KP * kCH + (size_t)kCH * kCH;
  if (bwd) f += ((size_t)color_blocks(n_max) + kCSplit) * P;
  return f;
}

int check_color(const ff_color_mlp_desc* m, const ff_color_rows* r) {
  if (!m || !r) return FF_ERR_ARG;
  if (r->n_max < 0) return FF_ERR_ARG;
  if (!m->w0 || !m->b0 || !m->w1 || !m->b1 || !m->w2) return FF_ERR_ARG;
  if (r->n_max > 0 && (!r->feat || !r->dirs || r->feat_stride < m->in_feat || r->dir_stride < 3))
    return FF_ERR_ARG;
  return FF_OK;
}

}  // namespace
}  // namespace ff

using namespace ff;

extern "C" {

size_t ff_color_mlp_workspace_bytes(const ff_color_mlp_desc* m, int64_t n_max, int backward) {
  if (!m || n_max < 0) return 0;
  return color_ws_floats(*m, n_max, backward != 0) * sizeof(float);
}

int ff_color_mlp_fwd(const ff_color_mlp_desc* m, const ff_color_rows* r, float* rgb, int64_t rgb_stride,
                     void* ws, size_t ws_bytes, void* stream) {
  int st = check_color(m, r);
  if (st) return st;
  if (r->n_max > 0 && (!rgb || rgb_stride < 3)) return FF_ERR_ARG;
  const ff_color_mlp_desc md = *m;
  const ff_color_rows rv = *r;
  return with_color(md, [&](auto VPEi) {
    constexpr int VPE = decltype(VPEi)::value;
    using S = ColorShape<31, 2, VPE>;
    if (rv.n_max == 0) return (int)FF_OK;
    if (!ws || ws_bytes < color_ws_floats(md, rv.n_max, false) * sizeof(float)) return (int)FF_ERR_WORKSPACE;
    hipStream_t s = (hipStream_t)stream;
    float* w0t = (float*)ws;
    float* w1t = w0t + (size_t)S::KP * kCH;
    const int nt = S::KP * kCH + kCH * kCH;
    hipLaunchKernelGGL(color_transpose_kernel, dim3((nt + 255) / 256), dim3(256), 0, s, md.w0, md.w1, S::KIN, S::KP,
                       w0t, w1t);
    hipLaunchKernelGGL((color_mlp_kernel<31, 2, VPE, false>), dim3(color_blocks(rv.n_max)), dim3(256), 0, s, md, rv,
                       w0t, w1t, rgb, rgb_stride, nullptr, 0, nullptr, 0, nullptr);
    return color_status();
  });
}

int ff_color_mlp_bwd(const ff_color_mlp_desc* m, const ff_color_rows* r, const float* d_rgb, int64_t d_rgb_stride,
                     float* d_feat, int64_t d_feat_stride, const ff_color_mlp_grads* g, void* ws, size_t ws_bytes,
                     void* stream) {
  int st = check_color(m, r);
  if (st) return st;
  if (!g || !g->w0 || !g->b0 || !g->w1 || !g->b1 || !g->w2) return FF_ERR_ARG;
  if (r->n_max > 0 && (!d_rgb || d_rgb_stride < 3 || !d_feat || d_feat_stride < m->in_feat)) return FF_ERR_ARG;
  const ff_color_mlp_desc md = *m;
  const ff_color_rows rv = *r;
  const ff_color_mlp_grads gv = *g;
  return with_color(md, [&](auto VPEi) {
    constexpr int VPE = decltype(VPEi)::value;
    using S = ColorShape<31, 2, VPE>;
    if (rv.n_max == 0) return (int)FF_OK;
    if (!ws || ws_bytes < color_ws_floats(md, rv.n_max, true) * sizeof(float)) return (int)FF_ERR_WORKSPACE;
    hipStream_t s = (hipStream_t)stream;
    float* w0t = (float*)ws;
    float* w1t = w0t + (size_t)S::KP * kCH;
    float* rows = w1t + (size_t)kCH * kCH;
    const unsigned nb = color_blocks(rv.n_max);
    float* part2 = rows + (size_t)nb * S::P;
    const int nt = S::KP * kCH + kCH * kCH;
    hipLaunchKernelGGL(color_transpose_kernel, dim3((nt + 255) / 256), dim3(256), 0, s, md.w0, md.w1, S::KIN, S::KP,
                       w0t, w1t);
    hipLaunchKernelGGL((color_mlp_kernel<31, 2, VPE, true>), dim3(nb), dim3(256), 0, s, md, rv, w0t, w1t, nullptr, 0,
                       d_rgb, d_rgb_stride, d_feat, d_feat_stride, rows);
    hipLaunchKernelGGL(color_reduce_stage1, dim3((S::P + 255) / 256, kCSplit), dim3(256), 0, s, rows, (int)nb, S::P,
                       part2);
    hipLaunchKernelGGL(color_reduce_stage2, dim3((S::P + 255) / 256), dim3(256), 0, s, part2, S::P, S::PB0, S::PW1,
                       S::PB1, S::PW2, gv);
    return color_status();
  });
}

int ff_dropout(const float* x, int64_t n, float p, const uint8_t* keep, uint64_t seed, uint64_t offset, float* out,
               uint8_t* keep_out, void* stream) {
  if (n < 0 || (n > 0 && (!x || !out)) || !(p >= 0.f && p < 1.f)) return FF_ERR_ARG;
  if (n == 0) return FF_OK;
  // key: splitmix64 of the seed, odd
  uint64_t z = seed + 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  z = z ^ (z >> 31);
  const uint64_t key = z | 1ull;
  const float keep_prob = 1.f - p;
  const float scale = 1.f / keep_prob;
  hipLaunchKernelGGL(dropout_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, x, n,
                     scale, keep_prob, keep, key, offset, out, keep_out);
  return color_status();
}

}  // extern "C"
