// FP32 matrix-core decoder for one wavefront of 64 samples (lane = sample):
// linear_mat 18 -> 64 -> O, O <= 32 (models/FactorFields.py:113-159, ReLU
// hidden, no output bias) on v_mfma_f32_32x32x2_f32 (FP32 operands, FP32
// accumulation), and its backward.  The FP32-table render kernels use it in
// place of the vector decoder, whose 3,264 weights per launch either sit in
// scalar registers (and spill) or stream through the scalar cache with one
// wave per SIMD waiting on every load.  Here each wave loads its weights once,
// in operand form, into vector registers (116 per lane).
//
// Orientation as in the BF16 decoder (ff_mfma.h): samples are the N
// dimension, H^T [64 x 64] = W0a [64 x 20] . X^T [20 x 64] with X the 18
// features, a constant-1 column (b0) and a zero column; out^T [32 x 64] =
// W1 [32 x 64] . relu(H^T).
//
// Operand layouts of v_mfma_f32_32x32x2_f32 (lane l, r = l % 32, q = l / 32):
//   A [32 x 2]: lane l holds A[r][q]
//   B [2 x 32]: lane l holds B[q][r]
//   C [32 x 32]: lane l, register i holds C[8(i/4) + 4q + i%4][r]
// Lane-per-sample vectors enter B operands through one v_permlane32_swap per
// K step: with P = element 2s and Q = element 2s+1 of every lane, the swap
// leaves P holding the operand of samples 0..31 and Q that of samples 32..63.
//
// K order of the products over hidden units: K step s (0..31), half q reads
// hidden unit f32_hid(s, q) = 32(s/16) + 8((s%16)/4) + 4q + s%4, which is
// register s%16 of layer 1's C tile s/16.  W1 and W0a^T are loaded in that
// order, so relu(H^T) and dH^T feed the next product from their registers.
//
// Weight gradients (K = samples) go through a per-wave LDS transpose, one
// 32-sample tile at a time: dout^T, X^T, relu(H)^T and dH^T rows of 32
// samples (stride kF32Row words).  K step s, half q takes sample
// 8(s/4) + 4q + s%4, so one 16-byte LDS read gives a lane four K steps.
#pragma once
#include "ff_mfma.h"

namespace ff {

__device__ __forceinline__ f32x16_t mfma_f32(float a, float b, const f32x16_t& c) {
  return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ int f32_hid(int s, int q) { return 32 * (s >> 4) + 8 * ((s & 15) >> 2) + 4 * q + (s & 3); }

constexpr int kF32K0 = 10;   // K steps of layer 1 (18 features + bias + zero = 20)
constexpr int kF32Row = 40;  // LDS row stride (words): 16-byte aligned rows, q halves 32 banks apart
// LDS words per wave: dout^T [32] | X^T [32] | relu(H)^T [64] | dH^T [64] rows.
constexpr int kF32LdsWords = (32 + 32 + 64 + 64) * kF32Row;

// Forward operands, FP32.
template <int O>
struct MfmaDecoderF32 {
  static_assert(O >= 1 && O <= 32, "layer 2 fills one 32-row tile");
  float w0a[2][kF32K0];  // A of layer 1: [hidden tile][K step]
  float w1a[32];         // A of layer 2: [K step] (rows >= O zero)
  __device__ __forceinline__ void load(const ff_mlp_desc& m) {
    const int lane = threadIdx.x & 63, r = lane & 31, q = lane >> 5;
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
      for (int ks = 0; ks < kF32K0; ++ks) w0a[mt][ks] = w0_aug(m, mt * 32 + r, 2 * ks + q);
#pragma unroll
    for (int s = 0; s < 32; ++s) {
      const float v = m.w1[(r < O ? r : O - 1) * kMfH + f32_hid(s, q)];
      w1a[s] = r < O ? v : 0.f;
    }
  }
};

// Backward operands: W1^T (A of dH^T = W1^T . dout^T) and W0a^T in the
// permuted hidden order (A of dX^T = W0a^T . dH^T; feature rows >= 18 zero).
template <int O>
struct MfmaDecoderF32Bwd {
  float w1t[2][16];
  float w0t[32];
  __device__ __forceinline__ void load(const ff_mlp_desc& m) {
    const int lane = threadIdx.x & 63, r = lane & 31, q = lane >> 5;
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
      for (int ks = 0; ks < 16; ++ks) {
        const int o = 2 * ks + q;
        const float v = m.w1[(o < O ? o : O - 1) * kMfH + mt * 32 + r];
        w1t[mt][ks] = o < O ? v : 0.f;
      }
#pragma unroll
    for (int s = 0; s < 32; ++s) {
      const float v = m.w0[f32_hid(s, q) * kMfIn + (r < kMfIn ? r : kMfIn - 1)];
      w0t[s] = r < kMfIn ? v : 0.f;
    }
  }
};

// Lane-per-sample v[2K] -> B operands b[K step][sample tile].
template <int K>
__device__ __forceinline__ void f32_pack_b(const float* v, float (&b)[K][2]) {
#pragma unroll
  for (int s = 0; s < K; ++s) {
    float p = v[2 * s], x = v[2 * s + 1];
    swap32(p, x);
    b[s][0] = p;
    b[s][1] = x;
  }
}

// Features (+ bias column, zero column) -> X operands.
__device__ __forceinline__ void f32_pack_x(const float* f, float (&xb)[kF32K0][2]) {
  float x[2 * kF32K0];
#pragma unroll
  for (int k = 0; k < 2 * kF32K0; ++k) x[k] = k < kMfIn ? f[k] : (k == kMfIn ? 1.f : 0.f);
  f32_pack_b<kF32K0>(x, xb);
}

// Forward of sample tile nt: hr[mt] = relu(H^T) tiles (C layout), c2 = out^T.
template <int O>
__device__ __forceinline__ void f32_fwd_tile(const MfmaDecoderF32<O>& W, const float (&xb)[kF32K0][2], int nt,
                                             f32x16_t (&hr)[2], f32x16_t& c2) {
#pragma unroll
  for (int mt = 0; mt < 2; ++mt) {
    f32x16_t acc = zero16();
#pragma unroll
    for (int ks = 0; ks < kF32K0; ++ks) acc = mfma_f32(W.w0a[mt][ks], xb[ks][nt], acc);
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = fmaxf(acc[i], 0.f);
    hr[mt] = acc;
  }
  c2 = zero16();
#pragma unroll
  for (int s = 0; s < 32; ++s) c2 = mfma_f32(W.w1a[s], hr[s >> 4][s & 15], c2);
}

// Forward: f [18] per lane -> out [32] per lane (O used).
template <int O>
__device__ __forceinline__ void mfma_f32_decoder_fwd(const MfmaDecoderF32<O>& W, const float* f, float* out) {
  float xb[kF32K0][2];
  f32_pack_x(f, xb);
  f32x16_t c2[2];
#pragma unroll
  for (int nt = 0; nt < 2; ++nt) {
    f32x16_t hr[2];
    f32_fwd_tile(W, xb, nt, hr, c2[nt]);
  }
  mf_unpack_c(c2, out);
}

// Training pass of one 64-sample tile on the matrix cores.  make_dout(out,
// dout) turns the lane's outputs into d out (32 values, zero for o >= O and
// for inactive lanes); df receives d features [18].  The weight gradients
// are added to acc.  Uses the wave's kF32LdsWords of LDS and ends with it free.
// For blocks of one wave (the LDS hand-offs are wave fences).
template <int O, class DoutFn>
__device__ __forceinline__ void mfma_f32_decoder_train(const MfmaDecoderF32<O>& W, const MfmaDecoderF32Bwd<O>& WB,
                                                       const float* f, DoutFn&& make_dout, float* df,
                                                       MfmaDecoderAcc<O>& acc, float* lds) {
  const int lane = threadIdx.x & 63, r = lane & 31, q = lane >> 5;
  float xb[kF32K0][2];
  f32_pack_x(f, xb);
  f32x16_t hr[2][2], c2[2];
#pragma unroll
  for (int nt = 0; nt < 2; ++nt) f32_fwd_tile(W, xb, nt, hr[nt], c2[nt]);
  float dout[32];
  {
    float out[32];
    mf_unpack_c(c2, out);
    make_dout(out, dout);
  }
  float db[16][2];
  f32_pack_b<16>(dout, db);
  float* DT = lds;
  float* XT = DT + 32 * kF32Row;
  float* HT = XT + 32 * kF32Row;
  float* GT = HT + 64 * kF32Row;
  f32x16_t dx[2];
#pragma unroll
  for (int nt = 0; nt < 2; ++nt) {
    // dH^T = W1^T . dout^T, masked by the ReLU pattern (relu(h) > 0 <=> h > 0)
    f32x16_t dh[2];
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
      f32x16_t a = zero16();
#pragma unroll
      for (int ks = 0; ks < 16; ++ks) a = mfma_f32(WB.w1t[mt][ks], db[ks][nt], a);
#pragma unroll
      for (int i = 0; i < 16; ++i) a[i] = hr[nt][mt][i] > 0.f ? a[i] : 0.f;
      dh[mt] = a;
    }
    // dX^T = W0a^T . dH^T
    f32x16_t a = zero16();
#pragma unroll
    for (int s = 0; s < 32; ++s) a = mfma_f32(WB.w0t[s], dh[s >> 4][s & 15], a);
    dx[nt] = a;
    // this sample tile's operands of the weight gradients, unit-major; X^T
    // rows 20..31 are zero (the dW0 product reads them as feature columns
    // 20..31, whose sums are not stored)
    for (int w = lane; w < (32 - 2 * kF32K0) * kF32Row; w += 64) XT[2 * kF32K0 * kF32Row + w] = 0.f;
#pragma unroll
    for (int ks = 0; ks < 16; ++ks) DT[(2 * ks + q) * kF32Row + r] = db[ks][nt];
#pragma unroll
    for (int ks = 0; ks < kF32K0; ++ks) XT[(2 * ks + q) * kF32Row + r] = xb[ks][nt];
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const int h = 32 * mt + 8 * (i >> 2) + 4 * q + (i & 3);
        HT[h * kF32Row + r] = hr[nt][mt][i];
        GT[h * kF32Row + r] = dh[mt][i];
      }
    wave_lds_fence();
    // dW1 [o][h] += dout^T . relu(H);  dW0a [h][k] += dH^T . X   (K = samples)
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      const int col = 8 * g + 4 * q;
      const float4 ad = *reinterpret_cast<const float4*>(DT + r * kF32Row + col);
      const float4 bx = *reinterpret_cast<const float4*>(XT + r * kF32Row + col);
      float4 bh[2], ag[2];
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        bh[t] = *reinterpret_cast<const float4*>(HT + (32 * t + r) * kF32Row + col);
        ag[t] = *reinterpret_cast<const float4*>(GT + (32 * t + r) * kF32Row + col);
      }
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        acc.gw1[t] = mfma_f32(ad.x, bh[t].x, acc.gw1[t]);
        acc.gw1[t] = mfma_f32(ad.y, bh[t].y, acc.gw1[t]);
        acc.gw1[t] = mfma_f32(ad.z, bh[t].z, acc.gw1[t]);
        acc.gw1[t] = mfma_f32(ad.w, bh[t].w, acc.gw1[t]);
        acc.gw0[t] = mfma_f32(ag[t].x, bx.x, acc.gw0[t]);
        acc.gw0[t] = mfma_f32(ag[t].y, bx.y, acc.gw0[t]);
        acc.gw0[t] = mfma_f32(ag[t].z, bx.z, acc.gw0[t]);
        acc.gw0[t] = mfma_f32(ag[t].w, bx.w, acc.gw0[t]);
      }
    }
    wave_lds_fence();
  }
  float dxs[32];
  mf_unpack_c(dx, dxs);
#pragma unroll
  for (int k = 0; k < kMfIn; ++k) df[k] = dxs[k];
}

}  // namespace ff
