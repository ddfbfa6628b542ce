// BF16 matrix-core decoder for one wavefront of 64 samples (lane = sample):
// linear_mat 18 -> 64 -> O, O <= 32 (models/FactorFields.py:113-159, ReLU
// hidden, no output bias; O = 28 for the SH colour head, 32 for the feature
// head of MLPRender_Fea) on v_mfma_f32_32x32x16_bf16 with FP32 accumulation,
// and its backward (data gradient + weight gradients).
//
// Orientation: samples are the N dimension.  H^T [64 x 64] = W0a [64 x 32] .
// X^T [32 x 64], where X is the 18 features plus a constant-1 column that
// carries b0 (K padded to 32); out^T [32 x 64] = W1 [32 x 64] . relu(H^T).
//
// Operand layouts of v_mfma_f32_32x32x16_bf16 (lane l, r = l % 32, q = l / 32):
//   A [32 x 16]: lane l holds A[r][8q .. 8q+7]
//   B [16 x 32]: lane l holds B[8q .. 8q+7][r]
//   C [32 x 32]: lane l, register i holds C[8(i/4) + 4q + i%4][r]
// Lane-per-sample vectors enter B operands through one v_permlane32_swap per
// dword: with P = elements 0..7 and Q = elements 8..15 of a 16-wide K step,
// the swap leaves P holding the operand of samples 0..31 and Q that of samples
// 32..63.  C tiles return to lane-per-sample form the same way.
//
// The C layout of layer 1 already is the B layout of layer 2 once the K order
// (hidden units) is permuted: K step ks, half q, element j reads hidden unit
// hid(ks, q, j) = 32(ks/2) + 8((ks%2)*2 + j/4) + 4q + j%4, which is register
// (ks%2)*8 + j of tile ks/2.  W1 (and W0^T in the backward) are loaded in that
// order, so no data moves between the two products.
#pragma once
#include "ff_common.h"

namespace ff {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8_t;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16_t;

constexpr int kMfIn = 18, kMfH = 64;

// LDS words the backward needs per wave: four bf16 matrices whose rows hold 64
// samples (+8 padding so that the 16-byte operand reads of 16 consecutive
// rows start in different banks): d out^T [32], X^T [32], relu(H)^T [64],
// dH^T [64].
constexpr int kMfRow = 72;
constexpr int kMfLdsWords = (32 + 32 + 64 + 64) * kMfRow / 2;

__device__ __forceinline__ f32x16_t mfma_bf16(const bf16x8_t& a, const bf16x8_t& b, const f32x16_t& c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ f32x16_t zero16() {
  f32x16_t z;
#pragma unroll
  for (int i = 0; i < 16; ++i) z[i] = 0.f;
  return z;
}

// Swap the upper half of a with the lower half of b (v_permlane32_swap).
__device__ __forceinline__ void swap32(float& a, float& b) {
  auto r = __builtin_amdgcn_permlane32_swap(__float_as_int(a), __float_as_int(b), false, false);
  a = __int_as_float(r[0]);
  b = __int_as_float(r[1]);
}

__device__ __forceinline__ void swap32(bf16x8_t& a, bf16x8_t& b) {
  typedef __attribute__((__vector_size__(4 * sizeof(int)))) int i32x4_t;
  i32x4_t x = __builtin_bit_cast(i32x4_t, a), y = __builtin_bit_cast(i32x4_t, b);
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    auto r = __builtin_amdgcn_permlane32_swap(x[w], y[w], false, false);
    x[w] = r[0];
    y[w] = r[1];
  }
  a = __builtin_bit_cast(bf16x8_t, x);
  b = __builtin_bit_cast(bf16x8_t, y);
}

__device__ __forceinline__ int mf_hid(int ks, int q, int j) {
  return 32 * (ks >> 1) + 8 * ((ks & 1) * 2 + (j >> 2)) + 4 * q + (j & 3);
}

// W0 with b0 as column 18 and zero columns 19..31.  The weight loads in this
// file are unconditional, at indices clamped into the matrix, and the zero
// padding is selected afterwards: a load under a per-lane condition becomes a
// branch that waits for its load before the next one issues, which serialised
// the whole operand set (about a hundred memory latencies per wave).
__device__ __forceinline__ float w0_aug(const ff_mlp_desc& m, int h, int k) {
  const float w = m.w0[h * kMfIn + (k < kMfIn ? k : kMfIn - 1)];
  const float b = m.b0[h];
  return k < kMfIn ? w : (k == kMfIn ? b : 0.f);
}

// Decoder weights in operand form (BF16, rounded from the FP32 masters).
template <int O>
struct MfmaDecoder {
  static_assert(O >= 1 && O <= 32, "layer 2 fills one 32-row tile");
  bf16x8_t w0a[2][2];  // A of layer 1: [hidden tile][K step]
  bf16x8_t w1a[4];     // A of layer 2: [K step] (rows >= O zero)
  __device__ __forceinline__ void load(const ff_mlp_desc& m) {
    const int lane = threadIdx.x & 63, r = lane & 31, q = lane >> 5;
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
      for (int ks = 0; ks < 2; ++ks)
#pragma unroll
        for (int j = 0; j < 8; ++j) w0a[mt][ks][j] = (__bf16)w0_aug(m, mt * 32 + r, ks * 16 + 8 * q + j);
#pragma unroll
    for (int ks = 0; ks < 4; ++ks)
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const float v = m.w1[(r < O ? r : O - 1) * kMfH + mf_hid(ks, q, j)];
        w1a[ks][j] = (__bf16)(r < O ? v : 0.f);
      }
  }
};

// Backward-only operands: W1^T (A of dH^T = W1^T . dout^T) and W0a^T in the
// permuted hidden order (A of dX^T = W0a^T . dH^T).
template <int O>
struct MfmaDecoderBwd {
  bf16x8_t w1t[2][2];
  bf16x8_t w0t[4];
  __device__ __forceinline__ void load(const ff_mlp_desc& m) {
    const int lane = threadIdx.x & 63, r = lane & 31, q = lane >> 5;
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
      for (int ks = 0; ks < 2; ++ks)
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const int o = ks * 16 + 8 * q + j;
          const float v = m.w1[(o < O ? o : O - 1) * kMfH + mt * 32 + r];
          w1t[mt][ks][j] = (__bf16)(o < O ? v : 0.f);
        }
#pragma unroll
    for (int ks = 0; ks < 4; ++ks)
#pragma unroll
      for (int j = 0; j < 8; ++j) w0t[ks][j] = (__bf16)w0_aug(m, mf_hid(ks, q, j), r);
  }
};

// Lane-per-sample vector v[32] (K = 32) -> B operands [K step][sample tile].
__device__ __forceinline__ void mf_pack_b(const float* v, bf16x8_t (&b)[2][2]) {
#pragma unroll
  for (int ks = 0; ks < 2; ++ks) {
    bf16x8_t p, q;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      p[j] = (__bf16)v[ks * 16 + j];
      q[j] = (__bf16)v[ks * 16 + 8 + j];
    }
    swap32(p, q);
    b[ks][0] = p;
    b[ks][1] = q;
  }
}

// C tiles of both sample tiles -> lane-per-sample v[32].
__device__ __forceinline__ void mf_unpack_c(f32x16_t (&c)[2], float* v) {
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    float a = c[0][i], b = c[1][i];
    swap32(a, b);
    v[8 * (i >> 2) + (i & 3)] = a;
    v[8 * (i >> 2) + 4 + (i & 3)] = b;
  }
}

// Forward for one sample tile nt: layer-1 tiles from the X operands, ReLU,
// layer 2 into c2.  With HT != nullptr (training) the ReLU outputs are also
// written to relu(H)^T in LDS and the ReLU pattern returned in msk (bit i of
// msk[mt] = pre-activation register i of hidden tile mt > 0).
template <int O>
__device__ __forceinline__ void mf_fwd_tile(const MfmaDecoder<O>& W, const bf16x8_t (&xb)[2][2], int nt,
                                            f32x16_t& c2, __bf16* HT, uint32_t* msk) {
  const int lane = threadIdx.x & 63, r = lane & 31, q = lane >> 5;
  bf16x8_t hb[4];
#pragma unroll
  for (int mt = 0; mt < 2; ++mt) {
    f32x16_t acc = zero16();
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) acc = mfma_bf16(W.w0a[mt][ks], xb[ks][nt], acc);
    uint32_t bits = 0;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const float v = fmaxf(acc[i], 0.f);
      bits |= uint32_t(acc[i] > 0.f) << i;
      const __bf16 hv = (__bf16)v;
      hb[2 * mt + (i >> 3)][i & 7] = hv;
      if (HT) HT[(32 * mt + 8 * (i >> 2) + 4 * q + (i & 3)) * kMfRow + 32 * nt + r] = hv;
    }
    if (msk) msk[mt] = bits;
  }
  c2 = zero16();
#pragma unroll
  for (int ks = 0; ks < 4; ++ks) c2 = mfma_bf16(W.w1a[ks], hb[ks], c2);
}

// Lane-per-sample features (+ the constant-1 bias column) -> X operands.
__device__ __forceinline__ void mf_pack_x(const float* f, bf16x8_t (&xb)[2][2]) {
  float x[32];
#pragma unroll
  for (int k = 0; k < 32; ++k) x[k] = k < kMfIn ? f[k] : (k == kMfIn ? 1.f : 0.f);
  mf_pack_b(x, xb);
}

// Forward: f [18] per lane -> out [32] per lane (O used).
template <int O>
__device__ __forceinline__ void mfma_decoder_fwd(const MfmaDecoder<O>& W, const float* f, float* out) {
  bf16x8_t xb[2][2];
  mf_pack_x(f, xb);
  f32x16_t c2[2];
#pragma unroll
  for (int nt = 0; nt < 2; ++nt) mf_fwd_tile(W, xb, nt, c2[nt], nullptr, nullptr);
  mf_unpack_c(c2, out);
}

// Per-wave weight-gradient accumulators (C layout):
//   gw1[nt]: dW1[o = 8(i/4) + 4q + i%4][hid = 32 nt + r]
//   gw0[mt]: dW0a[hid = 32 mt + 8(i/4) + 4q + i%4][k = r]  (k = 18: db0)
template <int O>
struct MfmaDecoderAcc {
  f32x16_t gw1[2];
  f32x16_t gw0[2];
  __device__ __forceinline__ void zero() {
    gw1[0] = gw1[1] = gw0[0] = gw0[1] = zero16();
  }
  // Partial row in the decoder-gradient layout dW0 [64][18] | db0 [64] | dW1 [O][64].
  __device__ __forceinline__ void store(float* row) const {
    const int lane = threadIdx.x & 63, r = lane & 31, q = lane >> 5;
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const int a = 8 * (i >> 2) + 4 * q + (i & 3);
        if (a < O) row[kMfH * kMfIn + kMfH + a * kMfH + 32 * t + r] = gw1[t][i];
        const int h = 32 * t + a;
        if (r < kMfIn) row[h * kMfIn + r] = gw0[t][i];
        else if (r == kMfIn) row[kMfH * kMfIn + h] = gw0[t][i];
      }
  }
};


// Training forward: as mfma_decoder_fwd, and also leaves relu(H)^T and X^T in
// LDS for the weight gradients and the ReLU pattern in msk[mt][nt].
template <int O>
__device__ __forceinline__ void mfma_decoder_fwd_train(const MfmaDecoder<O>& W, const float* f, float* out,
                                                       uint32_t (&msk)[2][2], uint32_t* lds) {
  const int lane = threadIdx.x & 63;
  __bf16* DT = reinterpret_cast<__bf16*>(lds);
  __bf16* XT = DT + 32 * kMfRow;
  __bf16* HT = XT + 32 * kMfRow;
#pragma unroll
  for (int k = 0; k < 32; ++k) XT[k * kMfRow + lane] = (__bf16)(k < kMfIn ? f[k] : (k == kMfIn ? 1.f : 0.f));
  bf16x8_t xb[2][2];
  mf_pack_x(f, xb);
  f32x16_t c2[2];
#pragma unroll
  for (int nt = 0; nt < 2; ++nt) {
    uint32_t m2[2];
    mf_fwd_tile(W, xb, nt, c2[nt], HT, m2);
    msk[0][nt] = m2[0];
    msk[1][nt] = m2[1];
  }
  mf_unpack_c(c2, out);
}

// Backward of mfma_decoder_fwd_train: df [18] per lane from dout [32] (0 for
// inactive lanes and o >= O): dH^T = W1^T . dout^T, ReLU mask, dX^T =
// W0a^T . dH^T; then the wave's weight gradients from the LDS operands
// (K = samples) are added to acc.  Ends with the LDS area free.  For blocks
// of one wave: the LDS hand-offs are wave fences, so outstanding atomics of
// an earlier scatter are not waited for here.
template <int O>
__device__ __forceinline__ void mfma_decoder_bwd(const MfmaDecoderBwd<O>& WB, const uint32_t (&msk)[2][2],
                                                 const float* dout, float* df, MfmaDecoderAcc<O>& acc,
                                                 uint32_t* lds) {
  const int lane = threadIdx.x & 63, r = lane & 31, q = lane >> 5;
  __bf16* DT = reinterpret_cast<__bf16*>(lds);
  __bf16* XT = DT + 32 * kMfRow;
  __bf16* HT = XT + 32 * kMfRow;
  __bf16* GT = HT + 64 * kMfRow;
#pragma unroll
  for (int o = 0; o < 32; ++o) DT[o * kMfRow + lane] = (__bf16)dout[o];
  bf16x8_t db[2][2];
  mf_pack_b(dout, db);
  f32x16_t dx[2];
#pragma unroll
  for (int nt = 0; nt < 2; ++nt) {
    f32x16_t dh[2];
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
      f32x16_t a = zero16();
#pragma unroll
      for (int ks = 0; ks < 2; ++ks) a = mfma_bf16(WB.w1t[mt][ks], db[ks][nt], a);
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        a[i] = ((msk[mt][nt] >> i) & 1) ? a[i] : 0.f;
        const int h = 32 * mt + 8 * (i >> 2) + 4 * q + (i & 3);
        GT[h * kMfRow + 32 * nt + r] = (__bf16)a[i];
      }
      dh[mt] = a;
    }
    f32x16_t a = zero16();
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
      bf16x8_t b;
#pragma unroll
      for (int j = 0; j < 8; ++j) b[j] = (__bf16)dh[ks >> 1][(ks & 1) * 8 + j];
      a = mfma_bf16(WB.w0t[ks], b, a);
    }
    dx[nt] = a;
  }
  float dxs[32];
  mf_unpack_c(dx, dxs);
#pragma unroll
  for (int k = 0; k < kMfIn; ++k) df[k] = dxs[k];
  wave_lds_fence();
  constexpr int kRow8 = kMfRow / 8;  // bf16x8 units per row
  const bf16x8_t* DT8 = reinterpret_cast<const bf16x8_t*>(DT);
  const bf16x8_t* XT8 = reinterpret_cast<const bf16x8_t*>(XT);
  const bf16x8_t* HT8 = reinterpret_cast<const bf16x8_t*>(HT);
  const bf16x8_t* GT8 = reinterpret_cast<const bf16x8_t*>(GT);
#pragma unroll
  for (int ks = 0; ks < 4; ++ks) {
    const int col = 2 * ks + q;  // samples 16 ks + 8 q .. + 7
    const bf16x8_t a_d = DT8[r * kRow8 + col];
    const bf16x8_t b_x = XT8[r * kRow8 + col];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      acc.gw1[t] = mfma_bf16(a_d, HT8[(32 * t + r) * kRow8 + col], acc.gw1[t]);
      acc.gw0[t] = mfma_bf16(GT8[(32 * t + r) * kRow8 + col], b_x, acc.gw0[t]);
    }
  }
  wave_lds_fence();
}

}  // namespace ff
