// Dictionary image-set field: image-set mode (models/FactorFields.py:286-288,
// :469-470) with a shared 2-D basis of 144 channels (basis_dims
// [32,32,32,16,16,16]), a per-image 3-D coefficient grid [n_images, h, w, 144]
// (get_coeff, :425-465) and a 144 -> 64 -> O decoder (MLPMixer, :113-159), in
// two forms: BF16 tables and BF16 decoder operands with FP32 accumulation
// (BF = true), or FP32 tables and FP32 decoder operands (BF = false).
//
// The FP32 form runs the same products on the FP32 matrix cores: each BF16
// instruction, whose lanes hold 8 consecutive K values, becomes 8 FP32
// instructions (v_mfma_f32_16x16x4_f32 for v_mfma_f32_16x16x32_bf16,
// v_mfma_f32_32x32x2_f32 for v_mfma_f32_32x32x16_bf16) that take element j of
// every lane's 8 values in step j -- the same operand registers, 8 floats
// instead of 8 bf16 per lane.  Its layer-1 and data-gradient operands (77 KB
// in FP32) do not fit in LDS beside the per-wave areas, so the backward reads
// them from a packed copy in the workspace (L2-resident) and the forward
// keeps only the layer-1 operands in LDS.
//
// Decomposition: a wave handles 16 pixels with 4 lanes per pixel; lane
// (pixel p, group g) owns the 4-channel pieces 16·mt + 4g .. +3 (mt = 0..8).
// That ownership is the accumulator layout of v_mfma_f32_16x16x32_bf16
// (lane l holds rows 4(l/16) .. +3 of column l%16), and, with the K order of
// each 32-wide step permuted to slot (ks, g, j) -> 16(2ks + j/4) + 4g + j%4,
// also its B-operand layout: the features, the hidden activations and the
// gradients feed the next product straight from registers.
//   forward:  H^T = W0a · X^T (W0a = [W0 | b0], X with a constant-1 column),
//             out^T = W1 · relu(H^T)
//   backward: dH^T = W1^T · dout^T (ReLU mask), dX^T = W0a^T · dH^T
// Decoder weight gradients take K = pixels: X, relu(H), dH and dout are
// transposed through LDS and accumulated by v_mfma_f32_32x32x16_bf16 in the
// accumulation registers across the wave's tiles; each wave writes one partial
// row, reduced afterwards in a fixed order.  Table gradients are staged in LDS
// and scattered so that every atomic instruction carries four full 64-byte
// lines (16 lanes on 16 consecutive channels of one texel); consecutive pixels
// on the same texel and corner are summed first and share one atomic.
#include <type_traits>

#include "ff_field_impl.h"

namespace ff {
namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 dbf8;
typedef __attribute__((__vector_size__(8 * sizeof(float)))) float df8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float df4;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float df16;

// Operand element and vector types of the two forms.
template <bool BF> struct DType;
template <> struct DType<true> {
  using E = __bf16;
  using V = dbf8;
  static constexpr int kRow = 24;  // LDS row of 16 pixels + 8 padding (16-byte aligned reads)
};
template <> struct DType<false> {
  using E = float;
  using V = df8;
  static constexpr int kRow = 20;  // 16 pixels + 4 padding
};
template <bool BF> using DE = typename DType<BF>::E;
template <bool BF> using DV = typename DType<BF>::V;
template <bool BF> __device__ __forceinline__ DE<BF> dcvt(float f) { return (DE<BF>)f; }

constexpr int kDC = 144;  // channels
constexpr int kDH = 64;   // hidden units
constexpr int kDPx = 16;  // pixels per wave tile
constexpr int kDMt = 9;   // 16-channel tiles of features (tile 9 carries the bias column)
constexpr int kDWaves = 4;
// Blocks per CU of the forward kernel (its waves per SIMD): a register bound
// the compiler meets with some spilling (232 B per lane for BF16).
constexpr int kDFwdBlocksPerCu = 2;
constexpr int kDOMax = 4;

__device__ __forceinline__ int dlevel(int mt) { return mt < 2 ? 0 : (mt < 4 ? 1 : (mt < 6 ? 2 : mt - 3)); }
__device__ __forceinline__ int doff(int L) { return L < 3 ? 32 * L : 96 + 16 * (L - 3); }
__device__ __forceinline__ int ddim(int L) { return L < 3 ? 32 : 16; }

__device__ __forceinline__ df4 mfma16(const dbf8& a, const dbf8& b, const df4& c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ df16 mfma32(const dbf8& a, const dbf8& b, const df16& c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}
// FP32: lane element j is K index (lane / 16) of step j (16x16x4) or (lane /
// 32) of step j (32x32x2); over the 8 steps the K order is the BF16 one.
__device__ __forceinline__ df4 mfma16(const df8& a, const df8& b, df4 c) {
#pragma unroll
  for (int j = 0; j < 8; ++j) c = __builtin_amdgcn_mfma_f32_16x16x4f32(a[j], b[j], c, 0, 0, 0);
  return c;
}
__device__ __forceinline__ df16 mfma32(const df8& a, const df8& b, df16 c) {
#pragma unroll
  for (int j = 0; j < 8; ++j) c = __builtin_amdgcn_mfma_f32_32x32x2f32(a[j], b[j], c, 0, 0, 0);
  return c;
}

// Orders this wave's LDS accesses across a hand-off between lanes: the wave's
// LDS operations execute in order, so waiting for its outstanding LDS traffic
// and keeping the compiler from moving memory accesses across is enough (the
// waves of a block run different numbers of tiles, so no block barrier).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0)
  __builtin_amdgcn_wave_barrier();
  asm volatile("" ::: "memory");
}

// K slot (ks, g, j) of a 32-wide step -> index.
__device__ __forceinline__ int dperm(int ks, int g, int j) { return 16 * (2 * ks + (j >> 2)) + 4 * g + (j & 3); }

// Weight loads are unconditional, at indices clamped into the matrix, with the
// zero padding selected afterwards, so that they are not serialised behind
// per-lane branches.
__device__ __forceinline__ float w0a(const ff_mlp_desc& m, int h, int k) {
  const float w = m.w0[h * kDC + (k < kDC ? k : kDC - 1)];
  const float b = m.b0[h];
  return k < kDC ? w : (k == kDC ? b : 0.f);
}

// Block-shared A operands (BF16): layer 1 [hidden tile][K step][lane] and the
// data gradient [channel tile][K step][lane].
struct DictW {
  dbf8 a1[4][5][64];
  dbf8 a4[kDMt][2][64];
};
// FP32 forward: layer-1 operands only (the data gradient is not needed).
struct DictW1F {
  df8 a1[4][5][64];
};
constexpr int kDOpA1 = 4 * 5 * 64, kDOpA4 = kDMt * 2 * 64;  // operand vectors per matrix
constexpr size_t kDPackBytes = (size_t)(kDOpA1 + kDOpA4) * sizeof(df8);

// Operand sources: a1(mt, ks, lane), a4(mt, ks, lane).
struct DictWLds {  // BF16, block-shared LDS
  const DictW& w;
  __device__ __forceinline__ dbf8 a1(int mt, int ks, int l) const { return w.a1[mt][ks][l]; }
  __device__ __forceinline__ dbf8 a4(int mt, int ks, int l) const { return w.a4[mt][ks][l]; }
};
struct DictW1Lds {  // FP32 forward: layer 1 from LDS
  const DictW1F& w;
  __device__ __forceinline__ df8 a1(int mt, int ks, int l) const { return w.a1[mt][ks][l]; }
};
struct DictWPack {  // FP32 backward: the packed copy in the workspace
  const df8* __restrict__ p;
  __device__ __forceinline__ df8 a1(int mt, int ks, int l) const { return p[(mt * 5 + ks) * 64 + l]; }
  __device__ __forceinline__ df8 a4(int mt, int ks, int l) const { return p[kDOpA1 + (mt * 2 + ks) * 64 + l]; }
};

// Per-wave LDS area: weight-gradient operands (phase A, bf16 rows of 16 pixels
// + 8 padding), then, aliased, the scatter staging (phase B).
template <bool BF>
struct DictPhaseA {
  static constexpr int R = DType<BF>::kRow;
  DE<BF> xt[160][R];
  DE<BF> gt[64][R];
  DE<BF> ht[64][R];
  DE<BF> dt[32][R];
};
constexpr int kDVal = kDC + 1;  // staging row (odd stride: conflict-free column writes)
struct DictPhaseB {
  float db[kDPx][kDVal];
  float dc[kDPx][kDVal];
};
template <bool BF>
union DictWave {
  DictPhaseA<BF> a;
  DictPhaseB b;
};

// Eight consecutive operand values (pixels) of an LDS row.
__device__ __forceinline__ dbf8 lds_v8(const __bf16* p) { return *reinterpret_cast<const dbf8*>(p); }
__device__ __forceinline__ df8 lds_v8(const float* p) {
  const float4 a = reinterpret_cast<const float4*>(p)[0], b = reinterpret_cast<const float4*>(p)[1];
  df8 v;
  v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
  v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
  return v;
}
// Corners of the wave's pixels, written during the gather and read by the
// scatter (not aliased with the areas above).
struct __attribute__((aligned(16))) DictCorners {
  int2 bc[6][kDPx][4];  // basis corners: (texel, weight bits)
  int2 cc[kDPx][8];     // coefficient corners
  int bhead[6][64];     // run heads per basis level
  int chead[2 * 64 + 4];  // run heads of the coefficient (corners 0-3, then 4-7), padded for pair reads
};

// Run heads of one table's (pixel, corner) pairs for corners kofs .. kofs+3,
// lane = 4·pixel + corner, given the lane's pair (t, w) and the previous
// pixel's pair on the same corner (tp, wp).  A pair is live when its weight is
// nonzero; a live pair whose predecessor is live on the same texel continues
// that run.  Each run's first pair is appended to `head` at `base` as
// pixel | corner << 4 | length << 8; returns the number of heads appended.
// A batch ordered by image and coefficient cell puts pixels that share texels
// next to each other, so one atomic then carries a whole run's sum; in any
// other order every live pair is its own run.
__device__ __forceinline__ int dict_runs(int t, float w, int tp, float wp, int kofs, int* head, int base) {
  constexpr uint64_t kEvery4 = 0x1111111111111111ull;
  const int lane = threadIdx.x & 63, px = lane >> 2, k = lane & 3;
  const bool live = w != 0.f;
  const bool member = live && px > 0 && wp != 0.f && tp == t;
  const bool is_head = live && !member;
  const uint64_t M = __ballot(member);
  const uint64_t H = __ballot(is_head);
  // members following this pair on the same corner: bits 4j + k of M, j > px
  const uint64_t after = px < kDPx - 1 ? ((M >> k) & kEvery4) >> (4 * (px + 1)) : 0ull;
  const int len = 1 + __builtin_ctzll(~after & kEvery4) / 4;
  const int pos = __builtin_amdgcn_mbcnt_hi((uint32_t)(H >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)H, 0u));
  if (is_head) head[base + pos] = px | ((k + kofs) << 4) | (len << 8);
  return __popcll(H);
}

// Operand vector e of [a1 (kDOpA1) | a4 (kDOpA4)], in a1's / a4's
// [mt][ks][lane] order: layer 1's A (W0a rows 16 mt + r, K slots of step ks)
// and the data gradient's A (W0a^T).
template <class V>
__device__ __forceinline__ V dict_operand(const ff_mlp_desc& m, int e) {
  V v;
  if (e < kDOpA1) {
    const int l = e & 63, ks = (e >> 6) % 5, mt = e / 320;
    const int r = l & 15, g = l >> 4;
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = (std::remove_reference_t<decltype(v[0])>)w0a(m, 16 * mt + r, dperm(ks, g, j));
  } else {
    const int e2 = e - kDOpA1;
    const int l = e2 & 63, ks = (e2 >> 6) & 1, mt = e2 >> 7;
    const int r = l & 15, g = l >> 4;
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = (std::remove_reference_t<decltype(v[0])>)w0a(m, dperm(ks, g, j), 16 * mt + r);
  }
  return v;
}

__device__ __forceinline__ void load_dict_weights(const ff_mlp_desc& m, DictW& W) {
  for (int e = threadIdx.x; e < kDOpA1 + kDOpA4; e += blockDim.x) {
    const dbf8 v = dict_operand<dbf8>(m, e);
    if (e < kDOpA1) (&W.a1[0][0][0])[e] = v;
    else (&W.a4[0][0][0])[e - kDOpA1] = v;
  }
}

__device__ __forceinline__ void load_dict_w1f(const ff_mlp_desc& m, DictW1F& W) {
  for (int e = threadIdx.x; e < kDOpA1; e += blockDim.x) (&W.a1[0][0][0])[e] = dict_operand<df8>(m, e);
}

// FP32 backward: the packed operand copy, written once per launch.
__global__ __launch_bounds__(256) void dict_pack_f32_kernel(ff_mlp_desc m, df8* __restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < kDOpA1 + kDOpA4) out[e] = dict_operand<df8>(m, e);
}

// Gathered values of one pixel's pieces (lane-owned channels).
struct DictPix {
  float bas[kDMt][4], cof[kDMt][4];
};

template <bool BF>
__device__ __forceinline__ void load4(const void* t, int64_t e, float* v) { load_ch<4, BF>(t, e, v); }

// Basis and coefficient values of the lane's pieces at pixel xv; the g = 0
// lane of each pixel also records the corners in K (when K is given).
template <bool BF>
__device__ __forceinline__ void dict_gather(const ff_field_desc& fd, const float* xv, int g, int px, DictPix& P,
                                            DictCorners* K) {
#pragma unroll
  for (int L = 0; L < 6; ++L) {
    const int R = fd.basis_reso[L];
    float pos[2];
    const int Rs[2] = {R, R};
#pragma unroll
    for (int a = 0; a < 2; ++a) pos[a] = unnorm_ac(map_coord(fd.mapping, xv[a], fd.map_origin[a], fd.map_scale[L]), R);
    Corners<2> bc;
    linear_corners<2>(pos, Rs, bc);
    if (K && g == 0) {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        K->bc[L][px][k] = make_int2(bc.tex[k], __float_as_int(bc.w[k]));
      }
    }
    const int dl = ddim(L);
#pragma unroll
    for (int mt = 0; mt < kDMt; ++mt) {
      if (dlevel(mt) != L) continue;
      const int base = 16 * mt - doff(L) + 4 * g;
      float b[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        float v[4];
        load4<BF>(fd.basis[L], (int64_t)bc.tex[k] * dl + base, v);
#pragma unroll
        for (int i = 0; i < 4; ++i) b[i] = fmaf(v[i], bc.w[k], b[i]);
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) P.bas[mt][i] = b[i];
    }
  }
  float pts[3], pos3[3];
  coeff_pts<3>(fd, xv, pts);
#pragma unroll
  for (int a = 0; a < 3; ++a) pos3[a] = unnorm_border(pts[a], fd.coeff_reso[a]);
  Corners<3> cc;
  linear_corners<3>(pos3, fd.coeff_reso, cc);
  if (K && g == 0) {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      K->cc[px][k] = make_int2(cc.tex[k], __float_as_int(cc.w[k]));
    }
  }
#pragma unroll
  for (int mt = 0; mt < kDMt; ++mt) {
    float c[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      if (cc.w[k] != 0.f) {
        float v[4];
        load4<BF>(fd.coeff[0], (int64_t)cc.tex[k] * kDC + 16 * mt + 4 * g, v);
#pragma unroll
        for (int i = 0; i < 4; ++i) c[i] = fmaf(v[i], cc.w[k], c[i]);
      }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) P.cof[mt][i] = c[i];
  }
}

// Decoder forward for the wave's 16 pixels: returns the layer-1 accumulators
// c1[hidden tile] and the output accumulator c2 (lane (p, g): outputs 4g..4g+3).
template <bool BF, class WR>
__device__ __forceinline__ void dict_decoder_fwd(const WR& W, const DV<BF> (&a2)[2], const DictPix& P, int g,
                                                 df4 (&c1)[4], df4& c2) {
  const int lane = threadIdx.x & 63;
  DV<BF> xb[5];
#pragma unroll
  for (int ks = 0; ks < 5; ++ks)
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int mt = 2 * ks + (j >> 2), i = j & 3;
      float f;
      if (mt < kDMt) f = P.bas[mt][i] * P.cof[mt][i];
      else f = (g == 0 && i == 0) ? 1.f : 0.f;  // column 144: the bias
      xb[ks][j] = dcvt<BF>(f);
    }
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
    df4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < 5; ++ks) acc = mfma16(W.a1(mt, ks, lane), xb[ks], acc);
    c1[mt] = acc;
  }
  df4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int ks = 0; ks < 2; ++ks) {
    DV<BF> hb;
#pragma unroll
    for (int j = 0; j < 8; ++j) hb[j] = dcvt<BF>(fmaxf(c1[2 * ks + (j >> 2)][j & 3], 0.f));
    acc = mfma16(a2[ks], hb, acc);
  }
  c2 = acc;
}

template <bool BF>
__device__ __forceinline__ void load_a2(const ff_mlp_desc& m, DV<BF> (&a2)[2]) {
  const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
#pragma unroll
  for (int ks = 0; ks < 2; ++ks)
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const float v = m.w1[(r < m.out_dim ? r : m.out_dim - 1) * kDH + dperm(ks, g, j)];
      a2[ks][j] = dcvt<BF>(r < m.out_dim ? v : 0.f);
    }
}

template <bool BF>
__device__ __forceinline__ void dict_fwd_body(const ff_field_desc& fd, const ff_mlp_desc& m,
                                              const float* __restrict__ x, int64_t n, float* __restrict__ out) {
  __shared__ std::conditional_t<BF, DictW, DictW1F> W;
  if constexpr (BF)
    load_dict_weights(m, W);
  else
    load_dict_w1f(m, W);
  __syncthreads();
  const std::conditional_t<BF, DictWLds, DictW1Lds> WR{W};
  DV<BF> a2[2];
  load_a2<BF>(m, a2);
  const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
  const int O = m.out_dim;
  const int64_t ntiles = (n + kDPx - 1) / kDPx;
  const int64_t wave = (int64_t)blockIdx.x * kDWaves + (threadIdx.x >> 6);
  for (int64_t tile = wave; tile < ntiles; tile += (int64_t)gridDim.x * kDWaves) {
    const int64_t p = tile * kDPx + r;
    const bool active = p < n;
    const int64_t pc = active ? p : n - 1;
    const float xv[3] = {x[pc * 3 + 0], x[pc * 3 + 1], x[pc * 3 + 2]};
    DictPix P;
    dict_gather<BF>(fd, xv, g, r, P, nullptr);
    df4 c1[4], c2;
    dict_decoder_fwd<BF>(WR, a2, P, g, c1, c2);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int o = 4 * g + i;
      if (active && o < O) out[p * O + o] = c2[i];
    }
  }
}

// The forward kernels: BF16 tables at kDFwdBlocksPerCu blocks per CU; the
// FP32 form keeps one block per CU (its working set would spill 888 B per
// lane at two).
__global__ __launch_bounds__(256, kDFwdBlocksPerCu) void dict_fwd_kernel_bf16(ff_field_desc fd, ff_mlp_desc m,
                                                                          const float* __restrict__ x,
                                                                          int64_t n, float* __restrict__ out) {
  dict_fwd_body<true>(fd, m, x, n, out);
}

__global__ __launch_bounds__(256) void dict_fwd_kernel_f32(ff_field_desc fd, ff_mlp_desc m,
                                                           const float* __restrict__ x, int64_t n,
                                                           float* __restrict__ out) {
  dict_fwd_body<false>(fd, m, x, n, out);
}

// Fused backward (mode 0: d_out given in dsrc) or regression training pass
// (mode 1: MSE, mode 2: L1 against dsrc).
template <bool BF>
__global__ __launch_bounds__(256) void dict_bwd_kernel(ff_field_desc fd, ff_mlp_desc m, ff_grads gin,
                                                       const float* __restrict__ x, int64_t n,
                                                       const float* __restrict__ dsrc, int mode,
                                                       float loss_scale, float* __restrict__ ws,
                                                       float* __restrict__ loss_sum, const df8* __restrict__ pack) {
  // BF16: both operand matrices in LDS; FP32: the packed copy (the LDS area
  // below is then the weights' size smaller)
  __shared__ std::conditional_t<BF, DictW, char> W;
  __shared__ __attribute__((aligned(16))) DictWave<BF> WV[kDWaves];
  __shared__ __attribute__((aligned(16))) DictCorners WK[kDWaves];
  if constexpr (BF) {
    load_dict_weights(m, W);
    __syncthreads();
  }
  const auto WR = [&] {
    if constexpr (BF) return DictWLds{W};
    else return DictWPack{pack};
  }();
  const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4, l32 = lane & 31, h32 = lane >> 5;
  const int wib = threadIdx.x >> 6;
  DictWave<BF>& S = WV[wib];
  DictCorners& K = WK[wib];
  const int O = m.out_dim;
  DV<BF> a2[2], a3[4];
  load_a2<BF>(m, a2);
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int o = 8 * g + j;
      const float v = m.w1[(o < O ? o : O - 1) * kDH + 16 * mt + r];
      a3[mt][j] = dcvt<BF>(o < O ? v : 0.f);
    }
  df16 acc0[5][2], acc1[2];
#pragma unroll
  for (int a = 0; a < 5; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int i = 0; i < 16; ++i) acc0[a][b][i] = 0.f;
#pragma unroll
  for (int b = 0; b < 2; ++b)
#pragma unroll
    for (int i = 0; i < 16; ++i) acc1[b][i] = 0.f;
  const ff_grads g2 = local_grads(gin);
  const float grad_scale = (float)(1.0 / ((double)n * O)) * loss_scale;
  float lsum = 0.f;
  const int64_t ntiles = (n + kDPx - 1) / kDPx;
  const int64_t wave = (int64_t)blockIdx.x * kDWaves + wib;
  const int64_t nwaves = (int64_t)gridDim.x * kDWaves;
  for (int64_t tile = wave; tile < ntiles; tile += nwaves) {
    const int64_t p = tile * kDPx + r;
    const bool active = p < n;
    const int64_t pc = active ? p : n - 1;
    const float xv[3] = {x[pc * 3 + 0], x[pc * 3 + 1], x[pc * 3 + 2]};
    DictPix P;
    dict_gather<BF>(fd, xv, g, r, P, &K);
    df4 c1[4], c2;
    dict_decoder_fwd<BF>(WR, a2, P, g, c1, c2);
    // output gradient (lane (p, g) holds outputs 4g .. 4g+3)
    float d[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int o = 4 * g + i;
      float v = 0.f;
      if (active && o < O) {
        if (mode == 0) {
          v = dsrc[pc * O + o];
        } else {
          const float e = c2[i] - dsrc[pc * O + o];
          if (mode == 1) {
            v = 2.f * e * grad_scale;
            lsum += e * e;
          } else {
            v = (e > 0.f ? 1.f : (e < 0.f ? -1.f : 0.f)) * grad_scale;
            lsum += fabsf(e);
          }
        }
      }
      d[i] = v;
    }
    // dH^T = W1^T . dout^T (outputs 0..3 sit in the g = 0 lanes), ReLU mask
    DV<BF> bd;
#pragma unroll
    for (int j = 0; j < 8; ++j) bd[j] = dcvt<BF>((g == 0 && j < 4) ? d[j] : 0.f);
    df4 dh[4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      df4 z = {0.f, 0.f, 0.f, 0.f};
      df4 t = mfma16(a3[mt], bd, z);
#pragma unroll
      for (int i = 0; i < 4; ++i) t[i] = c1[mt][i] > 0.f ? t[i] : 0.f;
      dh[mt] = t;
    }
    DV<BF> bdh[2];
#pragma unroll
    for (int ks = 0; ks < 2; ++ks)
#pragma unroll
      for (int j = 0; j < 8; ++j) bdh[ks][j] = dcvt<BF>(dh[2 * ks + (j >> 2)][j & 3]);
    // weight-gradient operands through LDS (phase A)
#pragma unroll
    for (int mt = 0; mt < 10; ++mt)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        float f;
        if (mt < kDMt) f = P.bas[mt][i] * P.cof[mt][i];
        else f = (g == 0 && i == 0) ? 1.f : 0.f;
        S.a.xt[16 * mt + 4 * g + i][r] = dcvt<BF>(active ? f : 0.f);
      }
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        S.a.ht[16 * mt + 4 * g + i][r] = dcvt<BF>(fmaxf(c1[mt][i], 0.f));
        S.a.gt[16 * mt + 4 * g + i][r] = dcvt<BF>(dh[mt][i]);
      }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      S.a.dt[4 * g + i][r] = dcvt<BF>(g == 0 ? d[i] : 0.f);
      S.a.dt[16 + 4 * g + i][r] = dcvt<BF>(0.f);
    }
    wave_lds_sync();
    {
      const DV<BF> bdt = lds_v8(&S.a.dt[l32][8 * h32]);
#pragma unroll
      for (int nh = 0; nh < 2; ++nh) {
        const DV<BF> bg = lds_v8(&S.a.gt[32 * nh + l32][8 * h32]);
#pragma unroll
        for (int mk = 0; mk < 5; ++mk) {
          const DV<BF> ax = lds_v8(&S.a.xt[32 * mk + l32][8 * h32]);
          acc0[mk][nh] = mfma32(ax, bg, acc0[mk][nh]);
        }
        const DV<BF> ah = lds_v8(&S.a.ht[32 * nh + l32][8 * h32]);
        acc1[nh] = mfma32(ah, bdt, acc1[nh]);
      }
    }
    // data gradient dX^T = W0a^T . dH^T: lane (p, g) receives its own pieces
    float df[kDMt][4];
#pragma unroll
    for (int mt = 0; mt < kDMt; ++mt) {
      df4 t = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int ks = 0; ks < 2; ++ks) t = mfma16(WR.a4(mt, ks, lane), bdh[ks], t);
#pragma unroll
      for (int i = 0; i < 4; ++i) df[mt][i] = t[i];
    }
    wave_lds_sync();
    // phase B: stage d(basis), d(coeff) and the corners, then line-wide atomics
#pragma unroll
    for (int mt = 0; mt < kDMt; ++mt)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int c = 16 * mt + 4 * g + i;
        S.b.db[r][c] = active ? df[mt][i] * P.cof[mt][i] : 0.f;
        S.b.dc[r][c] = active ? df[mt][i] * P.bas[mt][i] : 0.f;
      }
    wave_lds_sync();
    // run heads of all seven tables in one pass (lane = 4 * pixel + corner)
    int nbh[6], nch;
    {
      const int px = lane >> 2, k = lane & 3, pq = px > 0 ? px - 1 : 0;
      int2 bp[6], bq[6];
#pragma unroll
      for (int L = 0; L < 6; ++L) {
        bp[L] = K.bc[L][px][k];
        bq[L] = K.bc[L][pq][k];
      }
      const int2 c0 = K.cc[px][k], c1 = K.cc[px][k + 4], q0 = K.cc[pq][k], q1 = K.cc[pq][k + 4];
#pragma unroll
      for (int L = 0; L < 6; ++L)
        nbh[L] = dict_runs(bp[L].x, __int_as_float(bp[L].y), bq[L].x, __int_as_float(bq[L].y), 0, K.bhead[L], 0);
      const int n0 = dict_runs(c0.x, __int_as_float(c0.y), q0.x, __int_as_float(q0.y), 0, K.chead, 0);
      nch = n0 + dict_runs(c1.x, __int_as_float(c1.y), q1.x, __int_as_float(q1.y), 4, K.chead, n0);
    }
    wave_lds_sync();
    // scatter: lane group g takes items 16p + 4g .. +3 (item = run head x
    // 16-channel tile) in pass p; the heads come in one vector read and every
    // LDS read of a pass is issued before its first use
#pragma unroll
    for (int L = 0; L < 6; ++L) {
      const int dl = ddim(L), nb = dl / 16, items = nbh[L] * nb;
      float* dst = g2.basis[L];
      for (int i0 = 4 * g; i0 < items; i0 += 16) {
        int e[4];
        if (nb == 2) {
          const int2 h = *reinterpret_cast<const int2*>(&K.bhead[L][i0 >> 1]);
          e[0] = h.x; e[1] = h.x; e[2] = h.y; e[3] = h.y;
        } else {
          const int4 h = *reinterpret_cast<const int4*>(&K.bhead[L][i0]);
          e[0] = h.x; e[1] = h.y; e[2] = h.z; e[3] = h.w;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) e[u] = i0 + u < items ? e[u] : 0;
        int2 cp[4];
        float dv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int px = e[u] & 15, k = (e[u] >> 4) & 7, blk = nb == 2 ? u & 1 : 0;
          cp[u] = K.bc[L][px][k];
          dv[u] = S.b.db[px][doff(L) + 16 * blk + r];
        }
        float v[4];
        int lmax = 0;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          v[u] = __int_as_float(cp[u].y) * dv[u];
          lmax = max(lmax, e[u] >> 8);
        }
        // the rest of the four runs in lockstep (one LDS round trip per step)
        for (int j = 1; j < lmax; ++j) {
          float wq[4], dq[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int px = e[u] & 15, k = (e[u] >> 4) & 7, q = j < (e[u] >> 8) ? px + j : px;
            wq[u] = __int_as_float(K.bc[L][q][k].y);
            dq[u] = S.b.db[q][doff(L) + 16 * (nb == 2 ? u & 1 : 0) + r];
          }
#pragma unroll
          for (int u = 0; u < 4; ++u)
            if (j < (e[u] >> 8)) v[u] = fmaf(wq[u], dq[u], v[u]);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
          if (i0 + u < items && v[u] != 0.f)
            unsafeAtomicAdd(dst + (int64_t)cp[u].x * dl + 16 * (nb == 2 ? u & 1 : 0) + r, v[u]);
      }
    }
    {
      float* dst = g2.coeff[0];
      const int items = nch * kDMt;
      for (int i0 = 4 * g; i0 < items; i0 += 16) {
        const int h0 = i0 / kDMt;
        const int2 hp = *reinterpret_cast<const int2*>(&K.chead[h0 & ~1]);
        const int ha = (h0 & 1) ? hp.y : hp.x, hb = (h0 & 1) ? K.chead[h0 + 1] : hp.y;
        int e[4], c[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int t = i0 + u - h0 * kDMt;  // 0 .. 11: tile within head h0, or h0 + 1 past 8
          e[u] = i0 + u < items ? (t < kDMt ? ha : hb) : 0;
          c[u] = 16 * (t < kDMt ? t : t - kDMt) + r;
        }
        int2 cp[4];
        float dv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int px = e[u] & 15, k = (e[u] >> 4) & 7;
          cp[u] = K.cc[px][k];
          dv[u] = S.b.dc[px][c[u]];
        }
        float v[4];
        int lmax = 0;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          v[u] = __int_as_float(cp[u].y) * dv[u];
          lmax = max(lmax, e[u] >> 8);
        }
        for (int j = 1; j < lmax; ++j) {
          float wq[4], dq[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int px = e[u] & 15, k = (e[u] >> 4) & 7, q = j < (e[u] >> 8) ? px + j : px;
            wq[u] = __int_as_float(K.cc[q][k].y);
            dq[u] = S.b.dc[q][c[u]];
          }
#pragma unroll
          for (int u = 0; u < 4; ++u)
            if (j < (e[u] >> 8)) v[u] = fmaf(wq[u], dq[u], v[u]);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
          if (i0 + u < items && v[u] != 0.f) unsafeAtomicAdd(dst + (int64_t)cp[u].x * kDC + c[u], v[u]);
      }
    }
    wave_lds_sync();
  }
  // partial row: dW0 [64][144] | db0 [64] | dW1 [O][64]
  float* row = ws + wave * (int64_t)(kDH * kDC + kDH + O * kDH);
#pragma unroll
  for (int mk = 0; mk < 5; ++mk)
#pragma unroll
    for (int nh = 0; nh < 2; ++nh)
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const int k = 32 * mk + 8 * (i >> 2) + 4 * h32 + (i & 3), h = 32 * nh + l32;
        if (k < kDC) row[h * kDC + k] = acc0[mk][nh][i];
        else if (k == kDC) row[kDH * kDC + h] = acc0[mk][nh][i];
      }
#pragma unroll
  for (int mh = 0; mh < 2; ++mh)
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int h = 32 * mh + 8 * (i >> 2) + 4 * h32 + (i & 3), o = l32;
      if (o < O) row[kDH * kDC + kDH + o * kDH + h] = acc1[mh][i];
    }
  if (mode != 0) {
    lsum = wave_sum(lsum);
    if (lane == 0 && loss_sum) unsafeAtomicAdd(loss_sum, lsum);
  }
}

int dict_cu() {
  int dev = 0, cu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 256;
  if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cu <= 0) return 256;
  return cu;
}

// One block of four waves per CU is resident (LDS, registers).
unsigned dict_blocks(int64_t n) {
  const int64_t tiles = (n + kDPx - 1) / kDPx;
  const int64_t need = (tiles + kDWaves - 1) / kDWaves;
  const int64_t cap = dict_cu();
  return (unsigned)(need < cap ? (need > 0 ? need : 1) : cap);
}

// The forward kernel: kDFwdBlocksPerCu blocks per CU.
unsigned dict_fwd_blocks(int64_t n) {
  const int64_t tiles = (n + kDPx - 1) / kDPx;
  const int64_t need = (tiles + kDWaves - 1) / kDWaves;
  const int64_t cap = kDFwdBlocksPerCu * dict_cu();
  return (unsigned)(need < cap ? (need > 0 ? need : 1) : cap);
}

int dict_partial(int O) { return kDH * kDC + kDH + O * kDH; }

}  // namespace

bool dict_supported(const ff_field_desc& d, const ff_mlp_desc& m) {
  static const int dims[6] = {32, 32, 32, 16, 16, 16};
  // BF16 tables: BF16 decoder operands; FP32 tables: FP32 decoder operands
  if (d.table_dtype != FF_BF16 && d.table_dtype != FF_F32) return false;
  if (d.in_dim != 2 || d.x_dim != 3 || d.n_levels != 6 || d.n_channels != kDC) return false;
  if (d.basis_kind != FF_BASIS_GRID || d.basis_interp != FF_INTERP_BILINEAR) return false;
  if (d.coeff_kind != FF_COEFF_GRID || d.coeff_ndim != 3 || d.coeff_interp != FF_INTERP_BILINEAR) return false;
  for (int l = 0; l < 6; ++l)
    if (d.basis_dims[l] != dims[l] || !d.basis[l] || d.basis_reso[l] < 2) return false;
  if (!d.coeff[0]) return false;
  return m.n_layers == 2 && m.in_dim == kDC && m.hidden == kDH && m.out_dim >= 1 && m.out_dim <= kDOMax &&
         m.w0 && m.b0 && m.w1;
}

// Workspace: the partial rows, then (FP32 form) the packed decoder operands,
// 256-byte aligned.
size_t dict_rows_bytes(const ff_mlp_desc& m, int64_t n) {
  return ((size_t)dict_blocks(n) * kDWaves + kReduceSplit) * dict_partial(m.out_dim) * sizeof(float);
}

size_t dict_workspace_bytes(const ff_field_desc& d, const ff_mlp_desc& m, int64_t n) {
  if (n <= 0) return 0;
  const size_t rows = dict_rows_bytes(m, n);
  return d.table_dtype == FF_BF16 ? rows : (rows + 255) / 256 * 256 + kDPackBytes;
}

int launch_dict_fwd(const ff_field_desc& d, const ff_mlp_desc& m, const float* x, int64_t n, float* out,
                    hipStream_t s) {
  if (d.table_dtype == FF_BF16)
    hipLaunchKernelGGL(dict_fwd_kernel_bf16, dim3(dict_fwd_blocks(n)), dim3(64 * kDWaves), 0, s, d, m, x, n, out);
  else
    hipLaunchKernelGGL(dict_fwd_kernel_f32, dim3(dict_blocks(n)), dim3(64 * kDWaves), 0, s, d, m, x, n, out);
  return hipGetLastError() == hipSuccess ? FF_OK : FF_ERR_LAUNCH;
}

int launch_dict_bwd(const ff_field_desc& d, const ff_mlp_desc& m, const ff_grads& g, const float* x, int64_t n,
                    const float* dsrc, int mode, float loss_scale, void* ws, size_t ws_bytes, float* loss_sum,
                    hipStream_t s) {
  if (!ws || ws_bytes < dict_workspace_bytes(d, m, n)) return FF_ERR_WORKSPACE;
  const unsigned blocks = dict_blocks(n);
  if (d.table_dtype == FF_BF16) {
    hipLaunchKernelGGL(dict_bwd_kernel<true>, dim3(blocks), dim3(64 * kDWaves), 0, s, d, m, g, x, n, dsrc, mode,
                       loss_scale, (float*)ws, loss_sum, (const df8*)nullptr);
  } else {
    df8* pack = reinterpret_cast<df8*>(static_cast<char*>(ws) + (dict_rows_bytes(m, n) + 255) / 256 * 256);
    hipLaunchKernelGGL(dict_pack_f32_kernel, dim3((kDOpA1 + kDOpA4 + 255) / 256), dim3(256), 0, s, m, pack);
    hipLaunchKernelGGL(dict_bwd_kernel<false>, dim3(blocks), dim3(64 * kDWaves), 0, s, d, m, g, x, n, dsrc, mode,
                       loss_scale, (float*)ws, loss_sum, (const df8*)pack);
  }
  if (hipGetLastError() != hipSuccess) return FF_ERR_LAUNCH;
  return launch_reduce_partials((const float*)ws, (int)(blocks * kDWaves), 2, kDC, kDH, m.out_dim, g, s);
}

}  // namespace ff
