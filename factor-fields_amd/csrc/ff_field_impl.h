// Field evaluation kernels: grid_mapping + multi-level basis gather +
// coefficient gather (grid / VM) + coeff*basis + linear_mat decoder, forward
// and hand-written backward, for gfx950 (wave64).
//
// Reference semantics: models/FactorFields.py:11-33 (grid_mapping),
// :425-465 (get_coeff), :467-516 (get_basis), :523-533 (get_coding),
// :113-159 (MLPMixer / linear_mat).
//
// Work decomposition: one lane per point.  Each level's corners are a
// contiguous C-vector in the texel-major tables, so one point issues one
// vector load per corner; interpolation weights stay in registers and are
// recomputed in the backward pass instead of being stored.  The decoder runs
// on the same lane with its weights read through the scalar cache.  Weight
// gradients of the decoder are reduced per wave by transposing the per-point
// hidden gradients through LDS (lane = hidden unit afterwards) and
// accumulating in registers across the persistent loop; each wave writes one
// partial row into a workspace that a second kernel reduces deterministically.
#pragma once
#include "ff_common.h"
#include "ff_field_launch.h"
#include "ff_mfma_f32.h"

namespace ff {

// basis_dims [4,4,4,2,2,2]: the basis layout of the image, SDF and NeRF
// workloads (18 channels over 6 levels).
struct Spec18 {
  static constexpr int L = 6;
  static constexpr int C = 18;
  static constexpr int dim(int l) { return l < 3 ? 4 : 2; }
  static constexpr int off(int l) { return l < 3 ? 4 * l : 12 + 2 * (l - 3); }
};

constexpr int kMatMode[3][2] = {{0, 1}, {0, 2}, {1, 2}};  // models/FactorFields.py:213
constexpr int kVecMode[3] = {2, 1, 0};                     // models/FactorFields.py:214

template <int D, int CK>
struct XDim {
  static constexpr int value = (CK == CK_GRID3 || CK == CK_VM) ? 3 : D;
};

// ---------------------------------------------------------------- coding
// Basis level l at point x: bas[o .. o+dl) (models/FactorFields.py:487-490).
template <int D, int dl, bool BF>
__device__ __forceinline__ void basis_level_fwd(const ff_field_desc& fd, int l, const float* x,
                                                float* out) {
  const int R = fd.basis_reso[l];
  float pos[D];
  int Rs[D];
#pragma unroll
  for (int a = 0; a < D; ++a) {
    pos[a] = unnorm_ac(map_coord(fd.mapping, x[a], fd.map_origin[a], fd.map_scale[l]), R);
    Rs[a] = R;
  }
#pragma unroll
  for (int c = 0; c < dl; ++c) out[c] = 0.f;
  if (fd.basis_interp == FF_INTERP_BILINEAR) {
    // The x-neighbour corners (k, k+1) of each (y, z) come from one load of
    // two adjacent texels: the pair starts at xb = clamp(x0, 0, R-2) in the
    // row (row indices clamped into the table).  Corners, weights (in the
    // reference's product order) and the accumulation order are those of
    // linear_corners; a corner outside the grid has weight 0, whichever
    // texel of the pair it reads.
    float f0[D], w0[D], w1[D];
    int i0[D];
#pragma unroll
    for (int a = 0; a < D; ++a) {
      const float p = safe_coord(pos[a], R);
      f0[a] = floorf(p);
      i0[a] = (int)f0[a];
      w1[a] = p - f0[a];
      w0[a] = (f0[a] + 1.f) - p;
    }
    const int xb = min(max(i0[0], 0), R - 2);
    const bool sel = i0[0] == xb;  // x0 is the pair's first texel (else its second)
    const bool okx0 = i0[0] >= 0 && i0[0] < R, okx1 = i0[0] + 1 >= 0 && i0[0] + 1 < R;
    // BF16 tables: all rows' pair loads are issued before any is used (the
    // scheduler otherwise waits for each one: 45.2 -> 42.6 us for the render
    // forward's batch).  FP32 tables keep one loop that loads each pair just
    // before its use, which measured slightly faster for the FP32 kernels.
    constexpr int NP = 1 << (D - 1);
    float v2s[BF ? NP : 1][2 * dl];
    bool oks[BF ? NP : 1];
    if constexpr (BF) {
#pragma unroll
      for (int pr = 0; pr < NP; ++pr) {
        int row = 0;
        bool okr = true;
#pragma unroll
        for (int a = D - 1; a >= 1; --a) {
          const int ii = i0[a] + ((pr >> (a - 1)) & 1);
          okr = okr && ii >= 0 && ii < R;
          row = row * R + min(max(ii, 0), R - 1);
        }
        oks[pr] = okr;
        load_ch<2 * dl, BF>(fd.basis[l], ((int64_t)row * R + xb) * dl, v2s[pr]);
      }
    }
#pragma unroll
    for (int pr = 0; pr < NP; ++pr) {
      bool okr;
      const float* v2;
      float v2l[2 * dl];
      if constexpr (BF) {
        okr = oks[pr];
        v2 = v2s[pr];
      } else {
        int row = 0;
        okr = true;
#pragma unroll
        for (int a = D - 1; a >= 1; --a) {
          const int ii = i0[a] + ((pr >> (a - 1)) & 1);
          okr = okr && ii >= 0 && ii < R;
          row = row * R + min(max(ii, 0), R - 1);
        }
        load_ch<2 * dl, BF>(fd.basis[l], ((int64_t)row * R + xb) * dl, v2l);
        v2 = v2l;
      }
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int k = 2 * pr + h;
        float w = h ? w1[0] : w0[0];
#pragma unroll
        for (int a = 1; a < D; ++a) w = w * (((k >> a) & 1) ? w1[a] : w0[a]);
        w = (okr && (h ? okx1 : okx0)) ? w : 0.f;
        const float* v = ((h == 0) == sel) ? v2 : v2 + dl;
#pragma unroll
        for (int c = 0; c < dl; ++c) out[c] = fmaf(v[c], w, out[c]);
      }
    }
  } else {
    const int t = nearest_texel<D>(pos, Rs);
    if (t >= 0) load_ch<dl, BF>(fd.basis[l], (int64_t)t * dl, out);
  }
}


// Normalised coefficient coordinate (models/FactorFields.py:635-637).
template <int XD>
__device__ __forceinline__ void coeff_pts(const ff_field_desc& fd, const float* x, float* pts) {
#pragma unroll
  for (int a = 0; a < XD; ++a) pts[a] = (x[a] - fd.coeff_origin[a]) * fd.coeff_inv[a] - 1.f;
}

// Sample an ND-dim C-channel table at normalised pts (align_corners=False, border).
template <int ND, int C, bool BF>
__device__ __forceinline__ void grid_sample_border(const void* tab, const int* R, const float* pts,
                                                   int interp, float* out) {
  float pos[ND];
#pragma unroll
  for (int a = 0; a < ND; ++a) pos[a] = unnorm_border(pts[a], R[a]);
#pragma unroll
  for (int c = 0; c < C; ++c) out[c] = 0.f;
  if (interp == FF_INTERP_BILINEAR) {
    Corners<ND> cr;
    linear_corners<ND>(pos, R, cr);
#pragma unroll
    for (int k = 0; k < (1 << ND); ++k) {
      if (cr.w[k] != 0.f) {
        float v[C];
        load_ch<C, BF>(tab, (int64_t)cr.tex[k] * C, v);
#pragma unroll
        for (int c = 0; c < C; ++c) out[c] = fmaf(v[c], cr.w[k], out[c]);
      }
    }
  } else {
    const int t = nearest_texel<ND>(pos, R);
    if (t >= 0) load_ch<C, BF>(tab, (int64_t)t * C, out);
  }
}


// VM plane m and line m values at pts (3-D).
template <int C, bool BF>
__device__ __forceinline__ void vm_terms(const ff_field_desc& fd, int m, const float* pts,
                                         float* pv, float* lv) {
  const int a0 = kMatMode[m][0], a1 = kMatMode[m][1], av = kVecMode[m];
  const int Rp[2] = {fd.coeff_reso[a0], fd.coeff_reso[a1]};
  const float pp[2] = {pts[a0], pts[a1]};
  grid_sample_border<2, C, BF>(fd.coeff[m], Rp, pp, fd.coeff_interp, pv);
  const int Rl[2] = {1, fd.coeff_reso[av]};
  const float pl[2] = {0.f, pts[av]};
  grid_sample_border<2, C, BF>(fd.coeff[3 + m], Rl, pl, fd.coeff_interp, lv);
}

// get_coding for one point: bas, cof (models/FactorFields.py:523-533).
template <int D, int CK, bool BF, class S>
__device__ __forceinline__ void coding_fwd(const ff_field_desc& fd, const float* x, float* bas,
                                           float* cof) {
  constexpr int C = S::C;
  if (fd.basis_kind == FF_BASIS_GRID) {
    static_for<0, S::L>([&](auto LI) {
      constexpr int l = decltype(LI)::value;
      basis_level_fwd<D, S::dim(l), BF>(fd, l, x, bas + S::off(l));
    });
  }
  constexpr int XD = XDim<D, CK>::value;
  if constexpr (CK == CK_GRID2 || CK == CK_GRID3) {
    constexpr int ND = CK == CK_GRID2 ? 2 : 3;
    float pts[XD];
    coeff_pts<XD>(fd, x, pts);
    grid_sample_border<ND, C, BF>(fd.coeff[0], fd.coeff_reso, pts, fd.coeff_interp, cof);
  } else if constexpr (CK == CK_VM) {
    float pts[3];
    coeff_pts<3>(fd, x, pts);
#pragma unroll 1
    for (int m = 0; m < 3; ++m) {
      float pv[C], lv[C];
      vm_terms<C, BF>(fd, m, pts, pv, lv);
#pragma unroll
      for (int c = 0; c < C; ++c) {
        float t = pv[c] * lv[c];
        cof[c] = m == 0 ? t : cof[c] + t;
      }
    }
  }
}


// ------------------------------------------------------------ coalesced scatter
// Atomic throughput is bounded by cache-line requests, not by the lanes each
// request carries, and atomics to one address serialise, so the backward does
// not let every lane scatter its own point.  Each lane stages its corners
// (texel, weight) and gradient vector in a per-wave LDS area; for each corner
// slot k, consecutive points whose corner k is the same texel (samples along
// a ray, neighbouring query points) form a run, and only the first point of
// each run enters a head list, in point-major order, with the run's length.
// Groups of d consecutive lanes then take one head each, sum w * grad over its
// run in registers and add the texel's d contiguous channels with one atomic
// per channel.  Staging layout (words): G [64*d] | W [64*K] | T [64*K] |
// heads [64*K].  Every lane of the wave must call wave_scatter, with the same
// K and d, from a block of one wave (the staging area belongs to the wave).
constexpr int kStageWords = 64 * 18 + 64 * 8 * 3;

// Two loop forms with the same sums (each run added in point order):
// - serial: the prefix sum of head counts by shuffles; each pass of heads walks
//   its runs comparing texels; block barriers around the loop (their release
//   fence also waits for the wave's outstanding atomics); fastest measured in
//   the FP32 render backward;
// - batched: heads from per-slot ballots and mbcnt, each entry carrying its run
//   length; four passes of heads at a time with their LDS reads issued
//   together and the runs summed in lockstep; fastest measured in the BF16
//   matrix-core render backward and the proposal backward.
template <int K, int d>
__device__ __forceinline__ void wave_scatter_serial(float* stage, float* dst, const int* tex, const float* w,
                                             const float* g, bool active) {
  static_assert(d <= 64 && K <= 16 && 64 * d + 3 * 64 * K <= kStageWords, "staging area too small");
  const int lane = threadIdx.x & 63;
  float* G = stage;
  float* W = stage + 64 * d;
  int* T = reinterpret_cast<int*>(stage + 64 * d + 64 * K);
  int* H = T + 64 * K;
#pragma unroll
  for (int c = 0; c < d; ++c) G[lane * d + c] = active ? g[c] : 0.f;
  int hm = 0;  // corner slots where this point starts a run
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int t = active ? tex[k] : -1;  // inactive lanes: a texel no run shares
    W[lane * K + k] = active ? w[k] : 0.f;
    T[lane * K + k] = t;
    const int tp = __shfl_up(t, 1);
    hm |= int(t >= 0 && (lane == 0 || t != tp)) << k;
  }
  // heads in point-major order (the corners of a point stay adjacent, so x
  // neighbours share cache lines within one atomic instruction)
  const int cnt = __popc(hm);
  int incl = cnt;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int v = __shfl_up(incl, o);
    if (lane >= o) incl += v;
  }
  const int nh = __builtin_amdgcn_readfirstlane(__shfl(incl, 63));
  int pos = incl - cnt;
#pragma unroll
  for (int k = 0; k < K; ++k)
    if ((hm >> k) & 1) H[pos++] = (lane << 4) | k;
  __syncthreads();
  constexpr int PPI = 64 / d;  // heads per pass
  const int sub = lane / d, c = lane - sub * d;
  for (int base = 0; base < nh; base += PPI) {
    const int h = base + sub;
    if (sub < PPI && h < nh) {
      const int e = H[h];
      const int k = e & 15;
      int s = e >> 4;
      const int t = T[s * K + k];
      float acc = 0.f;
      do {
        acc = fmaf(W[s * K + k], G[s * d + c], acc);
        ++s;
      } while (s < 64 && T[s * K + k] == t);
      if (acc != 0.f) global_atomic_add(dst + (int64_t)t * d + c, acc);
    }
  }
  __syncthreads();
}

template <int K, int d>
__device__ __forceinline__ void wave_scatter_batched(float* stage, float* dst, const int* tex, const float* w,
                                             const float* g, bool active) {
  static_assert(d <= 64 && K <= 16 && 64 * d + 3 * 64 * K <= kStageWords, "staging area too small");
  constexpr int U = 4;         // passes of heads handled together
  constexpr int PPI = 64 / d;  // heads per pass
  const int lane = threadIdx.x & 63;
  float* G = stage;
  float* W = stage + 64 * d;
  int* T = reinterpret_cast<int*>(stage + 64 * d + 64 * K);
  int* H = T + 64 * K;
#pragma unroll
  for (int c = 0; c < d; ++c) G[lane * d + c] = active ? g[c] : 0.f;
  // run heads: per corner slot, ballots of heads and of run members; a head's
  // position is the number of heads of all slots in earlier lanes (mbcnt)
  // plus its own earlier slots; its length is 1 + the members right after it
  uint32_t pos = 0;
  int nh = 0, own = 0;
  uint64_t hk[K];
  int len[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int t = active ? tex[k] : -1;  // inactive lanes: a texel no run shares
    W[lane * K + k] = active ? w[k] : 0.f;
    T[lane * K + k] = t;
    const int tp = __shfl_up(t, 1);
    const bool member = t >= 0 && lane > 0 && t == tp;
    const uint64_t M = __ballot(member);
    hk[k] = __ballot(t >= 0 && !member);
    const uint64_t after = lane < 63 ? M >> (lane + 1) : 0ull;
    len[k] = 1 + __builtin_ctzll(~after);
    pos = __builtin_amdgcn_mbcnt_hi((uint32_t)(hk[k] >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)hk[k], pos));
    nh += __popcll(hk[k]);
  }
#pragma unroll
  for (int k = 0; k < K; ++k)
    if ((hk[k] >> lane) & 1) H[pos + own++] = lane | (k << 6) | (len[k] << 10);
  wave_lds_fence();
  const int sub = lane / d, c = lane - sub * d;
  for (int base = 0; base < nh; base += U * PPI) {
    int e[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int h = base + u * PPI + sub;
      const bool ok = sub < PPI && h < nh;
      const int v = H[ok ? h : 0];
      e[u] = ok ? v : 0;  // 0: no head (length 0)
    }
    float acc[U];
    int t[U], lmax = 0;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int s0 = e[u] & 63, k = (e[u] >> 6) & 15;
      acc[u] = fmaf(W[s0 * K + k], G[s0 * d + c], 0.f);
      t[u] = T[s0 * K + k];
      lmax = max(lmax, e[u] >> 10);
    }
    for (int j = 1; j < lmax; ++j) {
      float wq[U], gq[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int s0 = e[u] & 63, k = (e[u] >> 6) & 15, sq = j < (e[u] >> 10) ? s0 + j : s0;
        wq[u] = W[sq * K + k];
        gq[u] = G[sq * d + c];
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (j < (e[u] >> 10)) acc[u] = fmaf(wq[u], gq[u], acc[u]);
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if ((e[u] >> 10) > 0 && acc[u] != 0.f) global_atomic_add(dst + (int64_t)t[u] * d + c, acc[u]);
  }
  wave_lds_fence();
}

template <int K, int d, bool kBatched = false>
__device__ __forceinline__ void wave_scatter(float* stage, float* dst, const int* tex, const float* w,
                                             const float* g, bool active) {
  if constexpr (kBatched) wave_scatter_batched<K, d>(stage, dst, tex, w, g, active);
  else wave_scatter_serial<K, d>(stage, dst, tex, w, g, active);
}

// Level l of the basis: corners and scatter of grad[dl] (wave-wide).
template <int D, int dl, bool kB = false>
__device__ __forceinline__ void basis_level_bwd_wave(const ff_field_desc& fd, float* dst, int l, const float* x,
                                                     const float* grad, bool active, float* stage) {
  const int R = fd.basis_reso[l];
  float pos[D];
  int Rs[D];
#pragma unroll
  for (int a = 0; a < D; ++a) {
    pos[a] = unnorm_ac(map_coord(fd.mapping, x[a], fd.map_origin[a], fd.map_scale[l]), R);
    Rs[a] = R;
  }
  if (fd.basis_interp == FF_INTERP_BILINEAR) {
    Corners<D> cr;
    linear_corners<D>(pos, Rs, cr);
    wave_scatter<(1 << D), dl, kB>(stage, dst, cr.tex, cr.w, grad, active);
  } else {
    const int t = nearest_texel<D>(pos, Rs);
    const int tex[1] = {t >= 0 ? t : 0};
    const float w[1] = {t >= 0 ? 1.f : 0.f};
    wave_scatter<1, dl, kB>(stage, dst, tex, w, grad, active);
  }
}

// Border-padded C-channel table at normalised pts (wave-wide scatter).
template <int ND, int C, bool kB = false>
__device__ __forceinline__ void grid_scatter_wave(float* dst, const int* R, const float* pts, int interp,
                                                  const float* grad, bool active, float* stage) {
  float pos[ND];
#pragma unroll
  for (int a = 0; a < ND; ++a) pos[a] = unnorm_border(pts[a], R[a]);
  if (interp == FF_INTERP_BILINEAR) {
    Corners<ND> cr;
    linear_corners<ND>(pos, R, cr);
    wave_scatter<(1 << ND), C, kB>(stage, dst, cr.tex, cr.w, grad, active);
  } else {
    const int t = nearest_texel<ND>(pos, R);
    const int tex[1] = {t >= 0 ? t : 0};
    const float w[1] = {t >= 0 ? 1.f : 0.f};
    wave_scatter<1, C, kB>(stage, dst, tex, w, grad, active);
  }
}

// coding_bwd for a whole wave (every lane calls it; inactive lanes add nothing);
// kB selects the batched scatter loop.
template <int D, int CK, bool BF, class S, bool kB = false>
__device__ __forceinline__ void coding_bwd_wave(const ff_field_desc& fd, const ff_grads& g, const float* x,
                                                const float* dbas, const float* dcof, bool active,
                                                float* stage) {
  constexpr int C = S::C;
  if (fd.basis_kind == FF_BASIS_GRID) {
    static_for<0, S::L>([&](auto LI) {
      constexpr int l = decltype(LI)::value;
      basis_level_bwd_wave<D, S::dim(l), kB>(fd, g.basis[l], l, x, dbas + S::off(l), active, stage);
    });
  }
  constexpr int XD = XDim<D, CK>::value;
  if constexpr (CK == CK_GRID2 || CK == CK_GRID3) {
    constexpr int ND = CK == CK_GRID2 ? 2 : 3;
    float pts[XD];
    coeff_pts<XD>(fd, x, pts);
    grid_scatter_wave<ND, C, kB>(g.coeff[0], fd.coeff_reso, pts, fd.coeff_interp, dcof, active, stage);
  } else if constexpr (CK == CK_VM) {
    float pts[3];
    coeff_pts<3>(fd, x, pts);
#pragma unroll 1
    for (int m = 0; m < 3; ++m) {
      float pv[C], lv[C], t[C];
      vm_terms<C, BF>(fd, m, pts, pv, lv);
      const int a0 = kMatMode[m][0], a1 = kMatMode[m][1], av = kVecMode[m];
      const int Rp[2] = {fd.coeff_reso[a0], fd.coeff_reso[a1]};
      const float pp[2] = {pts[a0], pts[a1]};
#pragma unroll
      for (int c = 0; c < C; ++c) t[c] = dcof[c] * lv[c];
      grid_scatter_wave<2, C, kB>(g.coeff[m], Rp, pp, fd.coeff_interp, t, active, stage);
      const int Rl[2] = {1, fd.coeff_reso[av]};
      const float pl[2] = {0.f, pts[av]};
#pragma unroll
      for (int c = 0; c < C; ++c) t[c] = dcof[c] * pv[c];
      grid_scatter_wave<2, C, kB>(g.coeff[3 + m], Rl, pl, fd.coeff_interp, t, active, stage);
    }
  }
}

// VM coefficient lines accumulated in LDS: the three lines of length up to
// kVmLdsLine texels (C channels each) are small enough to sum in a block's
// LDS, [3][kVmLdsLine][C] floats, and reach the global gradients once per
// block (vm_lines_flush) instead of with every point's atomics.
constexpr int kVmLdsLine = 128;

// The line's corners and weights are those grid_scatter_wave forms for the
// [1, R] line table (the unit axis contributes weight 1), added with LDS
// atomics.
template <int C>
__device__ __forceinline__ void vm_line_lds_add(float* line, int R, float pt, int interp, const float* grad,
                                                bool active) {
  if (!active) return;
  const float p = safe_coord(unnorm_border(pt, R), R);
  if (interp == FF_INTERP_BILINEAR) {
    const float f0 = floorf(p);
    const int i0 = (int)f0;
    const float w1 = p - f0, w0 = (f0 + 1.f) - p;
    const bool ok0 = i0 >= 0 && i0 < R && w0 != 0.f, ok1 = i0 + 1 >= 0 && i0 + 1 < R && w1 != 0.f;
#pragma unroll
    for (int c = 0; c < C; ++c) {
      if (ok0) atomicAdd(line + i0 * C + c, w0 * grad[c]);
      if (ok1) atomicAdd(line + (i0 + 1) * C + c, w1 * grad[c]);
    }
  } else {
    const int t = (int)rintf(p);
    if (t < 0 || t >= R) return;
#pragma unroll
    for (int c = 0; c < C; ++c) atomicAdd(line + t * C + c, grad[c]);
  }
}

// coding_bwd_wave for the VM coefficient with its lines summed in LDS
// (lines: [3][kVmLdsLine][C], line reso <= kVmLdsLine); planes and basis as
// coding_bwd_wave.
template <int D, bool BF, class S, bool kB>
__device__ __forceinline__ void coding_bwd_wave_vm_lds(const ff_field_desc& fd, const ff_grads& g, const float* x,
                                                       const float* dbas, const float* dcof, bool active,
                                                       float* stage, float* lines) {
  constexpr int C = S::C;
  if (fd.basis_kind == FF_BASIS_GRID) {
    static_for<0, S::L>([&](auto LI) {
      constexpr int l = decltype(LI)::value;
      basis_level_bwd_wave<D, S::dim(l), kB>(fd, g.basis[l], l, x, dbas + S::off(l), active, stage);
    });
  }
  float pts[3];
  coeff_pts<3>(fd, x, pts);
#pragma unroll 1
  for (int m = 0; m < 3; ++m) {
    float pv[C], lv[C], t[C];
    vm_terms<C, BF>(fd, m, pts, pv, lv);
    const int a0 = kMatMode[m][0], a1 = kMatMode[m][1], av = kVecMode[m];
    const int Rp[2] = {fd.coeff_reso[a0], fd.coeff_reso[a1]};
    const float pp[2] = {pts[a0], pts[a1]};
#pragma unroll
    for (int c = 0; c < C; ++c) t[c] = dcof[c] * lv[c];
    grid_scatter_wave<2, C, kB>(g.coeff[m], Rp, pp, fd.coeff_interp, t, active, stage);
#pragma unroll
    for (int c = 0; c < C; ++c) t[c] = dcof[c] * pv[c];
    vm_line_lds_add<C>(lines + m * kVmLdsLine * C, fd.coeff_reso[av], pts[av], fd.coeff_interp, t, active);
  }
}

// feats = basis * coeff, with the 'none' variants of get_coding.
template <int C, int CK>
__device__ __forceinline__ void combine(const ff_field_desc& fd, const float* bas, const float* cof,
                                       float* f) {
  const bool hb = fd.basis_kind != FF_BASIS_NONE;
  const bool hc = CK != CK_NONE;
#pragma unroll
  for (int c = 0; c < C; ++c) f[c] = (hb && hc) ? bas[c] * cof[c] : (hc ? cof[c] : bas[c]);
}

template <int C, int CK>
__device__ __forceinline__ void combine_bwd(const ff_field_desc& fd, const float* bas,
                                            const float* cof, const float* df, float* dbas,
                                            float* dcof) {
  const bool hb = fd.basis_kind != FF_BASIS_NONE;
  const bool hc = CK != CK_NONE;
#pragma unroll
  for (int c = 0; c < C; ++c) {
    dbas[c] = (hb && hc) ? df[c] * cof[c] : (hc ? 0.f : df[c]);
    dcof[c] = (hb && hc) ? df[c] * bas[c] : (hc ? df[c] : 0.f);
  }
}

// ---------------------------------------------------------------- decoder
// Decoder weights are read through the constant address space so they are
// fetched with scalar loads (uniform across the wave) even in kernels that
// also issue atomics.
typedef const __attribute__((address_space(4))) float* cfp;

template <int IN, int H, int O, int NL>
__device__ __forceinline__ void mlp_fwd(const ff_mlp_desc& m, const float* f, float* h, float* out) {
  const cfp w0 = (cfp)m.w0;
  if constexpr (NL == 1) {
#pragma unroll
    for (int o = 0; o < O; ++o) {
      float a = 0.f;
#pragma unroll
      for (int i = 0; i < IN; ++i) a = fmaf(w0[o * IN + i], f[i], a);
      out[o] = a;
    }
  } else {
    const cfp b0 = (cfp)m.b0;
    const cfp w1 = (cfp)m.w1;
#pragma unroll
    for (int j = 0; j < H; ++j) {
      float a = b0[j];
#pragma unroll
      for (int i = 0; i < IN; ++i) a = fmaf(w0[j * IN + i], f[i], a);
      h[j] = fmaxf(a, 0.f);
    }
#pragma unroll
    for (int o = 0; o < O; ++o) {
      float a = 0.f;
#pragma unroll
      for (int j = 0; j < H; ++j) a = fmaf(w1[o * H + j], h[j], a);
      out[o] = a;
    }
  }
}

// ReLU pattern of the hidden layer as a bit mask (bit j = h[j] > 0).
template <int H>
__device__ __forceinline__ uint64_t relu_mask(const float* h) {
  uint64_t msk = 0;
#pragma unroll
  for (int j = 0; j < H; ++j) msk |= (uint64_t)(h[j] > 0.f) << j;
  return msk;
}

// d_out -> dh (NL == 2) and d_feats.
template <int IN, int H, int O, int NL>
__device__ __forceinline__ void mlp_bwd_local(const ff_mlp_desc& m, uint64_t msk, const float* dout,
                                              float* dh, float* df) {
  const cfp w0 = (cfp)m.w0;
  if constexpr (NL == 1) {
#pragma unroll
    for (int i = 0; i < IN; ++i) {
      float a = 0.f;
#pragma unroll
      for (int o = 0; o < O; ++o) a = fmaf(w0[o * IN + i], dout[o], a);
      df[i] = a;
    }
  } else {
    const cfp w1 = (cfp)m.w1;
#pragma unroll
    for (int j = 0; j < H; ++j) {
      float a = 0.f;
#pragma unroll
      for (int o = 0; o < O; ++o) a = fmaf(w1[o * H + j], dout[o], a);
      dh[j] = ((msk >> j) & 1) ? a : 0.f;
    }
#pragma unroll
    for (int i = 0; i < IN; ++i) {
      float a = 0.f;
#pragma unroll
      for (int j = 0; j < H; ++j) a = fmaf(w0[j * IN + i], dh[j], a);
      df[i] = a;
    }
  }
}

// Per-wave decoder weight-gradient accumulators.
//   NL == 2 (H == 64): lane j owns row j of dW0 (IN values), db0[j] and
//   column j of dW1 (O values).
//   NL == 1: every lane holds the wave total of dW0 (O*IN values).
template <int IN, int H, int O, int NL>
struct MlpAcc {
  static constexpr int kRow = NL == 1 ? O * IN : IN + 1 + O;
  float v[kRow];
  __device__ __forceinline__ void zero() {
#pragma unroll
    for (int k = 0; k < kRow; ++k) v[k] = 0.f;
  }
  // Number of floats one wave contributes to the workspace row.
  static constexpr int kPartial = NL == 1 ? O * IN : H * IN + H + O * H;
};

__device__ __forceinline__ float lane_bcast(float v, int p) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), p));
}

// dW1[o][j] += sum_p dout[p][o] * h[p][j]: h is transposed through the
// per-wave 64 x 65 LDS buffer so lane j reads column j; dout[p][o] is a
// lane broadcast.  Inactive lanes pass dout = 0.
template <int IN, int H, int O, int NL>
__device__ __forceinline__ void wave_acc_w1(MlpAcc<IN, H, O, NL>& acc, float* lds, const float* h,
                                            const float* dout) {
  static_assert(NL == 2 && H == kWave, "hidden width must equal the wave size");
  const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
  for (int j = 0; j < H; ++j) lds[lane * 65 + j] = h[j];
  __syncthreads();
#pragma unroll 2
  for (int p = 0; p < kWave; ++p) {
    const float hv = lds[p * 65 + lane];
#pragma unroll
    for (int o = 0; o < O; ++o) acc.v[IN + 1 + o] = fmaf(lane_bcast(dout[o], p), hv, acc.v[IN + 1 + o]);
  }
  __syncthreads();
}

// dW0[j][i] += sum_p dh[p][j] * f[p][i] and db0[j] += sum_p dh[p][j].
template <int IN, int H, int O, int NL>
__device__ __forceinline__ void wave_acc_w0(MlpAcc<IN, H, O, NL>& acc, float* lds, const float* f,
                                            const float* dh) {
  const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
  for (int j = 0; j < H; ++j) lds[lane * 65 + j] = dh[j];
  __syncthreads();
#pragma unroll 2
  for (int p = 0; p < kWave; ++p) {
    const float d = lds[p * 65 + lane];
#pragma unroll
    for (int i = 0; i < IN; ++i) acc.v[i] = fmaf(d, lane_bcast(f[i], p), acc.v[i]);
    acc.v[IN] += d;
  }
  __syncthreads();
}

// NL == 1: dW0[o][i] += sum over the wave of dout[o] * f[i].
template <int IN, int H, int O, int NL>
__device__ __forceinline__ void wave_acc_linear(MlpAcc<IN, H, O, NL>& acc, const float* f,
                                                const float* dout) {
#pragma unroll
  for (int o = 0; o < O; ++o)
#pragma unroll
    for (int i = 0; i < IN; ++i) acc.v[o * IN + i] += wave_sum(dout[o] * f[i]);
}

template <int IN, int H, int O, int NL>
__device__ __forceinline__ void mlp_acc_store(const MlpAcc<IN, H, O, NL>& acc, float* row) {
  const int lane = threadIdx.x & (kWave - 1);
  if constexpr (NL == 1) {
    if (lane == 0) {
#pragma unroll
      for (int k = 0; k < O * IN; ++k) row[k] = acc.v[k];
    }
  } else {
#pragma unroll
    for (int i = 0; i < IN; ++i) row[lane * IN + i] = acc.v[i];
    row[H * IN + lane] = acc.v[IN];
#pragma unroll
    for (int o = 0; o < O; ++o) row[H * IN + H + o * H + lane] = acc.v[IN + 1 + o];
  }
}

// Decoder forward, d_out from `make_dout(out, dout)`, decoder backward for one
// wave-tile; leaves d_feats in df and accumulates the weight gradients.
// Inactive lanes contribute nothing (their d_out is forced to zero).
template <int IN, int H, int O, int NL, class DoutFn>
__device__ __forceinline__ void mlp_train_tile_fn(const ff_mlp_desc& m, MlpAcc<IN, H, O, NL>& acc,
                                                  float* lds, const float* f, bool active,
                                                  DoutFn&& make_dout, float* df) {
  float h[NL == 2 ? H : 1], o[O], dout[O];
  mlp_fwd<IN, H, O, NL>(m, f, h, o);
  make_dout(o, dout);
#pragma unroll
  for (int k = 0; k < O; ++k) dout[k] = active ? dout[k] : 0.f;
  if constexpr (NL == 2) {
    const uint64_t msk = relu_mask<H>(h);
    wave_acc_w1<IN, H, O, NL>(acc, lds, h, dout);
    float dh[H];
    mlp_bwd_local<IN, H, O, NL>(m, msk, dout, dh, df);
    wave_acc_w0<IN, H, O, NL>(acc, lds, f, dh);
  } else {
    mlp_bwd_local<IN, H, O, NL>(m, 0, dout, nullptr, df);
    wave_acc_linear<IN, H, O, NL>(acc, f, dout);
  }
}

// mode 0: d_out read from dsrc; 1: MSE against dsrc; 2: L1 against dsrc.
template <int IN, int H, int O, int NL>
__device__ __forceinline__ void mlp_train_tile(const ff_mlp_desc& m, MlpAcc<IN, H, O, NL>& acc,
                                               float* lds, const float* f, bool active, int64_t pc,
                                               const float* __restrict__ dsrc, int mode,
                                               float grad_scale, float& lsum, float* df) {
  mlp_train_tile_fn<IN, H, O, NL>(
      m, acc, lds, f, active,
      [&](const float* o, float* dout) {
#pragma unroll
        for (int k = 0; k < O; ++k) {
          if (mode == 0) {
            dout[k] = dsrc[pc * O + k];
          } else {
            const float e = o[k] - dsrc[pc * O + k];
            if (mode == 1) {
              dout[k] = 2.f * e * grad_scale;
              lsum += active ? e * e : 0.f;
            } else {
              dout[k] = (e > 0.f ? 1.f : (e < 0.f ? -1.f : 0.f)) * grad_scale;
              lsum += active ? fabsf(e) : 0.f;
            }
          }
        }
      },
      df);
}

// ---------------------------------------------------------------- kernels
template <int D, int CK, bool BF, class S>
__global__ __launch_bounds__(256) void coding_fwd_kernel(ff_field_desc fd, const float* __restrict__ x,
                                                         int64_t n, float* __restrict__ feats,
                                                         float* __restrict__ coeff,
                                                         float* __restrict__ basis) {
  constexpr int C = S::C;
  constexpr int XD = XDim<D, CK>::value;
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  float xv[XD];
#pragma unroll
  for (int a = 0; a < XD; ++a) xv[a] = x[p * XD + a];
  float bas[C], cof[C], f[C];
  coding_fwd<D, CK, BF, S>(fd, xv, bas, cof);
  combine<C, CK>(fd, bas, cof, f);
#pragma unroll
  for (int c = 0; c < C; ++c) {
    if (feats) feats[p * C + c] = f[c];
    if (coeff) coeff[p * C + c] = CK == CK_NONE ? bas[c] : cof[c];
    if (basis) basis[p * C + c] = fd.basis_kind == FF_BASIS_NONE ? cof[c] : bas[c];
  }
}

template <int D, int CK, bool BF, class S>
__global__ __launch_bounds__(64) void coding_bwd_kernel(ff_field_desc fd, ff_grads g,
                                                        const float* __restrict__ x, int64_t n,
                                                        const float* __restrict__ dfeat,
                                                        const float* __restrict__ dcoeff,
                                                        const float* __restrict__ dbasis) {
  constexpr int C = S::C;
  constexpr int XD = XDim<D, CK>::value;
  __shared__ float stage[kStageWords];
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = p < n;
  const int64_t pc = active ? p : n - 1;
  float xv[XD];
#pragma unroll
  for (int a = 0; a < XD; ++a) xv[a] = x[pc * XD + a];
  float bas[C], cof[C], df[C], db[C], dc[C];
  const ff_grads gl = local_grads(g);
  coding_fwd<D, CK, BF, S>(fd, xv, bas, cof);
#pragma unroll
  for (int c = 0; c < C; ++c) df[c] = dfeat ? dfeat[pc * C + c] : 0.f;
  combine_bwd<C, CK>(fd, bas, cof, df, db, dc);
  // the returned coeff aliases basis when the coefficient is 'none' (and vice versa)
#pragma unroll
  for (int c = 0; c < C; ++c) {
    const float gc = dcoeff ? dcoeff[pc * C + c] : 0.f;
    const float gb = dbasis ? dbasis[pc * C + c] : 0.f;
    if (CK == CK_NONE) db[c] += gc + gb;
    else if (fd.basis_kind == FF_BASIS_NONE) dc[c] += gc + gb;
    else { dc[c] += gc; db[c] += gb; }
  }
  coding_bwd_wave<D, CK, BF, S>(fd, gl, xv, db, dc, active, stage);
}

// Fused get_coding + linear_mat forward.
template <int D, int CK, bool BF, class S, int H, int O, int NL>
__global__ __launch_bounds__(256) void field_fwd_kernel(ff_field_desc fd, ff_mlp_desc m,
                                                        const float* __restrict__ x, int64_t n,
                                                        float* __restrict__ out) {
  constexpr int C = S::C;
  constexpr int XD = XDim<D, CK>::value;
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  float xv[XD];
#pragma unroll
  for (int a = 0; a < XD; ++a) xv[a] = x[p * XD + a];
  float bas[C], cof[C], f[C], h[H], o[O];
  coding_fwd<D, CK, BF, S>(fd, xv, bas, cof);
  combine<C, CK>(fd, bas, cof, f);
  mlp_fwd<C, H, O, NL>(m, f, h, o);
#pragma unroll
  for (int k = 0; k < O; ++k) out[p * O + k] = o[k];
}

// Fused backward (mode 0: d_out given) or regression training pass
// (mode 1: MSE, mode 2: L1 against `target`).  Persistent: one wave per block,
// blocks stride over 64-point tiles.
template <int D, int CK, bool BF, class S, int H, int O, int NL>
__global__ __launch_bounds__(64) void field_bwd_kernel(ff_field_desc fd, ff_mlp_desc m, ff_grads g,
                                                       const float* __restrict__ x, int64_t n,
                                                       const float* __restrict__ dsrc, int mode,
                                                       float loss_scale, float* __restrict__ ws,
                                                       float* __restrict__ loss_sum) {
  constexpr int C = S::C;
  constexpr int XD = XDim<D, CK>::value;
  using Acc = MlpAcc<C, H, O, NL>;
  __shared__ float lds[kWave * 65 > kStageWords ? kWave * 65 : kStageWords];
  Acc acc;
  acc.zero();
  float lsum = 0.f;
  const int lane = threadIdx.x;
  const int64_t ntiles = (n + kWave - 1) / kWave;
  const float grad_scale = (float)(1.0 / ((double)n * O)) * loss_scale;
  const ff_grads gl = local_grads(g);
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int64_t p = tile * kWave + lane;
    const bool active = p < n;
    const int64_t pc = active ? p : n - 1;
    float xv[XD];
#pragma unroll
    for (int a = 0; a < XD; ++a) xv[a] = x[pc * XD + a];
    float bas[C], cof[C], f[C], df[C];
    coding_fwd<D, CK, BF, S>(fd, xv, bas, cof);
    combine<C, CK>(fd, bas, cof, f);
    mlp_train_tile<C, H, O, NL>(m, acc, lds, f, active, pc, dsrc, mode, grad_scale, lsum, df);
    float db[C], dc[C];
    combine_bwd<C, CK>(fd, bas, cof, df, db, dc);
    coding_bwd_wave<D, CK, BF, S>(fd, gl, xv, db, dc, active, lds);
  }
  mlp_acc_store<C, H, O, NL>(acc, ws + (int64_t)blockIdx.x * Acc::kPartial);
  if (mode != 0) {
    lsum = wave_sum(lsum);
    if (lane == 0 && loss_sum) unsafeAtomicAdd(loss_sum, lsum);
  }
}

// field_bwd_kernel with the decoder on the FP32 matrix cores (18 -> 64 -> O,
// O <= 32): weights in operand registers, loaded once per wave, instead of
// 1,200-3,300 weights read through the scalar cache (which the vector decoder
// spills to lanes as 1,400-3,200 scalar registers).  Same loss modes, same
// partial-row layout.
//
// VM coefficient (lines of up to kVmLdsLine texels): blocks of kVmBwdWaves
// waves, one per SIMD of a CU, that sum the three coefficient lines'
// gradients in the block's LDS and add them to the global gradients once at
// the end (the lines are 3 x 128 x 18 floats; every point touches two texels
// of each, which as global atomics were about a fifth of the kernel's line
// requests).  `waves` is the number of logical waves (partial rows); blocks
// cover ceil(waves / kVmBwdWaves) x kVmBwdWaves waves and the extra ones
// only help to clear and flush the lines.
constexpr int kVmBwdWaves = 4;

template <int CK>
constexpr int bwd_block_waves() { return CK == CK_VM ? kVmBwdWaves : 1; }

template <int D, int CK, bool BF, class S, int O>
__global__ __launch_bounds__(64 * bwd_block_waves<CK>()) void field_bwd_mfma_kernel(
    ff_field_desc fd, ff_mlp_desc m, ff_grads g, const float* __restrict__ x, int64_t n,
    const float* __restrict__ dsrc, int mode, float loss_scale, float* __restrict__ ws,
    float* __restrict__ loss_sum, int64_t waves) {
  constexpr int C = S::C;
  static_assert(C == kMfIn, "the matrix-core decoder takes 18 features");
  constexpr int XD = XDim<D, CK>::value;
  constexpr int NW = bwd_block_waves<CK>();
  constexpr bool kLines = CK == CK_VM;
  constexpr int kWaveWords = kF32LdsWords > kStageWords ? kF32LdsWords : kStageWords;
  __shared__ __attribute__((aligned(16))) float lds_all[NW][kWaveWords];
  __shared__ __attribute__((aligned(16))) float lines[kLines ? 3 * kVmLdsLine * C : 1];
  const int wid = (int)(threadIdx.x >> 6);
  float* lds = lds_all[wid];
  const bool use_lines = kLines && fd.coeff_reso[0] <= kVmLdsLine && fd.coeff_reso[1] <= kVmLdsLine &&
                         fd.coeff_reso[2] <= kVmLdsLine;
  if constexpr (kLines) {
    for (int i = threadIdx.x; i < 3 * kVmLdsLine * C; i += blockDim.x) lines[i] = 0.f;
    __syncthreads();
  }
  MfmaDecoderF32<O> W;
  W.load(m);
  MfmaDecoderF32Bwd<O> WB;
  WB.load(m);
  MfmaDecoderAcc<O> acc;
  acc.zero();
  float lsum = 0.f;
  const int lane = (int)(threadIdx.x & (kWave - 1));
  const int64_t wave = (int64_t)blockIdx.x * NW + wid;
  const int64_t ntiles = (n + kWave - 1) / kWave;
  const float grad_scale = (float)(1.0 / ((double)n * O)) * loss_scale;
  const ff_grads gl = local_grads(g);
  for (int64_t tile = wave; wave < waves && tile < ntiles; tile += waves) {
    const int64_t p = tile * kWave + lane;
    const bool active = p < n;
    const int64_t pc = active ? p : n - 1;
    float xv[XD];
#pragma unroll
    for (int a = 0; a < XD; ++a) xv[a] = x[pc * XD + a];
    float bas[C], cof[C], f[C], df[C];
    coding_fwd<D, CK, BF, S>(fd, xv, bas, cof);
    combine<C, CK>(fd, bas, cof, f);
    // mode 0: d_out read from dsrc; 1: MSE against dsrc; 2: L1 against dsrc
    mfma_f32_decoder_train(
        W, WB, f,
        [&](const float* o, float* dout) {
#pragma unroll
          for (int k = 0; k < 32; ++k) {
            float d = 0.f;
            if (k < O) {
              if (mode == 0) {
                d = dsrc[pc * O + k];
              } else {
                const float e = o[k] - dsrc[pc * O + k];
                if (mode == 1) {
                  d = 2.f * e * grad_scale;
                  lsum += active ? e * e : 0.f;
                } else {
                  d = (e > 0.f ? 1.f : (e < 0.f ? -1.f : 0.f)) * grad_scale;
                  lsum += active ? fabsf(e) : 0.f;
                }
              }
            }
            dout[k] = active ? d : 0.f;
          }
        },
        df, acc, lds);
    float db[C], dc[C];
    combine_bwd<C, CK>(fd, bas, cof, df, db, dc);
    if constexpr (kLines) {
      if (use_lines)
        coding_bwd_wave_vm_lds<D, BF, S, true>(fd, gl, xv, db, dc, active, lds, lines);
      else
        coding_bwd_wave<D, CK, BF, S, true>(fd, gl, xv, db, dc, active, lds);
    } else {
      coding_bwd_wave<D, CK, BF, S, true>(fd, gl, xv, db, dc, active, lds);
    }
  }
  if (wave < waves) acc.store(ws + wave * (kMfH * kMfIn + kMfH + O * kMfH));
  if (mode != 0) {
    lsum = wave_sum(lsum);
    if (lane == 0 && loss_sum) unsafeAtomicAdd(loss_sum, lsum);
  }
  if constexpr (kLines) {
    if (use_lines) {
      __syncthreads();
      for (int i = threadIdx.x; i < 3 * kVmLdsLine * C; i += blockDim.x) {
        const int mm = i / (kVmLdsLine * C), e = i - mm * kVmLdsLine * C;
        const float v = lines[i];
        if (e < fd.coeff_reso[kVecMode[mm]] * C && v != 0.f) global_atomic_add(gl.coeff[3 + mm] + e, v);
      }
    }
  }
}

// Fast decoder kernels (lane = point) for standalone linear_mat.
template <int IN, int H, int O, int NL>
__global__ __launch_bounds__(256) void mlp_fwd_kernel(ff_mlp_desc m, const float* __restrict__ f,
                                                      int64_t n, float* __restrict__ out) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  float fv[IN], h[H], o[O];
#pragma unroll
  for (int i = 0; i < IN; ++i) fv[i] = f[p * IN + i];
  mlp_fwd<IN, H, O, NL>(m, fv, h, o);
#pragma unroll
  for (int k = 0; k < O; ++k) out[p * O + k] = o[k];
}

template <int IN, int H, int O, int NL>
__global__ __launch_bounds__(64) void mlp_bwd_kernel(ff_mlp_desc m, const float* __restrict__ f,
                                                     int64_t n, const float* __restrict__ dsrc,
                                                     float* __restrict__ dfeat,
                                                     float* __restrict__ ws) {
  using Acc = MlpAcc<IN, H, O, NL>;
  __shared__ float lds[NL == 2 ? kWave * 65 : 1];
  Acc acc;
  acc.zero();
  float lsum = 0.f;
  const int lane = threadIdx.x;
  const int64_t ntiles = (n + kWave - 1) / kWave;
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int64_t p = tile * kWave + lane;
    const bool active = p < n;
    const int64_t pc = active ? p : n - 1;
    float fv[IN], df[IN];
#pragma unroll
    for (int i = 0; i < IN; ++i) fv[i] = f[pc * IN + i];
    mlp_train_tile<IN, H, O, NL>(m, acc, lds, fv, active, pc, dsrc, 0, 1.f, lsum, df);
    if (active && dfeat) {
#pragma unroll
      for (int i = 0; i < IN; ++i) dfeat[p * IN + i] = df[i];
    }
  }
  mlp_acc_store<IN, H, O, NL>(acc, ws + (int64_t)blockIdx.x * Acc::kPartial);
}

}  // namespace ff


// ---------------------------------------------------------------- launchers
namespace ff {

inline int launch_status() { return hipGetLastError() == hipSuccess ? FF_OK : FF_ERR_LAUNCH; }
inline unsigned blocks_for(int64_t n, int per) { return (unsigned)((n + per - 1) / per); }

template <int D, int CK, bool BF>
int launch_coding_fwd(const ff_field_desc& d, const float* x, int64_t n, float* f, float* c,
                      float* b, hipStream_t s) {
  hipLaunchKernelGGL((coding_fwd_kernel<D, CK, BF, Spec18>), dim3(blocks_for(n, 256)), dim3(256), 0,
                     s, d, x, n, f, c, b);
  return launch_status();
}

template <int D, int CK, bool BF>
int launch_coding_bwd(const ff_field_desc& d, const ff_grads& g, const float* x, int64_t n,
                      const float* df, const float* dc, const float* db, hipStream_t s) {
  hipLaunchKernelGGL((coding_bwd_kernel<D, CK, BF, Spec18>), dim3(blocks_for(n, 64)), dim3(64), 0,
                     s, d, g, x, n, df, dc, db);
  return launch_status();
}

template <int D, int CK, bool BF, int H, int O, int NL>
int launch_field_fwd(const ff_field_desc& d, const ff_mlp_desc& m, const float* x, int64_t n,
                     float* out, hipStream_t s) {
  hipLaunchKernelGGL((field_fwd_kernel<D, CK, BF, Spec18, H, O, NL>), dim3(blocks_for(n, 256)),
                     dim3(256), 0, s, d, m, x, n, out);
  return launch_status();
}

template <int D, int CK, bool BF, int H, int O, int NL>
int launch_field_bwd(const ff_field_desc& d, const ff_mlp_desc& m, const ff_grads& g, const float* x,
                     int64_t n, const float* dsrc, int mode, float loss_scale, float* ws,
                     unsigned blocks, float* loss_sum, hipStream_t s) {
  if constexpr (NL == 2 && H == kMfH && O <= 32) {
    constexpr int NW = bwd_block_waves<CK>();
    hipLaunchKernelGGL((field_bwd_mfma_kernel<D, CK, BF, Spec18, O>), dim3((blocks + NW - 1) / NW), dim3(64 * NW),
                       0, s, d, m, g, x, n, dsrc, mode, loss_scale, ws, loss_sum, (int64_t)blocks);
  } else
    hipLaunchKernelGGL((field_bwd_kernel<D, CK, BF, Spec18, H, O, NL>), dim3(blocks), dim3(64), 0, s, d,
                       m, g, x, n, dsrc, mode, loss_scale, ws, loss_sum);
  return launch_status();
}

template <int H, int O, int NL>
int launch_mlp_fwd(const ff_mlp_desc& m, const float* f, int64_t n, float* out, hipStream_t s) {
  hipLaunchKernelGGL((mlp_fwd_kernel<18, H, O, NL>), dim3(blocks_for(n, 256)), dim3(256), 0, s, m, f,
                     n, out);
  return launch_status();
}

template <int H, int O, int NL>
int launch_mlp_bwd(const ff_mlp_desc& m, const float* f, int64_t n, const float* dout, float* dfeat,
                   float* ws, unsigned blocks, hipStream_t s) {
  hipLaunchKernelGGL((mlp_bwd_kernel<18, H, O, NL>), dim3(blocks), dim3(64), 0, s, m, f, n, dout,
                     dfeat, ws);
  return launch_status();
}

}  // namespace ff
