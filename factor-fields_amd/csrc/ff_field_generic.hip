// Generic field and decoder kernels: any basis_dims (<= FF_MAX_LEVELS levels,
// any channel count) and any decoder shape (out_dim <= 32).  One lane per
// point with runtime channel loops; slower than the fused fast path, used for
// configurations that have no instantiation there.
#include "ff_field_impl.h"

namespace ff {

// ------------------------------------------------------- generic fallbacks
// Any basis_dims (<= FF_MAX_LEVELS levels, any C): one lane per point, runtime
// channel loops, results written straight to memory.
struct GenCoeff {
  int ck;
  Corners<3> c3;
  Corners<2> c2;
  int tn;  // nearest texel (grid) or -1
  Corners<2> pc[3], lc[3];
  int pn[3], ln[3];
};

template <int XD>
__device__ void gen_coeff_prepare(const ff_field_desc& fd, int ck, const float* x, GenCoeff& gc) {
  gc.ck = ck;
  if (ck == CK_NONE) return;
  float pts[3] = {0.f, 0.f, 0.f};
  coeff_pts<XD>(fd, x, pts);
  if (ck == CK_GRID2 || ck == CK_GRID3) {
    const int nd = ck == CK_GRID2 ? 2 : 3;
    float pos[3];
    for (int a = 0; a < nd; ++a) pos[a] = unnorm_border(pts[a], fd.coeff_reso[a]);
    if (fd.coeff_interp == FF_INTERP_BILINEAR) {
      if (nd == 2) linear_corners<2>(pos, fd.coeff_reso, gc.c2);
      else linear_corners<3>(pos, fd.coeff_reso, gc.c3);
    } else {
      gc.tn = nd == 2 ? nearest_texel<2>(pos, fd.coeff_reso) : nearest_texel<3>(pos, fd.coeff_reso);
    }
  } else {
    for (int m = 0; m < 3; ++m) {
      const int a0 = kMatMode[m][0], a1 = kMatMode[m][1], av = kVecMode[m];
      const int Rp[2] = {fd.coeff_reso[a0], fd.coeff_reso[a1]};
      const float pp[2] = {unnorm_border(pts[a0], Rp[0]), unnorm_border(pts[a1], Rp[1])};
      const int Rl[2] = {1, fd.coeff_reso[av]};
      const float pl[2] = {unnorm_border(0.f, 1), unnorm_border(pts[av], Rl[1])};
      if (fd.coeff_interp == FF_INTERP_BILINEAR) {
        linear_corners<2>(pp, Rp, gc.pc[m]);
        linear_corners<2>(pl, Rl, gc.lc[m]);
      } else {
        gc.pn[m] = nearest_texel<2>(pp, Rp);
        gc.ln[m] = nearest_texel<2>(pl, Rl);
      }
    }
  }
}

template <bool BF>
__device__ __forceinline__ float gen_tab(const void* t, int64_t e) {
  if constexpr (BF) return bf16_to_f32(reinterpret_cast<const uint16_t*>(t)[e]);
  else return reinterpret_cast<const float*>(t)[e];
}

template <bool BF, int NK>
__device__ __forceinline__ float gen_interp(const void* t, const Corners<NK == 4 ? 2 : 3>& cr, int C,
                                            int c) {
  float s = 0.f;
  for (int k = 0; k < NK; ++k)
    if (cr.w[k] != 0.f) s = fmaf(gen_tab<BF>(t, (int64_t)cr.tex[k] * C + c), cr.w[k], s);
  return s;
}

template <bool BF>
__device__ float gen_coeff_value(const ff_field_desc& fd, const GenCoeff& gc, int c) {
  const int C = fd.n_channels;
  const bool bil = fd.coeff_interp == FF_INTERP_BILINEAR;
  if (gc.ck == CK_GRID2) {
    if (bil) return gen_interp<BF, 4>(fd.coeff[0], gc.c2, C, c);
    return gc.tn >= 0 ? gen_tab<BF>(fd.coeff[0], (int64_t)gc.tn * C + c) : 0.f;
  }
  if (gc.ck == CK_GRID3) {
    if (bil) return gen_interp<BF, 8>(fd.coeff[0], gc.c3, C, c);
    return gc.tn >= 0 ? gen_tab<BF>(fd.coeff[0], (int64_t)gc.tn * C + c) : 0.f;
  }
  float s = 0.f;
  for (int m = 0; m < 3; ++m) {
    float pv, lv;
    if (bil) {
      pv = gen_interp<BF, 4>(fd.coeff[m], gc.pc[m], C, c);
      lv = gen_interp<BF, 4>(fd.coeff[3 + m], gc.lc[m], C, c);
    } else {
      pv = gc.pn[m] >= 0 ? gen_tab<BF>(fd.coeff[m], (int64_t)gc.pn[m] * C + c) : 0.f;
      lv = gc.ln[m] >= 0 ? gen_tab<BF>(fd.coeff[3 + m], (int64_t)gc.ln[m] * C + c) : 0.f;
    }
    const float t = pv * lv;
    s = m == 0 ? t : s + t;
  }
  return s;
}

template <int NK>
__device__ __forceinline__ void gen_scatter(float* t, const Corners<NK == 4 ? 2 : 3>& cr, int C, int c,
                                            float gr) {
  for (int k = 0; k < NK; ++k)
    if (cr.w[k] != 0.f) unsafeAtomicAdd(t + (int64_t)cr.tex[k] * C + c, cr.w[k] * gr);
}

template <bool BF>
__device__ void gen_coeff_scatter(const ff_field_desc& fd, const ff_grads& g, const GenCoeff& gc,
                                  int c, float gr) {
  const int C = fd.n_channels;
  const bool bil = fd.coeff_interp == FF_INTERP_BILINEAR;
  if (gc.ck == CK_GRID2 || gc.ck == CK_GRID3) {
    if (bil) {
      if (gc.ck == CK_GRID2) gen_scatter<4>(g.coeff[0], gc.c2, C, c, gr);
      else gen_scatter<8>(g.coeff[0], gc.c3, C, c, gr);
    } else if (gc.tn >= 0) {
      unsafeAtomicAdd(g.coeff[0] + (int64_t)gc.tn * C + c, gr);
    }
    return;
  }
  for (int m = 0; m < 3; ++m) {
    float pv, lv;
    if (bil) {
      pv = gen_interp<BF, 4>(fd.coeff[m], gc.pc[m], C, c);
      lv = gen_interp<BF, 4>(fd.coeff[3 + m], gc.lc[m], C, c);
      gen_scatter<4>(g.coeff[m], gc.pc[m], C, c, gr * lv);
      gen_scatter<4>(g.coeff[3 + m], gc.lc[m], C, c, gr * pv);
    } else {
      pv = gc.pn[m] >= 0 ? gen_tab<BF>(fd.coeff[m], (int64_t)gc.pn[m] * C + c) : 0.f;
      lv = gc.ln[m] >= 0 ? gen_tab<BF>(fd.coeff[3 + m], (int64_t)gc.ln[m] * C + c) : 0.f;
      if (gc.pn[m] >= 0) unsafeAtomicAdd(g.coeff[m] + (int64_t)gc.pn[m] * C + c, gr * lv);
      if (gc.ln[m] >= 0) unsafeAtomicAdd(g.coeff[3 + m] + (int64_t)gc.ln[m] * C + c, gr * pv);
    }
  }
}

// Generic get_coding forward / backward (MODE 0 fwd, 1 bwd).
template <int D, int XD, bool BF, int MODE>
__global__ __launch_bounds__(256) void coding_generic_kernel(
    ff_field_desc fd, ff_grads gin, int ck, const float* __restrict__ x, int64_t n,
    float* __restrict__ feats, float* __restrict__ coeff, float* __restrict__ basis,
    const float* __restrict__ dfeat, const float* __restrict__ dcoeff,
    const float* __restrict__ dbasis) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const ff_grads g = local_grads(gin);
  const int C = fd.n_channels;
  float xv[XD];
  for (int a = 0; a < XD; ++a) xv[a] = x[p * XD + a];
  GenCoeff gc;
  gen_coeff_prepare<XD>(fd, ck, xv, gc);
  const bool hb = fd.basis_kind != FF_BASIS_NONE;
  const bool hc = ck != CK_NONE;
  int off = 0;
  const int L = hb ? fd.n_levels : 1;
  for (int l = 0; l < L; ++l) {
    const int dl = hb ? fd.basis_dims[l] : C;
    const int R = hb ? fd.basis_reso[l] : 1;
    Corners<D> cr;
    int tn = -1;
    if (hb) {
      float pos[D];
      int Rs[D];
      for (int a = 0; a < D; ++a) {
        pos[a] = unnorm_ac(map_coord(fd.mapping, xv[a], fd.map_origin[a], fd.map_scale[l]), R);
        Rs[a] = R;
      }
      if (fd.basis_interp == FF_INTERP_BILINEAR) linear_corners<D>(pos, Rs, cr);
      else tn = nearest_texel<D>(pos, Rs);
    }
    for (int c = 0; c < dl; ++c) {
      const int ch = off + c;
      float bv = 0.f;
      if (hb) {
        if (fd.basis_interp == FF_INTERP_BILINEAR) {
          for (int k = 0; k < (1 << D); ++k)
            bv = fmaf(gen_tab<BF>(fd.basis[l], (int64_t)cr.tex[k] * dl + c), cr.w[k], bv);
        } else if (tn >= 0) {
          bv = gen_tab<BF>(fd.basis[l], (int64_t)tn * dl + c);
        }
      }
      const float cv = hc ? gen_coeff_value<BF>(fd, gc, ch) : 0.f;
      if constexpr (MODE == 0) {
        const float f = (hb && hc) ? bv * cv : (hc ? cv : bv);
        if (feats) feats[p * C + ch] = f;
        if (coeff) coeff[p * C + ch] = hc ? cv : bv;
        if (basis) basis[p * C + ch] = hb ? bv : cv;
      } else {
        const float df = dfeat ? dfeat[p * C + ch] : 0.f;
        const float gcf = dcoeff ? dcoeff[p * C + ch] : 0.f;
        const float gbs = dbasis ? dbasis[p * C + ch] : 0.f;
        float db, dc;
        if (hb && hc) { db = df * cv + gbs; dc = df * bv + gcf; }
        else if (hc) { db = 0.f; dc = df + gcf + gbs; }
        else { db = df + gcf + gbs; dc = 0.f; }
        if (hb && db != 0.f) {
          if (fd.basis_interp == FF_INTERP_BILINEAR) {
            for (int k = 0; k < (1 << D); ++k)
              if (cr.w[k] != 0.f)
                unsafeAtomicAdd(g.basis[l] + (int64_t)cr.tex[k] * dl + c, cr.w[k] * db);
          } else if (tn >= 0) {
            unsafeAtomicAdd(g.basis[l] + (int64_t)tn * dl + c, db);
          }
        }
        if (hc && dc != 0.f) gen_coeff_scatter<BF>(fd, g, gc, ch, dc);
      }
    }
    off += dl;
  }
}

// Generic decoder forward: one lane per point, O <= 32.
__global__ __launch_bounds__(256) void mlp_generic_fwd_kernel(ff_mlp_desc m,
                                                              const float* __restrict__ f, int64_t n,
                                                              float* __restrict__ out) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int IN = m.in_dim, H = m.hidden, O = m.out_dim;
  const float* fp = f + p * IN;
  float acc[32];
#pragma unroll
  for (int o = 0; o < 32; ++o) acc[o] = 0.f;
  if (m.n_layers == 1) {
    for (int i = 0; i < IN; ++i) {
      const float v = fp[i];
#pragma unroll
      for (int o = 0; o < 32; ++o)
        if (o < O) acc[o] = fmaf(m.w0[o * IN + i], v, acc[o]);
    }
  } else {
    for (int j = 0; j < H; ++j) {
      float a = m.b0[j];
      for (int i = 0; i < IN; ++i) a = fmaf(m.w0[j * IN + i], fp[i], a);
      a = fmaxf(a, 0.f);
#pragma unroll
      for (int o = 0; o < 32; ++o)
        if (o < O) acc[o] = fmaf(m.w1[o * H + j], a, acc[o]);
    }
  }
#pragma unroll
  for (int o = 0; o < 32; ++o)
    if (o < O) out[p * O + o] = acc[o];
}

// Generic decoder backward: per-point atomics into the weight gradients.
__global__ __launch_bounds__(256) void mlp_generic_bwd_kernel(ff_mlp_desc m, ff_grads g,
                                                              const float* __restrict__ f, int64_t n,
                                                              const float* __restrict__ dout,
                                                              float* __restrict__ dfeat) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int IN = m.in_dim, H = m.hidden, O = m.out_dim;
  const float* fp = f + p * IN;
  const float* dp = dout + p * O;
  if (m.n_layers == 1) {
    for (int i = 0; i < IN; ++i) {
      float a = 0.f;
      for (int o = 0; o < O; ++o) {
        a = fmaf(m.w0[o * IN + i], dp[o], a);
        if (g.w0) unsafeAtomicAdd(g.w0 + o * IN + i, dp[o] * fp[i]);
      }
      if (dfeat) dfeat[p * IN + i] = a;
    }
    return;
  }
  if (dfeat)
    for (int i = 0; i < IN; ++i) dfeat[p * IN + i] = 0.f;
  for (int j = 0; j < H; ++j) {
    float a = m.b0[j];
    for (int i = 0; i < IN; ++i) a = fmaf(m.w0[j * IN + i], fp[i], a);
    const float hj = fmaxf(a, 0.f);
    float d = 0.f;
    for (int o = 0; o < O; ++o) {
      d = fmaf(m.w1[o * H + j], dp[o], d);
      if (g.w1) unsafeAtomicAdd(g.w1 + o * H + j, dp[o] * hj);
    }
    d = a > 0.f ? d : 0.f;
    if (d != 0.f) {
      if (g.b0) unsafeAtomicAdd(g.b0 + j, d);
      for (int i = 0; i < IN; ++i) {
        if (g.w0) unsafeAtomicAdd(g.w0 + j * IN + i, d * fp[i]);
        if (dfeat) dfeat[p * IN + i] += m.w0[j * IN + i] * d;
      }
    }
  }
}

// Fast decoder launchers (standalone linear_mat) are instantiated here.
template int launch_mlp_fwd<64, 1, 1>(const ff_mlp_desc&, const float*, int64_t, float*, hipStream_t);
template int launch_mlp_fwd<64, 1, 2>(const ff_mlp_desc&, const float*, int64_t, float*, hipStream_t);
template int launch_mlp_fwd<64, 3, 2>(const ff_mlp_desc&, const float*, int64_t, float*, hipStream_t);
template int launch_mlp_fwd<64, 28, 2>(const ff_mlp_desc&, const float*, int64_t, float*, hipStream_t);
template int launch_mlp_bwd<64, 1, 1>(const ff_mlp_desc&, const float*, int64_t, const float*, float*,
                                      float*, unsigned, hipStream_t);
template int launch_mlp_bwd<64, 1, 2>(const ff_mlp_desc&, const float*, int64_t, const float*, float*,
                                      float*, unsigned, hipStream_t);
template int launch_mlp_bwd<64, 3, 2>(const ff_mlp_desc&, const float*, int64_t, const float*, float*,
                                      float*, unsigned, hipStream_t);
template int launch_mlp_bwd<64, 28, 2>(const ff_mlp_desc&, const float*, int64_t, const float*, float*,
                                       float*, unsigned, hipStream_t);

int launch_coding_generic(const ff_field_desc& d, const ff_grads& g, int ck, int bwd, const float* x,
                          int64_t n, float* feats, float* coeff, float* basis, const float* dfeat,
                          const float* dcoeff, const float* dbasis, hipStream_t s) {
  auto go = [&](auto Di, auto XDi, auto BFi) {
    constexpr int D = decltype(Di)::value, XD = decltype(XDi)::value;
    constexpr bool BF = decltype(BFi)::value;
    if (bwd)
      hipLaunchKernelGGL((coding_generic_kernel<D, XD, BF, 1>), dim3(blocks_for(n, 256)), dim3(256), 0,
                         s, d, g, ck, x, n, nullptr, nullptr, nullptr, dfeat, dcoeff, dbasis);
    else
      hipLaunchKernelGGL((coding_generic_kernel<D, XD, BF, 0>), dim3(blocks_for(n, 256)), dim3(256), 0,
                         s, d, g, ck, x, n, feats, coeff, basis, nullptr, nullptr, nullptr);
    return launch_status();
  };
  using I2 = std::integral_constant<int, 2>;
  using I3 = std::integral_constant<int, 3>;
  using T = std::true_type;
  using F = std::false_type;
  const bool bf = d.table_dtype == FF_BF16;
  if (d.in_dim == 2 && d.x_dim == 2) return bf ? go(I2{}, I2{}, T{}) : go(I2{}, I2{}, F{});
  if (d.in_dim == 2) return bf ? go(I2{}, I3{}, T{}) : go(I2{}, I3{}, F{});
  return bf ? go(I3{}, I3{}, T{}) : go(I3{}, I3{}, F{});
}

int launch_mlp_generic_fwd(const ff_mlp_desc& m, const float* f, int64_t n, float* out,
                           hipStream_t s) {
  hipLaunchKernelGGL(mlp_generic_fwd_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, m, f, n, out);
  return launch_status();
}

int launch_mlp_generic_bwd(const ff_mlp_desc& m, const ff_grads& g, const float* f, int64_t n,
                           const float* dout, float* dfeat, hipStream_t s) {
  hipLaunchKernelGGL(mlp_generic_bwd_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, m, g, f, n,
                     dout, dfeat);
  return launch_status();
}

}  // namespace ff
