// C ABI of the field / decoder path (include/ffields.h): argument checks,
// dispatch to the fast-path instantiations or the generic kernels, and the
// deterministic reduction of the per-wave decoder weight-gradient partials.
#include <type_traits>

#include "ff_field_launch.h"

namespace ff {


// Sum workspace rows [rows, P] into the decoder gradients in a fixed order:
// stage 1 sums row slices r = y, y + 32, ... into part2[y][k]; stage 2 sums the
// 32 slices.  Row layout (NL == 2): [dW0 (H*IN, torch row-major) | db0 (H) |
// dW1 (O*H)]; (NL == 1): [dW0 (O*IN)].
__global__ __launch_bounds__(256) void reduce_partials_stage1(const float* __restrict__ ws, int rows, int P,
                                                              float* __restrict__ part2) {
  __shared__ float part[4][64];
  const int k = blockIdx.x * 64 + (threadIdx.x & 63);
  const int grp = threadIdx.x >> 6;
  float s = 0.f;
  if (k < P) {
    // unrolled so the loads go out together; the sums keep their order
#pragma unroll 8
    for (int r = blockIdx.y + kReduceSplit * grp; r < rows; r += 4 * kReduceSplit) s += ws[(int64_t)r * P + k];
  }
  part[grp][threadIdx.x & 63] = s;
  __syncthreads();
  if (grp == 0 && k < P)
    part2[(int64_t)blockIdx.y * P + k] = (part[0][threadIdx.x] + part[1][threadIdx.x]) +
                                         (part[2][threadIdx.x] + part[3][threadIdx.x]);
}

__global__ __launch_bounds__(256) void reduce_partials_stage2(const float* __restrict__ part2, int P, int n_w0,
                                                              int n_b0, float* __restrict__ gw0,
                                                              float* __restrict__ gb0,
                                                              float* __restrict__ gw1) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= P) return;
  float t = 0.f;
#pragma unroll
  for (int y = 0; y < kReduceSplit; ++y) t += part2[(int64_t)y * P + k];
  if (k < n_w0) gw0[k] += t;
  else if (k < n_w0 + n_b0) gb0[k - n_w0] += t;
  else gw1[k - n_w0 - n_b0] += t;
}

namespace {

constexpr int kWaveSize = 64;
template <int V> using ic = std::integral_constant<int, V>;
template <bool V> using bc = std::integral_constant<bool, V>;

int api_status() { return hipGetLastError() == hipSuccess ? FF_OK : FF_ERR_LAUNCH; }

int cu_count() {
  static int cached[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (cached[dev] == 0) {
    int cu = 0;
    if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cu <= 0)
      cu = 256;
    cached[dev] = cu;
  }
  return cached[dev];
}

int resolve_ck(const ff_field_desc& d) {
  if (d.coeff_kind == FF_COEFF_GRID) return d.coeff_ndim == 2 ? CK_GRID2 : CK_GRID3;
  if (d.coeff_kind == FF_COEFF_VM) return CK_VM;
  return CK_NONE;
}

int check_desc(const ff_field_desc* d) {
  if (!d) return FF_ERR_ARG;
  if (d->in_dim != 2 && d->in_dim != 3) return FF_ERR_ARG;
  if (d->table_dtype != FF_F32 && d->table_dtype != FF_BF16) return FF_ERR_ARG;
  if (d->n_channels <= 0 || d->n_channels > 1024) return FF_ERR_ARG;
  if (d->basis_interp != FF_INTERP_BILINEAR && d->basis_interp != FF_INTERP_NEAREST) return FF_ERR_ARG;
  if (d->coeff_interp != FF_INTERP_BILINEAR && d->coeff_interp != FF_INTERP_NEAREST) return FF_ERR_ARG;
  if (d->basis_kind == FF_BASIS_GRID) {
    if (d->n_levels <= 0 || d->n_levels > FF_MAX_LEVELS) return FF_ERR_ARG;
    int s = 0;
    for (int l = 0; l < d->n_levels; ++l) {
      if (d->basis_dims[l] <= 0 || d->basis_reso[l] < 2 || !d->basis[l]) return FF_ERR_ARG;
      s += d->basis_dims[l];
    }
    if (s != d->n_channels) return FF_ERR_ARG;
    if (d->mapping < FF_MAP_SAWTOOTH || d->mapping > FF_MAP_X) return FF_ERR_ARG;
  } else if (d->basis_kind != FF_BASIS_NONE) {
    return FF_ERR_ARG;
  }
  const int ck = resolve_ck(*d);
  if (d->coeff_kind == FF_COEFF_GRID) {
    if (d->coeff_ndim != 2 && d->coeff_ndim != 3) return FF_ERR_ARG;
    if (!d->coeff[0]) return FF_ERR_ARG;
    for (int a = 0; a < d->coeff_ndim; ++a)
      if (d->coeff_reso[a] < 1) return FF_ERR_ARG;
  } else if (d->coeff_kind == FF_COEFF_VM) {
    if (d->in_dim != 3) return FF_ERR_ARG;
    for (int k = 0; k < 6; ++k)
      if (!d->coeff[k]) return FF_ERR_ARG;
    for (int a = 0; a < 3; ++a)
      if (d->coeff_reso[a] < 1) return FF_ERR_ARG;
  } else if (d->coeff_kind != FF_COEFF_NONE) {
    return FF_ERR_ARG;
  }
  if (ck == CK_NONE && d->basis_kind == FF_BASIS_NONE) return FF_ERR_ARG;
  const int xd = (ck == CK_GRID3 || ck == CK_VM) ? 3 : d->in_dim;
  if (d->x_dim != xd) return FF_ERR_ARG;
  if (ck == CK_GRID2 && d->in_dim != 2) return FF_ERR_ARG;
  return FF_OK;
}

int check_grads(const ff_field_desc& d, const ff_grads* g) {
  if (!g) return FF_ERR_ARG;
  if (d.basis_kind == FF_BASIS_GRID)
    for (int l = 0; l < d.n_levels; ++l)
      if (!g->basis[l]) return FF_ERR_ARG;
  const int nc = d.coeff_kind == FF_COEFF_VM ? 6 : (d.coeff_kind == FF_COEFF_GRID ? 1 : 0);
  for (int k = 0; k < nc; ++k)
    if (!g->coeff[k]) return FF_ERR_ARG;
  return FF_OK;
}

int check_mlp(const ff_mlp_desc* m) {
  if (!m || !m->w0) return FF_ERR_ARG;
  if (m->n_layers != 1 && m->n_layers != 2) return FF_ERR_ARG;
  if (m->in_dim <= 0 || m->out_dim <= 0 || m->out_dim > 32) return FF_ERR_ARG;
  if (m->n_layers == 2 && (m->hidden <= 0 || !m->b0 || !m->w1)) return FF_ERR_ARG;
  return FF_OK;
}

bool is_spec18(const ff_field_desc& d) {
  static const int dims[6] = {4, 4, 4, 2, 2, 2};
  if (d.n_channels != 18) return false;
  if (d.basis_kind == FF_BASIS_NONE) return true;
  if (d.n_levels != 6) return false;
  for (int l = 0; l < 6; ++l)
    if (d.basis_dims[l] != dims[l]) return false;
  return true;
}

// Field combinations with a fast path.
template <class F>
int with_field(const ff_field_desc& d, F&& f) {
  if (!is_spec18(d)) return FF_ERR_UNSUPPORTED;
  const int ck = resolve_ck(d);
  if (d.table_dtype == FF_F32) {
    if (d.in_dim == 2 && ck == CK_GRID2) return f(ic<2>{}, ic<CK_GRID2>{}, bc<false>{});
    if (d.in_dim == 2 && ck == CK_GRID3) return f(ic<2>{}, ic<CK_GRID3>{}, bc<false>{});
    if (d.in_dim == 3 && ck == CK_GRID3) return f(ic<3>{}, ic<CK_GRID3>{}, bc<false>{});
    if (d.in_dim == 3 && ck == CK_VM) return f(ic<3>{}, ic<CK_VM>{}, bc<false>{});
  } else {
    if (d.in_dim == 3 && ck == CK_GRID3) return f(ic<3>{}, ic<CK_GRID3>{}, bc<true>{});
  }
  return FF_ERR_UNSUPPORTED;
}

// Decoder shapes with a fast path: (NL, H, O) for in_dim 18.
template <class F>
int with_mlp(const ff_mlp_desc& m, int in_dim, F&& f) {
  if (m.in_dim != in_dim) return FF_ERR_ARG;
  if (in_dim != 18) return FF_ERR_UNSUPPORTED;
  if (m.n_layers == 1 && m.out_dim == 1) return f(ic<1>{}, ic<64>{}, ic<1>{});
  if (m.n_layers == 2 && m.hidden == 64) {
    if (m.out_dim == 1) return f(ic<2>{}, ic<64>{}, ic<1>{});
    if (m.out_dim == 3) return f(ic<2>{}, ic<64>{}, ic<3>{});
    if (m.out_dim == 28) return f(ic<2>{}, ic<64>{}, ic<28>{});
  }
  return FF_ERR_UNSUPPORTED;
}

int partial_floats(int NL, int IN, int H, int O) { return NL == 1 ? O * IN : H * IN + H + O * H; }

// The fused backward kernels run one 64-thread block per SIMD (register- and
// LDS-limited), so four blocks per CU are all that is resident; more blocks
// would only add decoder-gradient partial rows to reduce.
unsigned persistent_blocks(int64_t n) {
  const int64_t tiles = (n + kWaveSize - 1) / kWaveSize;
  const int64_t cap = (int64_t)cu_count() * 4;
  return (unsigned)(tiles < cap ? tiles : cap);
}

int reduce_partials(const float* ws, int rows, int NL, int IN, int H, int O, const ff_grads* g,
                    hipStream_t s) {
  const int P = partial_floats(NL, IN, H, O);
  const int n_w0 = NL == 1 ? O * IN : H * IN;
  const int n_b0 = NL == 1 ? 0 : H;
  float* part2 = const_cast<float*>(ws) + (int64_t)rows * P;   // workspace tail
  hipLaunchKernelGGL(reduce_partials_stage1, dim3((P + 63) / 64, kReduceSplit), dim3(256), 0, s, ws, rows,
                     P, part2);
  hipLaunchKernelGGL(reduce_partials_stage2, dim3((P + 255) / 256), dim3(256), 0, s, part2, P, n_w0, n_b0,
                     g->w0, g->b0, g->w1);
  return api_status();
}

int check_mlp_grads(const ff_mlp_desc& m, const ff_grads* g) {
  if (!g || !g->w0) return FF_ERR_ARG;
  if (m.n_layers == 2 && (!g->b0 || !g->w1)) return FF_ERR_ARG;
  return FF_OK;
}

int field_bwd_common(const ff_field_desc* desc, const ff_mlp_desc* mlp, const float* x, int64_t n,
                     const float* dsrc, int mode, float loss_scale, const ff_grads* grads,
                     float* loss_sum, void* ws, size_t ws_bytes, void* stream) {
  int st = check_desc(desc);
  if (st) return st;
  if ((st = check_mlp(mlp))) return st;
  if (n < 0 || (n > 0 && (!x || !dsrc))) return FF_ERR_ARG;
  if ((st = check_grads(*desc, grads))) return st;
  if ((st = check_mlp_grads(*mlp, grads))) return st;
  if (n == 0) return FF_OK;
  hipStream_t s = (hipStream_t)stream;
  const ff_field_desc& d = *desc;
  const ff_mlp_desc& m = *mlp;
  if (dict_supported(d, m))
    return launch_dict_bwd(d, m, *grads, x, n, dsrc, mode, loss_scale, ws, ws_bytes, loss_sum, s);
  return with_field(d, [&](auto Di, auto CKi, auto BFi) {
    constexpr int D = decltype(Di)::value, CK = decltype(CKi)::value;
    constexpr bool BF = decltype(BFi)::value;
    return with_mlp(m, d.n_channels, [&](auto NLi, auto Hi, auto Oi) {
      constexpr int NL = decltype(NLi)::value, H = decltype(Hi)::value, O = decltype(Oi)::value;
      const unsigned blocks = persistent_blocks(n);
      const size_t need = ((size_t)blocks + kReduceSplit) * partial_floats(NL, 18, H, O) * 4;
      if (!ws || ws_bytes < need) return (int)FF_ERR_WORKSPACE;
      int r = launch_field_bwd<D, CK, BF, H, O, NL>(d, m, *grads, x, n, dsrc, mode, loss_scale,
                                                    (float*)ws, blocks, loss_sum, s);
      if (r) return r;
      return reduce_partials((const float*)ws, (int)blocks, NL, 18, H, O, grads, s);
    });
  });
}

}  // namespace

int launch_reduce_partials(const float* ws, int rows, int NL, int IN, int H, int O, const ff_grads& g,
                           hipStream_t s) {
  return reduce_partials(ws, rows, NL, IN, H, O, &g, s);
}

}  // namespace ff

using namespace ff;

extern "C" {

const char* ff_status_str(int status) {
  switch (status) {
    case FF_OK: return "ok";
    case FF_ERR_ARG: return "invalid argument";
    case FF_ERR_UNSUPPORTED: return "configuration has no kernel instantiation";
    case FF_ERR_WORKSPACE: return "workspace missing or too small";
    case FF_ERR_LAUNCH: return "kernel launch failed";
    case FF_ERR_SCRATCH_LIMIT:
      return "request exceeds the static scratch of the allocation-free entry point";
    default: return "unknown status";
  }
}

int ff_version(void) { return 1; }

int ff_get_coding(const ff_field_desc* desc, const float* x, int64_t n, float* feats, float* coeff,
                  float* basis, void* stream) {
  int st = check_desc(desc);
  if (st) return st;
  if (n < 0 || (n > 0 && !x)) return FF_ERR_ARG;
  if (n == 0) return FF_OK;
  hipStream_t s = (hipStream_t)stream;
  const ff_field_desc& d = *desc;
  st = with_field(d, [&](auto Di, auto CKi, auto BFi) {
    return launch_coding_fwd<decltype(Di)::value, decltype(CKi)::value, decltype(BFi)::value>(
        d, x, n, feats, coeff, basis, s);
  });
  if (st != FF_ERR_UNSUPPORTED) return st;
  ff_grads none{};
  return launch_coding_generic(d, none, resolve_ck(d), 0, x, n, feats, coeff, basis, nullptr, nullptr,
                               nullptr, s);
}

int ff_get_coding_bwd(const ff_field_desc* desc, const float* x, int64_t n, const float* d_feats,
                      const float* d_coeff, const float* d_basis, const ff_grads* grads,
                      void* stream) {
  int st = check_desc(desc);
  if (st) return st;
  if (n < 0 || (n > 0 && !x)) return FF_ERR_ARG;
  if ((st = check_grads(*desc, grads))) return st;
  if (n == 0) return FF_OK;
  hipStream_t s = (hipStream_t)stream;
  const ff_field_desc& d = *desc;
  st = with_field(d, [&](auto Di, auto CKi, auto BFi) {
    return launch_coding_bwd<decltype(Di)::value, decltype(CKi)::value, decltype(BFi)::value>(
        d, *grads, x, n, d_feats, d_coeff, d_basis, s);
  });
  if (st != FF_ERR_UNSUPPORTED) return st;
  return launch_coding_generic(d, *grads, resolve_ck(d), 1, x, n, nullptr, nullptr, nullptr, d_feats,
                               d_coeff, d_basis, s);
}

int ff_linear_mat_fwd(const ff_mlp_desc* mlp, const float* feats, int64_t n, float* out,
                      void* stream) {
  int st = check_mlp(mlp);
  if (st) return st;
  if (n < 0 || (n > 0 && (!feats || !out))) return FF_ERR_ARG;
  if (n == 0) return FF_OK;
  hipStream_t s = (hipStream_t)stream;
  const ff_mlp_desc& m = *mlp;
  st = with_mlp(m, m.in_dim, [&](auto NLi, auto Hi, auto Oi) {
    return launch_mlp_fwd<decltype(Hi)::value, decltype(Oi)::value, decltype(NLi)::value>(m, feats, n,
                                                                                         out, s);
  });
  if (st != FF_ERR_UNSUPPORTED) return st;
  return launch_mlp_generic_fwd(m, feats, n, out, s);
}

size_t ff_linear_mat_bwd_workspace_bytes(const ff_mlp_desc* mlp, int64_t n) {
  if (check_mlp(mlp) || n <= 0) return 0;
  size_t bytes = 0;
  with_mlp(*mlp, mlp->in_dim, [&](auto NLi, auto Hi, auto Oi) {
    bytes = ((size_t)persistent_blocks(n) + kReduceSplit) *
            partial_floats(decltype(NLi)::value, 18, decltype(Hi)::value, decltype(Oi)::value) * 4;
    return 0;
  });
  return bytes;
}

int ff_linear_mat_bwd(const ff_mlp_desc* mlp, const float* feats, int64_t n, const float* d_out,
                      float* d_feats, const ff_grads* grads, void* ws, size_t ws_bytes,
                      void* stream) {
  int st = check_mlp(mlp);
  if (st) return st;
  if (n < 0 || (n > 0 && (!feats || !d_out))) return FF_ERR_ARG;
  if ((st = check_mlp_grads(*mlp, grads))) return st;
  if (n == 0) return FF_OK;
  hipStream_t s = (hipStream_t)stream;
  const ff_mlp_desc& m = *mlp;
  st = with_mlp(m, m.in_dim, [&](auto NLi, auto Hi, auto Oi) {
    constexpr int NL = decltype(NLi)::value, H = decltype(Hi)::value, O = decltype(Oi)::value;
    const unsigned blocks = persistent_blocks(n);
    const size_t need = ((size_t)blocks + kReduceSplit) * partial_floats(NL, 18, H, O) * 4;
    if (!ws || ws_bytes < need) return (int)FF_ERR_WORKSPACE;
    int r = launch_mlp_bwd<H, O, NL>(m, feats, n, d_out, d_feats, (float*)ws, blocks, s);
    if (r) return r;
    return reduce_partials((const float*)ws, (int)blocks, NL, 18, H, O, grads, s);
  });
  if (st != FF_ERR_UNSUPPORTED) return st;
  return launch_mlp_generic_bwd(m, *grads, feats, n, d_out, d_feats, s);
}

int ff_field_fwd(const ff_field_desc* desc, const ff_mlp_desc* mlp, const float* x, int64_t n,
                 float* out, void* stream) {
  int st = check_desc(desc);
  if (st) return st;
  if ((st = check_mlp(mlp))) return st;
  if (n < 0 || (n > 0 && (!x || !out))) return FF_ERR_ARG;
  if (n == 0) return FF_OK;
  hipStream_t s = (hipStream_t)stream;
  const ff_field_desc& d = *desc;
  const ff_mlp_desc& m = *mlp;
  if (dict_supported(d, m)) return launch_dict_fwd(d, m, x, n, out, s);
  return with_field(d, [&](auto Di, auto CKi, auto BFi) {
    constexpr int D = decltype(Di)::value, CK = decltype(CKi)::value;
    constexpr bool BF = decltype(BFi)::value;
    return with_mlp(m, d.n_channels, [&](auto NLi, auto Hi, auto Oi) {
      return launch_field_fwd<D, CK, BF, decltype(Hi)::value, decltype(Oi)::value,
                              decltype(NLi)::value>(d, m, x, n, out, s);
    });
  });
}

int ff_field_fused_supported(const ff_field_desc* desc, const ff_mlp_desc* mlp) {
  if (check_desc(desc) || check_mlp(mlp)) return 0;
  if (dict_supported(*desc, *mlp)) return 1;
  const int st = with_field(*desc, [&](auto, auto, auto) {
    return with_mlp(*mlp, desc->n_channels, [&](auto, auto, auto) { return (int)FF_OK; });
  });
  return st == FF_OK ? 1 : 0;
}

size_t ff_field_bwd_workspace_bytes(const ff_field_desc* desc, const ff_mlp_desc* mlp, int64_t n) {
  if (check_desc(desc) || check_mlp(mlp) || n <= 0) return 0;
  if (dict_supported(*desc, *mlp)) return dict_workspace_bytes(*desc, *mlp, n);
  return ff_linear_mat_bwd_workspace_bytes(mlp, n);
}

int ff_field_bwd(const ff_field_desc* desc, const ff_mlp_desc* mlp, const float* x, int64_t n,
                 const float* d_out, const ff_grads* grads, void* ws, size_t ws_bytes,
                 void* stream) {
  return field_bwd_common(desc, mlp, x, n, d_out, 0, 1.f, grads, nullptr, ws, ws_bytes, stream);
}

int ff_field_train_regression(const ff_field_desc* desc, const ff_mlp_desc* mlp, const float* x,
                              const float* target, int64_t n, int loss_kind, float loss_scale,
                              const ff_grads* grads, float* loss_sum, void* ws, size_t ws_bytes,
                              void* stream) {
  if (loss_kind != FF_LOSS_MSE && loss_kind != FF_LOSS_L1) return FF_ERR_ARG;
  return field_bwd_common(desc, mlp, x, n, target, loss_kind == FF_LOSS_MSE ? 1 : 2, loss_scale,
                          grads, loss_sum, ws, ws_bytes, stream);
}

}  // extern "C"
