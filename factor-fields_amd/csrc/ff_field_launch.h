// Launcher declarations for the field kernels.  Each fast-path combination is
// explicitly instantiated in its own translation unit (ff_field_inst_*.hip)
// so the library builds in parallel; the C ABI (ff_field_api.hip) only sees
// these declarations.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ffields.h"

namespace ff {

// Internal coefficient kinds (spatial dims resolved).
enum { CK_NONE = 0, CK_GRID2 = 1, CK_GRID3 = 2, CK_VM = 3 };

template <int D, int CK, bool BF>
int launch_coding_fwd(const ff_field_desc& d, const float* x, int64_t n, float* f, float* c,
                      float* b, hipStream_t s);
template <int D, int CK, bool BF>
int launch_coding_bwd(const ff_field_desc& d, const ff_grads& g, const float* x, int64_t n,
                      const float* df, const float* dc, const float* db, hipStream_t s);
template <int D, int CK, bool BF, int H, int O, int NL>
int launch_field_fwd(const ff_field_desc& d, const ff_mlp_desc& m, const float* x, int64_t n,
                     float* out, hipStream_t s);
template <int D, int CK, bool BF, int H, int O, int NL>
int launch_field_bwd(const ff_field_desc& d, const ff_mlp_desc& m, const ff_grads& g, const float* x,
                     int64_t n, const float* dsrc, int mode, float loss_scale, float* ws,
                     unsigned blocks, float* loss_sum, hipStream_t s);
template <int H, int O, int NL>
int launch_mlp_fwd(const ff_mlp_desc& m, const float* f, int64_t n, float* out, hipStream_t s);
template <int H, int O, int NL>
int launch_mlp_bwd(const ff_mlp_desc& m, const float* f, int64_t n, const float* dout, float* dfeat,
                   float* ws, unsigned blocks, hipStream_t s);

// Deterministic reduction of per-wave decoder weight-gradient partials into
// the gradient buffers (ff_field_api.hip).
// Dictionary image-set fast path (ff_dict.hip): image-set mode, basis_dims
// [32,32,32,16,16,16], BF16 tables, decoder 144 -> 64 -> O (O <= 4).
bool dict_supported(const ff_field_desc& d, const ff_mlp_desc& m);
size_t dict_workspace_bytes(const ff_field_desc& d, const ff_mlp_desc& m, int64_t n);
int launch_dict_fwd(const ff_field_desc& d, const ff_mlp_desc& m, const float* x, int64_t n, float* out,
                    hipStream_t s);
int launch_dict_bwd(const ff_field_desc& d, const ff_mlp_desc& m, const ff_grads& g, const float* x, int64_t n,
                    const float* dsrc, int mode, float loss_scale, void* ws, size_t ws_bytes, float* loss_sum,
                    hipStream_t s);

// Deterministic two-level sum of per-wave partial rows [rows][P] (P = decoder
// parameters of an NL-layer IN -> H -> O MLP) into g.w0 / g.b0 / g.w1; the
// workspace holds the rows followed by kReduceSplit rows of scratch.
constexpr int kReduceSplit = 32;
int launch_reduce_partials(const float* ws, int rows, int NL, int IN, int H, int O, const ff_grads& g,
                           hipStream_t s);

// Generic fallbacks (any level dims / decoder shape), ff_field_generic.hip.
int launch_coding_generic(const ff_field_desc& d, const ff_grads& g, int ck, int bwd, const float* x,
                          int64_t n, float* feats, float* coeff, float* basis, const float* dfeat,
                          const float* dcoeff, const float* dbasis, hipStream_t s);
int launch_mlp_generic_fwd(const ff_mlp_desc& m, const float* f, int64_t n, float* out,
                           hipStream_t s);
int launch_mlp_generic_bwd(const ff_mlp_desc& m, const ff_grads& g, const float* f, int64_t n,
                           const float* dout, float* dfeat, hipStream_t s);

}  // namespace ff
