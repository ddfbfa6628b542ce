/*
 * ffields.h -- C ABI of the MI355X (gfx950) field-evaluation and
 * volume-rendering library (libffields.so).
 *
 * The reference (autonomousvision/factor-fields) has no native layer: its hot
 * path is a chain of ATen calls inside the Python module
 * models/FactorFields.py.  Each entry point below replaces one of those
 * Python-level operations; the comment on each names the reference function
 * (file:line) whose semantics it implements.
 *
 * Conventions
 *   - Plain C11 types only.  All array arguments are DEVICE pointers.
 *   - Every entry point enqueues on `stream` (a hipStream_t passed as void*)
 *     and never synchronises the host.  Argument errors are reported
 *     synchronously as a nonzero status (see ff_status_str); launch errors as
 *     FF_ERR_LAUNCH.
 *   - The caller owns every buffer; the library keeps no allocation and no
 *     global mutable state.  Workspaces are passed in; size them with the
 *     matching *_workspace_bytes query.  The one exception is the
 *     allocation-free ff_render_field_bwd_noalloc, whose scratch is a static
 *     device array (see there).
 *   - Tables are texel-major: a grid of spatial size (D,)H,W with C channels is
 *     stored as [(D,)H,W,C] (the channels_last layout of the reference's
 *     [1,C,(D,)H,W] parameter), so one corner is one contiguous C-vector.
 *   - Gradient buffers are FP32 and are ACCUMULATED into (+=); zero them first.
 */
#ifndef FFIELDS_H
#define FFIELDS_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FF_MAX_LEVELS 8

enum ff_status {
  FF_OK = 0,
  FF_ERR_ARG = 1,        /* bad argument (null pointer, size, unsupported combination) */
  FF_ERR_UNSUPPORTED = 2,/* configuration without a kernel instantiation */
  FF_ERR_WORKSPACE = 3,  /* workspace missing or too small */
  FF_ERR_LAUNCH = 4,     /* hipGetLastError after the launch was not hipSuccess */
  FF_ERR_SCRATCH_LIMIT = 5 /* an allocation-free entry point would need more than its static scratch */
};

/* models/FactorFields.py:11-33 (grid_mapping) */
enum ff_mapping { FF_MAP_SAWTOOTH = 0, FF_MAP_TRIANGLE = 1, FF_MAP_SINC = 2, FF_MAP_X = 3 };
/* grid_sample mode: cfg.model.basis_mode / coef_mode */
enum ff_interp { FF_INTERP_BILINEAR = 0, FF_INTERP_NEAREST = 1 };
/* cfg.model.basis_type */
enum ff_basis_kind { FF_BASIS_NONE = 0, FF_BASIS_GRID = 1 };
/* cfg.model.coeff_type: GRID = models/FactorFields.py:432-434; VM = sum of
 * plane (x) line products over the three axis pairs (TensoRF-style, the
 * combination the reference cannot form with a grid basis) */
enum ff_coeff_kind { FF_COEFF_NONE = 0, FF_COEFF_GRID = 1, FF_COEFF_VM = 2 };
enum ff_dtype { FF_F32 = 0, FF_BF16 = 1 };

/*
 * Field description: everything get_coeff/get_basis need
 * (models/FactorFields.py:425-516 with the shapes from setup_params :262-307).
 */
typedef struct ff_field_desc {
  int32_t in_dim;        /* spatial dims of the basis grids (2 or 3) */
  int32_t x_dim;         /* coordinates per point: in_dim, or in_dim+1 in 'images' mode */
  int32_t n_levels;      /* L = len(basis_dims) */
  int32_t n_channels;    /* C = sum(basis_dims) */
  int32_t mapping;       /* enum ff_mapping */
  int32_t basis_kind;    /* enum ff_basis_kind */
  int32_t basis_interp;  /* enum ff_interp */
  int32_t coeff_kind;    /* enum ff_coeff_kind */
  int32_t coeff_interp;  /* enum ff_interp */
  int32_t coeff_ndim;    /* spatial dims of the coefficient grid (2 or 3) */
  int32_t table_dtype;   /* enum ff_dtype of basis/coeff tables (grads are always FP32) */
  int32_t basis_dims[FF_MAX_LEVELS];
  int32_t basis_reso[FF_MAX_LEVELS];
  int32_t coeff_reso[3]; /* extents along x, y, z (W, H, D) */
  float map_origin[3];   /* aabb[0] used by grid_mapping */
  float map_scale[FF_MAX_LEVELS]; /* FP32 max(aabb[1]-aabb[0]) / freq_bands[l] */
  float coeff_origin[3]; /* aabb[0] used by normalize_coord */
  float coeff_inv[3];    /* FP32 2 / (aabb[1]-aabb[0]) */
  const void* basis[FF_MAX_LEVELS]; /* level tables [R^in_dim, d_l] */
  const void* coeff[6];  /* GRID: [0]; VM: planes xy,xz,yz in [0..2], lines z,y,x in [3..5] */
} ff_field_desc;

/* models/FactorFields.py:113-159 (MLPMixer = linear_mat), pe = 0.
 * n_layers 1: out = W0 x.  n_layers 2: out = W1 relu(W0 x + b0).
 * Weights in torch.nn.Linear layout [out, in], FP32. */
typedef struct ff_mlp_desc {
  int32_t n_layers;
  int32_t in_dim;
  int32_t hidden;
  int32_t out_dim;
  const float* w0;
  const float* b0;
  const float* w1;
} ff_mlp_desc;

/* FP32 gradient accumulators matching ff_field_desc / ff_mlp_desc.
 * replicas > 1: every table gradient pointer is the first of `replicas`
 * copies spaced `replica_stride` floats apart (one flat arena), and each wave
 * of a scatter kernel adds into the copy of the XCD it runs on, so no cache
 * line is shared between the XCDs' L2 caches.  The copies are summed by
 * ff_reduce_replicas or by ff_adam_multi.  The decoder gradients w0/b0/w1
 * are never replicated. */
typedef struct ff_grads {
  float* basis[FF_MAX_LEVELS];
  float* coeff[6];
  float* w0;
  float* b0;
  float* w1;
  int32_t replicas;
  int64_t replica_stride;
} ff_grads;

const char* ff_status_str(int status);
int ff_version(void);

/* ---------------------------------------------------------------- field --
 * ff_get_coding: models/FactorFields.py:523-533 (get_coding), which calls
 * get_coeff :425-465 and get_basis :467-516.  x [n, x_dim].  Outputs (any may
 * be NULL): feats [n, C] = basis * coeff, coeff [n, C], basis [n, C].
 */
int ff_get_coding(const ff_field_desc* desc, const float* x, int64_t n,
                  float* feats, float* coeff, float* basis, void* stream);

/* Backward of ff_get_coding w.r.t. the tables.  d_feats [n, C] (may be NULL),
 * d_coeff [n, C] (gradient reaching the returned coeff; may be NULL),
 * d_basis [n, C] (gradient reaching a separately returned basis; may be NULL).
 * Interpolation weights are recomputed from x. */
int ff_get_coding_bwd(const ff_field_desc* desc, const float* x, int64_t n,
                      const float* d_feats, const float* d_coeff, const float* d_basis,
                      const ff_grads* grads, void* stream);

/* ---------------------------------------------------------- linear_mat --
 * models/FactorFields.py:144-159 (MLPMixer.forward).  feats [n, in], out [n, out].
 */
int ff_linear_mat_fwd(const ff_mlp_desc* mlp, const float* feats, int64_t n, float* out,
                      void* stream);
size_t ff_linear_mat_bwd_workspace_bytes(const ff_mlp_desc* mlp, int64_t n);
/* d_out [n, out] -> d_feats [n, in] (may be NULL) and accumulated weight grads. */
int ff_linear_mat_bwd(const ff_mlp_desc* mlp, const float* feats, int64_t n, const float* d_out,
                      float* d_feats, const ff_grads* grads, void* ws, size_t ws_bytes,
                      void* stream);

/* ---------------------------------------------------- fused field path --
 * get_coding followed by linear_mat in one kernel (the feats never leave
 * registers): the reference's `linear_mat(get_coding(x)[0])`
 * (scripts/2D_regression.ipynb:190-192, scripts/sdf_regression.ipynb:369-370,
 * models/FactorFields.py:874-875).  out [n, out_dim].
 */
int ff_field_fwd(const ff_field_desc* desc, const ff_mlp_desc* mlp, const float* x, int64_t n,
                 float* out, void* stream);
/* 1 if ff_field_fwd / ff_field_bwd / ff_field_train_regression have a kernel
 * for this (field, decoder) pair, else 0 (callers then compose ff_get_coding
 * and ff_linear_mat). */
int ff_field_fused_supported(const ff_field_desc* desc, const ff_mlp_desc* mlp);
size_t ff_field_bwd_workspace_bytes(const ff_field_desc* desc, const ff_mlp_desc* mlp, int64_t n);
/* Backward of ff_field_fwd given d_out [n, out_dim]; recomputes the forward. */
int ff_field_bwd(const ff_field_desc* desc, const ff_mlp_desc* mlp, const float* x, int64_t n,
                 const float* d_out, const ff_grads* grads, void* ws, size_t ws_bytes,
                 void* stream);

/* Regression training pass: forward, loss and backward fused
 * (scripts/sdf_regression.ipynb:369-379, scripts/2D_regression.ipynb:190-206).
 * loss_kind 0 = mean squared error, 1 = mean absolute error, both averaged
 * over n*out_dim and multiplied by loss_scale.  target [n, out_dim].
 * loss_sum (FP32 scalar, device) accumulates sum(err^2) or sum|err| unscaled,
 * may be NULL. */
enum ff_loss { FF_LOSS_MSE = 0, FF_LOSS_L1 = 1 };
int ff_field_train_regression(const ff_field_desc* desc, const ff_mlp_desc* mlp, const float* x,
                              const float* target, int64_t n, int loss_kind, float loss_scale,
                              const ff_grads* grads, float* loss_sum, void* ws, size_t ws_bytes,
                              void* stream);

/* ----------------------------------------------------------- optimiser --
 * torch.optim.Adam step (train_per_scene.py:132, betas (0.9, 0.99), eps 1e-8,
 * no weight decay) fused over one flat parameter group: param/grad/exp_avg/
 * exp_avg_sq [n] FP32; step is the 1-based step count after increment.
 * If shadow_bf16 is non-NULL the updated parameter is also written there as
 * BF16 (round to nearest even).  If zero_grad != 0 the grad is cleared. */
int ff_adam_step(float* param, float* grad, float* exp_avg, float* exp_avg_sq,
                 uint16_t* shadow_bf16, int64_t n, float lr, float beta1, float beta2, float eps,
                 int64_t step, int zero_grad, void* stream);

/* The same Adam step over several tensors (e.g. every parameter of
 * get_optparam_groups, models/FactorFields.py:541-554) in one launch; each
 * tensor carries its group's learning rate. */
typedef struct ff_adam_tensor {
  float* param;
  float* grad;
  float* exp_avg;
  float* exp_avg_sq;
  uint16_t* shadow_bf16; /* may be NULL */
  int64_t n;
  float lr;
  int32_t grad_replicas;  /* > 1: grad is the first of this many copies (summed here) */
  int64_t grad_stride;    /* floats between copies */
} ff_adam_tensor;
int ff_adam_multi(const ff_adam_tensor* tensors, int count, float beta1, float beta2, float eps,
                  int64_t step, int zero_grad, void* stream);

/* arena[i] = sum_r arena[r * stride + i] for i < n, copies 1.. cleared. */
int ff_reduce_replicas(float* arena, int32_t replicas, int64_t n, int64_t stride, void* stream);

/* FP32 -> BF16 (round to nearest even) copy, used to refresh BF16 table shadows. */
int ff_cast_bf16(const float* src, uint16_t* dst, int64_t n, void* stream);


/* ============================================================ rendering ==
 * Ray sampling with occupancy skipping: models/FactorFields.py:586-602
 * (sample_point), :91-110 (AlphaGridMask.sample_alpha) and the masking of
 * FactorFields.forward :859-867.  Surviving samples are compacted ray-major
 * into a packed list (the reference's boolean indexing xyz_sampled[ray_valid],
 * :874); the sample count per ray and the candidate index of every sample
 * reproduce the reference's masks exactly. */
typedef struct ff_ray_sampler {
  float aabb_min[3];     /* model.aabb[0] */
  float aabb_max[3];     /* model.aabb[1] */
  float step_size;       /* FP32 stepSize (update_renderParams :693-699) */
  int32_t n_samples;     /* candidates per ray (N_samples / nSamples), <= 2048 */
  float distance_scale;  /* cfg.renderer.distance_scale */
  const uint8_t* occ;    /* occupancy volume [Dz, Hy, Wx] of 0/1, NULL = no AlphaGridMask */
  int32_t occ_reso[3];   /* Wx, Hy, Dz */
  float occ_origin[3];   /* alphaMask.aabb[0] */
  float occ_inv[3];      /* FP32 1.0 / (aabb[1]-aabb[0]) * 2 (AlphaGridMask.invgridSize) */
} ff_ray_sampler;

/* Packed sample list (structure of arrays, capacity cap >= n_rays * n_samples). */
typedef struct ff_samples {
  float* xyz;            /* [cap, 3] sample positions */
  float* z;              /* [cap] t along the ray (z_vals) */
  float* dist;           /* [cap] (t[i+1]-t[i]) * distance_scale, 0 for the last candidate */
  int32_t* ray;          /* [cap] ray index */
  int32_t* idx;          /* [cap] candidate index along the ray */
  int32_t* ray_off;      /* [n_rays + 1] exclusive offsets; ray_off[n_rays] = total */
  int64_t cap;
  float* coding;         /* optional [2C, cap], channel-major: the basis values (rows 0..C-1)
                          * and coefficient values (rows C..2C-1) of every sample.  When
                          * non-NULL (BF16 tables, or FP32 tables with a grid coefficient),
                          * ff_render_field_fwd writes them and ff_render_field_bwd reads them
                          * instead of gathering again, so both calls must see the same
                          * samples and tables (the autograd saved features of the
                          * reference's get_coding, models/FactorFields.py:523-533).  With
                          * FP32 tables the backward overwrites them with their gradients.
                          * NULL (or a VM coefficient): the backward gathers again. */
} ff_samples;

/* rays [n_rays, 6] = (origin, direction); jitter [n_rays] U[0,1) per ray
 * (the reference's torch.rand_like(rng[:, [0]]), :595) or NULL for
 * is_train=False.  The total sample count stays on the device. */
int ff_sample_rays(const ff_ray_sampler* sampler, const float* rays, const float* jitter,
                   int64_t n_rays, const ff_samples* out, void* stream);

/* Density / colour head of the NeRF field (models/FactorFields.py:639-643,
 * basis2density).  color = FF_COLOR_SH: degree-2 spherical harmonics of
 * models/sh.py:34-64 on decoder outputs 1..27 followed by a sigmoid (decoder
 * out_dim 28).  color = FF_COLOR_FEATURES: the decoder outputs 1..out_dim-1
 * are the features the reference hands to its colour module
 * (feat[app_mask, 1:], models/FactorFields.py:883-885; MLPRender_Fea below),
 * with out_dim 32. */
enum ff_color_kind { FF_COLOR_SH = 0, FF_COLOR_FEATURES = 1 };
typedef struct ff_render_head {
  int32_t density_act;   /* 0 softplus, 1 relu */
  float density_shift;   /* cfg.renderer.density_shift */
  int32_t color;         /* enum ff_color_kind */
  int32_t feat_stride;   /* FEATURES: floats between the feature rows below (>= out_dim - 1) */
  float* feat;           /* FEATURES, forward: [cap, feat_stride] rows written per sample */
  const float* d_feat;   /* FEATURES, backward: rows of dL/dfeatures, indexed by app_pos */
  const int32_t* app_pos;/* FEATURES, backward: [cap] row of each sample in d_feat, -1 = no
                          * colour gradient (not an app sample) */
} ff_render_head;

/* Field + head over the packed samples (count read on the device):
 * sigma [cap], and rgb [cap, 3] (FF_COLOR_SH; ignored with FF_COLOR_FEATURES,
 * which writes head->feat instead).  rays supplies each sample's view
 * direction. */
int ff_render_field_fwd(const ff_field_desc* desc, const ff_mlp_desc* mlp,
                        const ff_render_head* head, const float* rays, int64_t n_rays,
                        const ff_samples* s, float* sigma, float* rgb, void* stream);
size_t ff_render_field_bwd_workspace_bytes(const ff_field_desc* desc, const ff_mlp_desc* mlp,
                                           int64_t cap);
/* d_sample [cap, 4] = (dL/dsigma, dL/drgb[3]) per sample -> table and decoder
 * grads.  With FF_COLOR_FEATURES only dL/dsigma is read from d_sample; the
 * feature gradients come from head->d_feat / head->app_pos. */
int ff_render_field_bwd(const ff_field_desc* desc, const ff_mlp_desc* mlp,
                        const ff_render_head* head, const float* rays, int64_t n_rays,
                        const ff_samples* s, const float* d_sample, const ff_grads* grads, void* ws,
                        size_t ws_bytes, void* stream);

/* Allocation-free variant of ff_render_field_bwd.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc).  The
 * decoder-gradient partial rows that ff_render_field_bwd keeps in the caller's
 * workspace live instead in a static device array of
 * ff_render_field_bwd_noalloc_limit_bytes() = 16 MiB, one per GPU, held from
 * the library's first use on that GPU until the process ends; no allocation
 * happens in the call.  Requests whose workspace
 * (ff_render_field_bwd_workspace_bytes) exceeds it return
 * FF_ERR_SCRATCH_LIMIT.  Same kernels and results as ff_render_field_bwd.
 * Every call on one GPU shares that GPU's array: calls that overlap in time on
 * one GPU (e.g. on two streams) race; serialise them. */
size_t ff_render_field_bwd_noalloc_limit_bytes(void);
int ff_render_field_bwd_noalloc(const ff_field_desc* desc, const ff_mlp_desc* mlp,
                                const ff_render_head* head, const float* rays, int64_t n_rays,
                                const ff_samples* s, const float* d_sample, const ff_grads* grads,
                                void* stream);

/* Alpha compositing, one wavefront per ray with a multiplicative prefix scan
 * for the transmittance: raw2alpha :82-88 and the composite of :878-896
 * (app mask weight > weight_thres, white background, clamp).  Outputs:
 * rgb_map [n_rays, 3] (clamped), rgb_raw [n_rays, 3] (before the clamp),
 * depth [n_rays], acc [n_rays], weights [cap] (may be NULL). */
int ff_composite_fwd(const ff_samples* s, int64_t n_rays, const float* sigma, const float* rgb,
                     int white_bg, float weight_thres, float* rgb_map, float* rgb_raw, float* depth,
                     float* acc, float* weights, void* stream);
/* Backward of ff_composite_fwd by a reverse affine scan (no stored
 * transmittance).  Either d_rgb_map [n_rays, 3] is given, or target
 * [n_rays, 3] (then the loss is mean((rgb_map - target)^2) * loss_scale,
 * train_per_scene.py:158, and loss_sum accumulates sum((rgb_map-target)^2)).
 * Writes d_sample [cap, 4]. */
int ff_composite_bwd(const ff_samples* s, int64_t n_rays, const float* sigma, const float* rgb,
                     int white_bg, float weight_thres, const float* rgb_raw,
                     const float* d_rgb_map, const float* target, float loss_scale,
                     float* loss_sum, float* d_sample, void* stream);

/* App samples of FactorFields.forward (models/FactorFields.py:878-885): the
 * samples whose compositing weight (raw2alpha :82-88 on sigma and s->dist)
 * exceeds weight_thres, the only ones the colour module evaluates.  One
 * wavefront per ray with the arithmetic of ff_composite_fwd, so both agree on
 * every sample.  Outputs (device): app_off [n_rays + 1] exclusive offsets per
 * ray (app_off[n_rays] = number of app samples), rows [cap] the app samples'
 * indices in ray-major order, app_pos [cap] (may be NULL) the position of
 * each live sample in rows or -1. */
int ff_select_app(const ff_samples* s, int64_t n_rays, const float* sigma, float weight_thres,
                  int32_t* app_off, int32_t* rows, int32_t* app_pos, void* stream);

/* ====================================================== colour module ==
 * MLPRender_Fea (models/FactorFields.py:162-203): the input is
 * [features, viewdir, PE(features, fea_pe), PE(viewdir, view_pe)] with
 * positional_encoding :74-79 (channel-major sin block, then the cos block),
 * followed by Linear(in_mlpC, hidden) + ReLU, Linear(hidden, hidden) + ReLU,
 * Linear(hidden, 3, no bias) and a sigmoid.  FP32 weights in torch layout.
 * Built for in_feat 31, fea_pe 2, view_pe 6 or 2, hidden 128, 3 layers (the
 * reference's configs); FF_ERR_UNSUPPORTED otherwise. */
typedef struct ff_color_mlp_desc {
  int32_t in_feat;       /* inChanel */
  int32_t view_pe;
  int32_t fea_pe;
  int32_t hidden;
  int32_t n_layers;
  const float* w0;       /* [hidden, in_mlpC] */
  const float* b0;       /* [hidden] */
  const float* w1;       /* [hidden, hidden] */
  const float* b1;       /* [hidden] */
  const float* w2;       /* [3, hidden] */
} ff_color_mlp_desc;

/* Which rows the colour module evaluates.  Row j (j < count) is sample
 * r = rows ? rows[j] : j; its features are feat[r * feat_stride + 0 ..
 * in_feat-1] and its view direction dirs[d * dir_stride + 0..2] with
 * d = dir_index ? dir_index[r] : r.  count is read on the device when
 * count_dev is given (capped at n_max), else count = n_max. */
typedef struct ff_color_rows {
  const float* feat;
  int64_t feat_stride;
  const float* dirs;
  int64_t dir_stride;
  const int32_t* dir_index;
  const int32_t* rows;
  const int32_t* count_dev;
  int64_t n_max;
} ff_color_rows;

typedef struct ff_color_mlp_grads {
  float* w0;
  float* b0;
  float* w1;
  float* b1;
  float* w2;
} ff_color_mlp_grads;

size_t ff_color_mlp_workspace_bytes(const ff_color_mlp_desc* m, int64_t n_max, int backward);
/* rgb of row j written at rgb[r * rgb_stride + 0..2] (r as above). */
int ff_color_mlp_fwd(const ff_color_mlp_desc* m, const ff_color_rows* r, float* rgb, int64_t rgb_stride,
                     void* ws, size_t ws_bytes, void* stream);
/* Backward: dL/drgb of row j read at d_rgb[r * d_rgb_stride + 0..2]; the
 * feature gradient of row j WRITTEN (not accumulated) at
 * d_feat[j * d_feat_stride + 0 .. in_feat-1] (row order j, so a packed
 * sample list gets a compact block); the weight gradients ACCUMULATED into g
 * after a fixed-order reduction.  The forward is recomputed. */
int ff_color_mlp_bwd(const ff_color_mlp_desc* m, const ff_color_rows* r, const float* d_rgb,
                     int64_t d_rgb_stride, float* d_feat, int64_t d_feat_stride,
                     const ff_color_mlp_grads* g, void* ws, size_t ws_bytes, void* stream);

/* ============================================================= dropout ==
 * torch.nn.functional.dropout(x, p) as linear_mat applies it to its input
 * when with_dropout and is_train (models/FactorFields.py:150-151):
 * out = x * keep / (1 - p).  keep comes from `keep` [n] (0/1, may be NULL)
 * or, when NULL, from a counter-based generator (the Squares RNG, Widynski
 * 2020) keyed by seed at counters offset .. offset + n - 1, kept when its
 * uniform draw in [0, 1) is below 1 - p.  keep_out [n] (may be NULL) receives
 * the mask.  The backward is the same call on the output gradient with the
 * same mask (or seed and offset). */
int ff_dropout(const float* x, int64_t n, float p, const uint8_t* keep, uint64_t seed, uint64_t offset,
               float* out, uint8_t* keep_out, void* stream);

/* Occupancy-grid maintenance (models/FactorFields.py:710-793).
 * ff_render_alpha: compute_alpha :710-727 over n points: alpha_accum[i] +=
 * 1 - exp(-sigma(x_i) * length), where sigma is 0 for points the current
 * occupancy volume (sampler->occ, tested with > 0 as the reference does; may
 * be NULL) marks empty.  Only the occupancy fields of `sampler` are used. */
int ff_render_alpha(const ff_field_desc* desc, const ff_mlp_desc* mlp, const ff_render_head* head,
                    const ff_ray_sampler* sampler, const float* xyz, int64_t n, float length,
                    float* alpha_accum, void* stream);
/* max_pool3d(kernel 3, stride 1, padding 1) of alpha [D, H, W] followed by
 * `>= thres` (updateAlphaMask :764-772 before the small-object filter):
 * out [D, H, W] uint8. alpha is divided by `times` and clamped to [0, 1]
 * first. */
int ff_occupancy_pool(const float* alpha, int32_t D, int32_t H, int32_t W, float times, float thres,
                      uint8_t* out, void* stream);

/* ========================================================= proposal path ==
 * Learned proposal sampler (no counterpart in the reference's training loop;
 * its resampling follows dataLoader/ray_utils.py:132-174 sample_pdf).  A
 * low-resolution density grid [Dz, Hy, Wx, 8] (trilinear, align_corners=False,
 * border) feeds an 8 -> 32 -> 1 MLP (ReLU, softplus(out + shift)); n_prop = 64
 * stratified samples over the ray/box interval are composited into weights,
 * and sample_pdf places n_fine samples along the ray.  One wavefront per ray,
 * one lane per proposal sample. */
typedef struct ff_proposal_desc {
  int32_t reso[3];        /* grid extents x, y, z */
  int32_t table_dtype;    /* FF_F32 or FF_BF16 */
  const void* grid;       /* [Dz, Hy, Wx, 8] texel-major */
  const float* w0;        /* [32, 8] */
  const float* b0;        /* [32] */
  const float* w1;        /* [1, 32] */
  float aabb_min[3];
  float aabb_max[3];
  float density_shift;
  float distance_scale;
  float near;             /* smallest t (sample_point's 0.05) */
  int32_t n_fine;         /* fine samples per ray, <= 64 */
} ff_proposal_desc;

/* Fine samples of every ray into `out` (ray-major; n_fine per ray that meets
 * the box, 0 otherwise).  prop_u [n_rays, 64] jitters the proposal samples in
 * their bins and fine_u [n_rays, n_fine] the sample_pdf draws (stratified:
 * u_k = (k + fine_u)/n_fine); NULL means the bin midpoints and
 * linspace(0, 1, n_fine) (sample_pdf det=True). */
int ff_proposal_sample(const ff_proposal_desc* p, const float* rays, const float* prop_u,
                       const float* fine_u, int64_t n_rays, const ff_samples* out, void* stream);

/* Backward of the proposal: interlevel bound loss between the proposal
 * weights and the field's fine weights (weights [cap] from ff_composite_fwd,
 * treated as constants): for each fine interval [t_k, t_k+1], bound_k = sum of
 * the proposal weights of the bins it overlaps and
 * loss_k = max(0, w_k - bound_k)^2 / (w_k + 1e-7), averaged over rays and
 * scaled by loss_scale.  Accumulates grid / MLP gradients (grid_grad with
 * `replicas` per-XCD copies spaced `replica_stride` floats; w0/b0/w1 grads
 * FP32, summed deterministically through the workspace) and loss_sum.
 * ws must hold ff_proposal_bwd_workspace_bytes(n_rays) bytes
 * (FF_ERR_WORKSPACE otherwise).  No counterpart in the reference (its
 * training loop has no proposal network). */
typedef struct ff_proposal_grads {
  float* grid;
  float* w0;
  float* b0;
  float* w1;
  int32_t replicas;
  int64_t replica_stride;
} ff_proposal_grads;
size_t ff_proposal_bwd_workspace_bytes(int64_t n_rays);
int ff_proposal_bwd(const ff_proposal_desc* p, const float* rays, const float* prop_u,
                    int64_t n_rays, const ff_samples* fine, const float* weights, float loss_scale,
                    const ff_proposal_grads* grads, float* loss_sum, void* ws, size_t ws_bytes,
                    void* stream);

/* The degree-2 SH colour head on its own (the render kernels evaluate it
 * inside ff_render_field_fwd): rgb[i, c] = sigmoid(eval_sh(2, sh[i, 9c..9c+8],
 * dirs[i])) with models/sh.py:34-85's polynomials (eval_sh, deg 2) and
 * MLPRender_Fea's output sigmoid (models/FactorFields.py:202).  Rows of dirs
 * and sh are dir_stride / sh_stride floats apart (>= 3 / >= 27); rgb [n, 3].
 * ff_sh_rgb_bwd writes d_sh [n, 27] (dense) from d_rgb [n, 3]. */
int ff_sh_rgb(const float* dirs, int64_t dir_stride, const float* sh, int64_t sh_stride, int64_t n, float* rgb,
              void* stream);
int ff_sh_rgb_bwd(const float* dirs, int64_t dir_stride, const float* sh, int64_t sh_stride, int64_t n,
                  const float* d_rgb, float* d_sh, void* stream);

/* ------------------------------------------------------------ camera rays
 * World-space rays of pinhole cameras, one thread per ray: the camera-space
 * direction through the pixel centre (x + 0.5, y + 0.5),
 *   ((x + 0.5 - cx) / fx, (y + 0.5 - cy) / fy, 1)      (get_ray_directions,
 *                                                      dataLoader/ray_utils.py:27-45)
 *   ((x + 0.5 - cx) / fx, -(y + 0.5 - cy) / fy, -1)    (blender = 1,
 *                                                      get_ray_directions_blender, :48-66)
 * optionally divided by its norm (normalize = 1: the blender loader's
 * directions / torch.norm(directions), dataLoader/blender.py:54-55), then
 * rotated by c2w[:3, :3] (rays_d = d @ c2w[:3,:3].T) with the origin
 * c2w[:3, 3] (get_rays, ray_utils.py:69-90).  rays [n, 6] = (o, d).
 * view/pix: n (camera, row-major pixel) pairs; both NULL: every pixel of
 * every camera in order (n must be n_cams * H * W), the concatenation the
 * blender loader builds (blender.py:86-96).  A pair outside the cameras or
 * the image gives a NaN ray. */
typedef struct ff_cameras {
  const float* c2w;   /* [n_cams, 3, 4] row-major camera-to-world */
  int32_t n_cams;
  int32_t H, W;
  float fx, fy;       /* focal lengths in pixels */
  float cx, cy;       /* principal point (the reference's default: W / 2, H / 2) */
  int32_t blender;
  int32_t normalize;
} ff_cameras;
int ff_camera_rays(const ff_cameras* cams, const int64_t* view, const int64_t* pix, int64_t n, float* rays,
                   void* stream);

/* Training batch gather (train_per_scene.py:151-152: allrays[ray_idx],
 * allrgbs[ray_idx]): out_rays[b] = rays[idx[b]] ([*, 6]) and, when rgbs and
 * out_rgbs are non-NULL, out_rgbs[b] = rgbs[idx[b]] ([*, 3]), b < B.  Exact
 * copies.  An index outside [0, n_total) gives NaN rows (nothing is read for
 * it). */
int ff_gather_rays(const float* rays, const float* rgbs, int64_t n_total, const int64_t* idx, int64_t B,
                   float* out_rays, float* out_rgbs, void* stream);

/* The proposal referee's squared error (no counterpart in the reference,
 * which has no proposal network): out[0] += sum_i (a[i] - b[i])^2 over n
 * floats, one block summing in a fixed order.  Used for the per-batch
 * comparison of the occupancy sampler's and the proposal's renders of the
 * same rays. */
int ff_sq_err_sum(const float* a, const float* b, int64_t n, float* out, void* stream);

/* The small-object filter of updateAlphaMask: skimage.morphology.
 * remove_small_objects(mask, min_size, connectivity=1) at
 * models/FactorFields.py:774-775.  mask [D, H, W] uint8 (non-zero = set);
 * out[i] = 1 where voxel i is set and its face-connected component has at
 * least min_size voxels, 0 elsewhere.  labels and sizes are caller-given
 * int32 work buffers of D * H * W entries; on return labels[i] is the
 * smallest index in voxel i's component (-1 for unset voxels) and
 * sizes[labels[i]] that component's voxel count.  D H W <= 2^31 - 1.
 * Five launches on `stream`. */
int ff_remove_small_components(const uint8_t* mask, int32_t D, int32_t H, int32_t W, double min_size,
                               int32_t* labels, int32_t* sizes, uint8_t* out, void* stream);

/* Isosurface extraction of a dense [A, B, C] FP32 volume (row-major, point
 * (i, j, k) at vol[(i * B + j) * C + k]) at `level`, in three launches on
 * `stream`.  Replaces skimage.measure.marching_cubes in
 * convert_sdf_samples_to_ply (utils.py:203-205; called from
 * train_per_scene.py:32-41) by marching tetrahedra: each cube splits into
 * the six tetrahedra around its (0,0,0)-(1,1,1) diagonal, each tetrahedron
 * straddling the level gives one or two triangles.  Inside is value > level
 * when inside_above != 0 (densities), value < level otherwise (distances).
 * A, B, C >= 2, each <= 2^20, (A B C)^2 < 9e18.
 *
 * ff_mt_count: counts[c] = triangles of cube c, c = (i * (B-1) + j) * (C-1)
 *   + k, for the (A-1)(B-1)(C-1) cubes.
 * ff_mt_emit: offsets[c] = exclusive prefix sum of counts; for each cube c
 *   of the list cubes[0 .. n_cubes) (every cube when cubes is NULL; listing
 *   only the cubes with triangles saves the others' threads) writes three
 *   edge keys per triangle at keys[3 * offsets[c] ...]: key = lo * (A B C)
 *   + hi for the edge between grid points lo < hi, ordered so that the
 *   normal points from the tetrahedron's inside corners to its outside ones.
 * ff_mt_verts: for n distinct keys, verts[i] = the FP64 point in index
 *   coordinates where the level crosses edge keys[i] (linear in the two
 *   endpoint values, from the lower point).  The caller deduplicates the
 *   emitted keys (a sort) and maps every triangle corner to its key's rank. */
int ff_mt_count(const float* vol, int64_t A, int64_t B, int64_t C, double level, int32_t inside_above,
                int32_t* counts, void* stream);
int ff_mt_emit(const float* vol, int64_t A, int64_t B, int64_t C, double level, int32_t inside_above,
               const int64_t* cubes, int64_t n_cubes, const int64_t* offsets, int64_t* keys, void* stream);
int ff_mt_verts(const float* vol, int64_t A, int64_t B, int64_t C, double level, const int64_t* keys, int64_t n,
                double* verts, void* stream);

/* Measurement only (no counterpart in the reference): the FP32 atomic
 * line-request rate that bounds every backward's table scatter.  Issues
 * n_instr wave instructions of global_atomic_add_f32 (+1.0), each touching
 * lines_per_instr (4, 16 or 64) distinct 64-byte lines of dst
 * ([n_lines * 16] floats), lines spread by a hash.  Time it with events on
 * `stream`; n_instr * lines_per_instr line requests. */
int ff_atomic_rate_probe(float* dst, int64_t n_lines, int64_t n_instr, int32_t lines_per_instr, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* FFIELDS_H */
