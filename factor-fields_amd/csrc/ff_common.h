// Device helpers shared by the ffields HIP kernels (gfx950, wave64).
//
// Arithmetic that decides masks or must match the reference bit-for-bit is
// written out operation by operation; the library is compiled with
// -ffp-contract=off and fused multiply-adds are requested explicitly (fmaf)
// only where the reference's own GPU path fuses them too.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>
#include <type_traits>
#include <utility>

#include "ffields.h"

namespace ff {

constexpr int kWave = 64;

// Orders a wave's LDS accesses across a hand-off between its lanes (blocks of
// one wave): a wave's LDS operations execute in order, so waiting for its
// outstanding LDS traffic and keeping the compiler from moving memory
// accesses across is enough.  Unlike a block barrier, whose release fence
// waits for every outstanding global store and atomic of the wave, it lets
// the atomics of a scatter drain while the wave computes on.
__device__ __forceinline__ void wave_lds_fence() {
  __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0)
  __builtin_amdgcn_wave_barrier();
  asm volatile("" ::: "memory");
}
constexpr float kPi = 3.14159265358979323846f;
constexpr float kHalfPi = 1.57079632679489661923f;

// ------------------------------------------------------------ static_for
template <int B, int E, class F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(std::forward<F>(f));
  }
}

// ------------------------------------------------------------ vector loads
typedef float f4u __attribute__((ext_vector_type(4), aligned(4)));
typedef float f2u __attribute__((ext_vector_type(2), aligned(4)));
typedef uint32_t u4u __attribute__((ext_vector_type(4), aligned(4)));
typedef uint32_t u2u __attribute__((ext_vector_type(2), aligned(4)));

__device__ __forceinline__ float bf16_lo(uint32_t w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float bf16_hi(uint32_t w) { return __uint_as_float(w & 0xffff0000u); }
__device__ __forceinline__ float bf16_to_f32(uint16_t v) { return __uint_as_float(((uint32_t)v) << 16); }

// Round-to-nearest-even float -> bf16 (NaN kept quiet).
__device__ __forceinline__ uint16_t f32_to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// Load N consecutive channels of a texel into registers (FP32 or BF16 table).
template <int N, bool BF>
__device__ __forceinline__ void load_ch(const void* __restrict__ base, int64_t elem, float* v) {
  if constexpr (!BF) {
    const float* p = reinterpret_cast<const float*>(base) + elem;
    constexpr int n4 = N / 4;
    static_for<0, n4>([&](auto K) {
      constexpr int k = decltype(K)::value;
      f4u t = reinterpret_cast<const f4u*>(p)[k];
      v[4 * k + 0] = t.x; v[4 * k + 1] = t.y; v[4 * k + 2] = t.z; v[4 * k + 3] = t.w;
    });
    constexpr int r = N - 4 * n4;
    if constexpr (r >= 2) {
      f2u t = *reinterpret_cast<const f2u*>(p + 4 * n4);
      v[4 * n4] = t.x; v[4 * n4 + 1] = t.y;
    }
    if constexpr (r == 1 || r == 3) v[N - 1] = p[N - 1];
  } else {
    const uint16_t* p = reinterpret_cast<const uint16_t*>(base) + elem;
    constexpr int n8 = N / 8;
    static_for<0, n8>([&](auto K) {
      constexpr int k = decltype(K)::value;
      u4u t = reinterpret_cast<const u4u*>(p)[k];
      v[8 * k + 0] = bf16_lo(t.x); v[8 * k + 1] = bf16_hi(t.x);
      v[8 * k + 2] = bf16_lo(t.y); v[8 * k + 3] = bf16_hi(t.y);
      v[8 * k + 4] = bf16_lo(t.z); v[8 * k + 5] = bf16_hi(t.z);
      v[8 * k + 6] = bf16_lo(t.w); v[8 * k + 7] = bf16_hi(t.w);
    });
    constexpr int r = N - 8 * n8;
    if constexpr (r >= 4) {
      u2u t = *reinterpret_cast<const u2u*>(p + 8 * n8);
      v[8 * n8 + 0] = bf16_lo(t.x); v[8 * n8 + 1] = bf16_hi(t.x);
      v[8 * n8 + 2] = bf16_lo(t.y); v[8 * n8 + 3] = bf16_hi(t.y);
    }
    constexpr int r4 = r >= 4 ? 8 * n8 + 4 : 8 * n8;
    if constexpr (N - r4 >= 2) {
      uint32_t t = *reinterpret_cast<const uint32_t*>(p + r4);
      v[r4] = bf16_lo(t); v[r4 + 1] = bf16_hi(t);
    }
    if constexpr ((N - r4) % 2 == 1) v[N - 1] = bf16_to_f32(p[N - 1]);
  }
}

// ------------------------------------------------------------ torch semantics
// torch.remainder for floating point: fmod, plus the divisor when the signs differ.
__device__ __forceinline__ float torch_remainder(float a, float b) {
  float m = fmodf(a, b);
  if (m != 0.f && ((b < 0.f) != (m < 0.f))) m += b;
  return m;
}

// torch's floor division (Python //) for floating point.
__device__ __forceinline__ float torch_floor_divide(float a, float b) {
  float mod = fmodf(a, b);
  float div = (a - mod) / b;
  if (mod != 0.f && ((b < 0.f) != (mod < 0.f))) div -= 1.f;
  float fd;
  if (div != 0.f) {
    fd = floorf(div);
    if (div - fd > 0.5f) fd += 1.f;
  } else {
    fd = copysignf(0.f, a / b);
  }
  return fd;
}

// grid_mapping (models/FactorFields.py:11-33) for one axis and one level.
__device__ __forceinline__ float map_coord(int kind, float p, float origin, float scale) {
  const float d = p - origin;
  const float half = scale * 0.5f;  // scale / 2, exact
  if (kind == FF_MAP_SAWTOOTH) {
    float u = torch_remainder(d, scale) / half - 1.f;
    return fminf(fmaxf(u, -1.f), 1.f);
  } else if (kind == FF_MAP_TRIANGLE) {
    float r = torch_remainder(d, scale);
    float q = torch_floor_divide(d, scale);
    float parity = torch_remainder(q, 2.f);
    float u = r / half - 1.f;
    return parity == 1.f ? -u : u;
  } else if (kind == FF_MAP_SINC) {
    return sinf(d / (scale / kPi) - kHalfPi);
  }
  return d / scale;  // FF_MAP_X
}

// grid_sample unnormalisation.  align_corners=True: ((u + 1) / 2) * (R - 1).
__device__ __forceinline__ float unnorm_ac(float u, int R) {
  return ((u + 1.f) * 0.5f) * (float)(R - 1);
}
// align_corners=False with border padding: ((u + 1) * R - 1) / 2 clipped to [0, R-1].
__device__ __forceinline__ float unnorm_border(float u, int R) {
  float v = ((u + 1.f) * (float)R - 1.f) * 0.5f;
  return fminf((float)(R - 1), fmaxf(v, 0.f));
}

// Keep far-out coordinates in an int-safe range; every corner of a clamped
// coordinate is out of bounds exactly when it was before clamping.
__device__ __forceinline__ float safe_coord(float i, int R) {
  return fminf(fmaxf(i, -2.f), (float)R + 1.f);
}

// Bi/trilinear corners in the reference's accumulation order
// (nw, ne, sw, se [, bottom layer]): corner k has x = k&1, y = (k>>1)&1, z = k>>2.
// Out-of-bounds corners get weight 0 and texel 0.
template <int ND>
struct Corners {
  int tex[1 << ND];
  float w[1 << ND];
};

template <int ND>
__device__ __forceinline__ void linear_corners(const float* pos, const int* R, Corners<ND>& c) {
  float f0[ND], w0[ND], w1[ND];
  int i0[ND];
#pragma unroll
  for (int a = 0; a < ND; ++a) {
    float p = safe_coord(pos[a], R[a]);
    f0[a] = floorf(p);
    i0[a] = (int)f0[a];
    w1[a] = p - f0[a];            // ix - ix_nw
    w0[a] = (f0[a] + 1.f) - p;    // ix_se - ix
  }
#pragma unroll
  for (int k = 0; k < (1 << ND); ++k) {
    float w = 1.f;
    bool ok = true;
    int t = 0;
#pragma unroll
    for (int a = ND - 1; a >= 0; --a) {
      const int bit = (k >> a) & 1;
      const int ii = i0[a] + bit;
      ok = ok && ii >= 0 && ii < R[a];
      t = t * R[a] + ii;
    }
    // product order of the reference: ((wx * wy) * wz)
    w = (k & 1) ? w1[0] : w0[0];
#pragma unroll
    for (int a = 1; a < ND; ++a) w = w * (((k >> a) & 1) ? w1[a] : w0[a]);
    c.w[k] = ok ? w : 0.f;
    c.tex[k] = ok ? t : 0;
  }
}

// Nearest texel (round half to even); returns -1 when out of bounds.
template <int ND>
__device__ __forceinline__ int nearest_texel(const float* pos, const int* R) {
  int t = 0;
  bool ok = true;
#pragma unroll
  for (int a = ND - 1; a >= 0; --a) {
    int ii = (int)rintf(safe_coord(pos[a], R[a]));
    ok = ok && ii >= 0 && ii < R[a];
    t = t * R[a] + ii;
  }
  return ok ? t : -1;
}

// Atomic add into a gradient table in device memory.  The cast tells the
// compiler the address is global, so it emits global_atomic_add_f32, which
// counts only against vmcnt.  Through a generic pointer it would emit a FLAT
// atomic, which also counts against lgkmcnt: every later wait for an LDS
// read (lgkmcnt) would then wait for the atomic's round trip to memory too.
typedef __attribute__((address_space(1))) float global_float;
__device__ __forceinline__ void global_atomic_add(float* p, float v) {
  __hip_atomic_fetch_add((global_float*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ------------------------------------------------------------ XCD-local gradients
// XCC_ID hardware register (which XCD this wave runs on).
__device__ __forceinline__ int xcc_id() { return __builtin_amdgcn_s_getreg(20 | (0 << 6) | (3 << 11)) & 0xf; }

// Table-gradient pointers of the replica owned by this wave's XCD.
__device__ __forceinline__ ff_grads local_grads(const ff_grads& g) {
  if (g.replicas <= 1) return g;
  ff_grads r = g;
  const int64_t off = (int64_t)(xcc_id() % g.replicas) * g.replica_stride;
#pragma unroll
  for (int l = 0; l < FF_MAX_LEVELS; ++l)
    if (r.basis[l]) r.basis[l] += off;
#pragma unroll
  for (int k = 0; k < 6; ++k)
    if (r.coeff[k]) r.coeff[k] += off;
  return r;
}

// ------------------------------------------------------------ wave reductions
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

}  // namespace ff
