// Small-object filter of the occupancy update (models/FactorFields.py:774-775:
// skimage.morphology.remove_small_objects(mask, min_size, connectivity=1)):
// label the face-connected components of a [D, H, W] mask, count their
// voxels, and keep the voxels of components with at least min_size voxels.
//
// Labelling is union-find with atomic linking (Playne and Hawick, "A New
// Algorithm for Parallel Connected-Component Labelling on GPUs", IEEE TPDS
// 2018): every label points to a voxel of the same component with an index
// no larger than its own, a root points to itself, and a union links the
// larger root below the smaller one with atomicMin.  A link that loses a race
// returns the label that beat it and the union retries from there; every
// retry lowers a + b, so each union ends after finitely many steps, and no
// lane ever waits on another.
//
//   cc_init:     labels[i] = i for mask voxels (-1 elsewhere), sizes[i] = 0
//   cc_merge:    union of each mask voxel with its -x, -y, -z mask neighbours
//   cc_compress: labels[i] = root of i (the component's smallest index)
//   cc_count:    sizes[root] += 1, one atomic per run of equal roots in a wave
//   cc_filter:   out[i] = mask[i] && sizes[labels[i]] >= min_size
#include "ff_common.h"

namespace ff {
namespace {

// label reads inside cc_merge see other workgroups' links as they land (a
// stale value is an older, larger label of the same component: still valid)
__device__ __forceinline__ int32_t cc_load(const int32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int32_t cc_find(const int32_t* L, int32_t a) {
  int32_t n = cc_load(L + a);
  while (n != a) {
    a = n;
    n = cc_load(L + a);
  }
  return a;
}

__device__ __forceinline__ void cc_union(int32_t* L, int32_t a, int32_t b) {
  for (;;) {
    a = cc_find(L, a);
    b = cc_find(L, b);
    if (a == b) return;
    if (a > b) {
      const int32_t t = a;
      a = b;
      b = t;
    }
    // link root b below a; old == b means b was still a root and now is not
    const int32_t old = atomicMin(L + b, a);
    if (old == b) return;
    b = old;
  }
}

__global__ __launch_bounds__(256) void cc_init_kernel(const uint8_t* __restrict__ mask, int64_t n,
                                                      int32_t* __restrict__ labels, int32_t* __restrict__ sizes) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  labels[i] = mask[i] ? (int32_t)i : -1;
  sizes[i] = 0;
}

__global__ __launch_bounds__(256) void cc_merge_kernel(const uint8_t* __restrict__ mask, int32_t D, int32_t H,
                                                       int32_t W, int32_t* labels) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = (int64_t)D * H * W;
  if (i >= n || !mask[i]) return;
  const int32_t x = (int32_t)(i % W), y = (int32_t)((i / W) % H), z = (int32_t)(i / ((int64_t)W * H));
  const int32_t me = (int32_t)i;
  if (x > 0 && mask[i - 1]) cc_union(labels, me, me - 1);
  if (y > 0 && mask[i - W]) cc_union(labels, me, me - W);
  if (z > 0 && mask[i - (int64_t)W * H]) cc_union(labels, me, me - W * H);
}

__global__ __launch_bounds__(256) void cc_compress_kernel(int64_t n, int32_t* labels) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t l = labels[i];
  if (l < 0 || l == (int32_t)i) return;
  // every write is a root of i's component, so a concurrent reader of
  // labels[i] sees either the old chain or the root
  labels[i] = cc_find(labels, l);
}

__global__ __launch_bounds__(256) void cc_count_kernel(const int32_t* __restrict__ labels, int64_t n,
                                                       int32_t* __restrict__ sizes) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int32_t lab = i < n ? labels[i] : -1;
  bool pending = lab >= 0;
  const int lane = (int)(threadIdx.x & (kWave - 1));
  // peel one distinct root per pass: its lanes add their count with one atomic
  while (__ballot(pending)) {
    if (pending) {
      const int32_t lead = __builtin_amdgcn_readfirstlane(lab);
      const uint64_t same = __ballot(lab == lead);
      if (lab == lead) {
        if (lane == __ffsll((unsigned long long)same) - 1) atomicAdd(sizes + lead, (int32_t)__popcll(same));
        pending = false;
      }
    }
  }
}

__global__ __launch_bounds__(256) void cc_filter_kernel(const int32_t* __restrict__ labels,
                                                        const int32_t* __restrict__ sizes, int64_t n,
                                                        double min_size, uint8_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t l = labels[i];
  out[i] = (l >= 0 && (double)sizes[l] >= min_size) ? 1 : 0;
}

}  // namespace
}  // namespace ff

using namespace ff;

extern "C" {

int ff_remove_small_components(const uint8_t* mask, int32_t D, int32_t H, int32_t W, double min_size,
                               int32_t* labels, int32_t* sizes, uint8_t* out, void* stream) {
  if (D < 0 || H < 0 || W < 0) return FF_ERR_ARG;
  const int64_t n = (int64_t)D * H * W;
  if (n == 0) return FF_OK;
  if (n > (int64_t)INT32_MAX) return FF_ERR_ARG;
  if (!mask || !labels || !sizes || !out || !(min_size == min_size)) return FF_ERR_ARG;
  const hipStream_t st = (hipStream_t)stream;
  const dim3 grid((unsigned)((n + 255) / 256)), block(256);
  hipLaunchKernelGGL(cc_init_kernel, grid, block, 0, st, mask, n, labels, sizes);
  hipLaunchKernelGGL(cc_merge_kernel, grid, block, 0, st, mask, D, H, W, labels);
  hipLaunchKernelGGL(cc_compress_kernel, grid, block, 0, st, n, labels);
  hipLaunchKernelGGL(cc_count_kernel, grid, block, 0, st, labels, n, sizes);
  hipLaunchKernelGGL(cc_filter_kernel, grid, block, 0, st, labels, sizes, n, min_size, out);
  return hipGetLastError() == hipSuccess ? FF_OK : FF_ERR_LAUNCH;
}

}  // extern "C"
