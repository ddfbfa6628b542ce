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
//   cc_init:     labels[i] = the first voxel of i's x-run within its wave's
//                64 voxels (-1 for unset voxels), sizes[i] = 0
//   cc_merge:    unions with the -x, -y, -z neighbours, one per boundary
//                stretch between two runs
//   cc_compress: labels[i] = root of i (the component's smallest index)
//   cc_count:    sizes[root] += 1: runs of equal roots in a wave, summed per
//                block in LDS, one global atomic per root and block
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

// Each set voxel starts labelled with the first voxel of its x-run inside
// its wave's 64 consecutive voxels (ballot of the set voxels; a run also
// starts at x = 0 and at the wave's first voxel).
__global__ __launch_bounds__(256) void cc_init_kernel(const uint8_t* __restrict__ mask, int32_t W, int64_t n,
                                                      int32_t* __restrict__ labels, int32_t* __restrict__ sizes) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool set = i < n && mask[i];
  const bool row0 = i < n && i % W == 0;
  const int lane = (int)(threadIdx.x & (kWave - 1));
  const uint64_t S = __ballot(set), R = __ballot(row0);
  const uint64_t heads = S & (~(S << 1) | R | 1ull);
  if (i >= n) return;
  sizes[i] = 0;
  if (!set) {
    labels[i] = -1;
    return;
  }
  const uint64_t upto = lane == 63 ? heads : heads & ((2ull << lane) - 1);
  labels[i] = (int32_t)(i - lane + (63 - __clzll(upto)));
}

// Unions: across x only where a run enters a wave's voxels from the previous
// wave's; across y (z) only at the first voxel of each stretch where a run
// lies next to a run of the row (slice) before -- the rest of the stretch
// is connected through that voxel and the two runs.
__global__ __launch_bounds__(256) void cc_merge_kernel(const uint8_t* __restrict__ mask, int32_t D, int32_t H,
                                                       int32_t W, int32_t* labels) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = (int64_t)D * H * W;
  if (i >= n || !mask[i]) return;
  const int64_t WH = (int64_t)W * H;
  const int32_t x = (int32_t)(i % W), y = (int32_t)((i / W) % H), z = (int32_t)(i / WH);
  const int32_t me = (int32_t)i;
  const bool left = x > 0 && mask[i - 1];
  if (left && (i & (kWave - 1)) == 0) cc_union(labels, me, me - 1);
  if (y > 0 && mask[i - W] && !(left && mask[i - 1 - W])) cc_union(labels, me, me - W);
  if (z > 0 && mask[i - WH] && !(left && mask[i - 1 - WH])) cc_union(labels, me, (int32_t)(i - WH));
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

// Voxel counts per root.  A component's voxels all add to one address, and
// adds to one address serialise, so the counts of a block are first summed
// in a small LDS table keyed by root (linear probing, kCcSlots slots); each
// table entry then takes one global atomic.  A root that finds no free slot
// after kCcSlots probes adds its count to global memory directly.
constexpr int kCcBlock = 1024;
constexpr int kCcSlots = 128;

__global__ __launch_bounds__(kCcBlock) void cc_count_kernel(const int32_t* __restrict__ labels, int64_t n,
                                                            int32_t* __restrict__ sizes) {
  __shared__ int32_t key[kCcSlots];
  __shared__ int32_t cnt[kCcSlots];
  if (threadIdx.x < kCcSlots) {
    key[threadIdx.x] = -1;
    cnt[threadIdx.x] = 0;
  }
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int32_t lab = i < n ? labels[i] : -1;
  bool pending = lab >= 0;
  const int lane = (int)(threadIdx.x & (kWave - 1));
  // peel one distinct root per pass: the first of its lanes adds the run's count
  while (__ballot(pending)) {
    if (pending) {
      const int32_t lead = __builtin_amdgcn_readfirstlane(lab);
      const uint64_t same = __ballot(lab == lead);
      if (lab == lead) {
        if (lane == __ffsll((unsigned long long)same) - 1) {
          const int32_t c = (int32_t)__popcll(same);
          int h = (int)(((uint32_t)lead * 2654435761u) >> 25) & (kCcSlots - 1);
          bool done = false;
          for (int probe = 0; probe < kCcSlots && !done; ++probe) {
            const int32_t prev = atomicCAS(&key[h], -1, lead);
            if (prev == -1 || prev == lead) {
              atomicAdd(&cnt[h], c);
              done = true;
            }
            h = (h + 1) & (kCcSlots - 1);
          }
          if (!done) atomicAdd(sizes + lead, c);
        }
        pending = false;
      }
    }
  }
  __syncthreads();
  if (threadIdx.x < kCcSlots && key[threadIdx.x] >= 0) atomicAdd(sizes + key[threadIdx.x], cnt[threadIdx.x]);
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
  hipLaunchKernelGGL(cc_init_kernel, grid, block, 0, st, mask, W, n, labels, sizes);
  hipLaunchKernelGGL(cc_merge_kernel, grid, block, 0, st, mask, D, H, W, labels);
  hipLaunchKernelGGL(cc_compress_kernel, grid, block, 0, st, n, labels);
  hipLaunchKernelGGL(cc_count_kernel, dim3((unsigned)((n + kCcBlock - 1) / kCcBlock)), dim3(kCcBlock), 0, st,
                     labels, n, sizes);
  hipLaunchKernelGGL(cc_filter_kernel, grid, block, 0, st, labels, sizes, n, min_size, out);
  return hipGetLastError() == hipSuccess ? FF_OK : FF_ERR_LAUNCH;
}

}  // extern "C"
