// Isosurface extraction of a dense volume for the mesh export
// (train_per_scene.py:32-41 -> utils.py:183-249, convert_sdf_samples_to_ply).
// The reference runs skimage's marching cubes; here every grid cube is split
// into the six tetrahedra that share its main diagonal (the same split in
// every cube, so the surface is closed) and each tetrahedron straddling the
// level contributes one or two triangles whose vertices lie on grid edges.
//
//   ff_mt_count: one thread per cube, the number of triangles it emits;
//   ff_mt_emit:  one thread per listed cube (those with triangles), at its
//                offset in the exclusive scan of the counts: three edge keys
//                per triangle (key = lo * N + hi
//                for grid points lo < hi, N the volume's point count), in an
//                order whose normal points from the tetrahedron's inside
//                corners to its outside ones;
//   ff_mt_verts: one thread per distinct edge key, the vertex where the
//                level crosses that edge, linear in the two endpoint values,
//                in index coordinates (float64).
// Inside is value > level (inside_above, densities) or value < level (signed
// distances).  Vertex positions and the orientation test are FP64 without
// fused multiply-adds, so they round exactly as the same formulas on the
// CPU do.  The caller deduplicates the keys (a sort) between emit and
// verts; shared edges then give shared vertices.
#include "ff_common.h"

namespace ff {
namespace {

// corner c of a cube sits at offset (c & 1, (c >> 1) & 1, c >> 2) on the
// volume's axes (0, 1, 2); tetrahedron t = corners 0, e_p, e_p + e_q, 7 for
// the axis order (p, q, r) number t of the six permutations
__constant__ int kTet[6][4] = {{0, 1, 3, 7}, {0, 1, 5, 7}, {0, 2, 3, 7}, {0, 2, 6, 7}, {0, 4, 5, 7}, {0, 4, 6, 7}};

// Triangles of a tetrahedron for its inside mask m (bit q: corner q inside):
// up to two, each as three (inside corner, outside corner) edges.
struct MtCase {
  int n;
  int e[2][3][2];
};

__device__ __forceinline__ MtCase mt_case(int m) {
  MtCase c{};
  int ins[4], out[4], ni = 0, no = 0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if ((m >> q) & 1) ins[ni++] = q;
    else out[no++] = q;
  }
  if (ni == 1) {
    c.n = 1;
    for (int k = 0; k < 3; ++k) { c.e[0][k][0] = ins[0]; c.e[0][k][1] = out[k]; }
  } else if (ni == 3) {
    c.n = 1;
    for (int k = 0; k < 3; ++k) { c.e[0][k][0] = ins[k]; c.e[0][k][1] = out[0]; }
  } else if (ni == 2) {
    c.n = 2;
    const int a = ins[0], b = ins[1], cc = out[0], d = out[1];
    const int t0[3][2] = {{a, cc}, {a, d}, {b, d}};
    const int t1[3][2] = {{a, cc}, {b, d}, {b, cc}};
    for (int k = 0; k < 3; ++k) {
      c.e[0][k][0] = t0[k][0]; c.e[0][k][1] = t0[k][1];
      c.e[1][k][0] = t1[k][0]; c.e[1][k][1] = t1[k][1];
    }
  }
  return c;
}

// q = x / d, r = x % d with 32-bit division when both operands fit in 32 bits
// (every volume up to 2^32 points), 64-bit otherwise.
__device__ __forceinline__ void divmod(int64_t x, int64_t d, int64_t& q, int64_t& r) {
  if (((uint64_t)x | (uint64_t)d) >> 32 == 0) {
    const uint32_t q32 = (uint32_t)x / (uint32_t)d;
    q = q32;
    r = (int64_t)((uint32_t)x - q32 * (uint32_t)d);
  } else {
    q = x / d;
    r = x - q * d;
  }
}

struct MtGrid {
  const float* v;
  int64_t A, B, C;
  double level;
  int inside_above;
  __device__ __forceinline__ int64_t pid(int64_t i, int64_t j, int64_t k) const { return (i * B + j) * C + k; }
  __device__ __forceinline__ bool inside(float x) const {
    return inside_above ? (double)x > level : (double)x < level;
  }
};

// The cube's origin (i, j, k), its 8 point ids and inside flags; returns
// false for cubes off the grid.
__device__ __forceinline__ bool mt_cube(const MtGrid& g, int64_t c, int64_t (&ijk)[3], int64_t (&pid)[8],
                                        bool (&in)[8]) {
  const int64_t nk = g.C - 1, nj = g.B - 1, ni = g.A - 1;
  if (c < 0 || c >= ni * nj * nk) return false;
  int64_t q, k, i, j;
  divmod(c, nk, q, k);
  divmod(q, nj, i, j);
  ijk[0] = i;
  ijk[1] = j;
  ijk[2] = k;
#pragma unroll
  for (int o = 0; o < 8; ++o) {
    pid[o] = g.pid(i + (o & 1), j + ((o >> 1) & 1), k + (o >> 2));
    in[o] = g.inside(g.v[pid[o]]);
  }
  return true;
}

__device__ __forceinline__ int tet_mask(const bool (&in)[8], int t) {
  int m = 0;
#pragma unroll
  for (int q = 0; q < 4; ++q) m |= int(in[kTet[t][q]]) << q;
  return m;
}

__device__ __forceinline__ void mt_coords(const MtGrid& g, int64_t p, double* x) {
  int64_t q, k, i, j;
  divmod(p, g.C, q, k);
  divmod(q, g.B, i, j);
  x[0] = (double)i;
  x[1] = (double)j;
  x[2] = (double)k;
}

// The level crossing on the edge between grid points p < q at index
// coordinates a, b.
__device__ __forceinline__ void mt_vertex_at(const MtGrid& g, int64_t p, int64_t q, const double* a,
                                             const double* b, double* out) {
#pragma clang fp contract(off)
  const double f0 = g.v[p], f1 = g.v[q];
  const double t = (g.level - f0) / (f1 - f0);
#pragma unroll
  for (int k = 0; k < 3; ++k) out[k] = a[k] + t * (b[k] - a[k]);
}

__device__ __forceinline__ void mt_vertex(const MtGrid& g, int64_t p, int64_t q, double* out) {
  double a[3], b[3];
  mt_coords(g, p, a);
  mt_coords(g, q, b);
  mt_vertex_at(g, p, q, a, b, out);
}

__global__ __launch_bounds__(256) void mt_count_kernel(MtGrid g, int32_t* __restrict__ counts) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t ijk[3], pid[8];
  bool in[8];
  if (!mt_cube(g, c, ijk, pid, in)) return;
  int n = 0;
#pragma unroll
  for (int t = 0; t < 6; ++t) {
    const int m = tet_mask(in, t);
    const int k = __popc(m);
    n += (k == 1 || k == 3) ? 1 : (k == 2 ? 2 : 0);
  }
  counts[c] = n;
}

// One thread per listed cube (cubes[i], or cube i when the list is NULL).
__global__ __launch_bounds__(256) void mt_emit_kernel(MtGrid g, const int64_t* __restrict__ cubes, int64_t n_cubes,
                                                      const int64_t* __restrict__ offsets,
                                                      int64_t* __restrict__ keys) {
  const int64_t ci = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (ci >= n_cubes) return;
  const int64_t c = cubes ? cubes[ci] : ci;
  int64_t ijk[3], pid[8];
  bool in[8];
  if (!mt_cube(g, c, ijk, pid, in)) return;
  int64_t o = offsets[c];
  const int64_t N = g.A * g.B * g.C;
  // index coordinates of the 8 corners (exact in FP64)
  double X[8][3];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    X[q][0] = (double)(ijk[0] + (q & 1));
    X[q][1] = (double)(ijk[1] + ((q >> 1) & 1));
    X[q][2] = (double)(ijk[2] + (q >> 2));
  }
  for (int t = 0; t < 6; ++t) {
    const int m = tet_mask(in, t);
    if (m == 0 || m == 15) continue;
    const MtCase cs = mt_case(m);
    // inside -> outside direction of the tetrahedron (corner centroids)
    double cin[3] = {0, 0, 0}, cout[3] = {0, 0, 0};
    int nin = 0;
    for (int q = 0; q < 4; ++q) {
      const double* x = X[kTet[t][q]];
      const bool iq = (m >> q) & 1;
      nin += iq;
      for (int a = 0; a < 3; ++a) (iq ? cin : cout)[a] += x[a];
    }
    double dir[3];
    for (int a = 0; a < 3; ++a) dir[a] = cout[a] / (4 - nin) - cin[a] / nin;
    for (int tr = 0; tr < cs.n; ++tr) {
#pragma clang fp contract(off)
      int64_t key[3];
      double p[3][3];
      for (int e = 0; e < 3; ++e) {
        const int cu = kTet[t][cs.e[tr][e][0]], cv = kTet[t][cs.e[tr][e][1]];
        const int64_t pu = pid[cu], pv = pid[cv];
        const bool uv = pu < pv;
        const int64_t lo = uv ? pu : pv, hi = uv ? pv : pu;
        key[e] = lo * N + hi;
        mt_vertex_at(g, lo, hi, X[uv ? cu : cv], X[uv ? cv : cu], p[e]);
      }
      const double u[3] = {p[1][0] - p[0][0], p[1][1] - p[0][1], p[1][2] - p[0][2]};
      const double w[3] = {p[2][0] - p[0][0], p[2][1] - p[0][1], p[2][2] - p[0][2]};
      const double nrm[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]};
      const bool flip = nrm[0] * dir[0] + nrm[1] * dir[1] + nrm[2] * dir[2] < 0.0;
      keys[3 * o + 0] = key[0];
      keys[3 * o + 1] = flip ? key[2] : key[1];
      keys[3 * o + 2] = flip ? key[1] : key[2];
      ++o;
    }
  }
}

__global__ __launch_bounds__(256) void mt_verts_kernel(MtGrid g, const int64_t* __restrict__ keys, int64_t n,
                                                       double* __restrict__ verts) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t N = g.A * g.B * g.C;
  int64_t lo, hi;
  divmod(keys[i], N, lo, hi);
  if (keys[i] < 0 || lo >= N) {
    for (int k = 0; k < 3; ++k) verts[i * 3 + k] = __builtin_nan("");
    return;
  }
  mt_vertex(g, lo, hi, verts + i * 3);
}

int mt_args(const float* vol, int64_t A, int64_t B, int64_t C) {
  if (!vol || A < 2 || B < 2 || C < 2) return FF_ERR_ARG;
  if (A > (int64_t)1 << 20 || B > (int64_t)1 << 20 || C > (int64_t)1 << 20) return FF_ERR_ARG;
  // edge keys lo * N + hi must fit in int64
  const double N = (double)A * (double)B * (double)C;
  if (N * N >= 9.0e18) return FF_ERR_ARG;
  return FF_OK;
}

int mt_status() { return hipGetLastError() == hipSuccess ? FF_OK : FF_ERR_LAUNCH; }

}  // namespace
}  // namespace ff

using namespace ff;

extern "C" {

int ff_mt_count(const float* vol, int64_t A, int64_t B, int64_t C, double level, int32_t inside_above,
                int32_t* counts, void* stream) {
  const int st = mt_args(vol, A, B, C);
  if (st) return st;
  if (!counts) return FF_ERR_ARG;
  const int64_t n = (A - 1) * (B - 1) * (C - 1);
  const MtGrid g{vol, A, B, C, level, inside_above};
  hipLaunchKernelGGL(mt_count_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, g,
                     counts);
  return mt_status();
}

int ff_mt_emit(const float* vol, int64_t A, int64_t B, int64_t C, double level, int32_t inside_above,
               const int64_t* cubes, int64_t n_cubes, const int64_t* offsets, int64_t* keys, void* stream) {
  const int st = mt_args(vol, A, B, C);
  if (st) return st;
  const int64_t n = (A - 1) * (B - 1) * (C - 1);
  if (!cubes) n_cubes = n;
  if (n_cubes < 0 || n_cubes > n) return FF_ERR_ARG;
  if (n_cubes == 0) return FF_OK;
  if (!offsets || !keys) return FF_ERR_ARG;
  const MtGrid g{vol, A, B, C, level, inside_above};
  hipLaunchKernelGGL(mt_emit_kernel, dim3((unsigned)((n_cubes + 255) / 256)), dim3(256), 0, (hipStream_t)stream, g,
                     cubes, n_cubes, offsets, keys);
  return mt_status();
}

int ff_mt_verts(const float* vol, int64_t A, int64_t B, int64_t C, double level, const int64_t* keys, int64_t n,
                double* verts, void* stream) {
  const int st = mt_args(vol, A, B, C);
  if (st) return st;
  if (n < 0) return FF_ERR_ARG;
  if (n == 0) return FF_OK;
  if (!keys || !verts) return FF_ERR_ARG;
  const MtGrid g{vol, A, B, C, level, 1};
  hipLaunchKernelGGL(mt_verts_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, g,
                     keys, n, verts);
  return mt_status();
}

}  // extern "C"
