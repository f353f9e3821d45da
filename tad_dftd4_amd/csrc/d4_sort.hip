// Spatial order for large molecules (gfx950).
//
// The atoms of a molecule are sorted by the Morton code of their position on
// a 1024^3 grid over the molecule's bounding box, padding atoms (Z = 0)
// last, and the real atoms are then refined into a k-d order: level by
// level, every segment longer than 8 atoms is sorted along the longest edge
// of its bounding box and split at a multiple of 8 (the left part takes
// 8 ceil(len / 16) atoms). The segments depend on the number of real atoms
// alone, so consecutive 8-atom groups are the leaves of a balanced k-d tree
// (compact boxes) and 16-atom groups are pairs of sibling leaves. The
// three-body kernels skip pair tiles and third-atom groups beyond the cutoff
// from the groups' bounding spheres. Every key carries the input index in
// its low 32 bits: the order is unique and the same on every run.
//
// All kernels run in sorted order; the C ABI gathers the inputs into the
// workspace before the first kernel and scatters the per-atom (and pair)
// results back to input order after the last one.

#include "d4_device.h"

namespace d4 {
namespace {

constexpr int kSortMax = 16384;  // atoms per molecule that fit the LDS sort
constexpr int kSortThreads = 1024;

__device__ __forceinline__ uint32_t spread3(uint32_t v) {  // 10 bits to every third bit
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000ffu;
  v = (v | (v << 8)) & 0x0300f00fu;
  v = (v | (v << 4)) & 0x030c30c3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

// One work-group per molecule: bounding box, keys, bitonic sort in LDS.
__global__ __launch_bounds__(kSortThreads) void k_sort(const int64_t* numbers,
                                                       const double* positions, Work w,
                                                       int P) {
  extern __shared__ uint64_t keys[];
  __shared__ double s_box[6][kSortThreads / 64];
  const int64_t m = blockIdx.x;
  const int64_t N = w.N;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t* Z = numbers + m * N;
  const double* R = positions + m * N * 3;
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int i = tid; i < N; i += kSortThreads) {
    if (Z[i] == 0) continue;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      lo[d] = fmin(lo[d], R[3 * i + d]);
      hi[d] = fmax(hi[d], R[3 * i + d]);
    }
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      lo[d] = fmin(lo[d], __shfl_xor(lo[d], off, 64));
      hi[d] = fmax(hi[d], __shfl_xor(hi[d], off, 64));
    }
    if (lane == 0) {
      s_box[d][wave] = lo[d];
      s_box[3 + d][wave] = hi[d];
    }
  }
  __syncthreads();
  double b0[3], scale[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    double l = INFINITY, h = -INFINITY;
    for (int k = 0; k < kSortThreads / 64; ++k) {
      l = fmin(l, s_box[d][k]);
      h = fmax(h, s_box[3 + d][k]);
    }
    b0[d] = l;
    scale[d] = h > l ? 1023.0 / (h - l) : 0.0;
  }
  for (int i = tid; i < P; i += kSortThreads) {
    uint64_t key = ~0ull;
    if (i < N) {
      uint32_t code = 1u << 30;  // padding after every real atom
      if (Z[i] != 0) {
        uint32_t c[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const double u = (R[3 * i + d] - b0[d]) * scale[d];
          c[d] = (uint32_t)fmin(fmax(u, 0.0), 1023.0);
        }
        code = spread3(c[0]) | (spread3(c[1]) << 1) | (spread3(c[2]) << 2);
      }
      key = ((uint64_t)code << 32) | (uint32_t)i;
    }
    keys[i] = key;
  }
  __syncthreads();
  for (int k = 2; k <= P; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int t = tid; t < P; t += kSortThreads) {
        const int u = t ^ j;
        if (u > t) {
          const uint64_t a = keys[t], b = keys[u];
          const bool up = (t & k) == 0;
          if ((a > b) == up) {
            keys[t] = b;
            keys[u] = a;
          }
        }
      }
      __syncthreads();
    }
  }
  for (int i = tid; i < N; i += kSortThreads) w.perm[m * N + i] = (int32_t)(keys[i] & 0xffffffffu);
}

// Segment `seg` of k-d level `level` over the real atoms [0, nreal): the
// left child of [a, b) is [a, a + 8 ceil((b - a) / 16)). Returns false for a
// segment that does not exist (an ancestor was a leaf of at most 8 atoms).
__device__ __forceinline__ bool kd_segment(int nreal, int level, int seg, int& a, int& b) {
  a = 0;
  b = nreal;
  for (int l = 0; l < level; ++l) {
    const int bit = (seg >> (level - 1 - l)) & 1;
    if (b - a <= 8) return false;  // inside a leaf: nothing to sort
    const int mid = a + 8 * ((b - a + 15) / 16);
    if (bit)
      a = mid;
    else
      b = mid;
  }
  return b - a > 8;
}

// One k-d level: work-group (m, seg) sorts its segment of w.perm along the
// longest edge of the segment's bounding box (bitonic sort of P keys in LDS).
__global__ __launch_bounds__(kSortThreads) void k_kd_level(const int64_t* numbers,
                                                           const double* positions, Work w,
                                                           int level, int P) {
  extern __shared__ uint64_t keys[];
  __shared__ double s_box[6][kSortThreads / 64];
  __shared__ int s_cnt[kSortThreads / 64];
  const int64_t nseg = 1ll << level;
  const int64_t m = blockIdx.x / nseg;
  const int seg = (int)(blockIdx.x - m * nseg);
  const int64_t N = w.N;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t* Z = numbers + m * N;
  const double* R = positions + m * N * 3;
  int32_t* perm = w.perm + m * N;
  // real atoms come first in the Morton order
  int cnt = 0;
  for (int i = tid; i < N; i += kSortThreads) cnt += Z[i] != 0 ? 1 : 0;
  for (int off = 32; off >= 1; off >>= 1) cnt += __shfl_xor(cnt, off, 64);
  if (lane == 0) s_cnt[wave] = cnt;
  __syncthreads();
  int nreal = 0;
  for (int k = 0; k < kSortThreads / 64; ++k) nreal += s_cnt[k];
  int a, b;
  if (!kd_segment(nreal, level, seg, a, b)) return;  // uniform over the work-group
  const int len = b - a;
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int i = tid; i < len; i += kSortThreads) {
    const int64_t src = min(max(perm[a + i], 0), (int32_t)N - 1);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      lo[d] = fmin(lo[d], R[3 * src + d]);
      hi[d] = fmax(hi[d], R[3 * src + d]);
    }
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      lo[d] = fmin(lo[d], __shfl_xor(lo[d], off, 64));
      hi[d] = fmax(hi[d], __shfl_xor(hi[d], off, 64));
    }
    if (lane == 0) {
      s_box[d][wave] = lo[d];
      s_box[3 + d][wave] = hi[d];
    }
  }
  __syncthreads();
  double l0[3], ext[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    double l = INFINITY, h = -INFINITY;
    for (int k = 0; k < kSortThreads / 64; ++k) {
      l = fmin(l, s_box[d][k]);
      h = fmax(h, s_box[3 + d][k]);
    }
    l0[d] = l;
    ext[d] = h - l;
  }
  const int ax = ext[0] >= ext[1] ? (ext[0] >= ext[2] ? 0 : 2) : (ext[1] >= ext[2] ? 1 : 2);
  const double scale = ext[ax] > 0.0 ? 4294967295.0 / ext[ax] : 0.0;
  for (int i = tid; i < P; i += kSortThreads) {
    uint64_t key = ~0ull;
    if (i < len) {
      const int32_t src = min(max(perm[a + i], 0), (int32_t)N - 1);
      const double u = (R[3 * (int64_t)src + ax] - l0[ax]) * scale;
      const uint32_t q = (uint32_t)fmin(fmax(u, 0.0), 4294967295.0);
      key = ((uint64_t)q << 32) | (uint32_t)src;
    }
    keys[i] = key;
  }
  __syncthreads();
  for (int k = 2; k <= P; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int t = tid; t < P; t += kSortThreads) {
        const int u = t ^ j;
        if (u > t) {
          const uint64_t x = keys[t], y = keys[u];
          const bool up = (t & k) == 0;
          if ((x > y) == up) {
            keys[t] = y;
            keys[u] = x;
          }
        }
      }
      __syncthreads();
    }
  }
  for (int i = tid; i < len; i += kSortThreads) perm[a + i] = (int32_t)(keys[i] & 0xffffffffu);
}

__global__ __launch_bounds__(256) void k_identity(Work w) {
  const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (g >= w.B * w.N) return;
  w.perm[g] = (int32_t)(g % w.N);
}

// perm entry of sorted slot g, clamped into the molecule (a workspace that
// no forward call has filled then yields wrong values but no stray access)
__device__ __forceinline__ int64_t slot_index(const Work& w, int64_t g) {
  return (int64_t)min(max(w.perm[g], 0), (int32_t)w.N - 1);
}

__global__ __launch_bounds__(256) void k_gather(Inputs in, Work w) {
  const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (g >= w.B * w.N) return;
  const int64_t m = g / w.N;
  const int64_t src = m * w.N + slot_index(w, g);
  w.snum[g] = in.numbers[src];
#pragma unroll
  for (int d = 0; d < 3; ++d) w.spos[3 * g + d] = in.positions[3 * src + d];
  if (in.q_in) w.sq[g] = in.q_in[src];
  if (in.r4r2_in) w.sr4[g] = in.r4r2_in[src];
  if (in.gE) w.sge[g] = in.gE[src];
}

__global__ __launch_bounds__(256) void k_scatter_atoms(Work w, const double* src, double* dst,
                                                       int comps) {
  const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (g >= w.B * w.N) return;
  const int64_t m = g / w.N;
  const int64_t d = m * w.N + slot_index(w, g);
  for (int c = 0; c < comps; ++c) dst[d * comps + c] = src[g * comps + c];
}

__global__ __launch_bounds__(256) void k_scatter_pairs(Work w, const double* src, double* dst) {
  const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t N = w.N;
  if (g >= w.B * N * N) return;
  const int64_t row = g / N;  // m * N + i
  const int64_t j = g - row * N;
  const int64_t m = row / N;
  dst[(m * N + slot_index(w, row)) * N + slot_index(w, m * N + j)] = src[g];
}

// Bounding sphere of each GS-atom group over its real atoms; a group without
// real atoms gets radius -1e300 so that no distance test keeps it.
template <int GS>
__global__ __launch_bounds__(256) void k_groups(Inputs in, Work w, double* out) {
  const int64_t NG = (w.N + GS - 1) / GS;
  const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (g >= w.B * NG) return;
  const int64_t m = g / NG;
  const int a0 = (int)(g - m * NG) * GS;
  const int n = w.natom[m];
  const int64_t* Z = in.numbers + m * w.N;
  const double* R = in.positions + m * w.N * 3;
  double c[3] = {0.0, 0.0, 0.0};
  int cnt = 0;
  for (int a = a0; a < min(a0 + GS, n); ++a) {
    if (Z[a] == 0) continue;
    ++cnt;
#pragma unroll
    for (int d = 0; d < 3; ++d) c[d] += R[3 * a + d];
  }
  double rad = -1e300;
  if (cnt > 0) {
#pragma unroll
    for (int d = 0; d < 3; ++d) c[d] /= cnt;
    rad = 0.0;
    for (int a = a0; a < min(a0 + GS, n); ++a) {
      if (Z[a] == 0) continue;
      const double dx = R[3 * a] - c[0], dy = R[3 * a + 1] - c[1], dz = R[3 * a + 2] - c[2];
      rad = fmax(rad, sqrt(dx * dx + dy * dy + dz * dz));
    }
  }
  double* S = out + 4 * g;
  S[0] = c[0];
  S[1] = c[1];
  S[2] = c[2];
  S[3] = rad;
}

inline unsigned atom_grid(const Work& w) { return (unsigned)((w.B * w.N + 255) / 256); }

}  // namespace

hipError_t launch_sort(hipStream_t s, const Inputs& in, const Work& w) {
  if (w.N > kSortMax) {
    hipLaunchKernelGGL(k_identity, dim3(atom_grid(w)), dim3(256), 0, s, w);
    return hipGetLastError();
  }
  int P = 1;
  while (P < w.N) P <<= 1;
  const size_t shmem = sizeof(uint64_t) * (size_t)P;
  hipError_t e = ensure_lds((const void*)k_sort, shmem);
  if (e == hipSuccess) e = ensure_lds((const void*)k_kd_level, shmem);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_sort, dim3((unsigned)w.B), dim3(kSortThreads), shmem, s, in.numbers,
                     in.positions, w, P);
  // k-d levels while the largest segment (the leftmost; it bounds the
  // segments of every molecule of the batch) is longer than 8 atoms
  int len = (int)w.N;
  for (int level = 0; len > 8; ++level) {
    int Pl = 1;
    while (Pl < len) Pl <<= 1;
    const int64_t blocks = w.B << level;
    hipLaunchKernelGGL(k_kd_level, dim3((unsigned)blocks), dim3(kSortThreads),
                       sizeof(uint64_t) * (size_t)Pl, s, in.numbers, in.positions, w, level, Pl);
    len = 8 * ((len + 15) / 16);
  }
  return hipGetLastError();
}

hipError_t launch_gather(hipStream_t s, const Inputs& in, const Work& w) {
  hipLaunchKernelGGL(k_gather, dim3(atom_grid(w)), dim3(256), 0, s, in, w);
  return hipGetLastError();
}

hipError_t launch_scatter_atoms(hipStream_t s, const Work& w, const double* src,
                                double* dst, int comps) {
  hipLaunchKernelGGL(k_scatter_atoms, dim3(atom_grid(w)), dim3(256), 0, s, w, src, dst,
                     comps);
  return hipGetLastError();
}

hipError_t launch_scatter_pairs(hipStream_t s, const Work& w, const double* src,
                                double* dst) {
  const int64_t total = w.B * w.N * w.N;
  hipLaunchKernelGGL(k_scatter_pairs, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
                     w, src, dst);
  return hipGetLastError();
}

hipError_t launch_groups(hipStream_t s, const Inputs& in, const Work& w) {
  const int64_t total = w.B * n_groups(w.N);
  hipLaunchKernelGGL(k_groups<kGroup>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
                     in, w, w.gsph);
  return hipGetLastError();
}

hipError_t launch_groups8(hipStream_t s, const Inputs& in, const Work& w) {
  const int64_t total = w.B * n_groups8(w.N);
  hipLaunchKernelGGL(k_groups<kAtm8>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
                     in, w, w.gsph8);
  return hipGetLastError();
}

}  // namespace d4
