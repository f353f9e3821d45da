// HIP kernels (gfx950) for the DFT-D4 energy and gradient path.
//
// FP64 throughout. Wave64 layout: "row" kernels give one wavefront to one
// atom row i of one molecule (lanes stride over partner atoms j and reduce
// with shuffles); the EEQ kernels give one work-group to one molecule; the
// three-body kernel gives one work-group to a 16 x 16 tile of atom pairs of
// one molecule and streams the third atom through LDS.
//
// All reductions have a fixed order that depends only on the molecule (not
// on its neighbours in the batch), and every output element has exactly one
// writer, so results are bitwise reproducible and independent of batching.

#include <math.h>

#include "d4_internal.h"

namespace d4 {
namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kEps = 2.220446049250313080847e-16;  // float64 epsilon
constexpr double kKcn = 7.5;                          // defaults.py:41
constexpr double kK4 = 4.10451;                       // defaults.py:44
constexpr double kK5 = 19.08857;                      // defaults.py:47
constexpr double kK6 = 2.0 * 11.28174 * 11.28174;     // defaults.py:50
constexpr double kCnMax = 8.0;                        // defaults.py:32
constexpr int kWaves = 4;                             // waves per 256-thread block
constexpr int kTile = 16;                             // ATM pair tile edge
constexpr int kZB = 32;                               // ATM third-atom rows per LDS round

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
  return v;
}

__device__ __forceinline__ int wave_or(int v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v |= __shfl_xor(v, off, 64);
  return v;
}

// Block-wide sum over 256 threads (4 waves), result broadcast to all.
__device__ __forceinline__ double block_sum(double v, double* scratch) {
  v = wave_sum(v);
  const int wave = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) scratch[wave] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < kWaves; ++k) s += scratch[k];
  return s;
}

// 1/2 (1 + erf(-k (r/rc - 1)))  (tad-mctc erf_count)
__device__ __forceinline__ double erf_count(double r, double rc) {
  return 0.5 * (1.0 + erf(-kKcn * (r / rc - 1.0)));
}

// d/dr of erf_count
__device__ __forceinline__ double erf_count_dr(double r, double rc) {
  const double x = r / rc - 1.0;
  return -(kKcn / rc) * exp(-kKcn * kKcn * x * x) / sqrt(kPi);
}

// Electronegativity factor of the D4 coordination number.
__device__ __forceinline__ double en_factor(double eni, double enj) {
  const double d = fabs(eni - enj) + kK5;
  return kK4 * exp(-(d * d) / kK6);
}

// Charge scaling function zeta and its derivative (model/base.py:308-365).
__device__ __forceinline__ double zeta_fn(double ga, double gam, double qref,
                                          double qmod) {
  if (qmod > 0.0) {
    const double scale = exp(gam * (1.0 - qref / (qmod - kEps)));
    return exp(ga * (1.0 - scale));
  }
  return exp(ga);
}

__device__ __forceinline__ double dzeta_fn(double ga, double gam, double qref,
                                           double qmod) {
  if (qmod > 0.0) {
    const double den = qmod - kEps;
    const double scale = exp(gam * (1.0 - qref / den));
    const double z = exp(ga * (1.0 - scale));
    return -ga * gam * scale * z * qref / (den * den);
  }
  return 0.0;
}

__host__ __device__ __forceinline__ int tri(int i) { return (i * (i + 1)) >> 1; }

// --------------------------------------------------------------------------
// prep: molecule extents and atomic-number validation (one wave / molecule)
// --------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_prep(const int64_t* __restrict__ numbers,
                                              Work w) {
  const int lane = threadIdx.x & 63;
  const int64_t m = (int64_t)blockIdx.x * kWaves + (threadIdx.x >> 6);
  if (m >= w.B) return;
  const int64_t* z = numbers + m * w.N;
  int last = -1, bad = 0;
  for (int64_t i = lane; i < w.N; i += 64) {
    const int64_t zi = z[i];
    if (zi != 0) last = (int)i;
    if (zi < 0 || zi >= kMaxZ) bad = 1;
  }
  last = wave_max(last);
  bad = wave_or(bad);
  if (lane == 0) {
    w.natom[m] = bad ? 0 : last + 1;
    w.status[m] = bad ? kStatusBadZ : 0;
  }
}

// --------------------------------------------------------------------------
// coordination numbers: D4 (EN-weighted) and EEQ (capped), one wave / atom
// --------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_cn(DevTables t, Params p, Inputs in,
                                            Work w) {
  const int lane = threadIdx.x & 63;
  const int64_t row = (int64_t)blockIdx.x * kWaves + (threadIdx.x >> 6);
  if (row >= w.B * w.N) return;
  const int64_t m = row / w.N;
  const int i = (int)(row - m * w.N);
  const int n = w.natom[m];
  const int64_t* Z = in.numbers + m * w.N;
  const double* R = in.positions + m * w.N * 3;
  const int zi = i < n ? (int)Z[i] : 0;
  if (zi == 0) {
    if (lane == 0) {
      w.cn[row] = 0.0;
      w.cnq[row] = 0.0;
      w.dcnq[row] = 0.0;
    }
    return;
  }
  const double xi = R[3 * i], yi = R[3 * i + 1], zzi = R[3 * i + 2];
  const double rci = t.rcov[zi], eni = t.en[zi];
  double s4 = 0.0, se = 0.0;
  for (int j = lane; j < n; j += 64) {
    const int zj = (int)Z[j];
    if (j == i || zj == 0) continue;
    const double dx = xi - R[3 * j], dy = yi - R[3 * j + 1], dz = zzi - R[3 * j + 2];
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    const double c = erf_count(r, rci + t.rcov[zj]);
    if (r <= p.cutcn) s4 += en_factor(eni, t.en[zj]) * c;
    if (r <= p.cuteeq) se += c;
  }
  s4 = wave_sum(s4);
  se = wave_sum(se);
  if (lane == 0) {
    w.cn[row] = s4;
    // smooth cap: log(1 + e^cmax) - log(1 + e^(cmax - cn))
    const double ex = exp(kCnMax - se);
    w.cnq[row] = log(1.0 + exp(kCnMax)) - log(1.0 + ex);
    w.dcnq[row] = ex / (1.0 + ex);
  }
}

// --------------------------------------------------------------------------
// EEQ charges: one work-group per molecule.
//
// The bordered system [[C, e], [e^T, 0]] [q; mu] = [b; Q] is solved with a
// Cholesky factorisation of the Coulomb block C (in LDS when it fits) and the
// Schur complement of the border: q = y - mu z with y = C^-1 b, z = C^-1 e,
// mu = (e.y - Q) / (e.z). A non-positive pivot flags the molecule for the
// pivoted-LU kernel below.
// --------------------------------------------------------------------------

__device__ __forceinline__ double eeq_rhs(const DevTables& t, int z, double cnq) {
  return -t.chi[z] + t.kcn[z] * sqrt(fmax(cnq, kEps));
}

__device__ __forceinline__ double eeq_offdiag(const DevTables& t, int zi, int zj,
                                              double r) {
  const double ri = t.rad[zi], rj = t.rad[zj];
  const double gamma = 1.0 / sqrt(ri * ri + rj * rj);
  return erf(r * gamma) / r;
}

__device__ __forceinline__ double eeq_diag(const DevTables& t, int z) {
  return t.eta[z] + sqrt(2.0 / kPi) / t.rad[z];
}

// ---- blocked Cholesky (16-column blocks) on a packed lower triangle -------
//
// Row i of the matrix starts at L + tri(i). For each block column kb:
//  1. wave 0 factorises the diagonal 16 x 16 block D in registers (lane r
//     holds row r; entries are broadcast with v_readlane, no LDS round trip)
//     and stores X = D^-1 (lower triangular) in its place;
//  2. every panel row below becomes a X^T (a = the row's block entries) and
//     the carried right-hand sides become X b: independent per thread;
//  3. the trailing triangle is updated in register tiles.
// The substitutions read X from the stored factor in the same way, so each
// solve step is a small parallel matrix-vector product between two barriers
// and no thread runs a serial dependent chain longer than 16 terms.
// All helpers are force-inlined so that pointers into LDS keep their
// address space (ds_* instructions rather than flat ones).

constexpr int kNB = 16;

__device__ __forceinline__ double lane_bcast(double v, int src) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), src);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), src);
  return __hiloint2double(hi, lo);
}

// Factor the diagonal block at kb (nb <= 16 rows) on wave 0 and overwrite it
// with the inverse of its Cholesky factor. Rows nb..15 are padded with the
// identity (their columns never mix with the real rows: their off-diagonal
// entries are zero and stay zero). Returns non-zero (uniform over the wave) if a pivot was not
// positive; the block is then left unspecified.
template <typename P>
__device__ __forceinline__ int chol_diag_block(P L, int kb, int nb) {
  const int lane = threadIdx.x & 63;
  double d[kNB];
#pragma unroll
  for (int c = 0; c < kNB; ++c) {
    double v = (c == lane) ? 1.0 : 0.0;
    if (lane < nb && c <= lane) v = L[tri(kb + lane) + kb + c];
    d[c] = v;
  }
  double invd[kNB];
  int bad = 0;
#pragma unroll
  for (int j = 0; j < kNB; ++j) {
    const double piv = lane_bcast(d[j], j);
    bad |= !(piv > 0.0);
    const double s = piv > 0.0 ? sqrt(piv) : 1.0;
    const double inv = 1.0 / s;
    invd[j] = inv;
    // Lanes above the diagonal (lane < j) carry junk from here on in d[j] and
    // in the columns they update; only lanes >= column are ever broadcast
    // (pivots and L[c][j] with c > j), so the junk is never read.
    d[j] = lane == j ? s : d[j] * inv;
#pragma unroll
    for (int c = j + 1; c < kNB; ++c) d[c] = fma(-d[j], lane_bcast(d[j], c), d[c]);
  }
  // lane r forms column r of X = D^-1 by forward substitution
  double x[kNB];
#pragma unroll
  for (int i = 0; i < kNB; ++i) {
    double s = (i == lane) ? 1.0 : 0.0;
#pragma unroll
    for (int k = 0; k < i; ++k) s = fma(-lane_bcast(d[k], i), x[k], s);
    x[i] = s * invd[i];
  }
  if (lane < nb) {
#pragma unroll
    for (int i = 0; i < kNB; ++i)
      if (i >= lane && i < nb) L[tri(kb + i) + kb + lane] = x[i];
  }
  return bad;
}

// Panel row i (full block, nb = 16): L[i][kb..kb+16) <- a X^T.
template <typename P>
__device__ __forceinline__ void panel_row(P L, int kb, int i) {
  P Li = L + tri(i) + kb;
  double a[kNB];
#pragma unroll
  for (int c = 0; c < kNB; ++c) a[c] = Li[c];
#pragma unroll
  for (int c = 0; c < kNB; ++c) {
    P Xc = L + tri(kb + c) + kb;
    double s = 0.0;
#pragma unroll
    for (int k = 0; k <= c; ++k) s = fma(a[k], Xc[k], s);
    Li[c] = s;
  }
}

// Block forward step on a vector: v[kb..kb+nb) <- X v[kb..kb+nb), copied to y.
template <typename P, typename V>
__device__ __forceinline__ void block_lower(P L, int kb, int nb, V v, double* y) {
  double a[kNB];
#pragma unroll
  for (int c = 0; c < kNB; ++c) a[c] = c < nb ? v[kb + c] : 0.0;
#pragma unroll
  for (int c = 0; c < kNB; ++c) {
    if (c < nb) {
      P Xc = L + tri(kb + c) + kb;
      double s = 0.0;
#pragma unroll
      for (int k = 0; k <= c; ++k) s = fma(a[k], Xc[k], s);
      v[kb + c] = s;
      y[c] = s;
    }
  }
}

// Trailing update of one T x T tile: rows r0 + T ti + a, columns r0 + T tj + b
// (lower part only) lose the product of their panel rows.
template <int T, typename P>
__device__ __forceinline__ void trailing_tile(P L, int kb, int r0, int ti, int tj, int n) {
  const int i0 = r0 + T * ti, j0 = r0 + T * tj;
  P pi[T], pj[T];
#pragma unroll
  for (int a = 0; a < T; ++a) {
    pi[a] = L + tri(min(i0 + a, n - 1)) + kb;
    pj[a] = L + tri(min(j0 + a, n - 1)) + kb;
  }
  double acc[T][T];
#pragma unroll
  for (int a = 0; a < T; ++a)
#pragma unroll
    for (int b = 0; b < T; ++b) acc[a][b] = 0.0;
#pragma unroll 1  // a rolled loop keeps only one chunk of rows in registers
  for (int c0 = 0; c0 < kNB; c0 += 4) {
    double xi[T][4], xj[T][4];
#pragma unroll
    for (int a = 0; a < T; ++a)
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        xi[a][c] = pi[a][c0 + c];
        xj[a][c] = pj[a][c0 + c];
      }
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
      for (int a = 0; a < T; ++a)
#pragma unroll
        for (int b = 0; b < T; ++b) acc[a][b] = fma(xi[a][c], xj[b][c], acc[a][b]);
  }
#pragma unroll
  for (int a = 0; a < T; ++a)
#pragma unroll
    for (int b = 0; b < T; ++b) {
      const int i = i0 + a, j = j0 + b;
      if (i < n && j <= i) L[tri(i) + j] -= acc[a][b];
    }
}

__device__ __forceinline__ void tile_index(int e, int& ti, int& tj) {
  int r = (int)((sqrtf(8.0f * (float)e + 1.0f) - 1.0f) * 0.5f);
  while (r * (r + 1) / 2 > e) --r;
  while ((r + 1) * (r + 2) / 2 <= e) ++r;
  ti = r;
  tj = e - r * (r + 1) / 2;
}

// Trailing triangle (rows and columns r0..n-1) in T x T tiles, plus the
// update of the two carried right-hand sides by the block's solution y.
template <int T, typename P, typename V>
__device__ __forceinline__ void trailing_update(P L, int kb, int n, V rb, V re,
                                                const double* yblk) {
  const int r0 = kb + kNB;
  const int mr = n - r0;
  const int mt = (mr + T - 1) / T;
  const int ntl = mt * (mt + 1) / 2;
  // right-hand-side rows first, on the threads that the tiles leave idle last
  const int first = (256 - (ntl + mr) % 256) % 256;
  for (int e = threadIdx.x - first; e < mr; e += 256) {
    if (e < 0) continue;
    const int i = r0 + e;
    P Pi = L + tri(i) + kb;
    double sb = 0.0, se = 0.0;
#pragma unroll
    for (int c = 0; c < kNB; ++c) {
      sb = fma(Pi[c], yblk[c], sb);
      se = fma(Pi[c], yblk[kNB + c], se);
    }
    rb[i] -= sb;
    re[i] -= se;
  }
  for (int e = threadIdx.x; e < ntl; e += 256) {
    int ti, tj;
    tile_index(e, ti, tj);
    trailing_tile<T>(L, kb, r0, ti, tj, n);
  }
}

// Blocked right-looking Cholesky of the assembled n x n lower triangle with
// two right-hand sides carried along (forward substitution fused): on return
// rb, re hold L^-1 b and L^-1 e and the diagonal blocks hold their inverses.
// yblk: 2 * kNB doubles of LDS.
// Returns non-zero if a pivot was not positive.
template <typename P, typename V>
__device__ __forceinline__ int chol_factor(P L, int n, V rb, V re, double* yblk,
                                           double* flag) {
  const int tid = threadIdx.x;
  for (int kb = 0; kb < n; kb += kNB) {
    const int nb = min(kNB, n - kb);
    if (tid < 64) {
      const int bad = chol_diag_block(L, kb, nb);
      if (tid == 0) flag[0] = bad ? 1.0 : 0.0;
    }
    __syncthreads();
    if (flag[0] != 0.0) return 1;  // uniform
    const int mr = n - kb - nb;    // > 0 only for full blocks
    for (int t2 = tid; t2 < mr + 2; t2 += 256) {
      if (t2 < mr)
        panel_row(L, kb, kb + nb + t2);
      else
        block_lower(L, kb, nb, t2 == mr ? rb : re, yblk + (t2 - mr) * kNB);
    }
    __syncthreads();
    if (mr > 0) {
      if (mr >= 48)
        trailing_update<4>(L, kb, n, rb, re, yblk);
      else
        trailing_update<2>(L, kb, n, rb, re, yblk);
    }
    __syncthreads();
  }
  return 0;
}

// Backward substitution L^T x = y with the stored factor; overwrites v.
// yblk: kNB doubles of LDS.
template <typename P, typename V>
__device__ __forceinline__ void chol_backward(P L, int n, V v, double* yblk) {
  const int tid = threadIdx.x;
  for (int kb = ((n - 1) / kNB) * kNB; kb >= 0; kb -= kNB) {
    const int nb = min(kNB, n - kb);
    if (tid < nb) {  // x_c = sum_{k >= c} X[k][c] v_k
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < kNB; ++k)
        if (k >= tid && k < nb) s = fma(L[tri(kb + k) + kb + tid], v[kb + k], s);
      yblk[tid] = s;
    }
    __syncthreads();
    if (tid < nb) v[kb + tid] = yblk[tid];
    for (int i = tid; i < kb; i += 256) {
      double s = 0.0;
#pragma unroll
      for (int cc = 0; cc < kNB; ++cc)
        if (cc < nb) s = fma(L[tri(kb + cc) + i], yblk[cc], s);  // last block may be partial
      v[i] -= s;
    }
    __syncthreads();
  }
}

// Forward substitution L y = v with the stored factor; overwrites v.
template <typename P, typename V>
__device__ __forceinline__ void chol_forward(P L, int n, V v, double* yblk) {
  const int tid = threadIdx.x;
  for (int kb = 0; kb < n; kb += kNB) {
    const int nb = min(kNB, n - kb);
    if (tid < nb) {  // y_c = sum_{k <= c} X[c][k] v_k
      P Xc = L + tri(kb + tid) + kb;
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < kNB; ++k)
        if (k <= tid) s = fma(Xc[k], v[kb + k], s);
      yblk[tid] = s;
    }
    __syncthreads();
    if (tid < nb) v[kb + tid] = yblk[tid];
    for (int i = kb + nb + tid; i < n; i += 256) {
      P Li = L + tri(i) + kb;
      double s = 0.0;
#pragma unroll
      for (int cc = 0; cc < kNB; ++cc) s = fma(Li[cc], yblk[cc], s);
      v[i] -= s;
    }
    __syncthreads();
  }
}

// LDS placement of the EEQ solve for a padded length N: the packed factor
// and `vectors` per-atom vectors (kEeqAllLds), the factor only
// (kEeqFactorLds), or nothing (kEeqGlobal, HBM workspace).
constexpr int kEeqGlobal = 0, kEeqFactorLds = 1, kEeqAllLds = 2;
constexpr size_t kLdsBytes = 160 * 1024 - 1024;  // dynamic LDS budget per work-group
constexpr int kEeqVectors = 6;                   // y_b, y_e and x, y, z, radius of each atom

inline int eeq_mode(int64_t N, int vectors) {
  const size_t f = sizeof(double) * (size_t)tri((int)N);
  if (f + sizeof(double) * vectors * (size_t)N <= kLdsBytes) return kEeqAllLds;
  if (f <= kLdsBytes) return kEeqFactorLds;
  return kEeqGlobal;
}

inline size_t eeq_shmem(int mode, int64_t N, int vectors) {
  if (mode == kEeqGlobal) return 0;
  return sizeof(double) * ((size_t)tri((int)N) + (mode == kEeqAllLds ? vectors * N : 0));
}

// Row of a strictly-lower packed index e: i (i - 1) / 2 <= e < i (i + 1) / 2.
__device__ __forceinline__ int strict_row(int e) {
  int i = (int)((sqrtf(8.0f * (float)e + 1.0f) + 1.0f) * 0.5f);
  while (i * (i - 1) / 2 > e) --i;
  while (i * (i + 1) / 2 <= e) ++i;
  return i;
}

// Two work-groups per CU fit the LDS of a 120-atom factor; keep the register
// use low enough for two waves per SIMD as well.
template <int MODE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_eeq_chol(
    DevTables t, Params p, Inputs in, Work w) {
  extern __shared__ double smem[];
  __shared__ double s_small[2 * kNB + 2 * kWaves + 2];  // two solution blocks, sums, flag
  const int64_t m = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t N = w.N;
  const int n = w.natom[m];
  const int64_t* Z = in.numbers + m * N;
  const double* R = in.positions + m * N * 3;
  double* q = w.q + m * N;

  if (in.q_in != nullptr) {  // charges supplied by the caller
    for (int i = tid; i < N; i += 256)
      q[i] = (i < n && Z[i] != 0) ? in.q_in[m * N + i] : 0.0;
    return;
  }
  for (int i = n + tid; i < N; i += 256) q[i] = 0.0;
  if (n == 0) return;

  double* gL = w.chol + (size_t)m * (size_t)tri((int)N);
  double* L = MODE == kEeqGlobal ? gL : smem;
  // y_b = L^-1 b is scratch in lam until the adjoint; y_e = L^-1 e is kept in ez
  double* rb = MODE == kEeqAllLds ? smem + tri((int)N) : w.lam + m * N;
  double* re = MODE == kEeqAllLds ? rb + N : w.ez + m * N;
  double* yblk = s_small;
  double* red = s_small + 2 * kNB;
  double* flag = red + 2 * kWaves;

  // assemble the lower triangle of C and the two right-hand sides
  if (MODE == kEeqAllLds) {
    // stage positions and radii (0 marks a padding atom), then fill the
    // strictly-lower triangle with all threads
    double* sx = re + N;
    double* sy = sx + N;
    double* sz = sy + N;
    double* srad = sz + N;
    for (int i = tid; i < n; i += 256) {
      const int zi = (int)Z[i];
      sx[i] = R[3 * i];
      sy[i] = R[3 * i + 1];
      sz[i] = R[3 * i + 2];
      srad[i] = zi != 0 ? t.rad[zi] : 0.0;
      L[tri(i) + i] = zi != 0 ? eeq_diag(t, zi) : 1.0;
      rb[i] = zi != 0 ? eeq_rhs(t, zi, w.cnq[m * N + i]) : 0.0;
      re[i] = zi != 0 ? 1.0 : 0.0;
    }
    __syncthreads();
    const int ne = n * (n - 1) / 2;
    for (int e = tid; e < ne; e += 256) {
      const int i = strict_row(e);
      const int j = e - i * (i - 1) / 2;
      const double ri = srad[i], rj = srad[j];
      double v = 0.0;
      if (ri != 0.0 && rj != 0.0) {
        const double dx = sx[i] - sx[j], dy = sy[i] - sy[j], dz = sz[i] - sz[j];
        const double r = sqrt(dx * dx + dy * dy + dz * dz);
        v = erf(r / sqrt(ri * ri + rj * rj)) / r;
      }
      L[tri(i) + j] = v;
    }
  } else {
    for (int i = wave; i < n; i += kWaves) {
      const int zi = (int)Z[i];
      const double xi = R[3 * i], yi = R[3 * i + 1], zzi = R[3 * i + 2];
      double* Li = L + tri(i);
      for (int j = lane; j <= i; j += 64) {
        const int zj = (int)Z[j];
        double v;
        if (j == i) {
          v = zi != 0 ? eeq_diag(t, zi) : 1.0;
        } else if (zi != 0 && zj != 0) {
          const double dx = xi - R[3 * j], dy = yi - R[3 * j + 1], dz = zzi - R[3 * j + 2];
          v = eeq_offdiag(t, zi, zj, sqrt(dx * dx + dy * dy + dz * dz));
        } else {
          v = 0.0;
        }
        Li[j] = v;
      }
      if (lane == 0) {
        rb[i] = zi != 0 ? eeq_rhs(t, zi, w.cnq[m * N + i]) : 0.0;
        re[i] = zi != 0 ? 1.0 : 0.0;
      }
    }
  }
  __syncthreads();

  if (p.force_lu || chol_factor(L, n, rb, re, yblk, flag)) {
    if (tid == 0) w.status[m] |= kStatusLU;
    return;
  }
  // e^T C^-1 b = y_e . y_b and e^T C^-1 e = y_e . y_e, so one backward
  // substitution of y_b - mu y_e gives q = C^-1 (b - mu e)
  double sbe = 0.0, see = 0.0;
  for (int i = tid; i < n; i += 256) {
    sbe = fma(re[i], rb[i], sbe);
    see = fma(re[i], re[i], see);
  }
  sbe = block_sum(sbe, red);
  see = block_sum(see, red + kWaves);
  const double mu = (sbe - in.charge[m]) / see;
  for (int i = tid; i < n; i += 256) rb[i] = fma(-mu, re[i], rb[i]);
  __syncthreads();
  chol_backward(L, n, rb, yblk);
  for (int i = tid; i < n; i += 256) q[i] = Z[i] != 0 ? rb[i] : 0.0;
  if (MODE == kEeqAllLds) {
    double* gz = w.ez + m * N;
    for (int i = tid; i < n; i += 256) gz[i] = re[i];
  }
  if (MODE != kEeqGlobal) {
    const int nt = tri(n);
    for (int k = tid; k < nt; k += 256) gL[k] = L[k];
  }
}

// Pivoted LU of the full bordered matrix in HBM, for molecules whose
// Cholesky failed (or all molecules under D4_FLAG_FORCE_LU).
__device__ void lu_assemble(const DevTables& t, const Inputs& in, const Work& w,
                            int64_t m, int n, double* A, int ld, double* rhs) {
  const int64_t N = w.N;
  const int64_t* Z = in.numbers + m * N;
  const double* R = in.positions + m * N * 3;
  for (int e = threadIdx.x; e < (n + 1) * (n + 1); e += 256) {
    const int i = e / (n + 1), j = e - i * (n + 1);
    double v;
    if (i == n && j == n) {
      v = 0.0;
    } else if (i == n || j == n) {
      const int k = i == n ? j : i;
      v = Z[k] != 0 ? 1.0 : 0.0;
    } else {
      const int zi = (int)Z[i], zj = (int)Z[j];
      if (i == j) {
        v = zi != 0 ? eeq_diag(t, zi) : 1.0;
      } else if (zi != 0 && zj != 0) {
        const double dx = R[3 * i] - R[3 * j], dy = R[3 * i + 1] - R[3 * j + 1],
                     dz = R[3 * i + 2] - R[3 * j + 2];
        v = eeq_offdiag(t, zi, zj, sqrt(dx * dx + dy * dy + dz * dz));
      } else {
        v = 0.0;
      }
    }
    A[(size_t)i * ld + j] = v;
  }
  if (rhs != nullptr) {
    for (int i = threadIdx.x; i < n; i += 256) {
      const int zi = (int)Z[i];
      rhs[i] = zi != 0 ? eeq_rhs(t, zi, w.cnq[m * N + i]) : 0.0;
    }
    if (threadIdx.x == 0) rhs[n] = in.charge[m];
  }
}

// Solve (P A = L U) x = b in place for b in HBM.
__device__ void lu_solve(const double* A, int ld, const int32_t* piv, int n1,
                         double* b) {
  const int tid = threadIdx.x;
  __syncthreads();
  if (tid == 0) {
    for (int k = 0; k < n1; ++k) {
      const int pk = piv[k];
      if (pk != k) {
        const double tmp = b[k];
        b[k] = b[pk];
        b[pk] = tmp;
      }
    }
  }
  for (int k = 0; k < n1; ++k) {  // unit lower
    __syncthreads();
    const double yk = b[k];
    for (int i = k + 1 + tid; i < n1; i += 256) b[i] -= A[(size_t)i * ld + k] * yk;
  }
  for (int k = n1 - 1; k >= 0; --k) {  // upper
    __syncthreads();
    const double xk = b[k] / A[(size_t)k * ld + k];
    for (int i = tid; i < k; i += 256) b[i] -= A[(size_t)i * ld + k] * xk;
  }
  __syncthreads();
  for (int i = tid; i < n1; i += 256) b[i] /= A[(size_t)i * ld + i];
  __syncthreads();
}

__global__ __launch_bounds__(256) void k_eeq_lu(DevTables t, Params p, Inputs in,
                                                Work w) {
  __shared__ double s_val[256];
  __shared__ int s_idx[256];
  const int64_t m = blockIdx.x;
  if (in.q_in != nullptr) return;
  if (!(w.status[m] & kStatusLU)) return;
  const int n = w.natom[m];
  const int n1 = n + 1;
  const int ld = (int)w.N + 1;
  const int tid = threadIdx.x;
  double* A = w.lu + (size_t)m * ld * ld;
  int32_t* piv = w.piv + m * ld;
  double* vec = w.vec + m * ld;
  lu_assemble(t, in, w, m, n, A, ld, vec);
  __syncthreads();
  int singular = 0;
  for (int k = 0; k < n1; ++k) {
    // pivot search in column k
    double best = -1.0;
    int bi = k;
    for (int i = k + tid; i < n1; i += 256) {
      const double v = fabs(A[(size_t)i * ld + k]);
      if (v > best) {
        best = v;
        bi = i;
      }
    }
    s_val[tid] = best;
    s_idx[tid] = bi;
    __syncthreads();
    for (int off = 128; off >= 1; off >>= 1) {
      if (tid < off) {
        const double o = s_val[tid + off];
        const int oi = s_idx[tid + off];
        if (o > s_val[tid] || (o == s_val[tid] && oi < s_idx[tid])) {
          s_val[tid] = o;
          s_idx[tid] = oi;
        }
      }
      __syncthreads();
    }
    const int pk = s_idx[0];
    const double pv = s_val[0];
    __syncthreads();
    if (!(pv > 0.0)) {
      singular = 1;
      break;
    }
    if (tid == 0) piv[k] = pk;
    if (pk != k) {
      for (int j = tid; j < n1; j += 256) {
        const double a = A[(size_t)k * ld + j];
        A[(size_t)k * ld + j] = A[(size_t)pk * ld + j];
        A[(size_t)pk * ld + j] = a;
      }
    }
    __syncthreads();
    const double inv = 1.0 / A[(size_t)k * ld + k];
    for (int i = k + 1 + tid; i < n1; i += 256) A[(size_t)i * ld + k] *= inv;
    __syncthreads();
    const int rem = n1 - k - 1;
    for (int e = tid; e < rem * rem; e += 256) {
      const int i = k + 1 + e / rem, j = k + 1 + e % rem;
      A[(size_t)i * ld + j] -= A[(size_t)i * ld + k] * A[(size_t)k * ld + j];
    }
    __syncthreads();
  }
  if (singular) {
    if (tid == 0) w.status[m] |= kStatusSingular;
    for (int i = tid; i < n; i += 256) w.q[m * w.N + i] = NAN;
    return;
  }
  lu_solve(A, ld, piv, n1, vec);
  for (int i = tid; i < n; i += 256)
    w.q[m * w.N + i] = in.numbers[m * w.N + i] != 0 ? vec[i] : 0.0;
}

// --------------------------------------------------------------------------
// Gaussian reference weights (model/d4.py:137-228) and their derivatives,
// one thread per atom
// --------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_weights(DevTables t, Params p, Work w,
                                                 Inputs in) {
  const int64_t row = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (row >= w.B * w.N) return;
  const int64_t m = row / w.N;
  const int i = (int)(row - m * w.N);
  const int z = i < w.natom[m] ? (int)in.numbers[row] : 0;
  double* gw = w.gw + row * kRefStride;
  double* gw0 = w.gw0 + row * kRefStride;
  double* dgw = w.dgw + row * kRefStride;
  double* dgw0 = w.dgw0 + row * kRefStride;
  double* dgwq = w.dgwq + row * kRefStride;
  const int nr = z != 0 ? t.nref[z] : 0;
  const double cn = z != 0 ? w.cn[row] : 0.0;
  const double q = z != 0 ? w.q[row] : 0.0;
  double expw[kMaxRef], dexpw[kMaxRef];
  double norm = 0.0, dnorm = 0.0;
#pragma unroll
  for (int a = 0; a < kMaxRef; ++a) {
    double e = 0.0, de = 0.0;
    if (a < nr) {
      // exp(-wf k d^2) evaluated as exp(-d^2)^(wf k), the reference's order
      // of operations (model/d4.py:174-188): identical rounding where the
      // Gaussians approach underflow
      const double d = cn - t.refcn[z * kMaxRef + a];
      const double tmp = exp(-(d * d));
      const int ng = t.ngw[z * kMaxRef + a];
      for (int k = 1; k <= ng; ++k) {
        const double kw = k * p.wf;
        const double ek = pow(tmp, kw);
        e += ek;
        de -= 2.0 * kw * d * ek;
      }
    }
    expw[a] = e;
    dexpw[a] = de;
    norm += e;
    dnorm += de;
  }
  const bool exceptional = !(norm > 0.0);
  const double maxcn = z != 0 ? t.maxcn[z] : 0.0;
  const double zeff = t.zeff[z];
  const double gam = t.gam[z] * p.gc;
#pragma unroll
  for (int a = 0; a < kRefStride; ++a) {
    double v = 0.0, v0 = 0.0, dv = 0.0, dv0 = 0.0, dq = 0.0;
    if (a < nr) {
      double g, dg;
      if (exceptional) {
        g = t.refcn[z * kMaxRef + a] == maxcn ? 1.0 : 0.0;
        dg = 0.0;
      } else {
        // form the ratio first: expw * dnorm underflows when the Gaussians
        // are far below 1 (CN several units from every reference CN)
        g = expw[a] / norm;
        dg = (dexpw[a] - g * dnorm) / norm;
      }
      const double qref = t.refq[z * kMaxRef + a] + zeff;
      const double zq = zeta_fn(p.ga, gam, qref, q + zeff);
      const double z0 = zeta_fn(p.ga, gam, qref, zeff);
      v = zq * g;
      v0 = z0 * g;
      dv = zq * dg;
      dv0 = z0 * dg;
      dq = dzeta_fn(p.ga, gam, qref, q + zeff) * g;
    }
    gw[a] = v;
    gw0[a] = v0;
    dgw[a] = dv;
    dgw0[a] = dv0;
    dgwq[a] = dq;
  }
}

// --------------------------------------------------------------------------
// pair pass: C6(q) and C6(0) contraction with the reference-C6 table, the
// rational-damped two-body term, the ATM pair factors; with GRAD also the
// two-body gradient pieces (one wave per atom row)
// --------------------------------------------------------------------------

template <bool GRAD>
__global__ __launch_bounds__(256) void k_pairs(DevTables t, Params p, Inputs in,
                                               Work w, double* grad_out) {
  const int lane = threadIdx.x & 63;
  const int64_t row = (int64_t)blockIdx.x * kWaves + (threadIdx.x >> 6);
  if (row >= w.B * w.N) return;
  const int64_t N = w.N;
  const int64_t m = row / N;
  const int i = (int)(row - m * N);
  const int n = w.natom[m];
  const int64_t* Z = in.numbers + m * N;
  const double* R = in.positions + m * N * 3;
  double* ptr = w.pt + row * N;
  double* ppr = w.pp + row * N;
  double* pkr = w.pk + row * N;
  const int zi = i < n ? (int)Z[i] : 0;
  if (zi == 0) {
    for (int j = lane; j < n; j += 64) {
      ptr[j] = 0.0;
      ppr[j] = 0.0;
      if (GRAD) pkr[j] = 0.0;
    }
    if (lane == 0) {
      w.e2[row] = 0.0;
      if (GRAD) {
        w.dldcn[row] = 0.0;
        w.dldq[row] = 0.0;
        grad_out[3 * row] = 0.0;
        grad_out[3 * row + 1] = 0.0;
        grad_out[3 * row + 2] = 0.0;
      }
    }
    return;
  }
  const double xi = R[3 * i], yi = R[3 * i + 1], zzi = R[3 * i + 2];
  const double r4i = in.r4r2_in ? in.r4r2_in[row] : t.r4r2[zi];
  const int nri = t.nref[zi];
  double gwi[kMaxRef], gw0i[kMaxRef], dgwi[kMaxRef], dgw0i[kMaxRef], dgwqi[kMaxRef];
#pragma unroll
  for (int a = 0; a < kMaxRef; ++a) {
    gwi[a] = w.gw[row * kRefStride + a];
    gw0i[a] = w.gw0[row * kRefStride + a];
    if (GRAD) {
      dgwi[a] = w.dgw[row * kRefStride + a];
      dgw0i[a] = w.dgw0[row * kRefStride + a];
      dgwqi[a] = w.dgwq[row * kRefStride + a];
    }
  }
  const double gi = GRAD ? (in.gE ? in.gE[row] : 1.0) : 0.0;
  const double alp3 = p.alp / 3.0;
  double acc_e = 0.0, acc_cn = 0.0, acc_q = 0.0, fx = 0.0, fy = 0.0, fz = 0.0;
  for (int j = lane; j < n; j += 64) {
    const int zj = (int)Z[j];
    if (j == i || zj == 0) {
      ptr[j] = 0.0;
      ppr[j] = 0.0;
      if (GRAD) pkr[j] = 0.0;
      continue;
    }
    const int64_t rj = m * N + j;
    const double dx = xi - R[3 * j], dy = yi - R[3 * j + 1], dz = zzi - R[3 * j + 2];
    const double r2 = dx * dx + dy * dy + dz * dz;
    const double r = sqrt(r2);
    const double r4j = in.r4r2_in ? in.r4r2_in[rj] : t.r4r2[zj];
    double gwj[kMaxRef], gw0j[kMaxRef];
#pragma unroll
    for (int b = 0; b < kMaxRef; ++b) {
      gwj[b] = w.gw[rj * kRefStride + b];
      gw0j[b] = w.gw0[rj * kRefStride + b];
    }
    const double* blk = t.rc6 + ((size_t)zi * kMaxZ + zj) * kRc6Block;
    double c6 = 0.0, c60 = 0.0, dc6cn = 0.0, dc6q = 0.0, dc60cn = 0.0;
    for (int a = 0; a < nri; ++a) {
      double va = 0.0, v0a = 0.0;
#pragma unroll
      for (int b = 0; b < kMaxRef; ++b) {
        const double rc = blk[a * kMaxRef + b];
        va += rc * gwj[b];
        v0a += rc * gw0j[b];
      }
      c6 += gwi[a] * va;
      c60 += gw0i[a] * v0a;
      if (GRAD) {
        dc6cn += dgwi[a] * va;
        dc6q += dgwqi[a] * va;
        dc60cn += dgw0i[a] * v0a;
      }
    }
    // rational damping, radii = 3 r4r2_i r4r2_j (twobody.py:147-157)
    const double qq = 3.0 * r4i * r4j;
    const double r0 = p.a1 * sqrt(qq) + p.a2;
    const double r6 = r2 * r2 * r2, r8 = r6 * r2;
    const double r06 = r0 * r0 * r0 * r0 * r0 * r0, r08 = r06 * r0 * r0;
    const double t6 = 1.0 / (r6 + r06);
    const double t8 = 1.0 / (r8 + r08);
    double de_dc6 = p.s6 * t6 + p.s8 * qq * t8;
    double t10 = 0.0, r10 = 0.0;
    if (p.s10on) {
      r10 = r8 * r2;
      t10 = 1.0 / (r10 + r08 * r0 * r0);
      de_dc6 += p.s10 * qq * qq * (49.0 / 40.0) * t10;
    }
    de_dc6 *= -0.5;
    const bool in2 = r <= p.cut2;
    if (in2) acc_e += de_dc6 * c6;
    // ATM pair factors (threebody.py:106-145, 241-256, 311-321)
    const double tf = pow(r0 / r, alp3);
    const double pf = sqrt(fabs(c60)) / (r2 * r);
    ptr[j] = tf;
    ppr[j] = pf;
    if (GRAD) {
      pkr[j] = c60 != 0.0 ? 0.5 * dc60cn / c60 : 0.0;
      if (in2) {
        const double gj = in.gE ? in.gE[rj] : 1.0;
        const double wij = gi + gj;
        acc_cn += wij * de_dc6 * dc6cn;
        acc_q += wij * de_dc6 * dc6q;
        // d/dr of 1/(r^n + r0^n) = -n r^(n-1) t^2
        double dedr = p.s6 * (-6.0 * r6 / r) * t6 * t6 +
                      p.s8 * qq * (-8.0 * r8 / r) * t8 * t8;
        if (p.s10on) dedr += p.s10 * qq * qq * (49.0 / 40.0) * (-10.0 * r10 / r) * t10 * t10;
        dedr *= -0.5 * c6;
        const double coef = wij * dedr / r;
        fx += coef * dx;
        fy += coef * dy;
        fz += coef * dz;
      }
    }
  }
  acc_e = wave_sum(acc_e);
  if (GRAD) {
    acc_cn = wave_sum(acc_cn);
    acc_q = wave_sum(acc_q);
    fx = wave_sum(fx);
    fy = wave_sum(fy);
    fz = wave_sum(fz);
  }
  if (lane == 0) {
    w.e2[row] = acc_e;
    if (GRAD) {
      w.dldcn[row] = acc_cn;
      w.dldq[row] = acc_q;
      grad_out[3 * row] = fx;
      grad_out[3 * row + 1] = fy;
      grad_out[3 * row + 2] = fz;
    }
  }
}

// --------------------------------------------------------------------------
// ATM three-body term. One work-group owns a 16 x 16 tile of atom pairs
// (x, y), x < y, of one molecule; each lane owns one pair and loops over all
// third atoms z, which are staged through LDS kZB rows at a time. A lane
// accumulates only quantities of its own pair, so there are no atomics:
//   pe[x][y], pe[y][x]  atom-energy partials of x and y,
//   pg[x][y] = pg[y][x] = dL / d r_xy^2,
//   ph[x][y] = ph[y][x] = dL / d log sqrt|C6(0)_xy|.
// The cutoff screens r_ij and r_jk of every ordered (i, j, k) term and never
// r_ik (threebody.py:153-160); per unordered triple this gives atom x the
// share e/6 [yz]([xy] + [xz]).
// --------------------------------------------------------------------------

template <bool ENERGY, bool GRAD>
__global__ __launch_bounds__(256) void k_atm(Params p, Inputs in, Work w,
                                             int tiles_per_mol, int64_t ntiles) {
  __shared__ double s_t[kZB][2 * kTile];
  __shared__ double s_p[kZB][2 * kTile];
  __shared__ double s_r[kZB][4];  // x, y, z, dL/dE (0 for padding)
  __shared__ int s_ok[kZB];

  // spread consecutive logical tiles over one XCD: the hardware deals
  // blocks round-robin to the 8 XCDs, so block b runs on XCD b % 8
  const int64_t G = gridDim.x;
  const int64_t per_xcd = G / 8;
  const int64_t b = blockIdx.x;
  const int64_t logical = (b % 8) * per_xcd + b / 8;
  if (logical >= ntiles) return;
  const int64_t m = logical / tiles_per_mol;
  const int tile = (int)(logical - m * tiles_per_mol);
  const int64_t N = w.N;
  const int n = w.natom[m];
  const int nb = (n + kTile - 1) / kTile;
  if (tile >= nb * (nb + 1) / 2) return;
  int by = (int)((sqrt(8.0 * tile + 1.0) - 1.0) * 0.5);
  while (by * (by + 1) / 2 > tile) --by;
  while ((by + 1) * (by + 2) / 2 <= tile) ++by;
  const int bx = tile - by * (by + 1) / 2;

  const int tx = threadIdx.x & (kTile - 1), ty = threadIdx.x >> 4;
  const int x = bx * kTile + tx, y = by * kTile + ty;
  const int64_t* Z = in.numbers + m * N;
  const double* R = in.positions + m * N * 3;
  const double* pt = w.pt + m * N * N;
  const double* pp = w.pp + m * N * N;
  const bool valid = x < y && y < n && Z[x] != 0 && Z[y] != 0;

  double px = 0, py = 0, pz = 0, qx = 0, qy = 0, qz = 0, a = 1.0;
  double txy = 0.0, pxy = 0.0, gx = 0.0, gy = 0.0;
  bool mxy = false;
  if (valid) {
    px = R[3 * x];
    py = R[3 * x + 1];
    pz = R[3 * x + 2];
    qx = R[3 * y];
    qy = R[3 * y + 1];
    qz = R[3 * y + 2];
    const double dx = px - qx, dy = py - qy, dz = pz - qz;
    a = dx * dx + dy * dy + dz * dz;
    txy = pt[(int64_t)x * N + y];
    pxy = pp[(int64_t)x * N + y];
    mxy = a <= p.cut3sq;
    if (GRAD) {
      gx = in.gE ? in.gE[m * N + x] : 1.0;
      gy = in.gE ? in.gE[m * N + y] : 1.0;
    }
  }
  const double s9 = p.s9, alp = p.alp;
  double ex = 0.0, ey = 0.0, accg = 0.0, acch = 0.0;

  for (int z0 = 0; z0 < n; z0 += kZB) {
    __syncthreads();
    for (int e = threadIdx.x; e < kZB * 4 * kTile; e += 256) {
      const int zl = e / (4 * kTile), c = e - zl * (4 * kTile);
      const int z = z0 + zl;
      const int half = c & (2 * kTile - 1);  // 0..31
      const int col = half < kTile ? bx * kTile + half : by * kTile + half - kTile;
      double v = 0.0;
      if (z < n && col < n) v = (c < 2 * kTile ? pt : pp)[(int64_t)z * N + col];
      if (c < 2 * kTile)
        s_t[zl][half] = v;
      else
        s_p[zl][half] = v;
    }
    if (threadIdx.x < kZB) {
      const int zl = threadIdx.x, z = z0 + zl;
      const bool ok = z < n && Z[z] != 0;
      s_ok[zl] = ok;
      s_r[zl][0] = ok ? R[3 * z] : 0.0;
      s_r[zl][1] = ok ? R[3 * z + 1] : 0.0;
      s_r[zl][2] = ok ? R[3 * z + 2] : 0.0;
      s_r[zl][3] = (GRAD && ok) ? (in.gE ? in.gE[m * N + z] : 1.0) : 0.0;
    }
    __syncthreads();
    if (!valid) continue;
    const int zmax = min(kZB, n - z0);
    for (int zl = 0; zl < zmax; ++zl) {
      const int z = z0 + zl;
      const bool vz = s_ok[zl] && z != x && z != y;
      const double rzx = s_r[zl][0], rzy = s_r[zl][1], rzz = s_r[zl][2];
      const double dxx = px - rzx, dxy = py - rzy, dxz = pz - rzz;
      const double dyx = qx - rzx, dyy = qy - rzy, dyz = qz - rzz;
      const double bb = vz ? dxx * dxx + dxy * dxy + dxz * dxz : 1.0;  // r_xz^2
      const double cc = vz ? dyx * dyx + dyy * dyy + dyz * dyz : 1.0;  // r_yz^2
      const double T = txy * s_t[zl][tx] * s_t[zl][kTile + ty];
      const double P = vz ? pxy * s_p[zl][tx] * s_p[zl][kTile + ty] : 0.0;
      const double X = a + bb - cc, Y = a - bb + cc, W3 = -a + bb + cc;
      const double s = X * Y * W3;
      const double abc = a * bb * cc;
      const double opT = 1.0 + 6.0 * T;
      const double q0 = 1.0 / (abc * opT);
      const double f = abc * q0;      // 1 / (1 + 6T)
      const double iabc = opT * q0;   // 1 / (abc)
      const double g = 0.375 * s * iabc + 1.0;
      const double en = s9 * P * f * g;
      const bool mxz = bb <= p.cut3sq, myz = cc <= p.cut3sq;
      if (ENERGY) {
        ex += (mxy && myz) ? en : 0.0;
        ey += (mxy && mxz) ? en : 0.0;
      }
      if (GRAD) {
        const double gz = s_r[zl][3];
        const double wgt =
            (gx * (myz ? (double)(mxy + mxz) : 0.0) + gy * (mxz ? (double)(mxy + myz) : 0.0) +
             gz * (mxy ? (double)(mxz + myz) : 0.0)) *
            (1.0 / 6.0);
        const double sa = Y * W3 + X * W3 - X * Y;  // ds/da
        const double ia = bb * cc * iabc;           // 1/a
        const double dea =
            s9 * P * f * (g * ia * (T * f * alp - 1.5) + 0.375 * (sa - s * ia) * iabc);
        accg += wgt * dea;
        acch += wgt * en;
      }
    }
  }
  if (!valid) return;
  const int64_t ixy = (m * N + x) * N + y, iyx = (m * N + y) * N + x;
  if (ENERGY) {
    w.pe[ixy] = ex * (1.0 / 6.0);
    w.pe[iyx] = ey * (1.0 / 6.0);
  }
  if (GRAD) {
    w.pg[ixy] = accg;
    w.pg[iyx] = accg;
    w.ph[ixy] = acch;
    w.ph[iyx] = acch;
  }
}

// Row reductions of the ATM pair outputs: atom energies and, with GRAD, the
// CN chain through C6(q=0) and the explicit ATM gradient.
template <bool ENERGY, bool GRAD>
__global__ __launch_bounds__(256) void k_atm_rows(Inputs in, Work w,
                                                  double* energy_out,
                                                  double* grad_out) {
  const int lane = threadIdx.x & 63;
  const int64_t row = (int64_t)blockIdx.x * kWaves + (threadIdx.x >> 6);
  if (row >= w.B * w.N) return;
  const int64_t N = w.N;
  const int64_t m = row / N;
  const int i = (int)(row - m * N);
  const int n = w.natom[m];
  const int64_t* Z = in.numbers + m * N;
  const double* R = in.positions + m * N * 3;
  const int zi = i < n ? (int)Z[i] : 0;
  if (zi == 0) {
    if (lane == 0 && ENERGY) energy_out[row] = 0.0;
    return;
  }
  const double xi = R[3 * i], yi = R[3 * i + 1], zzi = R[3 * i + 2];
  double e3 = 0.0, dcn = 0.0, fx = 0.0, fy = 0.0, fz = 0.0;
  const double* per = w.pe + row * N;
  const double* pgr = w.pg + row * N;
  const double* phr = w.ph + row * N;
  const double* pkr = w.pk + row * N;
  for (int j = lane; j < n; j += 64) {
    if (j == i || Z[j] == 0) continue;
    if (ENERGY) e3 += per[j];
    if (GRAD) {
      dcn += phr[j] * pkr[j];
      const double c = 2.0 * pgr[j];
      fx += c * (xi - R[3 * j]);
      fy += c * (yi - R[3 * j + 1]);
      fz += c * (zzi - R[3 * j + 2]);
    }
  }
  if (ENERGY) e3 = wave_sum(e3);
  if (GRAD) {
    dcn = wave_sum(dcn);
    fx = wave_sum(fx);
    fy = wave_sum(fy);
    fz = wave_sum(fz);
  }
  if (lane == 0) {
    if (ENERGY) energy_out[row] = w.e2[row] + e3;
    if (GRAD) {
      w.dldcn[row] += dcn;
      grad_out[3 * row] += fx;
      grad_out[3 * row + 1] += fy;
      grad_out[3 * row + 2] += fz;
    }
  }
}

// --------------------------------------------------------------------------
// EEQ adjoint: lambda = A^-1 [dL/dq; 0] (A symmetric), then
// beta_i = lambda_i d b_i / d raw-CN_i. One work-group per molecule.
// --------------------------------------------------------------------------

template <int MODE>
__global__ __launch_bounds__(256) void k_adjoint(DevTables t, Params p, Inputs in,
                                                 Work w) {
  extern __shared__ double smem[];
  __shared__ double red[2 * kWaves];
  __shared__ double yblk[kNB];
  const int64_t m = blockIdx.x;
  const int64_t N = w.N;
  const int n = w.natom[m];
  const int tid = threadIdx.x;
  const int64_t* Z = in.numbers + m * N;
  double* lam = w.lam + m * N;
  double* beta = w.beta + m * N;
  if (in.q_in != nullptr || n == 0 || (w.status[m] & kStatusSingular)) {
    for (int i = tid; i < N; i += 256) {
      lam[i] = 0.0;
      beta[i] = 0.0;
    }
    return;
  }
  for (int i = n + tid; i < N; i += 256) {
    lam[i] = 0.0;
    beta[i] = 0.0;
  }
  const double* g = w.dldq + m * N;
  if (w.status[m] & kStatusLU) {
    const int ld = (int)N + 1;
    const double* A = w.lu + (size_t)m * ld * ld;
    const int32_t* piv = w.piv + m * ld;
    double* vec = w.vec + m * ld;
    for (int i = tid; i < n; i += 256) vec[i] = Z[i] != 0 ? g[i] : 0.0;
    if (tid == 0) vec[n] = 0.0;
    lu_solve(A, ld, piv, n + 1, vec);
    for (int i = tid; i < n; i += 256) lam[i] = Z[i] != 0 ? vec[i] : 0.0;
  } else {
    double* gL = w.chol + (size_t)m * (size_t)tri((int)N);
    double* L = MODE == kEeqGlobal ? gL : smem;
    double* v = MODE == kEeqAllLds ? smem + tri((int)N) : lam;
    const double* ye = w.ez + m * N;
    if (MODE != kEeqGlobal) {
      const int nt = tri(n);
      for (int k = tid; k < nt; k += 256) L[k] = gL[k];
    }
    for (int i = tid; i < n; i += 256) v[i] = Z[i] != 0 ? g[i] : 0.0;
    __syncthreads();
    // lambda = C^-1 (g - nu e), nu = e^T C^-1 g / e^T C^-1 e, with the
    // forward-substituted y_e = L^-1 e kept by the charge kernel
    chol_forward(L, n, v, yblk);
    double sw = 0.0, sz = 0.0;
    for (int i = tid; i < n; i += 256) {
      sw = fma(ye[i], v[i], sw);
      sz = fma(ye[i], ye[i], sz);
    }
    sw = block_sum(sw, red);
    sz = block_sum(sz, red + kWaves);
    const double nu = sw / sz;
    for (int i = tid; i < n; i += 256) v[i] = fma(-nu, ye[i], v[i]);
    __syncthreads();
    chol_backward(L, n, v, yblk);
    for (int i = tid; i < n; i += 256) lam[i] = Z[i] != 0 ? v[i] : 0.0;
  }
  __syncthreads();
  for (int i = tid; i < n; i += 256) {
    const int zi = (int)Z[i];
    double b = 0.0;
    if (zi != 0) {
      const double c = w.cnq[m * N + i];
      const double dsq = c > kEps ? 0.5 / sqrt(c) : 0.0;
      b = lam[i] * t.kcn[zi] * dsq * w.dcnq[m * N + i];
    }
    beta[i] = b;
  }
}

// --------------------------------------------------------------------------
// chain pass: gradient through the D4 CN, the EEQ CN and the EEQ Coulomb
// matrix (one wave per atom row)
// --------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_chain(DevTables t, Params p, Inputs in,
                                               Work w, double* grad_out) {
  const int lane = threadIdx.x & 63;
  const int64_t row = (int64_t)blockIdx.x * kWaves + (threadIdx.x >> 6);
  if (row >= w.B * w.N) return;
  const int64_t N = w.N;
  const int64_t m = row / N;
  const int i = (int)(row - m * N);
  const int n = w.natom[m];
  const int64_t* Z = in.numbers + m * N;
  const double* R = in.positions + m * N * 3;
  const int zi = i < n ? (int)Z[i] : 0;
  if (zi == 0) return;
  const bool eeq = in.q_in == nullptr;
  const double xi = R[3 * i], yi = R[3 * i + 1], zzi = R[3 * i + 2];
  const double rci = t.rcov[zi], eni = t.en[zi], radi = t.rad[zi];
  const double dcni = w.dldcn[row], bi = w.beta[row], li = w.lam[row], qi = w.q[row];
  double fx = 0.0, fy = 0.0, fz = 0.0;
  for (int j = lane; j < n; j += 64) {
    const int zj = (int)Z[j];
    if (j == i || zj == 0) continue;
    const int64_t rj = m * N + j;
    const double dx = xi - R[3 * j], dy = yi - R[3 * j + 1], dz = zzi - R[3 * j + 2];
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    const double dc = erf_count_dr(r, rci + t.rcov[zj]);
    double coef = 0.0;
    if (r <= p.cutcn) coef += (dcni + w.dldcn[rj]) * en_factor(eni, t.en[zj]) * dc;
    if (eeq) {
      if (r <= p.cuteeq) coef += (bi + w.beta[rj]) * dc;
      const double radj = t.rad[zj];
      const double gm = 1.0 / sqrt(radi * radi + radj * radj);
      const double dA = (2.0 * gm / sqrt(kPi)) * exp(-gm * gm * r * r) / r -
                        erf(gm * r) / (r * r);
      coef -= (li * w.q[rj] + w.lam[rj] * qi) * dA;
    }
    coef /= r;
    fx += coef * dx;
    fy += coef * dy;
    fz += coef * dz;
  }
  fx = wave_sum(fx);
  fy = wave_sum(fy);
  fz = wave_sum(fz);
  if (lane == 0) {
    grad_out[3 * row] += fx;
    grad_out[3 * row + 1] += fy;
    grad_out[3 * row + 2] += fz;
  }
}

// --------------------------------------------------------------------------
// properties: C6 matrix with charge-dependent weights and static
// polarisabilities (disp.py:189-197), one wave per atom row
// --------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_properties(DevTables t, Inputs in, Work w,
                                                    double* c6_out,
                                                    double* alpha_out) {
  const int lane = threadIdx.x & 63;
  const int64_t row = (int64_t)blockIdx.x * kWaves + (threadIdx.x >> 6);
  if (row >= w.B * w.N) return;
  const int64_t N = w.N;
  const int64_t m = row / N;
  const int i = (int)(row - m * N);
  const int n = w.natom[m];
  const int64_t* Z = in.numbers + m * N;
  const int zi = i < n ? (int)Z[i] : 0;
  double gwi[kMaxRef];
#pragma unroll
  for (int a = 0; a < kMaxRef; ++a) gwi[a] = w.gw[row * kRefStride + a];
  if (c6_out) {
    for (int j = lane; j < N; j += 64) {
      const int zj = j < n ? (int)Z[j] : 0;
      double c6 = 0.0;
      if (zi != 0 && zj != 0) {
        const int64_t rj = m * N + j;
        const double* blk = t.rc6 + ((size_t)zi * kMaxZ + zj) * kRc6Block;
        const int nri = t.nref[zi];
        for (int a = 0; a < nri; ++a) {
          double va = 0.0;
#pragma unroll
          for (int b = 0; b < kMaxRef; ++b) va += blk[a * kMaxRef + b] * w.gw[rj * kRefStride + b];
          c6 += gwi[a] * va;
        }
      }
      c6_out[row * N + j] = c6;
    }
  }
  if (alpha_out && lane == 0) {
    double al = 0.0;
    if (zi != 0) {
      const int nri = t.nref[zi];
      for (int a = 0; a < nri; ++a) al += gwi[a] * t.alpha0[zi * kMaxRef + a];
    }
    alpha_out[row] = al;
  }
}

__global__ __launch_bounds__(256) void k_copy(const double* src, double* dst,
                                              int64_t count) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (k < count) dst[k] = src[k];
}

inline unsigned rows_grid(const Work& w) {
  return (unsigned)((w.B * w.N + kWaves - 1) / kWaves);
}

}  // namespace

// ------------------------------------------------------------------ launchers

hipError_t launch_prep(hipStream_t s, const Inputs& in, const Work& w) {
  const unsigned grid = (unsigned)((w.B + kWaves - 1) / kWaves);
  hipLaunchKernelGGL(k_prep, dim3(grid), dim3(256), 0, s, in.numbers, w);
  return hipGetLastError();
}

hipError_t launch_cn(hipStream_t s, const DevTables& t, const Params& p,
                     const Inputs& in, const Work& w) {
  hipLaunchKernelGGL(k_cn, dim3(rows_grid(w)), dim3(256), 0, s, t, p, in, w);
  return hipGetLastError();
}

// Launch a per-molecule EEQ kernel with `shmem` bytes of dynamic LDS, raising
// the kernel's LDS limit first when needed (once per kernel and size).
template <typename... Args>
hipError_t launch_lds(void (*kernel)(Args...), size_t& configured, unsigned grid,
                      size_t shmem, hipStream_t s, Args... args) {
  if (shmem > configured) {
    hipError_t e = hipFuncSetAttribute((const void*)kernel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem);
    if (e != hipSuccess) return e;
    configured = shmem;
  }
  hipLaunchKernelGGL(kernel, dim3(grid), dim3(256), shmem, s, args...);
  return hipGetLastError();
}

hipError_t launch_eeq(hipStream_t s, const DevTables& t, const Params& p,
                      const Inputs& in, const Work& w, bool /*keep_factor*/) {
  static size_t configured[3] = {0, 0, 0};
  const int mode = eeq_mode(w.N, kEeqVectors);
  const size_t shmem = eeq_shmem(mode, w.N, kEeqVectors);
  const unsigned grid = (unsigned)w.B;
  hipError_t e;
  if (mode == kEeqAllLds)
    e = launch_lds(k_eeq_chol<kEeqAllLds>, configured[2], grid, shmem, s, t, p, in, w);
  else if (mode == kEeqFactorLds)
    e = launch_lds(k_eeq_chol<kEeqFactorLds>, configured[1], grid, shmem, s, t, p, in, w);
  else
    e = launch_lds(k_eeq_chol<kEeqGlobal>, configured[0], grid, shmem, s, t, p, in, w);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_eeq_lu, dim3(grid), dim3(256), 0, s, t, p, in, w);
  return hipGetLastError();
}

hipError_t launch_weights(hipStream_t s, const DevTables& t, const Params& p,
                          const Work& w, const Inputs& in) {
  const unsigned grid = (unsigned)((w.B * w.N + 255) / 256);
  hipLaunchKernelGGL(k_weights, dim3(grid), dim3(256), 0, s, t, p, w, in);
  return hipGetLastError();
}

hipError_t launch_pairs(hipStream_t s, const DevTables& t, const Params& p,
                        const Inputs& in, const Work& w, bool grad,
                        double* grad_out) {
  if (grad)
    hipLaunchKernelGGL(k_pairs<true>, dim3(rows_grid(w)), dim3(256), 0, s, t, p, in, w,
                       grad_out);
  else
    hipLaunchKernelGGL(k_pairs<false>, dim3(rows_grid(w)), dim3(256), 0, s, t, p, in, w,
                       grad_out);
  return hipGetLastError();
}

hipError_t launch_atm(hipStream_t s, const Params& p, const Inputs& in,
                      const Work& w, bool energy, bool grad) {
  const int nb = (int)((w.N + kTile - 1) / kTile);
  const int tiles = nb * (nb + 1) / 2;
  const int64_t ntiles = w.B * tiles;
  const int64_t grid = (ntiles + 7) / 8 * 8;  // multiple of 8 for the XCD remap
  if (grid == 0) return hipSuccess;
  if (energy && grad)
    hipLaunchKernelGGL((k_atm<true, true>), dim3((unsigned)grid), dim3(256), 0, s, p,
                       in, w, tiles, ntiles);
  else if (energy)
    hipLaunchKernelGGL((k_atm<true, false>), dim3((unsigned)grid), dim3(256), 0, s, p,
                       in, w, tiles, ntiles);
  else if (grad)
    hipLaunchKernelGGL((k_atm<false, true>), dim3((unsigned)grid), dim3(256), 0, s, p,
                       in, w, tiles, ntiles);
  return hipGetLastError();
}

hipError_t launch_atm_rows(hipStream_t s, const Inputs& in, const Work& w,
                           bool energy, bool grad, double* energy_out,
                           double* grad_out) {
  if (energy && grad)
    hipLaunchKernelGGL((k_atm_rows<true, true>), dim3(rows_grid(w)), dim3(256), 0, s,
                       in, w, energy_out, grad_out);
  else if (energy)
    hipLaunchKernelGGL((k_atm_rows<true, false>), dim3(rows_grid(w)), dim3(256), 0, s,
                       in, w, energy_out, grad_out);
  else if (grad)
    hipLaunchKernelGGL((k_atm_rows<false, true>), dim3(rows_grid(w)), dim3(256), 0, s,
                       in, w, energy_out, grad_out);
  return hipGetLastError();
}

hipError_t launch_adjoint(hipStream_t s, const DevTables& t, const Params& p,
                          const Inputs& in, const Work& w) {
  static size_t configured[3] = {0, 0, 0};
  const int mode = eeq_mode(w.N, 1);
  const size_t shmem = eeq_shmem(mode, w.N, 1);
  const unsigned grid = (unsigned)w.B;
  if (mode == kEeqAllLds)
    return launch_lds(k_adjoint<kEeqAllLds>, configured[2], grid, shmem, s, t, p, in, w);
  if (mode == kEeqFactorLds)
    return launch_lds(k_adjoint<kEeqFactorLds>, configured[1], grid, shmem, s, t, p, in, w);
  return launch_lds(k_adjoint<kEeqGlobal>, configured[0], grid, shmem, s, t, p, in, w);
}

hipError_t launch_chain(hipStream_t s, const DevTables& t, const Params& p,
                        const Inputs& in, const Work& w, double* grad_out) {
  hipLaunchKernelGGL(k_chain, dim3(rows_grid(w)), dim3(256), 0, s, t, p, in, w,
                     grad_out);
  return hipGetLastError();
}

hipError_t launch_properties(hipStream_t s, const DevTables& t, const Inputs& in,
                             const Work& w, double* c6_out, double* alpha_out) {
  hipLaunchKernelGGL(k_properties, dim3(rows_grid(w)), dim3(256), 0, s, t, in, w,
                     c6_out, alpha_out);
  return hipGetLastError();
}

hipError_t launch_copy(hipStream_t s, const double* src, double* dst, int64_t count) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_copy, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, src,
                     dst, count);
  return hipGetLastError();
}

}  // namespace d4
