// HIP kernels (gfx950) for the EEQ charges and their adjoint.
//
// Molecules whose packed Coulomb factor fits in LDS are solved by one
// work-group each; larger molecules use the tiled multi-work-group path.
// Every output element has one writer and every reduction a fixed order.

#include <map>
#include <mutex>
#include <utility>

#include "d4_device.h"

namespace d4 {
namespace {

// --------------------------------------------------------------------------
// EEQ charges: one work-group per molecule.
//
// The bordered system [[C, e], [e^T, 0]] [q; mu] = [b; Q] is solved with a
// Cholesky factorisation of the Coulomb block C (in LDS when it fits) and the
// Schur complement of the border: q = y - mu z with y = C^-1 b, z = C^-1 e,
// mu = (e.y - Q) / (e.z). A non-positive pivot flags the molecule for the
// pivoted-LU kernel below.
// --------------------------------------------------------------------------

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
    // 1 / sqrt(piv): hardware estimate and two Newton steps (the diagonal
    // of the factor itself is never needed, only its reciprocal)
    const double pv = piv > 0.0 ? piv : 1.0;
    double inv = __builtin_amdgcn_rsq(pv);
    inv = inv * fma(-0.5 * pv * inv, inv, 1.5);
    inv = inv * fma(-0.5 * pv * inv, inv, 1.5);
    const double s = pv * inv;
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
template <int T, bool RHS, typename P, typename V>
__device__ __forceinline__ void trailing_update(P L, int kb, int n, V rb, V re,
                                                const double* yblk) {
  const int r0 = kb + kNB;
  const int mr = n - r0;
  const int mt = (mr + T - 1) / T;
  const int ntl = mt * (mt + 1) / 2;
  // right-hand-side rows first, on the threads that the tiles leave idle last
  const int first = (256 - (ntl + mr) % 256) % 256;
  for (int e = threadIdx.x - first; RHS && e < mr; e += 256) {
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

// Blocked right-looking Cholesky of the assembled n x n lower triangle; the
// diagonal blocks end up holding their inverses. With RHS, two right-hand
// sides are carried along (forward substitution fused): on return rb, re
// hold L^-1 b and L^-1 e. yblk: 2 * kNB doubles of LDS, flag: one double.
// Returns non-zero if a pivot was not positive.
template <bool RHS, typename P, typename V>
__device__ __forceinline__ int chol_factor(P L, int n, V rb, V re, double* yblk,
                                           double* flag) {
  const int tid = threadIdx.x;
  constexpr int nrhs = RHS ? 2 : 0;
  for (int kb = 0; kb < n; kb += kNB) {
    const int nb = min(kNB, n - kb);
    if (tid < 64) {
      const int bad = chol_diag_block(L, kb, nb);
      if (tid == 0) flag[0] = bad ? 1.0 : 0.0;
    }
    __syncthreads();
    if (flag[0] != 0.0) return 1;  // uniform
    const int mr = n - kb - nb;    // > 0 only for full blocks
    for (int t2 = tid; t2 < mr + nrhs; t2 += 256) {
      if (t2 < mr)
        panel_row(L, kb, kb + nb + t2);
      else if (RHS)
        block_lower(L, kb, nb, t2 == mr ? rb : re, yblk + (t2 - mr) * kNB);
    }
    __syncthreads();
    if (mr > 0) {
      if (mr >= 48)
        trailing_update<4, RHS>(L, kb, n, rb, re, yblk);
      else
        trailing_update<2, RHS>(L, kb, n, rb, re, yblk);
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

__device__ void eeq_lu_molecule(const DevTables& t, const Inputs& in, const Work& w, int64_t m,
                                double* s_val, int* s_idx);

// Charges of molecule m by one work-group (smem: the dynamic LDS of the
// launch; s_small: two solution blocks, sums and a flag).
template <int MODE>
__device__ __forceinline__ void eeq_chol_molecule(const DevTables& t, const Params& p,
                                                  const Inputs& in, const Work& w, int64_t m,
                                                  double* smem, double* s_small) {
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
    // four independent elements per thread and iteration (the erf chains
    // overlap)
    for (int e0 = tid; e0 < ne; e0 += 4 * 256) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int e = e0 + 256 * u;
        if (e < ne) {
          const int i = strict_row(e);
          const int j = e - i * (i - 1) / 2;
          const double ri = srad[i], rj = srad[j];
          double v = 0.0;
          if (ri != 0.0 && rj != 0.0) {
            const double dx = sx[i] - sx[j], dy = sy[i] - sy[j], dz = sz[i] - sz[j];
            v = eeq_offdiag_r2(rsqrt_nt(ri * ri + rj * rj), dx * dx + dy * dy + dz * dz);
          }
          L[tri(i) + j] = v;
        }
      }
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
          v = eeq_offdiag(t, zi, zj, dx * dx + dy * dy + dz * dz);
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

  if (p.force_lu || chol_factor<true>(L, n, rb, re, yblk, flag)) {
    // a pivot <= 0 (or forced): the pivoted LU of the bordered system in the
    // workspace, by this work-group (no separate launch behind this one)
    if (tid == 0) w.status[m] |= kStatusLU;
    eeq_lu_molecule(t, in, w, m, red, reinterpret_cast<int*>(red + kWaves));
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

// Two work-groups per CU fit the LDS of a 120-atom factor; keep the register
// use low enough for two waves per SIMD as well. With `paired` (factors that
// take a whole CU's LDS), work-group b solves the b-th largest and then the
// b-th smallest molecule: half the CUs hold the solve, and the others are
// free from the start for the three-body kernel on the companion stream.
template <int MODE, bool PAIRED>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_eeq_chol(
    DevTables t, Params p, Inputs in, Work w) {
  extern __shared__ double smem[];
  __shared__ double s_small[2 * kNB + 2 * kWaves + 2];
  const int64_t b = blockIdx.x;
  if (!PAIRED) {
    // smallest molecules first: their solves end early and leave their CUs
    // to the three-body side on the companion stream (the batch order and
    // largest first were both slower beside it)
    eeq_chol_molecule<MODE>(t, p, in, w, w.eqord[w.B - 1 - b], smem, s_small);
    return;
  }
  const int64_t b2 = w.B - 1 - b;
  for (int k = 0; k < (b2 > b ? 2 : 1); ++k) {
    if (k) __syncthreads();  // the LDS is reused
    eeq_chol_molecule<MODE>(t, p, in, w, w.eqord[k ? b2 : b], smem, s_small);
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
        v = eeq_offdiag(t, zi, zj, dx * dx + dy * dy + dz * dz);
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

// Largest |value| of the work-group and its index, the lowest index among
// equal values (the same pivot whatever the reduction order); s_val, s_idx:
// kWaves entries of LDS.
__device__ __forceinline__ void block_argmax(double& v, int& i, double* s_val, int* s_idx) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    const double ov = __shfl_xor(v, off, 64);
    const int oi = __shfl_xor(i, off, 64);
    if (ov > v || (ov == v && oi < i)) {
      v = ov;
      i = oi;
    }
  }
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    s_val[wave] = v;
    s_idx[wave] = i;
  }
  __syncthreads();
  v = s_val[0];
  i = s_idx[0];
  for (int k = 1; k < kWaves; ++k) {
    if (s_val[k] > v || (s_val[k] == v && s_idx[k] < i)) {
      v = s_val[k];
      i = s_idx[k];
    }
  }
  __syncthreads();
}

// Pivoted LU of the bordered system of molecule m (work-group wide, 256
// threads); s_val, s_idx: kWaves entries of LDS.
__device__ void eeq_lu_molecule(const DevTables& t, const Inputs& in, const Work& w, int64_t m,
                                double* s_val, int* s_idx) {
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
    block_argmax(best, bi, s_val, s_idx);
    const int pk = bi;
    const double pv = best;
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

// The molecules whose Cholesky pivot was <= 0 (or all, when forced): a few
// work-groups walk the batch, so the launch costs little when none needs it.
__global__ __launch_bounds__(256) void k_eeq_lu(DevTables t, Params p, Inputs in, Work w) {
  __shared__ double s_val[kWaves];
  __shared__ int s_idx[kWaves];
  if (in.q_in != nullptr) return;
  for (int64_t m = blockIdx.x; m < w.B; m += gridDim.x) {
    if (!(w.status[m] & kStatusLU)) continue;  // uniform over the work-group
    eeq_lu_molecule(t, in, w, m, s_val, s_idx);
    __syncthreads();
  }
}

// ==========================================================================
// Tiled EEQ for molecules whose packed factor does not fit in LDS.
//
// The matrix lives in HBM in the LU workspace as row-major rows of length N:
// rows 0..n-1 hold the lower triangle of C, row n holds b and row n+1 holds
// e. With the two border rows below the matrix, the forward substitution of
// b and e is just the panel and trailing steps applied to those rows. Per
// 64-column block step:
//   k_tf_diag   one work-group factors the 64 x 64 diagonal block in LDS
//               (16-column blocked Cholesky), forms X = D^-1 and stores it in
//               place of D;
//   k_tf_panel  64-row tiles below the block (and the border rows) become
//               A X^T;
//   k_tf_update 64 x 64 tiles of the trailing triangle lose P_I P_J^T.
// The substitutions then run one launch per block step: every work-group of
// a molecule recomputes the block's 64-entry solution from the same inputs
// and work-group 0 alone stores it, so no work-group reads what another
// writes within a launch.
// ==========================================================================

constexpr int kTB = 64;
constexpr int kEeqTiled = 3;  // adjoint mode: lambda already solved by the tiled kernels

// The tiled solve is a chain of a few hundred short launches that runs beside
// the screened three-body kernel's waves on the same SIMDs; its waves take
// issue priority over them, so the chain is not slowed to a third of its
// speed (the three-body waves use the cycles it leaves).
__device__ __forceinline__ void tf_priority() { __builtin_amdgcn_s_setprio(2); }

__device__ __forceinline__ double* tf_matrix(const Work& w, int64_t m) {
  return w.lu + (size_t)m * (size_t)(w.N + 1) * (size_t)(w.N + 1);
}

// Row r of an inclusive lower packed index: tri(r) <= e < tri(r + 1).
__device__ __forceinline__ int incl_row(int e) { return strict_row(e) - 1; }

// Molecule m still factorising at block kb (charges not given, no failure).
__device__ __forceinline__ bool tf_active(const Inputs& in, const Work& w, int64_t m,
                                          int kb) {
  return in.q_in == nullptr && kb < w.natom[m] && !(w.status[m] & kStatusLU);
}

// Assemble the lower triangle of C and the border rows b, e; one wave per row.
__global__ __launch_bounds__(256) void k_tf_assemble(DevTables t, Params p, Inputs in,
                                                     Work w) {
  const int lane = threadIdx.x & 63;
  const int64_t N = w.N, rows = N + 2;
  const int64_t g = (int64_t)blockIdx.x * kWaves + (threadIdx.x >> 6);
  if (g >= w.B * rows) return;
  const int64_t m = g / rows;
  const int i = (int)(g - m * rows);
  const int n = w.natom[m];
  if (in.q_in != nullptr || i >= n + 2) return;
  if (p.force_lu) {  // testing: every system goes to the pivoted-LU kernel
    if (i == 0 && lane == 0) w.status[m] |= kStatusLU;
    return;
  }
  const int64_t* Z = in.numbers + m * N;
  const double* R = in.positions + m * N * 3;
  double* Ai = tf_matrix(w, m) + (size_t)i * (size_t)N;
  if (i < n) {
    const int zi = (int)Z[i];
    const double xi = R[3 * i], yi = R[3 * i + 1], zzi = R[3 * i + 2];
    for (int j = lane; j <= i; j += 64) {
      const int zj = (int)Z[j];
      double v;
      if (j == i) {
        v = zi != 0 ? eeq_diag(t, zi) : 1.0;
      } else if (zi != 0 && zj != 0) {
        const double dx = xi - R[3 * j], dy = yi - R[3 * j + 1], dz = zzi - R[3 * j + 2];
        v = eeq_offdiag(t, zi, zj, dx * dx + dy * dy + dz * dz);
      } else {
        v = 0.0;
      }
      Ai[j] = v;
    }
  } else {
    const bool rhs = i == n;
    for (int j = lane; j < n; j += 64) {
      const int zj = (int)Z[j];
      Ai[j] = zj == 0 ? 0.0 : (rhs ? eeq_rhs(t, zj, w.cnq[m * N + j]) : 1.0);
    }
  }
}

// Factor the diagonal block at kb and replace it by the inverse of its factor.
// The tiled kernels take their LDS as dynamic shared memory and at most 64
// VGPRs (eight waves per SIMD as far as registers go): a work-group then fits
// beside the three-body kernel's waves (two per SIMD with 220 VGPRs) on any
// CU instead of waiting until one is nearly empty.
constexpr size_t kTfDiagLds = sizeof(double) * (kTB * (kTB + 1) / 2 + kTB * kTB + kNB * kTB +
                                                2 * kNB + 2);
constexpr size_t kTfTileLds = sizeof(double) * 2 * kTB * (kTB + 1);

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void k_tf_diag(
    Inputs in, Work w, int kb) {
  tf_priority();
  extern __shared__ double s_tf[];
  double* F = s_tf;                                             // [tri(kTB)]
  auto X = reinterpret_cast<double (*)[kTB]>(F + kTB * (kTB + 1) / 2);  // [kTB][kTB]
  auto T = reinterpret_cast<double (*)[kTB]>(&X[kTB][0]);      // [kNB][kTB]
  double* s_small = &T[kNB][0];                                 // [2 kNB + 2]
  const int64_t m = blockIdx.x;
  if (!tf_active(in, w, m, kb)) return;
  const int tid = threadIdx.x;
  const int64_t N = w.N;
  const int n = w.natom[m];
  const int nb = min(kTB, n - kb);
  double* A = tf_matrix(w, m);
  for (int e = tid; e < tri(nb); e += 256) {
    const int r = incl_row(e), c = e - tri(r);
    F[e] = A[(size_t)(kb + r) * N + kb + c];
  }
  __syncthreads();
  if (chol_factor<false>(F, nb, (double*)nullptr, (double*)nullptr, s_small,
                         s_small + 2 * kNB)) {
    if (tid == 0) w.status[m] |= kStatusLU;
    return;
  }
  // X = F^-1 by block forward substitution over the 16-row sub-blocks, whose
  // own inverses the factorisation left on the diagonal of F:
  // X_s = X16_s (I_s - F_{s,<s} X_{<s})
  for (int r0 = 0; r0 < nb; r0 += kNB) {
    const int rn = min(kNB, nb - r0);
    const int ncol = r0 + rn;
    for (int e = tid; e < rn * ncol; e += 256) {
      const int r = e / ncol, c = e - r * ncol;
      const double* Fr = F + tri(r0 + r);
      double sacc = (r0 + r == c) ? 1.0 : 0.0;
      for (int k = c; k < r0; ++k) sacc = fma(-Fr[k], X[k][c], sacc);
      T[r][c] = sacc;
    }
    __syncthreads();
    for (int e = tid; e < rn * ncol; e += 256) {
      const int r = e / ncol, c = e - r * ncol;
      const double* Xr = F + tri(r0 + r) + r0;
      double sacc = 0.0;
      for (int rr = 0; rr <= r; ++rr) sacc = fma(Xr[rr], T[rr][c], sacc);
      X[r0 + r][c] = c <= r0 + r ? sacc : 0.0;
    }
    __syncthreads();
  }
  for (int e = tid; e < tri(nb); e += 256) {
    const int r = incl_row(e), c = e - tri(r);
    A[(size_t)(kb + r) * N + kb + c] = X[r][c];
  }
}

// Panel tile: rows kb + nb + 64 y ... of the matrix and border become A X^T.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void k_tf_panel(
    Inputs in, Work w, int kb) {
  tf_priority();
  extern __shared__ double s_tf[];
  auto Xs = reinterpret_cast<double (*)[kTB + 1]>(s_tf);  // [kTB][kTB + 1]
  auto As = Xs + kTB;                                      // [kTB][kTB + 1]
  const int64_t m = blockIdx.x;
  if (!tf_active(in, w, m, kb)) return;
  const int tid = threadIdx.x;
  const int64_t N = w.N;
  const int n = w.natom[m];
  const int nb = min(kTB, n - kb);
  const int r0 = kb + nb + (int)blockIdx.y * kTB;
  const int rows = min(kTB, n + 2 - r0);
  if (rows <= 0) return;
  double* A = tf_matrix(w, m);
  for (int e = tid; e < kTB * kTB; e += 256) {
    const int r = e >> 6, c = e & (kTB - 1);
    Xs[r][c] = (r < nb && c <= r) ? A[(size_t)(kb + r) * N + kb + c] : 0.0;
    As[r][c] = (r < rows && c < nb) ? A[(size_t)(r0 + r) * N + kb + c] : 0.0;
  }
  __syncthreads();
  const int tr = tid >> 4, tc = tid & 15;
  double acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
  for (int k = 0; k < kTB; ++k) {
    double av[4], xv[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      av[a] = As[4 * tr + a][k];
      xv[a] = Xs[4 * tc + a][k];
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] = fma(av[a], xv[b], acc[a][b]);
  }
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int r = 4 * tr + a, c = 4 * tc + b;
      if (r < rows && c < nb) A[(size_t)(r0 + r) * N + kb + c] = acc[a][b];
    }
}

// Trailing tile (I, J): rows r0 + 64 I.. (matrix and border), columns
// r0 + 64 J.. of the matrix, r0 = kb + 64, lose P_I P_J^T.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void k_tf_update(
    Inputs in, Work w, int kb) {
  tf_priority();
  extern __shared__ double s_tf[];
  auto Pi = reinterpret_cast<double (*)[kTB + 1]>(s_tf);  // [kTB][kTB + 1]
  auto Pj = Pi + kTB;                                      // [kTB][kTB + 1]
  const int64_t m = blockIdx.x;
  if (!tf_active(in, w, m, kb)) return;
  const int tid = threadIdx.x;
  const int64_t N = w.N;
  const int n = w.natom[m];
  const int r0 = kb + kTB;
  if (r0 >= n) return;
  const int nI = (n + 2 - r0 + kTB - 1) / kTB, nJ = (n - r0 + kTB - 1) / kTB;
  const int e = (int)blockIdx.y;
  if (e >= nI * (nI + 1) / 2) return;
  const int I = incl_row(e), J = e - tri(I);
  if (J >= nJ) return;
  const int i0 = r0 + I * kTB, j0 = r0 + J * kTB;
  double* A = tf_matrix(w, m);
  for (int x = tid; x < kTB * kTB; x += 256) {
    const int r = x >> 6, c = x & (kTB - 1);
    Pi[r][c] = (i0 + r < n + 2) ? A[(size_t)(i0 + r) * N + kb + c] : 0.0;
    Pj[r][c] = (j0 + r < n) ? A[(size_t)(j0 + r) * N + kb + c] : 0.0;
  }
  __syncthreads();
  const int tr = tid >> 4, tc = tid & 15;
  double acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
  for (int k = 0; k < kTB; ++k) {
    double iv[4], jv[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      iv[a] = Pi[4 * tr + a][k];
      jv[a] = Pj[4 * tc + a][k];
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] = fma(iv[a], jv[b], acc[a][b]);
  }
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int i = i0 + 4 * tr + a, j = j0 + 4 * tc + b;
      if (i < n + 2 && j < n && j <= i) A[(size_t)i * N + j] -= acc[a][b];
    }
}

// After the factorisation: charges given or padding, then mu and the
// right-hand side y_b - mu y_e of the backward substitution (in vec); y_e is
// kept for the adjoint.
__global__ __launch_bounds__(256) void k_tf_reduce(Inputs in, Work w) {
  tf_priority();
  __shared__ double red[2 * kWaves];
  const int64_t m = blockIdx.x;
  const int tid = threadIdx.x;
  const int64_t N = w.N;
  const int n = w.natom[m];
  const int64_t* Z = in.numbers + m * N;
  double* q = w.q + m * N;
  if (in.q_in != nullptr) {
    for (int i = tid; i < N; i += 256)
      q[i] = (i < n && Z[i] != 0) ? in.q_in[m * N + i] : 0.0;
    return;
  }
  for (int i = n + tid; i < N; i += 256) q[i] = 0.0;
  if (n == 0 || (w.status[m] & kStatusLU)) return;
  const double* A = tf_matrix(w, m);
  const double* yb = A + (size_t)n * N;
  const double* ye = A + (size_t)(n + 1) * N;
  double sbe = 0.0, see = 0.0;
  for (int i = tid; i < n; i += 256) {
    sbe = fma(ye[i], yb[i], sbe);
    see = fma(ye[i], ye[i], see);
  }
  sbe = block_sum(sbe, red);
  see = block_sum(see, red + kWaves);
  const double mu = (sbe - in.charge[m]) / see;
  double* v = w.vec + m * (N + 1);
  double* gz = w.ez + m * N;
  for (int i = tid; i < n; i += 256) {
    v[i] = fma(-mu, ye[i], yb[i]);
    gz[i] = ye[i];
  }
}

// One backward-substitution step L^T x = v at block kb (descending kb):
// x_kb = X^T v_kb is stored to dst (zero on padding atoms) by work-group 0;
// work-group y updates v[j] for j in [256 y, 256 y + 256) below kb.
__global__ __launch_bounds__(256) void k_tf_bsolve(Inputs in, Work w, int kb,
                                                   double* dst_base) {
  tf_priority();
  __shared__ double xs[kTB];
  const int64_t m = blockIdx.x;
  const int y = (int)blockIdx.y;
  const int tid = threadIdx.x;
  const int64_t N = w.N;
  const int n = w.natom[m];
  if (in.q_in != nullptr || kb >= n || (w.status[m] & kStatusLU)) return;
  const int j = y * 256 + tid;
  if (y > 0 && y * 256 >= kb) return;
  const int nb = min(kTB, n - kb);
  const double* A = tf_matrix(w, m);
  double* v = w.vec + m * (N + 1);
  if (tid < nb) {
    double sacc = 0.0;
    for (int k = tid; k < nb; ++k) sacc = fma(A[(size_t)(kb + k) * N + kb + tid], v[kb + k], sacc);
    xs[tid] = sacc;
  }
  __syncthreads();
  if (y == 0 && tid < nb) {
    const int64_t* Z = in.numbers + m * N;
    dst_base[m * N + kb + tid] = Z[kb + tid] != 0 ? xs[tid] : 0.0;
  }
  if (j < kb) {
    double sacc = 0.0;
    for (int c = 0; c < nb; ++c) sacc = fma(A[(size_t)(kb + c) * N + j], xs[c], sacc);
    v[j] -= sacc;
  }
}

// One forward-substitution step L y = v at block kb (ascending kb): y_kb =
// X v_kb is stored to dst by work-group 0; work-group y updates the rows
// kb + nb + 256 y .. below the block.
__global__ __launch_bounds__(256) void k_tf_fsolve(Inputs in, Work w, int kb,
                                                   double* dst_base) {
  tf_priority();
  __shared__ double ys[kTB];
  const int64_t m = blockIdx.x;
  const int y = (int)blockIdx.y;
  const int tid = threadIdx.x;
  const int64_t N = w.N;
  const int n = w.natom[m];
  if (in.q_in != nullptr || kb >= n || (w.status[m] & (kStatusLU | kStatusSingular))) return;
  const int nb = min(kTB, n - kb);
  const int j = kb + nb + y * 256 + tid;
  if (y > 0 && kb + nb + y * 256 >= n) return;
  const double* A = tf_matrix(w, m);
  double* v = w.vec + m * (N + 1);
  if (tid < nb) {
    const double* Xr = A + (size_t)(kb + tid) * N + kb;
    double sacc = 0.0;
    for (int k = 0; k <= tid; ++k) sacc = fma(Xr[k], v[kb + k], sacc);
    ys[tid] = sacc;
  }
  __syncthreads();
  if (y == 0 && tid < nb) dst_base[m * N + kb + tid] = ys[tid];
  if (j < n) {
    const double* Aj = A + (size_t)j * N + kb;
    double sacc = 0.0;
    for (int c = 0; c < nb; ++c) sacc = fma(Aj[c], ys[c], sacc);
    v[j] -= sacc;
  }
}

// Adjoint set-up for the tiled solve: v = dL/dq on real atoms.
__global__ __launch_bounds__(256) void k_tf_adj_init(Inputs in, Work w) {
  tf_priority();
  const int64_t m = blockIdx.x;
  const int64_t N = w.N;
  const int n = w.natom[m];
  if (in.q_in != nullptr || (w.status[m] & (kStatusLU | kStatusSingular))) return;
  const int64_t* Z = in.numbers + m * N;
  double* v = w.vec + m * (N + 1);
  for (int i = threadIdx.x; i < n; i += 256) v[i] = Z[i] != 0 ? w.dldq[m * N + i] : 0.0;
}

// Between the two adjoint substitutions: w' = L^-1 g is in lam; v = w' - nu y_e
// with nu = y_e . w' / y_e . y_e.
__global__ __launch_bounds__(256) void k_tf_adj_mid(Inputs in, Work w) {
  tf_priority();
  __shared__ double red[2 * kWaves];
  const int64_t m = blockIdx.x;
  const int64_t N = w.N;
  const int n = w.natom[m];
  if (in.q_in != nullptr || n == 0 || (w.status[m] & (kStatusLU | kStatusSingular))) return;
  const int tid = threadIdx.x;
  const double* wl = w.lam + m * N;
  const double* ye = w.ez + m * N;
  double sw = 0.0, sz = 0.0;
  for (int i = tid; i < n; i += 256) {
    sw = fma(ye[i], wl[i], sw);
    sz = fma(ye[i], ye[i], sz);
  }
  sw = block_sum(sw, red);
  sz = block_sum(sz, red + kWaves);
  const double nu = sw / sz;
  double* v = w.vec + m * (N + 1);
  for (int i = tid; i < n; i += 256) v[i] = fma(-nu, ye[i], wl[i]);
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
  const int64_t m = w.eqord[blockIdx.x];  // largest molecules first
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
  } else if (MODE != kEeqTiled) {
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

// Order of the molecules for the one-work-group-per-molecule adjoint kernel:
// by atom count, largest first (counting sort; atom counts above kOrdMax
// share the first bin), so that the longest solves start first and the short
// ones fill in behind them (the adjoint of the 1024 x 20-120 batch alone 67 ->
// 48 us). Which work-group solves which molecule does not change any result;
// the order within a bin is arbitrary. (The charge kernel keeps the batch
// order: taking the largest first there made it faster alone, 217 -> 153 us,
// but the two-stream step slower, 1.10 -> 1.13 ms, as it then crowds the
// three-body pair pass that runs beside it.)
constexpr int kOrdMax = 256;

__global__ __launch_bounds__(1024) void k_eeq_order(Work w) {
  __shared__ int hist[kOrdMax + 1];
  __shared__ int start[kOrdMax + 1];
  const int tid = threadIdx.x;
  for (int c = tid; c <= kOrdMax; c += 1024) hist[c] = 0;
  __syncthreads();
  for (int64_t m = tid; m < w.B; m += 1024) atomicAdd(&hist[min(w.natom[m], kOrdMax)], 1);
  __syncthreads();
  if (tid == 0) {
    int acc = 0;
    for (int c = kOrdMax; c >= 0; --c) {
      start[c] = acc;
      acc += hist[c];
    }
  }
  __syncthreads();
  for (int64_t m = tid; m < w.B; m += 1024) {
    const int r = atomicAdd(&start[min(w.natom[m], kOrdMax)], 1);
    w.eqord[r] = (int32_t)m;
  }
}

}  // namespace

hipError_t ensure_lds(const void* kernel, size_t bytes) {
  // the limit is a property of the kernel on one device: remember what each
  // (kernel, device) was given
  static std::mutex mu;
  static std::map<std::pair<const void*, int>, size_t> given;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> lock(mu);
  size_t& have = given[{kernel, dev}];
  if (bytes <= have) return hipSuccess;
  e = hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (e == hipSuccess) have = bytes;
  return e;
}

// Launch a per-molecule EEQ kernel with `shmem` bytes of dynamic LDS, raising
// the kernel's LDS limit first when needed.
template <typename... Args>
hipError_t launch_lds(void (*kernel)(Args...), unsigned grid, size_t shmem, hipStream_t s,
                      Args... args) {
  hipError_t e = ensure_lds((const void*)kernel, shmem);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(kernel, dim3(grid), dim3(256), shmem, s, args...);
  return hipGetLastError();
}

// Molecules of this padded length take the tiled path (or every size, for
// testing).
inline bool eeq_tiled(const Params& p, int64_t N) {
  return p.force_tiled || eeq_mode(N, 0) == kEeqGlobal;
}

hipError_t launch_eeq_tiled(hipStream_t s, const DevTables& t, const Params& p,
                            const Inputs& in, const Work& w) {
  const int N = (int)w.N;
  const unsigned B = (unsigned)w.B;
  const unsigned rows = (unsigned)((w.B * (w.N + 2) + kWaves - 1) / kWaves);
  hipLaunchKernelGGL(k_tf_assemble, dim3(rows), dim3(256), 0, s, t, p, in, w);
  hipError_t e = ensure_lds((const void*)k_tf_diag, kTfDiagLds);
  if (e == hipSuccess) e = ensure_lds((const void*)k_tf_panel, kTfTileLds);
  if (e == hipSuccess) e = ensure_lds((const void*)k_tf_update, kTfTileLds);
  if (e != hipSuccess) return e;
  for (int kb = 0; kb < N; kb += kTB) {
    hipLaunchKernelGGL(k_tf_diag, dim3(B), dim3(256), kTfDiagLds, s, in, w, kb);
    const unsigned ptiles = (unsigned)((N - kb + 1 + kTB - 1) / kTB);
    hipLaunchKernelGGL(k_tf_panel, dim3(B, ptiles), dim3(256), kTfTileLds, s, in, w, kb);
    if (kb + kTB < N) {
      const int nI = (N + 2 - kb - kTB + kTB - 1) / kTB;
      hipLaunchKernelGGL(k_tf_update, dim3(B, (unsigned)(nI * (nI + 1) / 2)), dim3(256),
                         kTfTileLds, s, in, w, kb);
    }
  }
  hipLaunchKernelGGL(k_tf_reduce, dim3(B), dim3(256), 0, s, in, w);
  for (int kb = ((N - 1) / kTB) * kTB; kb >= 0; kb -= kTB) {
    const unsigned chunks = (unsigned)max(1, (kb + 255) / 256);
    hipLaunchKernelGGL(k_tf_bsolve, dim3(B, chunks), dim3(256), 0, s, in, w, kb, w.q);
  }
  return hipGetLastError();
}

constexpr unsigned kLuGroups = 64;  // work-groups of the LU fallback launch

hipError_t launch_eeq(hipStream_t s, const DevTables& t, const Params& p,
                      const Inputs& in, const Work& w, bool paired) {
  const unsigned grid = (unsigned)w.B;
  if (grid == 0) return hipSuccess;
  // molecule order of the adjoint (the backward reads the order the forward
  // left in the workspace)
  hipLaunchKernelGGL(k_eeq_order, dim3(1), dim3(1024), 0, s, w);
  hipError_t e;
  const bool tiled = eeq_tiled(p, w.N);
  if (tiled) {
    e = launch_eeq_tiled(s, t, p, in, w);
  } else {
    const int mode = eeq_mode(w.N, kEeqVectors);
    const size_t shmem = eeq_shmem(mode, w.N, kEeqVectors);
    if (mode == kEeqAllLds)
      e = launch_lds(k_eeq_chol<kEeqAllLds, false>, grid, shmem, s, t, p, in, w);
    else if (paired)  // one work-group per CU: two molecules each (sizes from k_eeq_order)
      e = launch_lds(k_eeq_chol<kEeqFactorLds, true>, (grid + 1) / 2, shmem, s, t, p, in, w);
    else
      e = launch_lds(k_eeq_chol<kEeqFactorLds, false>, grid, shmem, s, t, p, in, w);
  }
  if (e != hipSuccess) return e;
  if (tiled)  // (the LDS kernels solve their failed systems themselves)
    hipLaunchKernelGGL(k_eeq_lu, dim3(grid < kLuGroups ? grid : kLuGroups), dim3(256), 0, s, t,
                       p, in, w);
  return hipGetLastError();
}

hipError_t launch_adjoint(hipStream_t s, const DevTables& t, const Params& p,
                          const Inputs& in, const Work& w) {
  const unsigned grid = (unsigned)w.B;
  if (eeq_tiled(p, w.N)) {
    const int N = (int)w.N;
    hipLaunchKernelGGL(k_tf_adj_init, dim3(grid), dim3(256), 0, s, in, w);
    for (int kb = 0; kb < N; kb += kTB) {
      const unsigned chunks = (unsigned)max(1, (N - kb + 255) / 256);
      hipLaunchKernelGGL(k_tf_fsolve, dim3(grid, chunks), dim3(256), 0, s, in, w, kb, w.lam);
    }
    hipLaunchKernelGGL(k_tf_adj_mid, dim3(grid), dim3(256), 0, s, in, w);
    for (int kb = ((N - 1) / kTB) * kTB; kb >= 0; kb -= kTB) {
      const unsigned chunks = (unsigned)max(1, (kb + 255) / 256);
      hipLaunchKernelGGL(k_tf_bsolve, dim3(grid, chunks), dim3(256), 0, s, in, w, kb, w.lam);
    }
    hipLaunchKernelGGL(k_adjoint<kEeqTiled>, dim3(grid), dim3(256), 0, s, t, p, in, w);
    return hipGetLastError();
  }
  const int mode = eeq_mode(w.N, 1);
  const size_t shmem = eeq_shmem(mode, w.N, 1);
  if (mode == kEeqAllLds)
    return launch_lds(k_adjoint<kEeqAllLds>, grid, shmem, s, t, p, in, w);
  return launch_lds(k_adjoint<kEeqFactorLds>, grid, shmem, s, t, p, in, w);
}

}  // namespace d4
