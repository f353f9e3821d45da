// HIP kernels (gfx950) for the DFT-D4 energy and gradient path.
//
// FP64 throughout. Wave64 layout: "row" kernels give 16, 32 or 64 lanes to
// one atom row i of one molecule (lanes stride over partner atoms j and
// reduce with shuffles); the three-body kernels give a work-group to tiles
// of atom pairs of one molecule (16 x 16 per work-group, or 8 x 8 per wave)
// and stream the third atom through LDS.
// The EEQ charge solve and its adjoint live in d4_eeq.hip.
//
// All reductions have a fixed order that depends only on the molecule (not
// on its neighbours in the batch), and every output element has exactly one
// writer, so results are bitwise reproducible and independent of batching.

#include <type_traits>

#include "d4_device.h"

#define D4_PRAGMA(x) _Pragma(#x)
#define D4_UNROLL(n) D4_PRAGMA(unroll n)

namespace d4 {
namespace {

// Row kernels give a group of RL lanes to one atom row (RL = 16, 32 or 64,
// chosen from the padded length): short rows keep the lanes busy with a
// quarter wave per row, long rows take a full wave.
template <int RL>
__device__ __forceinline__ double row_sum(double v) {
#pragma unroll
  for (int off = RL / 2; off >= 1; off >>= 1) v += __shfl_xor(v, off, RL);
  return v;
}

inline int row_lanes(int64_t N) { return N <= 160 ? 16 : (N <= 384 ? 32 : 64); }

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

template <int RL>
__global__ __launch_bounds__(256) void k_cn(DevTables t, Params p, Inputs in,
                                            Work w) {
  const int lane = threadIdx.x & (RL - 1);
  const int64_t row = (int64_t)blockIdx.x * (256 / RL) + threadIdx.x / RL;
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
  double s4 = 0.0, se = 0.0;
  for (int j = lane; j < n; j += RL) {
    const int zj = (int)Z[j];
    if (j == i || zj == 0) continue;
    const double dx = xi - R[3 * j], dy = yi - R[3 * j + 1], dz = zzi - R[3 * j + 2];
    const double r2 = dx * dx + dy * dy + dz * dz;
    const double r = r2 * rsqrt_nt(r2);
    const PairConst pc = pair_const(t, zi, zj);
    const double c = erf_count(r, pc.irc);
    if (r <= p.cutcn) s4 += pc.enf * c;
    if (r <= p.cuteeq) se += c;
  }
  s4 = row_sum<RL>(s4);
  se = row_sum<RL>(se);
  if (lane == 0) {
    w.cn[row] = s4;
    // smooth cap: log(1 + e^cmax) - log(1 + e^(cmax - cn))
    const double ex = exp(kCnMax - se);
    w.cnq[row] = log(1.0 + exp(kCnMax)) - log(1.0 + ex);
    w.dcnq[row] = ex / (1.0 + ex);
  }
}

// --------------------------------------------------------------------------
// Gaussian reference weights (model/d4.py:137-228) and their derivatives,
// eight lanes per atom: lane a owns reference slot a (slot 7 only pads)
// --------------------------------------------------------------------------

// A double moved between lanes by one DPP control (both halves).
template <int CTRL>
__device__ __forceinline__ double dpp_double(double v) {
  // every pattern used here is a full permutation of its lane group, so no
  // lane keeps an old value (bound_ctrl; no copy of the source first)
  const int lo = __double2loint(v), hi = __double2hiint(v);
  return __hiloint2double(__builtin_amdgcn_mov_dpp(hi, CTRL, 0xF, 0xF, true),
                          __builtin_amdgcn_mov_dpp(lo, CTRL, 0xF, 0xF, true));
}

constexpr int kDppXor1 = 0xB1;        // quad_perm [1,0,3,2]: lane ^ 1
constexpr int kDppXor2 = 0x4E;        // quad_perm [2,3,0,1]: lane ^ 2
constexpr int kDppXor3 = 0x1B;        // quad_perm [3,2,1,0]: lane ^ 3
constexpr int kDppHalfMirror = 0x141;  // row_half_mirror: lane ^ 7 within 8 lanes

// Sum over each group of 8 lanes.
__device__ __forceinline__ double sum_lanes8(double v) {
  v += dpp_double<kDppXor1>(v);
  v += dpp_double<kDppXor2>(v);
  v += dpp_double<kDppHalfMirror>(v);  // the other quad of the group
  return v;
}

// Coordination numbers of molecules of up to 256 atoms, each unordered pair
// once: the counting function and the electronegativity factor are
// symmetric, so both atoms of a pair get the same terms. One work-group of
// kCnWaves independent waves per molecule over the 8 x 8 tiles (bx <= by),
// row and column sums in a fixed butterfly order into per-wave accumulators
// in LDS, added in wave order at the end. The work-group first does what
// k_prep and k_zslots do for its molecule (the extent n from the last
// non-zero Z, the Z range check and, with ZS, the element slots), so the
// step starts with one launch instead of three.
constexpr int kCnWaves = 8;

template <bool ZS>
__global__ __launch_bounds__(64 * kCnWaves) void k_cn_sym(DevTables t, Params p, Inputs in,
                                                            Work w) {
  extern __shared__ double s_dyn[];  // [n][4] x, y, z, Z; [kCnWaves][2][n] sums
  __shared__ int s_red[2][kCnWaves];
  __shared__ int s_present[kMaxZ];
  constexpr int NT = 64 * kCnWaves;
  const int64_t m = blockIdx.x;
  const int64_t N = w.N;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  int n;
  {
    const int64_t* z = in.numbers + m * N;
    int last = -1, bad = 0;
    for (int64_t i = tid; i < N; i += NT) {
      const int64_t zi = z[i];
      if (zi != 0) last = (int)i;
      if (zi < 0 || zi >= kMaxZ) bad = 1;
    }
    if (ZS)
      for (int e = tid; e < kMaxZ; e += NT) s_present[e] = 0;
    last = wave_max(last);
    bad = wave_or(bad);
    if (lane == 0) {
      s_red[0][wv] = last;
      s_red[1][wv] = bad;
    }
    __syncthreads();
    for (int k = 1; k < kCnWaves; ++k) {
      last = max(last, s_red[0][k]);
      bad |= s_red[1][k];
    }
    last = max(last, s_red[0][0]);
    bad |= s_red[1][0];
    n = bad ? 0 : last + 1;
    if (tid == 0) {
      w.natom[m] = n;
      w.status[m] = bad ? kStatusBadZ : 0;
    }
    if (ZS) {
      // element slots: the molecule's distinct elements in increasing Z
      for (int i = tid; i < n; i += NT) {
        const int64_t zi = z[i];
        if (zi > 0) s_present[zi] = 1;
      }
      __syncthreads();
      if (wv == 0) {
        const unsigned long long lo = __ballot(s_present[lane] != 0);
        const unsigned long long hi = __ballot(lane + 64 < kMaxZ && s_present[lane + 64] != 0);
        const int count = __popcll(lo) + __popcll(hi);
        const bool ok = count <= kZSlots;
        for (int zz = lane; zz < kMaxZ; zz += 64) {
          const unsigned long long below =
              zz < 64 ? (lo & ((1ull << zz) - 1ull))
                      : (hi & ((1ull << (zz - 64)) - 1ull));
          const int sl = (zz < 64 ? 0 : __popcll(lo)) + __popcll(below);
          const bool here = ok && zz > 0 && s_present[zz] != 0;
          if (here) w.zlist[m * kZSlots + sl] = zz;
          w.zslot[m * kMaxZ + zz] = here ? sl : -1;
        }
        if (lane == 0) w.nzs[m] = ok ? count : -1;
      }
    }
  }
  const int tx = lane & 7, ty = lane >> 3;
  const int64_t* Z = in.numbers + m * N;
  const double* R = in.positions + m * N * 3;
  double* s_at = s_dyn;
  double* s_f = s_dyn + (size_t)4 * n;
  for (int i = tid; i < n; i += NT) {
    double* at = s_at + 4 * i;
    at[0] = R[3 * i];
    at[1] = R[3 * i + 1];
    at[2] = R[3 * i + 2];
    at[3] = (double)Z[i];
  }
  for (int e = tid; e < kCnWaves * 2 * n; e += NT) s_f[e] = 0.0;
  {
    // the molecule's bounding box (the three-body kernel reads its squared
    // diagonal: a molecule inside the cutoff sphere has no screening)
    __shared__ double s_bb[kCnWaves][6];
    double lo0 = 1e300, lo1 = 1e300, lo2 = 1e300, hi0 = -1e300, hi1 = -1e300, hi2 = -1e300;
    for (int i = tid; i < n; i += NT) {
      if (Z[i] == 0) continue;
      lo0 = fmin(lo0, R[3 * i]);
      hi0 = fmax(hi0, R[3 * i]);
      lo1 = fmin(lo1, R[3 * i + 1]);
      hi1 = fmax(hi1, R[3 * i + 1]);
      lo2 = fmin(lo2, R[3 * i + 2]);
      hi2 = fmax(hi2, R[3 * i + 2]);
    }
    const double bb[6] = {wave_max(hi0), wave_max(-lo0), wave_max(hi1),
                          wave_max(-lo1), wave_max(hi2), wave_max(-lo2)};
    if (lane == 0)
      for (int k = 0; k < 6; ++k) s_bb[wv][k] = bb[k];
    __syncthreads();
    if (tid == 0) {
      double b[6];
      for (int k = 0; k < 6; ++k) {
        b[k] = s_bb[0][k];
        for (int q = 1; q < kCnWaves; ++q) b[k] = fmax(b[k], s_bb[q][k]);
      }
      const double d0 = b[0] + b[1], d1 = b[2] + b[3], d2 = b[4] + b[5];
      w.bbox2[m] = d0 * d0 + d1 * d1 + d2 * d2;
    }
  }
  const int nb = (n + 7) >> 3;
  const int tiles = nb * (nb + 1) / 2;
  double* f = s_f + (size_t)wv * 2 * n;
  for (int tl = wv; tl < tiles; tl += kCnWaves) {
    int by = (int)((sqrt(8.0 * tl + 1.0) - 1.0) * 0.5);
    while (by * (by + 1) / 2 > tl) --by;
    while ((by + 1) * (by + 2) / 2 <= tl) ++by;
    const int bx = tl - by * (by + 1) / 2;
    const int x = 8 * bx + tx, y = 8 * by + ty;
    double v[2] = {0.0, 0.0};  // D4 CN term, EEQ CN term
    if (x < y && y < n) {
      const double* ax = s_at + 4 * x;
      const double* ay = s_at + 4 * y;
      const int zx = (int)ax[3], zy = (int)ay[3];
      if (zx != 0 && zy != 0) {
        const double dx = ax[0] - ay[0], dy = ax[1] - ay[1], dz = ax[2] - ay[2];
        const double r2 = dx * dx + dy * dy + dz * dz;
        const double r = r2 * rsqrt_nt(r2);
        const PairConst pc = pair_const(t, zx, zy);
        const double c = erf_count(r, pc.irc);
        if (r <= p.cutcn) v[0] = pc.enf * c;
        if (r <= p.cuteeq) v[1] = c;
      }
    }
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const double rs = sum_lanes8(v[c]);  // y atom 8 by + ty
      double cs = v[c];                    // x atom 8 bx + tx
      cs += __shfl_xor(cs, 8);
      cs += __shfl_xor(cs, 16);
      cs += __shfl_xor(cs, 32);
      if (tx == 0 && y < n) f[c * n + y] += rs;
      if (ty == 0 && x < n) f[c * n + x] += cs;
    }
  }
  __syncthreads();
  for (int i = tid; i < N; i += NT) {
    const int64_t row = m * N + i;
    if (i >= n || s_at[4 * i + 3] == 0.0) {
      w.cn[row] = 0.0;
      w.cnq[row] = 0.0;
      w.dcnq[row] = 0.0;
      continue;
    }
    double s4 = 0.0, se = 0.0;
    for (int k = 0; k < kCnWaves; ++k) {
      s4 += s_f[(size_t)k * 2 * n + i];
      se += s_f[(size_t)k * 2 * n + n + i];
    }
    w.cn[row] = s4;
    // smooth cap: log(1 + e^cmax) - log(1 + e^(cmax - cn))
    const double ex = exp(kCnMax - se);
    w.cnq[row] = log(1.0 + exp(kCnMax)) - log(1.0 + ex);
    w.dcnq[row] = ex / (1.0 + ex);
  }
}

// part kPairsAtm writes the q = 0 weights (gw0, dgw0), kPairsTwoBody those at
// the EEQ charges (gw, dgw, dgwq), kPairsBoth all of them.
// CV (D4, part q = 0 or part q alone): also the atom's C6 contraction
// vectors, for each element slot e of its molecule v[a] = sum_b
// RC6[Z_e][Z_j][a][b] gw_j[b] (a < 7; entry 7 is zero) from the weights just
// formed (lane a walks the slots), which the pair pass reads in place of the
// per-pair 7 x 7 reference blocks.
template <bool CV>
__global__ __launch_bounds__(256) void k_weights(DevTables t, Params p, Work w,
                                                 Inputs in, int part) {
  const bool atm = part != kPairsTwoBody, two = part != kPairsAtm;
  const int64_t tid = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t row = tid >> 3;
  const int a = (int)(tid & 7);
  if (row >= w.B * w.N) return;  // uniform over each group of 8 lanes
  const int64_t m = row / w.N;
  const int i = (int)(row - m * w.N);
  const int z = i < w.natom[m] ? (int)in.numbers[row] : 0;
  const int nr = z != 0 ? t.nref[z] : 0;
  const double cn = z != 0 ? w.cn[row] : 0.0;
  const double q = (z != 0 && two) ? w.q[row] : 0.0;
  const bool used = a < nr;
  const double zeff = t.zeff[z];
  const double gam = t.gam[z] * p.gc;
  const double qref = used ? t.refq[z * kMaxRef + a] + zeff : 0.0;
  // charge scaling of the part in hand: zeta(q) and dzeta/dq, zeta(0)
  double zq = 0.0, dzq = 0.0, z0 = 0.0;
  if (used && two) zeta_dzeta(p.ga, gam, qref, q + zeff, zq, dzq);
  if (used && atm) z0 = zeta_fn(p.ga, gam, qref, zeff);
  const int64_t o = row * kRefStride + a;
  if (p.d4s) {
    // D4S: the Gaussian part depends on the partner atom and is formed in
    // the pair pass; keep the charge scaling zeta(q), zeta(0), dzeta/dq of
    // every reference (model/d4s.py:240-263)
    if (two) {
      w.gw[o] = zq;
      w.dgw[o] = 0.0;
      w.dgwq[o] = dzq;
    }
    if (atm) {
      w.gw0[o] = z0;
      w.dgw0[o] = 0.0;
    }
    return;
  }
  double e = 0.0, de = 0.0;
  if (used) {
    // the reference forms exp(-d^2)^(wf k) (model/d4.py:174-188); one
    // exponential of -wf k d^2 is the same value to ~1e-13 relative, and
    // underflows where it does (both round one exponent of nearly the same
    // argument)
    const double d = cn - t.refcn[z * kMaxRef + a];
    const double d2 = d * d;
    const int ng = t.ngw[z * kMaxRef + a];
    for (int k = 1; k <= ng; ++k) {
      const double kw = k * p.wf;
      const double ek = exp(-kw * d2);
      e += ek;
      de -= 2.0 * kw * d * ek;
    }
  }
  const double norm = sum_lanes8(e);
  const double dnorm = sum_lanes8(de);
  const bool exceptional = !(norm > 0.0);
  double g = 0.0, dg = 0.0;
  if (used) {
    if (exceptional) {
      g = t.refcn[z * kMaxRef + a] == t.maxcn[z] ? 1.0 : 0.0;
    } else {
      // form the ratio first: expw * dnorm underflows when the Gaussians
      // are far below 1 (CN several units from every reference CN)
      g = e / norm;
      dg = (de - g * dnorm) / norm;
    }
  }
  if (two) {
    w.gw[o] = zq * g;
    w.dgw[o] = zq * dg;
    w.dgwq[o] = dzq * g;
  }
  if (atm) {
    w.gw0[o] = z0 * g;
    w.dgw0[o] = z0 * dg;
  }
  if (CV) {
    // every lane of the atom's 8-lane group gathers the atom's weights of
    // the part (uniform per group: the group's lanes are all active here)
    const double gsel = atm ? z0 * g : zq * g;
    const int base = (threadIdx.x & 63) & ~7;
    double gg[kMaxRef];
#pragma unroll
    for (int b = 0; b < kMaxRef; ++b) gg[b] = __shfl(gsel, base + b, 64);
    const int nz = w.nzs[m];
    if (z == 0 || nz <= 0) return;
    double* out = (atm ? w.cv0 : w.cv) + row * kZSlots * 8 + a;
    for (int e = 0; e < nz; ++e) {
      double v = 0.0;
      if (a < kMaxRef) {
        const double* blk =
            t.rc6 + ((size_t)w.zlist[m * kZSlots + e] * kMaxZ + z) * kRc6Block + a * kMaxRef;
#pragma unroll
        for (int b = 0; b < kMaxRef; ++b) v += blk[b] * gg[b];
      }
      out[e * 8] = v;
    }
  }
}

// --------------------------------------------------------------------------
// D4S Gaussian weights (model/d4s.py:178-238): normalised weights of an atom
// of element z with coordination number cn for the pairwise factor wf of one
// partner, exp(-wf k (cn - cn_ref)^2) summed over k = 1..ngw, and with DCN
// their derivatives with respect to cn. When every Gaussian underflows the
// weight goes to the reference(s) with the largest reference CN.
// --------------------------------------------------------------------------

template <bool DCN>
__device__ __forceinline__ void pair_gauss(const DevTables& t, int z, int nr, double cn,
                                           double wf, double* g, double* dg) {
  double e[kMaxRef], de[kMaxRef];
  double norm = 0.0, dnorm = 0.0;
#pragma unroll
  for (int a = 0; a < kMaxRef; ++a) {
    e[a] = 0.0;
    de[a] = 0.0;
    if (a < nr) {
      const double d = cn - t.refcn[z * kMaxRef + a];
      const double tmp = exp(-(d * d) * wf);
      double sum = tmp, ksum = tmp;  // sum_k tmp^k and sum_k k tmp^k
      if (t.ngw[z * kMaxRef + a] == 3) {
        const double t2 = tmp * tmp, t3 = t2 * tmp;
        sum = tmp + t2 + t3;
        ksum = fma(3.0, t3, fma(2.0, t2, tmp));
      }
      e[a] = sum;
      if (DCN) de[a] = -2.0 * wf * d * ksum;
      norm += sum;
      if (DCN) dnorm += de[a];
    }
  }
  const bool exceptional = !(norm > 0.0);
  // one reciprocal per atom, unless the norm is so small that 1/norm
  // overflows (Gaussians near underflow): then divide per reference
  const bool tiny = !(norm >= 0x1p-1000);
  const double maxcn = t.maxcn[z];
  const double rnorm = tiny ? 0.0 : 1.0 / norm;
#pragma unroll
  for (int a = 0; a < kMaxRef; ++a) {
    double v = 0.0, dv = 0.0;
    if (a < nr) {
      if (exceptional) {
        v = t.refcn[z * kMaxRef + a] == maxcn ? 1.0 : 0.0;
      } else if (tiny) {
        v = e[a] / norm;
        if (DCN) dv = (de[a] - v * dnorm) / norm;
      } else {
        v = e[a] * rnorm;
        if (DCN) dv = (de[a] - v * dnorm) * rnorm;
      }
    }
    g[a] = v;
    if (DCN) dg[a] = dv;
  }
}

// --------------------------------------------------------------------------
// pair pass: C6(q) and C6(0) contraction with the reference-C6 table, the
// rational-damped two-body term, the ATM pair factors; with GRAD also the
// two-body gradient pieces. With S (D4S model) the weights of both atoms are
// formed per pair from the charge scaling that k_weights left and the pair's
// Gaussian weights. PART selects the three-body pair factors (C6 at q = 0:
// independent of the charges), the two-body term (C6 at the EEQ charges) or
// both in one pass.
// --------------------------------------------------------------------------

constexpr int kPartBoth = 0;
constexpr int kPartAtm = 1;  // t, p, d log sqrt|C6(0)| / d CN (and sqrt r4r2 for PG)
constexpr int kPartE2 = 2;   // two-body energy, dL/dCN, dL/dq, forces, PG partials

// Element slots of a molecule (one wave per molecule): its distinct elements
// in ascending order; a molecule with more than kZSlots elements gets none
// (nzs = -1) and its pair passes contract the reference C6 blocks per pair.
__global__ __launch_bounds__(64) void k_zslots(Inputs in, Work w) {
  __shared__ int present[kMaxZ];
  const int64_t m = blockIdx.x;
  const int lane = threadIdx.x;
  const int n = w.natom[m];
  const int64_t* Z = in.numbers + m * w.N;
  for (int z = lane; z < kMaxZ; z += 64) present[z] = 0;
  __syncthreads();
  for (int i = lane; i < n; i += 64) {
    const int64_t z = Z[i];
    if (z > 0 && z < kMaxZ) present[z] = 1;
  }
  __syncthreads();
  if (lane != 0) return;
  int count = 0;
  for (int z = 1; z < kMaxZ; ++z) count += present[z];
  const bool ok = count <= kZSlots;
  int s = 0;
  for (int z = 0; z < kMaxZ; ++z) {
    int slot = -1;
    if (ok && z > 0 && present[z]) {
      w.zlist[m * kZSlots + s] = z;
      slot = s++;
    }
    w.zslot[m * kMaxZ + z] = slot;
  }
  w.nzs[m] = ok ? count : -1;
}

// Atom i of a pair (the atom whose sums are accumulated).
struct PairAtom {
  int z, nr;
  double x, y, zc, r4, cn, g;
  // D4: weights and their derivatives; D4S: zeta(q), zeta(0), dzeta/dq
  double gw[kMaxRef], gw0[kMaxRef], dgw[kMaxRef], dgw0[kMaxRef], dgwq[kMaxRef];
};

struct PairSums {
  double e = 0.0, cn = 0.0, q = 0.0, fx = 0.0, fy = 0.0, fz = 0.0;
  // parameter gradients of sum_i gE_i E2_i: d/ds6, d/ds8, d/ds10, d/da1, d/da2
  double s6 = 0.0, s8 = 0.0, s10 = 0.0, a1 = 0.0, a2 = 0.0;
};

template <bool GRAD, bool S, int PART>
__device__ __forceinline__ void load_pair_atom(const DevTables& t, const Work& w,
                                               const Inputs& in, int64_t row, int zi,
                                               const double* R, int i, PairAtom& A) {
  constexpr bool ATMF = PART != kPartE2, E2 = PART != kPartAtm;
  A.z = zi;
  A.nr = t.nref[zi];
  A.x = R[3 * i];
  A.y = R[3 * i + 1];
  A.zc = R[3 * i + 2];
  A.r4 = in.r4r2_in ? in.r4r2_in[row] : t.r4r2[zi];
  A.cn = S ? w.cn[row] : 0.0;
  A.g = (GRAD && E2) ? (in.gE ? in.gE[row] : 1.0) : 0.0;
#pragma unroll
  for (int a = 0; a < kMaxRef; ++a) {
    A.gw[a] = E2 ? w.gw[row * kRefStride + a] : 0.0;
    A.gw0[a] = ATMF ? w.gw0[row * kRefStride + a] : 0.0;
    A.dgw[a] = (GRAD && E2 && !S) ? w.dgw[row * kRefStride + a] : 0.0;
    A.dgw0[a] = (GRAD && ATMF && !S) ? w.dgw0[row * kRefStride + a] : 0.0;
    A.dgwq[a] = (GRAD && E2) ? w.dgwq[row * kRefStride + a] : 0.0;
  }
}

// One pair (i, j): accumulates atom i's two-body energy, dL/dCN_i, dL/dq_i,
// forces and parameter partials, and returns the ATM pair factors
// tf = (R0/r)^(alp/3), pf = sqrt|C6(0)|/r^3 and pk = d log sqrt|C6(0)| / d CN_i.
template <bool GRAD, bool PG, bool S, int PART>
__device__ __forceinline__ void pair_eval(const DevTables& t, const Params& p, const PairAtom& A,
                                          int zj, double xj, double yj, double zzj, double r4j,
                                          double cnj, double gj, const double* gwjp,
                                          const double* gw0jp, const double* vj,
                                          PairSums& acc, double& tf, double& pf, double& pk) {
  constexpr bool ATMF = PART != kPartE2, E2 = PART != kPartAtm;
  const double dx = A.x - xj, dy = A.y - yj, dz = A.zc - zzj;
  const double r2 = dx * dx + dy * dy + dz * dz;
  const double ir = rsqrt_nt(r2);
  const double r = r2 * ir;
  // D4 split passes of molecules with element slots: the partner's weights
  // come precontracted with the reference C6 block of this atom's element
  const bool pv = !S && PART != kPartBoth && vj != nullptr;
  double gwj[kMaxRef], gw0j[kMaxRef];
#pragma unroll
  for (int b = 0; b < kMaxRef; ++b) {
    gwj[b] = (E2 && !pv) ? gwjp[b] : 0.0;
    gw0j[b] = (ATMF && !pv) ? gw0jp[b] : 0.0;
  }
  // D4S: Gaussian weights of both atoms for this pair; atom n's weights for
  // partner m use wf(Z_n, Z_m) (model/d4s.py:196). The weights of atom i
  // are zeta_i[a] g[a] (A.gw, A.gw0, A.dgwq hold zeta).
  double g[kMaxRef], dg[kMaxRef];
  if (S) {
    pair_gauss<GRAD>(t, A.z, A.nr, A.cn, t.wfpair[A.z * kMaxZ + zj], g, dg);
    double gjw[kMaxRef];
    pair_gauss<false>(t, zj, t.nref[zj], cnj, t.wfpair[zj * kMaxZ + A.z], gjw, nullptr);
#pragma unroll
    for (int b = 0; b < kMaxRef; ++b) {
      gwj[b] *= gjw[b];
      gw0j[b] *= gjw[b];
    }
  }
  const double* blk = t.rc6 + ((size_t)A.z * kMaxZ + zj) * kRc6Block;
  double c6 = 0.0, c60 = 0.0, dc6cn = 0.0, dc6q = 0.0, dc60cn = 0.0;
  if (pv) {
    // v[a] = sum_b RC6[Z_i][Z_j][a][b] gw_j[b], summed as below; the weights
    // of reference systems a >= nref are zero, so the sums match bit for bit
    const double2* v2 = reinterpret_cast<const double2*>(vj);
    const double2 v01 = v2[0], v23 = v2[1], v45 = v2[2], v67 = v2[3];
    const double v[kMaxRef] = {v01.x, v01.y, v23.x, v23.y, v45.x, v45.y, v67.x};
#pragma unroll
    for (int a = 0; a < kMaxRef; ++a) {
      if (E2) c6 += A.gw[a] * v[a];
      if (ATMF) c60 += A.gw0[a] * v[a];
      if (GRAD && E2) {
        dc6cn += A.dgw[a] * v[a];
        dc6q += A.dgwq[a] * v[a];
      }
      if (GRAD && ATMF) dc60cn += A.dgw0[a] * v[a];
    }
  }
#pragma unroll
  for (int a = 0; a < kMaxRef; ++a) {
    if (pv || a >= A.nr) continue;
    double va = 0.0, v0a = 0.0;
#pragma unroll
    for (int b = 0; b < kMaxRef; ++b) {
      const double rc = blk[a * kMaxRef + b];
      if (E2) va += rc * gwj[b];
      if (ATMF) v0a += rc * gw0j[b];
    }
    if (S) {
      if (E2) c6 += A.gw[a] * g[a] * va;
      if (ATMF) c60 += A.gw0[a] * g[a] * v0a;
      if (GRAD && E2) {
        dc6cn += A.gw[a] * dg[a] * va;
        dc6q += A.dgwq[a] * g[a] * va;
      }
      if (GRAD && ATMF) dc60cn += A.gw0[a] * dg[a] * v0a;
    } else {
      if (E2) c6 += A.gw[a] * va;
      if (ATMF) c60 += A.gw0[a] * v0a;
      if (GRAD && E2) {
        dc6cn += A.dgw[a] * va;
        dc6q += A.dgwq[a] * va;
      }
      if (GRAD && ATMF) dc60cn += A.dgw0[a] * v0a;
    }
  }
  // rational damping, radii = 3 r4r2_i r4r2_j (twobody.py:147-157)
  const double qq = 3.0 * A.r4 * r4j;
  const double r0 = p.a1 * sqrt(qq) + p.a2;
  tf = pf = pk = 0.0;
  if (ATMF) {
    // ATM pair factors (threebody.py:106-145, 241-256, 311-321); r0 > 0, so
    // (r0 / r)^(alp / 3) = exp(alp / 3 log(r0 / r))
    tf = exp((p.alp * (1.0 / 3.0)) * log(r0 * ir));
    pf = sqrt(fabs(c60)) * (ir * ir * ir);
    if (GRAD) pk = c60 != 0.0 ? 0.5 * dc60cn / c60 : 0.0;
  }
  if (!E2) return;
  const double r6 = r2 * r2 * r2, r8 = r6 * r2;
  const double r06 = r0 * r0 * r0 * r0 * r0 * r0, r08 = r06 * r0 * r0;
  const double t6 = rcp_nt(r6 + r06);  // r0^6 > 0
  const double t8 = rcp_nt(r8 + r08);
  double de_dc6 = p.s6 * t6 + p.s8 * qq * t8;
  double t10 = 0.0, r10 = 0.0;
  if (p.s10on) {
    r10 = r8 * r2;
    t10 = rcp_nt(r10 + r08 * r0 * r0);
    de_dc6 += p.s10 * qq * qq * (49.0 / 40.0) * t10;
  }
  de_dc6 *= -0.5;
  const bool in2 = r <= p.cut2;
  if (in2) acc.e += de_dc6 * c6;
  if (PG && in2) {
    const double ce = -0.5 * c6 * A.g;
    acc.s6 = fma(ce, t6, acc.s6);
    acc.s8 = fma(ce * qq, t8, acc.s8);
    // d/dr0 of s6 t6 + s8 qq t8 (+ s10 ...): d t_n / d r0 = -n r0^(n-1) t_n^2
    double dr0 = p.s6 * (-6.0 * r06 / r0) * t6 * t6 + p.s8 * qq * (-8.0 * r08 / r0) * t8 * t8;
    if (p.s10on) {
      acc.s10 = fma(ce * qq * qq * (49.0 / 40.0), t10, acc.s10);
      dr0 += p.s10 * qq * qq * (49.0 / 40.0) * (-10.0 * r08 * r0) * t10 * t10;
    }
    acc.a1 = fma(ce * dr0, sqrt(qq), acc.a1);
    acc.a2 = fma(ce, dr0, acc.a2);
  }
  if (GRAD && in2) {
    const double wij = A.g + gj;
    acc.cn += wij * de_dc6 * dc6cn;
    acc.q += wij * de_dc6 * dc6q;
    // d/dr of 1/(r^n + r0^n) = -n r^(n-1) t^2
    double dedr = p.s6 * (-6.0 * r6 * ir) * t6 * t6 + p.s8 * qq * (-8.0 * r8 * ir) * t8 * t8;
    if (p.s10on) dedr += p.s10 * qq * qq * (49.0 / 40.0) * (-10.0 * r10 * ir) * t10 * t10;
    dedr *= -0.5 * c6;
    const double coef = wij * dedr * ir;
    acc.fx += coef * dx;
    acc.fy += coef * dy;
    acc.fz += coef * dz;
  }
}

template <bool GRAD, bool PG, int PART>
__device__ __forceinline__ void store_pair_atom(const Work& w, int64_t row, double r4,
                                                const PairSums& acc, double* grad_out) {
  constexpr bool ATMF = PART != kPartE2, E2 = PART != kPartAtm;
  if (PG && ATMF) w.rt4[row] = sqrt(r4);
  if (!E2) return;
  w.e2[row] = acc.e;
  if (GRAD) {
    w.dldcn[row] = acc.cn;
    w.dldq[row] = acc.q;
    grad_out[3 * row] = acc.fx;
    grad_out[3 * row + 1] = acc.fy;
    grad_out[3 * row + 2] = acc.fz;
  }
  if (PG) {
    w.pg2[row * 5 + 0] = acc.s6;
    w.pg2[row * 5 + 1] = acc.s8;
    w.pg2[row * 5 + 2] = acc.s10;
    w.pg2[row * 5 + 3] = acc.a1;
    w.pg2[row * 5 + 4] = acc.a2;
  }
}

// RL lanes per atom row i walk the partners j; pt[i][j], pp[i][j] and
// pk[i][j] are written along the row.
template <bool GRAD, bool PG, bool S, int RL, int PART>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_pairs(
    DevTables t, Params p, Inputs in, Work w, double* grad_out) {
  constexpr bool ATMF = PART != kPartE2, E2 = PART != kPartAtm;
  const int lane = threadIdx.x & (RL - 1);
  const int64_t row = (int64_t)blockIdx.x * (blockDim.x / RL) + threadIdx.x / RL;
  if (row >= w.B * w.N) return;
  const int64_t N = w.N;
  const int64_t m = row / N;
  const int i = (int)(row - m * N);
  const int n = w.natom[m];
  const int64_t* Z = in.numbers + m * N;
  const double* R = in.positions + m * N * 3;
  double* __restrict__ ptr = w.pt + row * N;
  double* __restrict__ ppr = w.pp + row * N;
  double* __restrict__ pkr = w.pk + row * N;
  const double* __restrict__ gw = w.gw;
  const double* __restrict__ gw0 = w.gw0;
  const int zi = i < n ? (int)Z[i] : 0;
  if (zi == 0) {
    // rows past the molecule are never read (readers stop at row n)
    if (ATMF && i < n) {
      for (int j = lane; j < n; j += RL) {
        ptr[j] = 0.0;
        ppr[j] = 0.0;
        if (GRAD) pkr[j] = 0.0;
      }
    }
    if (lane == 0) store_pair_atom<GRAD, PG, PART>(w, row, 0.0, PairSums{}, grad_out);
    return;
  }
  PairAtom A;
  load_pair_atom<GRAD, S, PART>(t, w, in, row, zi, R, i, A);
  PairSums acc;
  const int slot = (!S && PART != kPartBoth) ? w.zslot[m * kMaxZ + zi] : -1;
  const double* cvp = PART == kPartAtm ? w.cv0 : w.cv;
  // the partner's Z is loaded one iteration ahead: the table reads of a pair
  // (reference C6 block, r4r2) depend on it
  int znext = lane < n ? (int)Z[lane] : 0;
  for (int j = lane; j < n; j += RL) {
    const int zj = znext;
    znext = j + RL < n ? (int)Z[j + RL] : 0;
    if (j == i || zj == 0) {
      if (ATMF) {
        ptr[j] = 0.0;
        ppr[j] = 0.0;
        if (GRAD) pkr[j] = 0.0;
      }
      continue;
    }
    const int64_t rj = m * N + j;
    const double r4j = in.r4r2_in ? in.r4r2_in[rj] : t.r4r2[zj];
    const double cnj = S ? w.cn[rj] : 0.0;
    const double gj = (GRAD && E2) ? (in.gE ? in.gE[rj] : 1.0) : 0.0;
    double tf, pf, pk;
    const double* vj = slot >= 0 ? cvp + (rj * kZSlots + slot) * 8 : nullptr;
    pair_eval<GRAD, PG, S, PART>(t, p, A, zj, R[3 * j], R[3 * j + 1], R[3 * j + 2], r4j, cnj,
                                 gj, gw + rj * kRefStride, gw0 + rj * kRefStride, vj, acc, tf,
                                 pf, pk);
    if (ATMF) {
      ptr[j] = tf;
      ppr[j] = pf;
      if (GRAD) pkr[j] = pk;
    }
  }
  if (E2) {
    acc.e = row_sum<RL>(acc.e);
    if (GRAD) {
      acc.cn = row_sum<RL>(acc.cn);
      acc.q = row_sum<RL>(acc.q);
      acc.fx = row_sum<RL>(acc.fx);
      acc.fy = row_sum<RL>(acc.fy);
      acc.fz = row_sum<RL>(acc.fz);
    }
    if (PG) {
      acc.s6 = row_sum<RL>(acc.s6);
      acc.s8 = row_sum<RL>(acc.s8);
      acc.s10 = row_sum<RL>(acc.s10);
      acc.a1 = row_sum<RL>(acc.a1);
      acc.a2 = row_sum<RL>(acc.a2);
    }
  }
  if (lane == 0) store_pair_atom<GRAD, PG, PART>(w, row, A.r4, acc, grad_out);
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
//
// Screening by groups: a triple counts only if two of its three edges are
// within the cutoff c3. With bounding spheres of the 16-atom groups (the
// tile edges), a tile whose two groups are farther apart than 2 c3 has no
// counted triple at all, and a third-atom group is streamed only if at least
// two of the three group-to-group lower distance bounds are within c3. The
// kept groups are listed in ascending order, so the third atoms are visited
// in the same order with or without screening and skipped terms are exact
// zeros: screening does not change a single bit of the result.
// --------------------------------------------------------------------------

constexpr int kMaxGroups = 2880;  // 46000 atoms / 16

__device__ __forceinline__ double sphere_gap(const double* a, const double* b) {
  const double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
  return sqrt(dx * dx + dy * dy + dz * dz) - a[3] - b[3];
}

template <bool ENERGY, bool GRAD, bool PG>
__global__ __launch_bounds__(256) void k_atm(Params p, Inputs in, Work w,
                                             int tiles_per_mol, int64_t ntiles) {
  __shared__ double s_t[kZB][2 * kTile];
  __shared__ double s_p[kZB][2 * kTile];
  __shared__ double s_r[kZB][4];  // x, y, z, dL/dE (0 for padding)
  __shared__ int s_ok[kZB];
  __shared__ int s_z[kZB];
  __shared__ int s_glist[kMaxGroups];
  __shared__ int s_cnt[kWaves];
  __shared__ double s_rt[kZB];  // sqrt(r4r2) of the third atoms (PG)
  __shared__ double s_pg[4][kWaves];

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

  // group screening (uniform over the work-group)
  const double c3 = sqrt(p.cut3sq) * (1.0 + 1e-10) + 1e-10;  // margin for rounding in the bounds
  const double* S = w.gsph + 4 * m * n_groups(N);
  const double gxy = sphere_gap(S + 4 * bx, S + 4 * by);
  if (gxy > 2.0 * c3) {  // no counted triple has this pair
    if (valid) {
      const int64_t ixy = (m * N + x) * N + y, iyx = (m * N + y) * N + x;
      if (ENERGY) w.pe[ixy] = w.pe[iyx] = 0.0;
      if (GRAD) w.pg[ixy] = w.pg[iyx] = w.ph[ixy] = w.ph[iyx] = 0.0;
    }
    if (PG && threadIdx.x < 4) w.pg3[logical * 4 + threadIdx.x] = 0.0;
    return;
  }
  // parameter gradients: s9, alp, a1, a2 (each triple is seen from its three
  // pairs, so the sums are divided by 3 at the end)
  double g_s9 = 0.0, g_alp = 0.0, g_a1 = 0.0, g_a2 = 0.0;
  const double a1p = p.a1, a2p = p.a2, sq3 = 1.7320508075688772;
  double rx4 = 0.0, ry4 = 0.0, Axy = 0.0, Bxy = 0.0;
  if (PG && valid) {
    rx4 = w.rt4[m * N + x];
    ry4 = w.rt4[m * N + y];
    const double sig = sq3 * rx4 * ry4;
    Bxy = 1.0 / fma(a1p, sig, a2p);
    Axy = sig * Bxy;
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int total = 0;
  for (int base = 0; base < nb; base += 256) {
    const int gz = base + threadIdx.x;
    bool keep = false;
    if (gz < nb) {
      const int near = (gxy <= c3) + (sphere_gap(S + 4 * bx, S + 4 * gz) <= c3) +
                       (sphere_gap(S + 4 * by, S + 4 * gz) <= c3);
      keep = near >= 2;
    }
    const unsigned long long mask = __ballot(keep);
    if (lane == 0) s_cnt[wave] = __popcll(mask);
    __syncthreads();
    int off = total;
    for (int k = 0; k < wave; ++k) off += s_cnt[k];
    if (keep) s_glist[off + __popcll(mask & ((1ull << lane) - 1ull))] = gz;
    for (int k = 0; k < kWaves; ++k) total += s_cnt[k];
    __syncthreads();
  }

  for (int gi = 0; gi < total; gi += 2) {
    __syncthreads();
    for (int e = threadIdx.x; e < kZB * 4 * kTile; e += 256) {
      const int zl = e / (4 * kTile), c = e - zl * (4 * kTile);
      const int gs = gi + (zl >> 4);
      const int z = gs < total ? s_glist[gs] * kTile + (zl & (kTile - 1)) : n;
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
      const int zl = threadIdx.x;
      const int gs = gi + (zl >> 4);
      const int z = gs < total ? s_glist[gs] * kTile + (zl & (kTile - 1)) : n;
      const bool ok = z < n && Z[z] != 0;
      s_z[zl] = z;
      s_ok[zl] = ok;
      s_r[zl][0] = ok ? R[3 * z] : 0.0;
      s_r[zl][1] = ok ? R[3 * z + 1] : 0.0;
      s_r[zl][2] = ok ? R[3 * z + 2] : 0.0;
      s_r[zl][3] = (GRAD && ok) ? (in.gE ? in.gE[m * N + z] : 1.0) : 0.0;
      if (PG) s_rt[zl] = ok ? w.rt4[m * N + z] : 0.0;
    }
    __syncthreads();
    if (!valid) continue;
    // slots up to the last atom of the molecule (only the last group is partial)
    const int z0 = s_glist[gi] * kTile;
    const int zmax = gi + 1 < total ? kTile + min(kTile, n - s_glist[gi + 1] * kTile)
                                    : min(kTile, n - z0);
    for (int zl = 0; zl < zmax; ++zl) {
      const int z = s_z[zl];
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
        if (PG) {
          const double e1 = P * f * g;  // energy without s9
          const double dT = -6.0 * f * en * wgt;
          const double TdT = T * dT;
          g_s9 = fma(wgt, e1, g_s9);
          g_alp = fma(TdT, T > 0.0 ? log(T) / alp : 0.0, g_alp);
          const double sxz = sq3 * rx4 * s_rt[zl], syz = sq3 * ry4 * s_rt[zl];
          const double bxz = 1.0 / fma(a1p, sxz, a2p), byz = 1.0 / fma(a1p, syz, a2p);
          g_a1 = fma(TdT * alp / 3.0, Axy + sxz * bxz + syz * byz, g_a1);
          g_a2 = fma(TdT * alp / 3.0, Bxy + bxz + byz, g_a2);
        }
      }
    }
  }
  if (PG) {
    __syncthreads();
    const double third = 1.0 / 3.0;
    const double v4[4] = {g_s9 * third, g_alp * third, g_a1 * third, g_a2 * third};
    for (int k = 0; k < 4; ++k) {
      const double v = wave_sum(v4[k]);
      if ((threadIdx.x & 63) == 0) s_pg[k][threadIdx.x >> 6] = v;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
      double v = 0.0;
      for (int q = 0; q < kWaves; ++q) v += s_pg[threadIdx.x][q];
      w.pg3[logical * 4 + threadIdx.x] = v;
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
// ATM three-body term with each unordered triple evaluated once (molecules
// of up to kAtm1MaxN atoms).
//
// Work-group (m, s) takes every S-th pair tile (bx <= by) of molecule m,
// S = atm1_splits(n). Lane (tx, ty) owns the pair x = 16 bx + tx <
// y = 16 by + ty and visits the third atoms z > y in chunks of 16. Inside a
// chunk the 16 lanes of a row visit the chunk on a skewed schedule,
// z = z0 + (k + tx) mod 16 at step k, and hand their running sums for z one
// lane down after every step (a DPP row rotate): after 16 steps every lane
// holds its row's complete sums for one z. Contributions of x and y stay in
// the lane. Row, tile and split sums are added in a fixed order, so there
// are no atomics and the result depends on the molecule alone.
//
// Per triple (threebody.py:54-163 with the partition of the asymmetric
// screening): atom x gets e/6 [yz]([xy] + [xz]) (y, z alike), and the
// gradient and CN chain use the weight W = sum over the three atoms of
// dL/dE_atom times its share.
// --------------------------------------------------------------------------

// Lane i of each 16-lane row receives the value of lane i + 1 (mod 16).
__device__ __forceinline__ double row_rotate_down(double v) {
  const int lo = __double2loint(v), hi = __double2hiint(v);
  return __hiloint2double(__builtin_amdgcn_update_dpp(hi, hi, 0x12F, 0xF, 0xF, false),
                          __builtin_amdgcn_update_dpp(lo, lo, 0x12F, 0xF, 0xF, false));
}

// 1 / d by the hardware estimate (relative error below 2^-24 on gfx950) and
// one Newton step: relative error below 3e-15.
__device__ __forceinline__ double rcp_newton1(double d) {
  const double r = __builtin_amdgcn_rcp(d);
  return fma(r, fma(-d, r, 1.0), r);
}


// Pair data of (z, c) staged for the one-pass kernel, c an atom of the tile.
struct ZPair {
  double t;    // (R0 / r)^(alp / 3)
  double p;    // sqrt|C6(q = 0)| / r^3
  double kz;   // d log sqrt|C6(0)| / d CN_z
  double kc;   // d log sqrt|C6(0)| / d CN_c
};

#ifndef D4_ATM1_UNROLL
#define D4_ATM1_UNROLL 2
#endif
#ifndef D4_ATM1_WAVES
#define D4_ATM1_WAVES 1
#endif

template <bool ENERGY, bool GRAD, bool PG>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(D4_ATM1_WAVES))) void k_atm1(
    Params p, Inputs in, Work w, int S, int64_t nblocks) {
  constexpr int NA = kAtm1Acc;
  constexpr double kFar = 1.0e6;  // position of padding third atoms (bohr)
  // [z][x block | y block] pair records. The row stride is 67 16-byte units
  // (odd): the 16 lanes of a row read 16 different rows z on the skewed
  // schedule, and an odd stride keeps their x- and y-column reads on
  // distinct LDS banks.
  constexpr int kZcRow = 134;                   // doubles per z row (33 records + 16 B)
  __shared__ __align__(16) double s_zc[kTile * kZcRow];
  __shared__ __align__(16) double s_r[kTile * 6];  // z position, dL/dE_z (48-byte stride)
  __shared__ double s_rt[kTile];                // sqrt(r4r2_z) (parameter gradients)
  __shared__ double s_red[kTile][kTile][NA];
  __shared__ double s_acc[kAtm1MaxN][NA];

  const int64_t b = blockIdx.x;
  // block b runs on XCD b % 8. Molecule m owns blocks m S .. m S + S - 1 and
  // its split s runs in block m S + (s + r) mod S, on XCD (m + s) mod 8 (up to
  // the wrap-around): consecutive splits of a molecule go to different XCDs,
  // and split 0 (and the other low splits, which every molecule has) of
  // consecutive molecules too. r = m mod S when S is a multiple of 8, else
  // r = m (1 - S) mod 8 (both measured: the first is up to 20 % faster where
  // it applies, where it does not it would undo the rotation m S gives).
  if (b >= nblocks) return;
  const int64_t m = b / S;
  const int slot = (int)(b - m * S);
#ifdef D4_ATM1_NOROT
  const int s = slot;
#else
  const int r = (S & 7) == 0 ? (int)(m % S) : (int)(m % 8) * (((1 - S) % 8 + 8) % 8) % 8;
  const int s = (slot + S - r % S) % S;
#endif
  const int64_t N = w.N;
  const int n = w.natom[m];
  const int Sm = atm1_splits(n);  // this molecule's split count (S is the batch maximum)
  if (s >= Sm) return;
  const int nb = (n + kTile - 1) / kTile;
  const int tiles = nb * (nb + 1) / 2;
  const int tid = threadIdx.x, tx = tid & (kTile - 1), ty = tid >> 4;
  const int64_t* Z = in.numbers + m * N;
  const double* R = in.positions + m * N * 3;
  const double* pt = w.pt + m * N * N;
  const double* pp = w.pp + m * N * N;
  const double* pk = w.pk + m * N * N;
  const double c3sq = p.cut3sq, s9 = p.s9, alp = p.alp;
  for (int e = tid; e < n * NA; e += 256) (&s_acc[0][0])[e] = 0.0;
  // parameter gradients (times 6): s9, alp, a1, a2
  double g_s9 = 0.0, g_alp = 0.0, g_a1 = 0.0, g_a2 = 0.0;
  // molecules whose bounding box fits in the cutoff sphere have every pair
  // within the cutoff: all screening weights are 2 and their masks drop out
  bool full;
  {
    double lo0 = 1e300, lo1 = 1e300, lo2 = 1e300, hi0 = -1e300, hi1 = -1e300, hi2 = -1e300;
    for (int i = tid & 63; i < n; i += 64) {
      if (Z[i] == 0) continue;
      lo0 = fmin(lo0, R[3 * i]);
      hi0 = fmax(hi0, R[3 * i]);
      lo1 = fmin(lo1, R[3 * i + 1]);
      hi1 = fmax(hi1, R[3 * i + 1]);
      lo2 = fmin(lo2, R[3 * i + 2]);
      hi2 = fmax(hi2, R[3 * i + 2]);
    }
    const double d0 = wave_max(hi0) + wave_max(-lo0), d1 = wave_max(hi1) + wave_max(-lo1),
                 d2 = wave_max(hi2) + wave_max(-lo2);
    full = d0 * d0 + d1 * d1 + d2 * d2 <= c3sq;  // the same in every wave
  }
  const double a1p = p.a1, a2p = p.a2, sq3 = 1.7320508075688772;

  for (int t = s; t < tiles; t += Sm) {
    int by = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
    while (by * (by + 1) / 2 > t) --by;
    while ((by + 1) * (by + 2) / 2 <= t) ++by;
    const int bx = t - by * (by + 1) / 2;
    const int x = bx * kTile + tx, y = by * kTile + ty;
    const bool valid = x < y && y < n && Z[x] != 0 && Z[y] != 0;
    // lanes without a pair sit far from every atom (no zero distance; their
    // pair factor is zero)
    double rx0 = -kFar, rx1 = -kFar, rx2 = -kFar, ry0 = -kFar, ry1 = -kFar, ry2 = -kFar;
    double a = 1.0, ia = 1.0, txy = 0.0, pxy = 0.0, pxy0 = 0.0, kxy = 0.0, kyx = 0.0;
    double gx = 0.0, gy = 0.0, mxy = 0.0;
    if (valid) {
      rx0 = R[3 * x];
      rx1 = R[3 * x + 1];
      rx2 = R[3 * x + 2];
      ry0 = R[3 * y];
      ry1 = R[3 * y + 1];
      ry2 = R[3 * y + 2];
      const double d0 = rx0 - ry0, d1 = rx1 - ry1, d2 = rx2 - ry2;
      a = d0 * d0 + d1 * d1 + d2 * d2;
      ia = 1.0 / a;
      txy = pt[(int64_t)x * N + y];
      pxy0 = pp[(int64_t)x * N + y];
      pxy = s9 * pxy0;  // s9 folded into the pair factor
      mxy = a <= c3sq ? 1.0 : 0.0;
      if (GRAD) {
        kxy = pk[(int64_t)x * N + y];
        kyx = pk[(int64_t)y * N + x];
        gx = in.gE ? in.gE[m * N + x] : 1.0;
        gy = in.gE ? in.gE[m * N + y] : 1.0;
      }
    }
    const double dxy0 = rx0 - ry0, dxy1 = rx1 - ry1, dxy2 = rx2 - ry2;
    const double asq = a * a;
    double rx4 = 0.0, ry4 = 0.0, Axy = 0.0, Bxy = 0.0;
    if (PG && valid) {
      rx4 = w.rt4[m * N + x];
      ry4 = w.rt4[m * N + y];
      const double sig = sq3 * rx4 * ry4;
      Bxy = 1.0 / fma(a1p, sig, a2p);
      Axy = sig * Bxy;
    }
    const double gxy = gx + gy;
    double ex = 0.0, ey = 0.0, cx = 0.0, cy = 0.0, sca = 0.0;
    double fx0 = 0.0, fx1 = 0.0, fx2 = 0.0, fy0 = 0.0, fy1 = 0.0, fy2 = 0.0;

    for (int z0 = by * kTile; z0 < n; z0 += kTile) {
      __syncthreads();
      for (int e = tid; e < kTile * 2 * kTile; e += 256) {
        const int zl = e >> 5, c = e & 31;
        const int z = z0 + zl;
        const int col = c < kTile ? bx * kTile + c : by * kTile + c - kTile;
        ZPair v = {0.0, 0.0, 0.0, 0.0};
        if (z < n && col < n) {
          const int64_t idx = (int64_t)z * N + col;
          v.t = pt[idx];
          v.p = pp[idx];
          if (GRAD) {
            v.kz = pk[idx];
            v.kc = pk[(int64_t)col * N + z];
          }
        }
        double2* dst = reinterpret_cast<double2*>(s_zc + zl * kZcRow + 4 * c);
        dst[0] = make_double2(v.t, v.p);
        dst[1] = make_double2(v.kz, v.kc);
      }
      if (tid < kTile) {
        const int z = z0 + tid;
        const bool ok = z < n && Z[z] != 0;
        double* r = s_r + 6 * tid;
        r[0] = ok ? R[3 * z] : kFar;
        r[1] = ok ? R[3 * z + 1] : kFar;
        r[2] = ok ? R[3 * z + 2] : kFar;
        r[3] = (GRAD && ok) ? (in.gE ? in.gE[m * N + z] : 1.0) : 0.0;
        if (PG) s_rt[tid] = ok ? w.rt4[m * N + z] : 0.0;
      }
      __syncthreads();

      double rE = 0.0, rD = 0.0, rF0 = 0.0, rF1 = 0.0, rF2 = 0.0;  // ring sums
      // FULL: every pair within the cutoff (weights sx = sy = sz = 2, the
      // factor 2 applied when the sums are stored); FIRST: the chunk of y's
      // own block, where third atoms up to y must be masked
      auto steps = [&](auto full_c, auto first_c) {
        constexpr bool FULL = decltype(full_c)::value, FIRST = decltype(first_c)::value;
        D4_UNROLL(D4_ATM1_UNROLL)
        for (int k = 0; k < kTile; ++k) {
          const int zl = (k + tx) & (kTile - 1);
          const double2 rz01 = *reinterpret_cast<const double2*>(s_r + 6 * zl);
          const double2 rz23 = *reinterpret_cast<const double2*>(s_r + 6 * zl + 2);
          const double4 rz = make_double4(rz01.x, rz01.y, rz23.x, rz23.y);
          const double2* px = reinterpret_cast<const double2*>(s_zc + zl * kZcRow + 4 * tx);
          const double2* py =
              reinterpret_cast<const double2*>(s_zc + zl * kZcRow + 4 * (kTile + ty));
          const double2 x01 = px[0], y01 = py[0];
          const double2 x23 = GRAD ? px[1] : make_double2(0.0, 0.0);
          const double2 y23 = GRAD ? py[1] : make_double2(0.0, 0.0);
          const double u0 = rx0 - rz.x, u1 = rx1 - rz.y, u2 = rx2 - rz.z;
          const double v0 = ry0 - rz.x, v1 = ry1 - rz.y, v2 = ry2 - rz.z;
          double bb = fma(u2, u2, fma(u1, u1, u0 * u0));  // r_xz^2
          double cc = fma(v2, v2, fma(v1, v1, v0 * v0));  // r_yz^2
          double P = pxy * x01.y * y01.y;                  // s9 is in pxy
          double P0 = PG ? pxy0 * x01.y * y01.y : 0.0;     // without s9
          if (FIRST) {
            // third atoms up to y carry no weight (z may coincide with x or y)
            const bool on = z0 + zl > y;
            bb = on ? bb : 1.0;
            cc = on ? cc : 1.0;
            P = on ? P : 0.0;
            if (PG) P0 = on ? P0 : 0.0;
          }
          const double T = txy * x01.x * y01.x;
          const double X = a + bb - cc, Y = a - bb + cc, W3 = bb + cc - a;
          // energy alone: s = (a^2 - (c - b)^2) W3 (= X Y W3) and
          // e = P (a b c + 3/8 s) / (a b c (1 + 6 T)) with one Newton step on the
          // reciprocal (f, g, G2 and K below are then unused)
          const double sabc =
              GRAD ? X * Y * W3 : fma(bb - cc, cc - bb, asq) * ((bb + cc) - a);
          const double bc = bb * cc;
          const double abc = a * bc;
          const double opT = fma(6.0, T, 1.0);
          const double q0 = GRAD ? rcp_nt(abc * opT) : rcp_newton1(abc * opT);
          const double f = abc * q0;     // 1 / (1 + 6 T)
          const double iabc = opT * q0;  // 1 / (a b c)
          const double G2 = 0.375 * iabc;
          const double G2s = G2 * sabc;
          const double g = 1.0 + G2s;
          const double K = P * f;
          const double e = GRAD ? K * g : (P * q0) * fma(0.375, sabc, abc);
          double sx, sy, sz;
          if (FULL) {
            sx = sy = sz = 2.0;
          } else {
            const double mxz = bb <= c3sq ? 1.0 : 0.0, myz = cc <= c3sq ? 1.0 : 0.0;
            const double m2 = mxz * myz;
            sx = fma(myz, mxy, m2);
            sy = fma(mxz, mxy, m2);
            sz = mxy * (mxz + myz);
          }
          if (ENERGY) {
            if (FULL) {
              ex += e;  // x and y alike
              rE += e;
            } else {
              ex = fma(sx, e, ex);
              ey = fma(sy, e, ey);
              rE = fma(sz, e, rE);
            }
          }
          if (GRAD) {
            const double W6 = FULL ? 2.0 * (gxy + rz.w) : fma(gx, sx, fma(gy, sy, rz.w * sz));
            const double We = W6 * e;
            const double H = fma(g, fma(alp * T, f, -1.5), -G2s);
            // ds/da, ds/db, ds/dc from A = Y W3, B = X W3, C = X Y
            const double A = Y * W3, Bq = X * W3, C = X * Y;
            const double Ssum = A + Bq + C;
            const double KW = K * W6;  // half of the dL/de-weighted factor; x2 at the end
            const double al = KW * H, be = KW * G2;
            const double ali = al * a * iabc;  // al / (b c)
            const double ca = fma(al, ia, be * fma(-2.0, C, Ssum));
            const double cb = fma(ali, cc, be * fma(-2.0, Bq, Ssum));
            const double cg = fma(ali, bb, be * fma(-2.0, A, Ssum));
            // x gets ca (x - y) + cb (x - z), y gets cg (y - z) - ca (x - y):
            // the (x - y) part is summed as a scalar
            sca += ca;
            fx0 = fma(cb, u0, fx0);
            fx1 = fma(cb, u1, fx1);
            fx2 = fma(cb, u2, fx2);
            fy0 = fma(cg, v0, fy0);
            fy1 = fma(cg, v1, fy1);
            fy2 = fma(cg, v2, fy2);
            rF0 = fma(-cb, u0, fma(-cg, v0, rF0));
            rF1 = fma(-cb, u1, fma(-cg, v1, rF1));
            rF2 = fma(-cb, u2, fma(-cg, v2, rF2));
            cx = fma(We, kxy + x23.y, cx);
            cy = fma(We, kyx + y23.y, cy);
            rD = fma(We, x23.x + y23.x, rD);
            if (PG) {
              const double TdT = -6.0 * f * We * T;  // dL6/dT times T
              g_s9 = fma(W6, P0 * f * g, g_s9);
              g_alp = fma(TdT, T > 0.0 ? log(T) / alp : 0.0, g_alp);
              const double sxz = sq3 * rx4 * s_rt[zl], syz = sq3 * ry4 * s_rt[zl];
              const double bxz = rcp_nt(fma(a1p, sxz, a2p)), byz = rcp_nt(fma(a1p, syz, a2p));
              g_a1 = fma(TdT * alp / 3.0, Axy + sxz * bxz + syz * byz, g_a1);
              g_a2 = fma(TdT * alp / 3.0, Bxy + bxz + byz, g_a2);
            }
          }
          if (ENERGY) rE = row_rotate_down(rE);
          if (GRAD) {
            rF0 = row_rotate_down(rF0);
            rF1 = row_rotate_down(rF1);
            rF2 = row_rotate_down(rF2);
            rD = row_rotate_down(rD);
          }
        }
      };
      const bool first = z0 == by * kTile;
      if (full) {
        if (first)
          steps(std::true_type{}, std::true_type{});
        else
          steps(std::true_type{}, std::false_type{});
      } else {
        if (first)
          steps(std::false_type{}, std::true_type{});
        else
          steps(std::false_type{}, std::false_type{});
      }
      if (full) rE *= 2.0;
      // lane tx of every row now holds that row's sums for z0 + tx (the
      // gradient parts were accumulated at half weight)
      s_red[ty][tx][0] = rE;
      s_red[ty][tx][1] = 2.0 * rF0;
      s_red[ty][tx][2] = 2.0 * rF1;
      s_red[ty][tx][3] = 2.0 * rF2;
      s_red[ty][tx][4] = rD;
      __syncthreads();
      if (tid < kTile * NA) {
        const int zl = tid & (kTile - 1), c = tid >> 4;
        if ((c == 0 ? ENERGY : GRAD) && z0 + zl < n) {
          double sum = 0.0;
          for (int r = 0; r < kTile; ++r) sum += s_red[r][zl][c];
          s_acc[z0 + zl][c] += sum;
        }
      }
    }
    // tile sums of the x atoms (over the rows) and the y atoms (over the columns)
    if (full) {
      ex *= 2.0;
      ey = ex;
    }
    fx0 = fma(sca, dxy0, fx0);
    fx1 = fma(sca, dxy1, fx1);
    fx2 = fma(sca, dxy2, fx2);
    fy0 = fma(-sca, dxy0, fy0);
    fy1 = fma(-sca, dxy1, fy1);
    fy2 = fma(-sca, dxy2, fy2);
    __syncthreads();
    s_red[ty][tx][0] = ex;
    s_red[ty][tx][1] = 2.0 * fx0;
    s_red[ty][tx][2] = 2.0 * fx1;
    s_red[ty][tx][3] = 2.0 * fx2;
    s_red[ty][tx][4] = cx;
    __syncthreads();
    if (tid < kTile * NA) {
      const int col = tid & (kTile - 1), c = tid >> 4;
      if ((c == 0 ? ENERGY : GRAD) && bx * kTile + col < n) {
        double sum = 0.0;
        for (int r = 0; r < kTile; ++r) sum += s_red[r][col][c];
        s_acc[bx * kTile + col][c] += sum;
      }
    }
    __syncthreads();
    s_red[ty][tx][0] = ey;
    s_red[ty][tx][1] = 2.0 * fy0;
    s_red[ty][tx][2] = 2.0 * fy1;
    s_red[ty][tx][3] = 2.0 * fy2;
    s_red[ty][tx][4] = cy;
    __syncthreads();
    if (tid < kTile * NA) {
      const int row = tid & (kTile - 1), c = tid >> 4;
      if ((c == 0 ? ENERGY : GRAD) && by * kTile + row < n) {
        double sum = 0.0;
        for (int q = 0; q < kTile; ++q) sum += s_red[row][q][c];
        s_acc[by * kTile + row][c] += sum;
      }
    }
  }
  __syncthreads();
  double* part = w.apart + (m * (int64_t)S + s) * NA * N;
  for (int e = tid; e < NA * n; e += 256) {
    const int c = e / n, i = e - c * n;
    part[c * N + i] = s_acc[i][c];
  }
  if (PG) {
    const double sixth = 1.0 / 6.0;
    const double v4[4] = {g_s9 * sixth, g_alp * sixth, g_a1 * sixth, g_a2 * sixth};
    double* red = &s_red[0][0][0];
    __syncthreads();
    for (int k = 0; k < 4; ++k) {
      const double v = wave_sum(v4[k]);
      if ((tid & 63) == 0) red[k * kWaves + (tid >> 6)] = v;
    }
    __syncthreads();
    if (tid < 4) {
      double v = 0.0;
      for (int q = 0; q < kWaves; ++q) v += red[tid * kWaves + q];
      w.pg3[(m * (int64_t)S + s) * 4 + tid] = v;
    }
  }
}

// --------------------------------------------------------------------------
// One-pass three-body kernel with 8 x 8 pair tiles, one wave per work-group
// (molecules of up to kAtm1MaxN atoms). Same per-triple arithmetic and
// partition as k_atm1; the smaller tiles waste fewer slots on the diagonal
// tiles, on third atoms z <= y and past the end of the molecule.
//
// Every triple of the 8-atom blocks of molecule m is counted once, by one of
// two kinds of tiles, and work-group (m, s) takes every S-th tile (S =
// atm8_splits(n)):
//  - W-tile t (the first ceil(nb / 2) tiles): the 28 pairs u < v inside
//    block 2 t (lanes 0..27) and inside block 2 t + 1 (lanes 32..59) against
//    the third atoms of every block; in the pair's own block only z > v
//    counts. This takes the triples with two or three atoms in one block.
//  - pair tile (bx, by), bx < by <= nb - 2: lane (tx, ty) = (lane & 7,
//    lane >> 3) owns x = 8 bx + tx, y = 8 by + ty and takes the third atoms
//    of the blocks above by (three distinct blocks).
// A triple's energy and gradient are symmetric in which two atoms form the
// pair, so the roles follow the tile. Compared with tiles bx <= by over the
// third atoms z > y, this drops the half-empty diagonal tiles and the
// masked first chunks: 19% fewer lane-slots on a 1024 x 20-120 batch.
// The third atoms come in chunks of 8. Inside a chunk lane tx of a row (lanes
// 8 r .. 8 r + 7) handles z = z0 + (tx ^ k) at step k, and the running sums
// for z move to the lane that handles z at the
// next step: lane L takes the sums of lane L ^ (k ^ (k + 1)), a swap inside
// each quad (xor 1 or 3) or a mirror of the 8-lane half row (xor 7), single
// DPP moves; one more xor 7 after the last step (so the moves repeat every
// four steps) leaves lane tx with its row's sums for z0 + tx.
// Every sum is added in a fixed order (no atomics).
// --------------------------------------------------------------------------

// Lane L of each 8-lane group receives the value of lane L ^ d (d = 1, 3, 7).
template <int D>
__device__ __forceinline__ double xor_lanes8(double v) {
  return dpp_double<D == 1 ? kDppXor1 : (D == 3 ? kDppXor3 : kDppHalfMirror)>(v);
}

// Running sums for the third atoms of a chunk, handed around the 8-lane ring.
struct Atm8Ring {
  double E, D, F0, F1, F2;  // energy share, dL/dCN, gradient (half weight)
};

template <int D>
__device__ __forceinline__ void ring_xor8(Atm8Ring& r, bool energy, bool grad) {
  if (energy) r.E = xor_lanes8<D>(r.E);
  if (grad) {
    r.D = xor_lanes8<D>(r.D);
    r.F0 = xor_lanes8<D>(r.F0);
    r.F1 = xor_lanes8<D>(r.F1);
    r.F2 = xor_lanes8<D>(r.F2);
  }
}


constexpr double kAtmFar = 1.0e6;  // position of lanes and third atoms without an atom (bohr)
// Staged chunk: row zl (1024 bytes) of third atom z0 + zl holds 32-byte
// records: c = 0..15 the pair records (t, p, kz, kc) against the x block
// (c < 8) and the y block, c = 16 the third atom's position and dL/dE, c = 17
// its sqrt(r4r2) (parameter gradients). Record c of row zl sits at byte
// (zl << 10) | ((c << 5) ^ (zl << 4)): the row bits and the record bits do
// not overlap, so with zl = tx ^ k the address is a lane constant xor a step
// constant (one v_xor per read, no multiply), and the (zl << 4) skew puts
// the eight lanes of a row on distinct 16-byte bank groups.
constexpr int kAtm8Row = 128;  // doubles per row
constexpr int kAtm8RecZ = 16, kAtm8RecRt = 17;
__device__ __forceinline__ uint32_t atm8_rec(int zl, int c) {
  return ((uint32_t)zl << 10) | (((uint32_t)c << 5) ^ ((uint32_t)zl << 4));
}
template <typename T>
__device__ __forceinline__ const T* atm8_at(const double* s_zc, uint32_t off) {
  return reinterpret_cast<const T*>(reinterpret_cast<const char*>(s_zc) + off);
}

// The pair (x, y) a lane owns. Lanes without a pair sit far from every atom
// (no zero distance) with zero pair factors.
struct Atm8Pair {
  double rx0, rx1, rx2, ry0, ry1, ry2;  // positions
  double a, ia, a2, asq;                // r_xy^2, 1 / r_xy^2, 2 r_xy^2, r_xy^4
  double txy, pxy, pxy0;                // t_xy, s9 p_xy, p_xy
  double kxy, kyx;                      // d log sqrt|C6(0)_xy| / d CN_x, / d CN_y
  double gx, gy, gxy;                   // dL/dE of x, y and their sum
  double mxy;                           // 1 if r_xy is within the cutoff
  double rx4, ry4, Axy, Bxy;            // parameter gradients
};

template <bool GRAD, bool PG>
__device__ __forceinline__ Atm8Pair atm8_pair(const Params& p, const Inputs& in, const Work& w,
                                              int64_t m, int x, int y, bool valid) {
  const int64_t N = w.N;
  const double* R = in.positions + m * N * 3;
  Atm8Pair q;
  q.rx0 = q.rx1 = q.rx2 = q.ry0 = q.ry1 = q.ry2 = -kAtmFar;
  q.a = q.ia = q.asq = 1.0;
  q.a2 = 2.0;
  q.txy = q.pxy = q.pxy0 = q.kxy = q.kyx = 0.0;
  q.gx = q.gy = q.mxy = 0.0;
  q.rx4 = q.ry4 = q.Axy = q.Bxy = 0.0;
  if (valid) {
    const double* pt = w.pt + m * N * N;
    const double* pp = w.pp + m * N * N;
    const double* pk = w.pk + m * N * N;
    q.rx0 = R[3 * x];
    q.rx1 = R[3 * x + 1];
    q.rx2 = R[3 * x + 2];
    q.ry0 = R[3 * y];
    q.ry1 = R[3 * y + 1];
    q.ry2 = R[3 * y + 2];
    const double d0 = q.rx0 - q.ry0, d1 = q.rx1 - q.ry1, d2 = q.rx2 - q.ry2;
    q.a = d0 * d0 + d1 * d1 + d2 * d2;
    q.ia = 1.0 / q.a;
    q.a2 = 2.0 * q.a;
    q.asq = q.a * q.a;
    q.txy = pt[(int64_t)x * N + y];
    q.pxy0 = pp[(int64_t)x * N + y];
    q.pxy = p.s9 * q.pxy0;  // s9 folded into the pair factor
    q.mxy = q.a <= p.cut3sq ? 1.0 : 0.0;
    if (GRAD) {
      q.kxy = pk[(int64_t)x * N + y];
      q.kyx = pk[(int64_t)y * N + x];
      q.gx = in.gE ? in.gE[m * N + x] : 1.0;
      q.gy = in.gE ? in.gE[m * N + y] : 1.0;
    }
    if (PG) {
      q.rx4 = w.rt4[m * N + x];
      q.ry4 = w.rt4[m * N + y];
      const double sig = 1.7320508075688772 * q.rx4 * q.ry4;
      q.Bxy = 1.0 / fma(p.a1, sig, p.a2);
      q.Axy = sig * q.Bxy;
    }
  }
  q.gxy = q.gx + q.gy;
  return q;
}

// Sums a lane keeps over its chunks: shares of x and y, their gradient parts
// (the (x - y) part as the scalar sca) and the parameter gradients (x 6).
struct Atm8Sums {
  double ex, ey, cx, cy, sca, fx0, fx1, fx2, fy0, fy1, fy2;
  double sal, sw;  // per tile: sums of al (x 1 / r_xy^2 into sca) and of We (x kxy, kyx)
  double g_s9, g_alp, g_a1, g_a2;
};

// Records of one chunk (8 third atoms x 16 columns, two per lane) and the
// third-atom data (lanes 0..7), held in registers until staged.
struct Atm8Chunk {
  double t[2], p[2], kz[2], kc[2], r[5];
};

// Load the chunk of third atoms z0 .. z0 + 7 against the x block starting at
// atom x0 and the y block starting at y0.
template <bool GRAD, bool PG>
__device__ __forceinline__ void atm8_fetch(Atm8Chunk& c, const Inputs& in, const Work& w,
                                           int64_t m, int n, int z0, int x0, int y0, int lane) {
  constexpr int T8 = kAtm8;
  const int64_t N = w.N;
  const double* pt = w.pt + m * N * N;
  const double* pp = w.pp + m * N * N;
  const double* pk = w.pk + m * N * N;
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const int e = lane + 64 * r;
    const int zl = e >> 4, cl = e & 15;
    const int z = z0 + zl;
    const int col = cl < T8 ? x0 + cl : y0 + cl - T8;
    c.t[r] = c.p[r] = c.kz[r] = c.kc[r] = 0.0;
    if (z < n && col < n) {
      const int64_t idx = (int64_t)z * N + col;
      c.t[r] = pt[idx];
      c.p[r] = pp[idx];
      if (GRAD) {
        c.kz[r] = pk[idx];
        c.kc[r] = pk[(int64_t)col * N + z];
      }
    }
  }
  if (lane < T8) {
    const int z = z0 + lane;
    const bool ok = z < n && in.numbers[m * N + z] != 0;
    const double* R = in.positions + m * N * 3;
    c.r[0] = ok ? R[3 * z] : kAtmFar;
    c.r[1] = ok ? R[3 * z + 1] : kAtmFar;
    c.r[2] = ok ? R[3 * z + 2] : kAtmFar;
    c.r[3] = (GRAD && ok) ? (in.gE ? in.gE[m * N + z] : 1.0) : 0.0;
    c.r[4] = (PG && ok) ? w.rt4[m * N + z] : 0.0;
  }
}

template <bool PG>
__device__ __forceinline__ void atm8_stage(const Atm8Chunk& c, double* s_zc, int lane) {
  char* base = reinterpret_cast<char*>(s_zc);
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const int e = lane + 64 * r;
    const int zl = e >> 4, cl = e & 15;
    double2* dst = reinterpret_cast<double2*>(base + atm8_rec(zl, cl));
    dst[0] = make_double2(c.t[r], c.p[r]);
    dst[1] = make_double2(c.kz[r], c.kc[r]);
  }
  if (lane < kAtm8) {
    double2* r = reinterpret_cast<double2*>(base + atm8_rec(lane, kAtm8RecZ));
    r[0] = make_double2(c.r[0], c.r[1]);
    r[1] = make_double2(c.r[2], c.r[3]);
    if (PG) *reinterpret_cast<double*>(base + atm8_rec(lane, kAtm8RecRt)) = c.r[4];
  }
}

// The 8 steps of one staged chunk (third atoms z0 .. z0 + 7) for the pair of
// lane (tx, ty). FULL: every pair within the cutoff (weights sx = sy = sz =
// 2; the factor 2 of the energies is applied by the caller); FIRST: the chunk
// of y's own block, where third atoms up to y carry no weight. On return lane
// tx of every row holds that row's ring sums for z0 + tx.
template <bool ENERGY, bool GRAD, bool PG, bool FULL, bool FIRST>
__device__ __forceinline__ void atm8_chunk(const Params& p, const Atm8Pair& q, Atm8Sums& u,
                                           Atm8Ring& rr, const double* s_zc, int tx, int ty,
                                           int z0, int y, int cx = -1, int cy = -1,
                                           bool own = true) {
  constexpr int T8 = kAtm8;
  // lane parts of the record addresses (the step part is (k << 10) | (k << 4))
  const uint32_t lrow = ((uint32_t)tx << 10) | ((uint32_t)tx << 4);
  const uint32_t ox = lrow ^ ((uint32_t)(cx < 0 ? tx : cx) << 5);
  const uint32_t oy = lrow ^ ((uint32_t)(cy < 0 ? T8 + ty : cy) << 5);
  const uint32_t oz = lrow ^ ((uint32_t)kAtm8RecZ << 5);
  const uint32_t ort = lrow ^ ((uint32_t)kAtm8RecRt << 5);
  const double c3sq = p.cut3sq, alp = p.alp, a1p = p.a1, a2p = p.a2, sq3 = 1.7320508075688772;
  // alp T f = alp / 6 (1 - f), f = 1 / (1 + 6 T): H below needs no alp T
  const double alp6 = alp / 6.0, h0 = alp6 - 2.5;
  // four steps per iteration (two with parameter gradients, whose extra
  // sums would not fit the registers), so every ring move below is static
#ifndef D4_ATM8_UNROLL
#define D4_ATM8_UNROLL 4
#endif
  constexpr int kUnroll = PG ? 2 : D4_ATM8_UNROLL;
  D4_UNROLL(kUnroll)
  for (int k = 0; k < T8; ++k) {
    const uint32_t ks = ((uint32_t)k << 10) | ((uint32_t)k << 4);
    const double2* pz = atm8_at<double2>(s_zc, oz ^ ks);
    const double2 rz01 = pz[0], rz23 = pz[1];
    const double4 rz = make_double4(rz01.x, rz01.y, rz23.x, rz23.y);
    // records of x and y against z = z0 + (tx ^ k)
    const double2* px = atm8_at<double2>(s_zc, ox ^ ks);
    const double2* py = atm8_at<double2>(s_zc, oy ^ ks);
    const double2 x01 = px[0], y01 = py[0];
    const double2 x23 = GRAD ? px[1] : make_double2(0.0, 0.0);
    const double2 y23 = GRAD ? py[1] : make_double2(0.0, 0.0);
    const double u0 = q.rx0 - rz.x, u1 = q.rx1 - rz.y, u2 = q.rx2 - rz.z;
    const double v0 = q.ry0 - rz.x, v1 = q.ry1 - rz.y, v2 = q.ry2 - rz.z;
    double bb = fma(u2, u2, fma(u1, u1, u0 * u0));  // r_xz^2
    double cc = fma(v2, v2, fma(v1, v1, v0 * v0));  // r_yz^2
    double P = q.pxy * x01.y * y01.y;                // s9 is in pxy
    double P0 = PG ? q.pxy0 * x01.y * y01.y : 0.0;   // without s9
    if (FIRST) {
      // third atoms up to y carry no weight (z may coincide with x or y);
      // `own`: the chunk is y's block for this lane
      const bool on = !own || z0 + (tx ^ k) > y;
      bb = on ? bb : 1.0;
      cc = on ? cc : 1.0;
      P = on ? P : 0.0;
      if (PG) P0 = on ? P0 : 0.0;
    }
    const double a = q.a;
    const double T = q.txy * x01.x * y01.x;
    // s = X Y W3 with X = a + b - c, Y = a - b + c, W3 = b + c - a (b = r_xz^2,
    // c = r_yz^2); X + Y = 2 a and Y - X = 2 (c - b) give its derivatives
    const double dcb = cc - bb;
    const double X = a - dcb, Y = a + dcb, W3 = (bb + cc) - a;
    const double C = GRAD ? X * Y : fma(-dcb, dcb, q.asq);  // a^2 - (c - b)^2 = X Y
    const double sabc = C * W3;
    const double bc = bb * cc;
    const double abc = a * bc;
    const double opT = fma(6.0, T, 1.0);
    const double q0 = rcp_newton1(abc * opT);
    const double f = abc * q0;     // 1 / (1 + 6 T)
    const double iabc = opT * q0;  // 1 / (a b c)
    const double G2 = 0.375 * iabc;
    const double g = fma(G2, sabc, 1.0);
    const double K = P * f;
    // energy alone: e = P (a b c + 3/8 s) / (a b c (1 + 6 T)), four operations
    // after the reciprocal instead of seven (f, g, G2 and K are then unused)
    const double e = GRAD ? K * g : (P * q0) * fma(0.375, sabc, abc);
    double sx, sy, sz;
    if (FULL) {
      sx = sy = sz = 2.0;
    } else {
      const double mxz = bb <= c3sq ? 1.0 : 0.0, myz = cc <= c3sq ? 1.0 : 0.0;
      const double m2 = mxz * myz;
      sx = fma(myz, q.mxy, m2);
      sy = fma(mxz, q.mxy, m2);
      sz = q.mxy * (mxz + myz);
    }
    if (ENERGY) {
      if (FULL) {
        u.ex += e;  // x and y alike
        rr.E += e;
      } else {
        u.ex = fma(sx, e, u.ex);
        u.ey = fma(sy, e, u.ey);
        rr.E = fma(sz, e, rr.E);
      }
    }
    if (GRAD) {
      // the fused energy + gradient call has dL/dE = 1 for every atom (its
      // triples without atoms have e = 0, whatever the weight)
      constexpr bool kUnit = ENERGY && GRAD && !PG;
      // with unit weights and every pair within the cutoff the weight is the
      // constant 6: the caller applies it to the staged gradient sums
      constexpr bool kSix = kUnit && FULL;
      const double W6 = kSix ? 1.0
                             : kUnit ? sx + sy + sz
                                     : (FULL ? 2.0 * (q.gxy + rz.w)
                                             : fma(q.gx, sx, fma(q.gy, sy, rz.w * sz)));
      const double We = W6 * e;
      // H = g (alp T f - 3/2) - (g - 1)
      const double H = fma(g, fma(-alp6, f, h0), 1.0);
      // ds/da = 2 a W3 - X Y, ds/db = X Y + 2 (c - b) W3, ds/dc = X Y - 2 (c - b) W3
      const double dW = dcb * W3;
      const double dsa = fma(q.a2, W3, -C), dsb = fma(2.0, dW, C), dsc = fma(-2.0, dW, C);
      const double KW = K * W6;  // half of the dL/de-weighted factor; x2 at the end
      const double al = KW * H, be = KW * G2;
      const double ali = al * a * iabc;  // al / (b c)
      const double cb = fma(ali, cc, be * dsb);
      const double cg = fma(ali, bb, be * dsc);
      // x gets ca (x - y) + cb (x - z), y gets cg (y - z) - ca (x - y):
      // the (x - y) part is summed as a scalar
      u.sal += al;  // ca = al / a + be dsa, the 1 / a applied per tile
      u.sca = fma(be, dsa, u.sca);
      u.fx0 = fma(cb, u0, u.fx0);
      u.fx1 = fma(cb, u1, u.fx1);
      u.fx2 = fma(cb, u2, u.fx2);
      u.fy0 = fma(cg, v0, u.fy0);
      u.fy1 = fma(cg, v1, u.fy1);
      u.fy2 = fma(cg, v2, u.fy2);
      rr.F0 = fma(-cb, u0, fma(-cg, v0, rr.F0));
      rr.F1 = fma(-cb, u1, fma(-cg, v1, rr.F1));
      rr.F2 = fma(-cb, u2, fma(-cg, v2, rr.F2));
      u.sw += We;  // times kxy and kyx per tile
      u.cx = fma(We, x23.y, u.cx);
      u.cy = fma(We, y23.y, u.cy);
      rr.D = fma(We, x23.x + y23.x, rr.D);
      if (PG) {
        const double TdT = -6.0 * f * We * T;  // dL6/dT times T
        u.g_s9 = fma(W6, P0 * f * g, u.g_s9);
        u.g_alp = fma(TdT, T > 0.0 ? log(T) / alp : 0.0, u.g_alp);
        const double rtz = *atm8_at<double>(s_zc, ort ^ ks);
        const double sxz = sq3 * q.rx4 * rtz, syz = sq3 * q.ry4 * rtz;
        const double bxz = rcp_nt(fma(a1p, sxz, a2p)), byz = rcp_nt(fma(a1p, syz, a2p));
        u.g_a1 = fma(TdT * alp / 3.0, q.Axy + sxz * bxz + syz * byz, u.g_a1);
        u.g_a2 = fma(TdT * alp / 3.0, q.Bxy + bxz + byz, u.g_a2);
      }
    }
    // hand the sums for z to the lane that takes z at the next step:
    // k ^ (k + 1) is 1, 3, 1, 7, 1, 3, 1, and one more xor 7 after the last
    // step (the same four moves in every group of four steps)
    if ((k & 1) == 0)
      ring_xor8<1>(rr, ENERGY, GRAD);
    else if ((k & 3) == 1)
      ring_xor8<3>(rr, ENERGY, GRAD);
    else
      ring_xor8<7>(rr, ENERGY, GRAD);
  }
}

// Fold the (x - y) force scalar of a tile into the x and y gradient sums.
__device__ __forceinline__ void atm8_fold_sca(const Atm8Pair& q, Atm8Sums& u) {
  const double d0 = q.rx0 - q.ry0, d1 = q.rx1 - q.ry1, d2 = q.rx2 - q.ry2;
  const double sca = fma(u.sal, q.ia, u.sca);
  u.fx0 = fma(sca, d0, u.fx0);
  u.fx1 = fma(sca, d1, u.fx1);
  u.fx2 = fma(sca, d2, u.fx2);
  u.fy0 = fma(-sca, d0, u.fy0);
  u.fy1 = fma(-sca, d1, u.fy1);
  u.fy2 = fma(-sca, d2, u.fy2);
  u.cx = fma(q.kxy, u.sw, u.cx);
  u.cy = fma(q.kyx, u.sw, u.cy);
  u.sca = u.sal = u.sw = 0.0;
}

// Stage five per-lane values (energy share, gradient at half weight, dL/dCN)
// in s_red for the sums over rows or columns of the tile; gs scales the
// gradient parts (6 where the chunks left out the constant weight 6, else 1).
template <int NA>
__device__ __forceinline__ void atm8_stage_sums(double (*s_red)[64], int lane, double e,
                                                double f0, double f1, double f2, double c,
                                                double gs = 1.0) {
  const double g2 = 2.0 * gs;
  s_red[0][lane] = e;
  s_red[1][lane] = g2 * f0;
  s_red[2][lane] = g2 * f1;
  s_red[3][lane] = g2 * f2;
  s_red[4][lane] = gs * c;
}

// Parameter-gradient partials of a wave: s9, alp, a1, a2 (the sums carry a
// factor 6 from the screening weights).
__device__ __forceinline__ void atm8_store_pgrad(const Atm8Sums& u, int lane, double* out) {
  const double sixth = 1.0 / 6.0;
  const double v4[4] = {u.g_s9 * sixth, u.g_alp * sixth, u.g_a1 * sixth, u.g_a2 * sixth};
  for (int k = 0; k < 4; ++k) {
    const double v = wave_sum(v4[k]);
    if (lane == 0) out[k] = v;
  }
}

template <bool ENERGY, bool GRAD, bool PG>
#ifndef D4_ATM8_WAVES
#define D4_ATM8_WAVES 2
#endif
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(D4_ATM8_WAVES))) void k_atm8(
    Params p, Inputs in, Work w, int S, int64_t nblocks) {
  constexpr int NA = kAtm1Acc;
  constexpr int T8 = kAtm8;
  __shared__ __align__(16) double s_zc[T8 * kAtm8Row];
  __shared__ double s_red[NA][64];
  __shared__ double s_red2[NA][64];
  extern __shared__ double s_acc[];             // [NA][N] atom sums of this split

  // block b runs on XCD b mod 8. Molecules go to the XCDs in runs of eight
  // (molecules 8c .. 8c + 7 to XCD c mod 8), all splits of a molecule in
  // consecutive blocks of its XCD: neighbouring molecules share an XCD's L2,
  // and a batch sorted by size still spreads its large molecules over all
  // XCDs (with whole eighths of the batch per XCD, a sorted 1024 x 20-120
  // batch took 1.86 ms per step instead of 1.10)
  const int64_t b = blockIdx.x, j = b >> 3, l = j / S;
  const int64_t m = (((l >> 3) << 3) + (b & 7)) * 8 + (l & 7);
  if (m >= nblocks) return;  // nblocks: the number of molecules
  const int s = (int)(j - l * S);
  const int64_t N = w.N;
  const int n = w.natom[m];
  const int Sm = atm8_splits(n);  // this molecule's split count (S is the batch maximum)
  if (s >= Sm) return;
  const int nb = (n + T8 - 1) / T8;
  const int nw = (nb + 1) / 2;  // W-tiles, then the pair tiles bx < by <= nb - 2
  const int tiles = nw + (nb > 1 ? (nb - 1) * (nb - 2) / 2 : 0);
  const int lane = threadIdx.x, tx = lane & (T8 - 1), ty = lane >> 3;
  const int64_t* Z = in.numbers + m * N;
  const double c3sq = p.cut3sq;
  for (int e = lane; e < NA * n; e += 64) s_acc[e] = 0.0;
  Atm8Sums u = {};
  // molecules whose bounding box fits in the cutoff sphere have every pair
  // within the cutoff: all screening weights are 2 and their masks drop out
  // (the squared box diagonal comes from the coordination-number pass)
  const bool full = w.bbox2[m] <= c3sq;
  // the weight 6 of the unit-weight chunks of such molecules (see atm8_chunk)
  const double gs = (ENERGY && GRAD && !PG && full) ? 6.0 : 1.0;

  // W-tile lane roles: lane (h, p) = (lane >> 5, lane & 31) owns the pair
  // p = v (v - 1) / 2 + u (u < v) of block 2 t + h; lanes p >= 28 are idle
  // (they read the records of the pair (0, 1), with zero pair factors)
  const int wh = lane >> 5, wp = lane & 31;
  const bool wok = wp < 28;
  int wv = 1;
  while (wok && (wv + 1) * wv / 2 <= wp) ++wv;
  const int wu = wok ? wp - wv * (wv - 1) / 2 : 0;

  for (int t = s; t < tiles; t += Sm) {
    const bool wt = t < nw;
    int bx, by, x, y, cx, cy, z1;  // staged blocks, pair, its record columns, first chunk
    if (wt) {
      bx = 2 * t;
      by = bx + 1;
      x = T8 * (bx + wh) + wu;
      y = T8 * (bx + wh) + wv;
      cx = T8 * wh + wu;
      cy = T8 * wh + wv;
      z1 = 0;
    } else {
      const int t2 = t - nw;  // = by (by - 1) / 2 + bx
      by = 1;
      while ((by + 1) * by / 2 <= t2) ++by;
      bx = t2 - by * (by - 1) / 2;
      x = bx * T8 + tx;
      y = by * T8 + ty;
      cx = tx;
      cy = T8 + ty;
      z1 = (by + 1) * T8;
    }
    const bool valid = (wok || !wt) && x < y && y < n && Z[x] != 0 && Z[y] != 0;
    const Atm8Pair q = atm8_pair<GRAD, PG>(p, in, w, m, x, y, valid);
    u.ex = u.ey = u.cx = u.cy = u.sca = 0.0;
    u.fx0 = u.fx1 = u.fx2 = u.fy0 = u.fy1 = u.fy2 = 0.0;

    Atm8Chunk ch;
    atm8_fetch<GRAD, PG>(ch, in, w, m, n, z1, bx * T8, by * T8, lane);
    for (int z0 = z1; z0 < n; z0 += T8) {
      __syncthreads();
      atm8_stage<PG>(ch, s_zc, lane);
      __syncthreads();
      if (z0 + T8 < n)  // in flight during this chunk
        atm8_fetch<GRAD, PG>(ch, in, w, m, n, z0 + T8, bx * T8, by * T8, lane);

      Atm8Ring rr = {0.0, 0.0, 0.0, 0.0, 0.0};
      // the two chunks of a W-tile's own blocks: third atoms up to y of the
      // lane's own block carry no weight
      const bool first = wt && (z0 >> 4) == t;
      const bool own = z0 == T8 * (bx + wh);
      if (full) {
        if (first)
          atm8_chunk<ENERGY, GRAD, PG, true, true>(p, q, u, rr, s_zc, tx, ty, z0, y,
                                                   cx, cy, own);
        else
          atm8_chunk<ENERGY, GRAD, PG, true, false>(p, q, u, rr, s_zc, tx, ty, z0, y,
                                                    cx, cy);
      } else {
        if (first)
          atm8_chunk<ENERGY, GRAD, PG, false, true>(p, q, u, rr, s_zc, tx, ty, z0, y,
                                                    cx, cy, own);
        else
          atm8_chunk<ENERGY, GRAD, PG, false, false>(p, q, u, rr, s_zc, tx, ty, z0,
                                                     y, cx, cy);
      }
      if (full) rr.E *= 2.0;
      // lane tx of every row now holds that row's sums for z0 + tx (the
      // gradient parts at half weight); add the 8 rows in order
      atm8_stage_sums<NA>(s_red, lane, rr.E, rr.F0, rr.F1, rr.F2, rr.D, gs);
      __syncthreads();
      if (lane < T8 * NA) {
        const int l = lane & (T8 - 1), c = lane >> 3;  // column l holds z0 + l
        const int z = z0 + l;
        if ((c == 0 ? ENERGY : GRAD) && z < n) {
          double sum = 0.0;
          for (int r = 0; r < T8; ++r) sum += s_red[c][r * T8 + l];
          s_acc[c * n + z] += sum;
        }
      }
    }
    // tile sums of the x and the y atoms
    if (full) {
      u.ex *= 2.0;
      u.ey = u.ex;
    }
    atm8_fold_sca(q, u);
    __syncthreads();
    atm8_stage_sums<NA>(s_red, lane, u.ex, u.fx0, u.fx1, u.fx2, u.cx, gs);
    atm8_stage_sums<NA>(s_red2, lane, u.ey, u.fy0, u.fy1, u.fy2, u.cy, gs);
    __syncthreads();
    if (wt) {
      // atom a of block 2 t + hh: x of the pairs (a, v > a), y of (u < a, a)
      for (int e = lane; e < 2 * T8 * NA; e += 64) {
        const int c = e >> 4, a16 = e & 15, hh = a16 >> 3, a = a16 & 7;
        const int atom = 2 * T8 * t + a16;
        if ((c == 0 ? ENERGY : GRAD) && atom < n) {
          const double* sx = s_red[c] + 32 * hh;
          const double* sy = s_red2[c] + 32 * hh + a * (a - 1) / 2;
          double sum = 0.0;
          for (int v = a + 1; v < T8; ++v) sum += sx[v * (v - 1) / 2 + a];
          for (int u0 = 0; u0 < a; ++u0) sum += sy[u0];
          s_acc[c * n + atom] += sum;
        }
      }
    } else if (lane < T8 * NA) {
      // x atoms over the rows, y atoms over the columns
      const int l = lane & (T8 - 1), c = lane >> 3;
      if ((c == 0 ? ENERGY : GRAD) && bx * T8 + l < n) {
        double sum = 0.0;
        for (int r = 0; r < T8; ++r) sum += s_red[c][r * T8 + l];
        s_acc[c * n + bx * T8 + l] += sum;
      }
      if ((c == 0 ? ENERGY : GRAD) && by * T8 + l < n) {
        double sum = 0.0;
        for (int q2 = 0; q2 < T8; ++q2) sum += s_red2[c][l * T8 + q2];
        s_acc[c * n + by * T8 + l] += sum;
      }
    }
  }
  __syncthreads();
  double* part = w.apart + (m * (int64_t)S + s) * NA * N;
  for (int e = lane; e < NA * n; e += 64) {
    const int c = e / n, i = e - c * n;
    part[c * N + i] = s_acc[e];
  }
  if (PG) atm8_store_pgrad(u, lane, w.pg3 + (m * (int64_t)S + s) * 4);
}

// --------------------------------------------------------------------------
// One-pass three-body kernel for large systems (the atoms in spatial order):
// the 8 x 8 evaluation above with group screening. Work item (m, gx, s)
// owns the x atoms of 8-atom group gx and takes every S-th of the pair tiles
// (gx, gy), gy >= gx, whose groups can hold a counted triple (bounding
// spheres within 2 c3); per tile it streams the third-atom groups gz >= gy
// with at least two of the three group-to-group gaps within c3. Screened-out
// triples contribute exact zeros, so the screening changes no bit of the
// result beyond the order of the sums.
//
// Outputs: each item adds its sums for the atoms of group g (x, y or z) to
// its own slot of g in the window buffer: slots of group g are contiguous,
// one per item (gx <= g, s), in item order, so k_atm8s_rows adds them in a
// fixed order. The number of splits S of x-group gx depends on the
// molecule's atom count alone (more for the first groups, whose forward
// neighbourhood is largest), and heavy items are dispatched first.
// --------------------------------------------------------------------------

// Slot tables of molecule m: [0] = largest split count, [1 + g] = first slot
// of items of x-group g in the list of any later group (g = 0 .. GN), then
// [GN + 2 + g] = first slot of group g in the window buffer (g = 0 .. GN).
__device__ __forceinline__ const int64_t* atm8s_table(const Work& w, int64_t m) {
  return w.wix + m * atm8s_table_len(w.N);
}

// Build the slot tables of every molecule (one thread per molecule: O(G)).
__global__ __launch_bounds__(64) void k_atm8s_index(Work w) {
  const int64_t m = (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (m >= w.B) return;
  const int n = w.natom[m];
  const int G = (n + kAtm8 - 1) / kAtm8;
  const int64_t GN = n_groups8(w.N);
  int64_t* t = w.wix + m * atm8s_table_len(w.N);
  const int smax = atm8s_smax(n);
  t[0] = smax;
  int64_t pre = 0, base = 0;
  for (int g = 0; g <= GN; ++g) {
    t[1 + g] = pre;
    t[GN + 2 + g] = base;
    if (g < G) {
      pre += atm8s_splits(G, g, smax);
      base += pre;  // group g has one slot per item of x-groups 0 .. g
    }
  }
}

// Near bits of the 8-atom groups: bit j of row i of molecule m is set when a
// real atom of group i lies within the three-body cutoff (with a margin for
// rounding) of a real atom of group j. A triple contributes only through an
// atom whose two edges are both within the cutoff, so a triple of atoms in
// groups x, y, z contributes only when two of near(x, y), near(x, z) and
// near(y, z) are set: the exact group test of the screened kernel (the
// bounding spheres keep about a quarter more third-atom groups). One wave
// per (molecule, group i, word of 64 groups j).
__global__ __launch_bounds__(64) void k_atm8s_near(Params p, Inputs in, Work w) {
  const int64_t GN = n_groups8(w.N), WN = near_words8(w.N);
  const int64_t b = blockIdx.x;
  const int64_t m = b / (GN * WN);
  const int64_t r = b - m * GN * WN;
  const int i = (int)(r / WN), word = (int)(r - (int64_t)i * WN);
  const int lane = threadIdx.x;
  const int n = w.natom[m];
  const int64_t* Z = in.numbers + m * w.N;
  const double* R = in.positions + m * w.N * 3;
  const double c3 = sqrt(p.cut3sq) * (1.0 + 1e-10) + 1e-10;
  const double c3sq = c3 * c3;
  double xi[kAtm8], yi[kAtm8], zi[kAtm8];
  bool ri[kAtm8];
#pragma unroll
  for (int a = 0; a < kAtm8; ++a) {
    const int at = kAtm8 * i + a;
    ri[a] = at < n && Z[at] != 0;
    xi[a] = ri[a] ? R[3 * at] : 0.0;
    yi[a] = ri[a] ? R[3 * at + 1] : 0.0;
    zi[a] = ri[a] ? R[3 * at + 2] : 0.0;
  }
  const int j = word * 64 + lane;
  bool near = false;
  if (j < GN) {
    for (int c = 0; c < kAtm8; ++c) {
      const int at = kAtm8 * j + c;
      if (at >= n || Z[at] == 0) continue;
      const double x = R[3 * at], y = R[3 * at + 1], z = R[3 * at + 2];
#pragma unroll
      for (int a = 0; a < kAtm8; ++a) {
        const double dx = x - xi[a], dy = y - yi[a], dz = z - zi[a];
        near |= ri[a] && dx * dx + dy * dy + dz * dz <= c3sq;
      }
    }
  }
  const unsigned long long bits = __ballot(near);
  if (lane == 0) w.gnear[(m * GN + i) * WN + word] = bits;
}

template <bool ENERGY, bool GRAD, bool PG>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void k_atm8s(
    Params p, Inputs in, Work w, int64_t nblocks) {
  constexpr int NA = kAtm1Acc;
  constexpr int T8 = kAtm8;
  constexpr int KS = kAtm8sMaxSplits;
  constexpr int kWin = NA * T8;  // doubles per slot: [NA][8 atoms]
  __shared__ __align__(16) double s_zc[T8 * kAtm8Row];
  __shared__ double s_red[NA][64];

  const int64_t b = blockIdx.x;
  if (b >= nblocks) return;
  const int lane = threadIdx.x, tx = lane & (T8 - 1), ty = lane >> 3;
  const int64_t N = w.N;
  const int64_t per_mol = atm8s_blocks_per_mol(N);
  const int64_t m = b / per_mol;
  // block r of a molecule runs on XCD r % 8: x-group gx stays on XCD gx % 8
  // with all its splits, and the heavy low groups come first
  const int r = (int)(b - m * per_mol);
  const int s = (r >> 3) % KS, gx = ((r >> 3) / KS) * 8 + (r & 7);
  const int n = w.natom[m];
  const int G = (n + T8 - 1) / T8;
  const int64_t* tab = atm8s_table(w, m);
  const int smax = (int)tab[0];
  const int S = gx < G ? atm8s_splits(G, gx, smax) : 0;
  if (s >= S) {
    if (PG && lane < 4) w.pg3[b * 4 + lane] = 0.0;
    return;
  }
  const int64_t GN = n_groups8(N);
  const int64_t slot = tab[1 + gx] + s;  // this item's slot in the list of every group >= gx
  const int64_t* gbase = tab + GN + 2;
  double* win = w.pg + m * atm8s_stride(N);
  auto slot_of = [&](int g) { return win + (gbase[g] + slot) * kWin; };
  for (int g = gx; g < G; ++g)
    if (lane < kWin) slot_of(g)[lane] = 0.0;
  __syncthreads();

  const int64_t* Z = in.numbers + m * N;
  const double* S8 = w.gsph8 + 4 * m * GN;
  const double* SX = S8 + 4 * gx;
  const double c3 = sqrt(p.cut3sq) * (1.0 + 1e-10) + 1e-10;  // margin for rounding in the bounds
  const int64_t WN = near_words8(N);
  const uint64_t* NEAR = w.gnear + m * GN * WN;
  const uint64_t* NX = NEAR + gx * WN;
  const int x = gx * T8 + tx;
  Atm8Sums u = {};
  int kept = 0;  // pair tiles of this x-group seen so far
  for (int ybase = gx; ybase < G; ybase += 64) {
    const int cand = ybase + lane;
    const bool ykeep = cand < G && sphere_gap(SX, S8 + 4 * cand) <= 2.0 * c3;
    unsigned long long ymask = __ballot(ykeep);
    while (ymask) {
      const int gy = ybase + __builtin_ctzll(ymask);
      ymask &= ymask - 1;
      if (kept++ % S != s) continue;
      const uint64_t* NY = NEAR + gy * WN;
      const bool nxy = (NX[gy >> 6] >> (gy & 63)) & 1ull;
      const int y = gy * T8 + ty;
      const bool valid = x < y && y < n && Z[x] != 0 && Z[y] != 0;
      const Atm8Pair q = atm8_pair<GRAD, PG>(p, in, w, m, x, y, valid);
      u.ey = u.cy = u.sca = 0.0;
      u.fy0 = u.fy1 = u.fy2 = 0.0;

      // third-atom groups gz >= gy in ascending order, 64 per word of near
      // bits: two of near(x, y), near(x, z), near(y, z) set (groups past the
      // molecule have no bits)
      int zw = gy >> 6;
      auto zscan = [&]() {
        const unsigned long long a = NX[zw], c = NY[zw];
        unsigned long long k = nxy ? (a | c) : (a & c);
        if (zw == (gy >> 6)) k &= ~0ull << (gy & 63);
        return k;
      };
      unsigned long long zmask = zscan();
      auto next_z = [&]() {
        while (true) {
          if (zmask) {
            const int g = zw * 64 + __builtin_ctzll(zmask);
            zmask &= zmask - 1;
            return g;
          }
          ++zw;
          if (zw * 64 >= G) return G;
          zmask = zscan();
        }
      };
      Atm8Chunk ch;
      int gz = next_z();
      if (gz < G) atm8_fetch<GRAD, PG>(ch, in, w, m, n, gz * T8, gx * T8, gy * T8, lane);
      while (gz < G) {
        __syncthreads();
        atm8_stage<PG>(ch, s_zc, lane);
        __syncthreads();
        const int gzn = next_z();
        if (gzn < G)  // in flight during this chunk
          atm8_fetch<GRAD, PG>(ch, in, w, m, n, gzn * T8, gx * T8, gy * T8, lane);
        // lane c * 8 + l adds the sums of component c for z = 8 gz + l
        double* wz = slot_of(gz);
        const int widx = lane;
        const double old = lane < kWin ? wz[widx] : 0.0;
        Atm8Ring rr = {0.0, 0.0, 0.0, 0.0, 0.0};
        if (gz == gy)
          atm8_chunk<ENERGY, GRAD, PG, false, true>(p, q, u, rr, s_zc, tx, ty,
                                                    gz * T8, y);
        else
          atm8_chunk<ENERGY, GRAD, PG, false, false>(p, q, u, rr, s_zc, tx, ty,
                                                     gz * T8, y);
        atm8_stage_sums<NA>(s_red, lane, rr.E, rr.F0, rr.F1, rr.F2, rr.D);
        __syncthreads();
        if (lane < kWin) {
          const int l = lane & (T8 - 1), c = lane >> 3;
          if (c == 0 ? ENERGY : GRAD) {
            double sum = 0.0;
            for (int r2 = 0; r2 < T8; ++r2) sum += s_red[c][r2 * T8 + l];
            wz[widx] = old + sum;
          }
        }
        gz = gzn;
      }
      // the y atoms' tile sums (over the columns)
      atm8_fold_sca(q, u);
      __syncthreads();
      atm8_stage_sums<NA>(s_red, lane, u.ey, u.fy0, u.fy1, u.fy2, u.cy);
      __syncthreads();
      if (lane < kWin) {
        const int row = lane & (T8 - 1), c = lane >> 3;
        if (c == 0 ? ENERGY : GRAD) {
          double sum = 0.0;
          for (int q2 = 0; q2 < T8; ++q2) sum += s_red[c][row * T8 + q2];
          double* wy = slot_of(gy);
          wy[lane] += sum;
        }
      }
    }
  }
  // the x atoms' sums over all tiles (over the rows)
  __syncthreads();
  atm8_stage_sums<NA>(s_red, lane, u.ex, u.fx0, u.fx1, u.fx2, u.cx);
  __syncthreads();
  if (lane < kWin) {
    const int col = lane & (T8 - 1), c = lane >> 3;
    if (c == 0 ? ENERGY : GRAD) {
      double sum = 0.0;
      for (int r2 = 0; r2 < T8; ++r2) sum += s_red[c][r2 * T8 + col];
      slot_of(gx)[lane] += sum;
    }
  }
  if (PG) atm8_store_pgrad(u, lane, w.pg3 + b * 4);
}

// Atom sums of the screened one-pass kernel: one work-group per 8-atom group
// adds the group's slots in slot order (thread (j, v) takes slots j, j + J,
// ..., then the J partial sums are added in order).
template <bool ENERGY, bool GRAD>
__global__ __launch_bounds__(1024) void k_atm8s_rows(Inputs in, Work w, double* energy_out,
                                                     double* grad_out) {
  constexpr int NA = kAtm1Acc, T8 = kAtm8, kWin = NA * T8, J = 1024 / kWin;
  __shared__ double s_sum[J][kWin];
  const int64_t N = w.N;
  const int64_t GN = n_groups8(N);
  const int64_t m = blockIdx.x / GN;
  const int g = (int)(blockIdx.x - m * GN);
  const int n = w.natom[m];
  const int G = (n + T8 - 1) / T8;
  const int tid = threadIdx.x;
  const int j = tid / kWin, v = tid - j * kWin;
  double acc = 0.0;
  if (g < G && j < J) {
    const int64_t* tab = atm8s_table(w, m);
    const int64_t cnt = tab[1 + g + 1];  // items of x-groups 0 .. g
    const double* src = w.pg + m * atm8s_stride(N) + tab[GN + 2 + g] * kWin + v;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int64_t k = j;
    for (; k + 3 * J < cnt; k += 4 * J) {
      a0 += src[k * kWin];
      a1 += src[(k + J) * kWin];
      a2 += src[(k + 2 * J) * kWin];
      a3 += src[(k + 3 * J) * kWin];
    }
    for (; k < cnt; k += J) a0 += src[k * kWin];
    acc = (a0 + a1) + (a2 + a3);
  }
  if (j < J) s_sum[j][v] = acc;
  __syncthreads();
  if (tid < kWin) {
    double t = 0.0;
    for (int q = 0; q < J; ++q) t += s_sum[q][tid];
    const int c = tid >> 3, l = tid & (T8 - 1);
    const int i = g * T8 + l;
    if (i < N) {
      const int64_t row = m * N + i;
      const bool real = i < n && in.numbers[row] != 0;
      const double sixth = 1.0 / 6.0;
      if (c == 0) {
        if (ENERGY) energy_out[row] = real ? w.e2[row] + t * sixth : 0.0;
      } else if (GRAD && real) {
        if (c <= 3)
          grad_out[3 * row + c - 1] += t * sixth;
        else
          w.dldcn[row] += t * sixth;
      }
    }
  }
}

// Atom sums of the one-pass three-body kernel over its splits: energies
// (with the two-body part), gradient and dL/dCN. Eight lanes per atom: lane
// l adds splits l, l + 8, ..., then the eight partial sums are added in a
// fixed order (DPP), so the sum depends on the molecule's atom count alone.
template <bool ENERGY, bool GRAD>
__global__ __launch_bounds__(256) void k_atm1_rows(Inputs in, Work w, int tile8,
                                                   double* energy_out, double* grad_out) {
  const int64_t tid = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t row = tid >> 3;
  const int l = (int)(tid & 7);
  if (row >= w.B * w.N) return;  // uniform over each group of 8 lanes
  const int64_t N = w.N;
  const int64_t m = row / N;
  const int i = (int)(row - m * N);
  const int n = w.natom[m];
  const int zi = i < n ? (int)in.numbers[row] : 0;
  if (zi == 0) {
    if (ENERGY && l == 0) energy_out[row] = 0.0;
    return;
  }
  const int S = one_pass_splits(n, tile8 != 0);
  const int64_t stride = (int64_t)one_pass_splits((int)N, tile8 != 0) * kAtm1Acc * N;
  const double* part = w.apart + m * stride + i;
  double acc[kAtm1Acc] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int s = l; s < S; s += 8)
#pragma unroll
    for (int c = 0; c < kAtm1Acc; ++c) acc[c] += part[(s * kAtm1Acc + c) * N];
#pragma unroll
  for (int c = 0; c < kAtm1Acc; ++c) acc[c] = sum_lanes8(acc[c]);
  if (l != 0) return;
  const double sixth = 1.0 / 6.0;
  if (ENERGY) energy_out[row] = w.e2[row] + acc[0] * sixth;
  if (GRAD) {
    grad_out[3 * row] += acc[1] * sixth;
    grad_out[3 * row + 1] += acc[2] * sixth;
    grad_out[3 * row + 2] += acc[3] * sixth;
    w.dldcn[row] += acc[4] * sixth;
  }
}

// --------------------------------------------------------------------------
// chain pass: gradient through the D4 CN, the EEQ CN and the EEQ Coulomb
// matrix (one wave per atom row)
// --------------------------------------------------------------------------

template <int RL>
__global__ __launch_bounds__(256) void k_chain(DevTables t, Params p, Inputs in,
                                               Work w, double* grad_out) {
  const int lane = threadIdx.x & (RL - 1);
  const int64_t row = (int64_t)blockIdx.x * (256 / RL) + threadIdx.x / RL;
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
  const double dcni = w.dldcn[row], bi = w.beta[row], li = w.lam[row], qi = w.q[row];
  double fx = 0.0, fy = 0.0, fz = 0.0;
  for (int j = lane; j < n; j += RL) {
    const int zj = (int)Z[j];
    if (j == i || zj == 0) continue;
    const int64_t rj = m * N + j;
    const double dx = xi - R[3 * j], dy = yi - R[3 * j + 1], dz = zzi - R[3 * j + 2];
    const double r2 = dx * dx + dy * dy + dz * dz;
    const double ir = rsqrt_nt(r2);
    const double r = r2 * ir;
    const PairConst pc = pair_const(t, zi, zj);
    const double dc = erf_count_dr(r, pc.irc);
    double coef = 0.0;
    if (r <= p.cutcn) coef += (dcni + w.dldcn[rj]) * pc.enf * dc;
    if (eeq) {
      if (r <= p.cuteeq) coef += (bi + w.beta[rj]) * dc;
      // d/dr erf(gamma r) / r
      const double gm = pc.gam;
      const double dA =
          ir * fma(2.0 * kInvSqrtPi * gm, exp(-gm * gm * r2), -erf(gm * r) * ir);
      coef -= (li * w.q[rj] + w.lam[rj] * qi) * dA;
    }
    coef *= ir;
    fx += coef * dx;
    fy += coef * dy;
    fz += coef * dz;
  }
  fx = row_sum<RL>(fx);
  fy = row_sum<RL>(fy);
  fz = row_sum<RL>(fz);
  if (lane == 0) {
    grad_out[3 * row] += fx;
    grad_out[3 * row + 1] += fy;
    grad_out[3 * row + 2] += fz;
  }
}

// --------------------------------------------------------------------------
// chain pass for molecules of up to kAtm1MaxN atoms: the same pair terms as
// k_chain, each unordered pair once. The coefficient of a pair is symmetric
// (dL/dCN_i + dL/dCN_j, beta_i + beta_j, lam_i q_j + lam_j q_i and the
// element-pair constants), so x gets +c (r_x - r_y) and y gets -c (r_x - r_y).
// One work-group per molecule; each of its kChainWaves waves takes every
// kChainWaves-th 8 x 8 tile (bx <= by) on its own (no work-group barrier in
// the loop): row sums by DPP within the 8-lane groups, column sums by lane
// shuffles across them, both in a fixed butterfly order, added into the
// wave's own per-atom accumulators in LDS; the waves' accumulators are added
// in wave order at the end (no atomics). Atom data are staged in LDS once.
// --------------------------------------------------------------------------

constexpr int kChainAtom = 8;   // doubles per staged atom: x, y, z, dL/dCN, beta, lam, q, Z
constexpr int kChainWaves = 8;  // waves per work-group

__global__ __launch_bounds__(64 * kChainWaves) void k_chain_sym(DevTables t, Params p,
                                                                  Inputs in, Work w,
                                                                  double* grad_out) {
  extern __shared__ double s_dyn[];  // [n][kChainAtom] atoms, [kChainWaves][3][n] sums
  const int64_t m = blockIdx.x;
  const int64_t N = w.N;
  const int n = w.natom[m];
  if (n < 2) return;
  constexpr int NT = 64 * kChainWaves;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int tx = lane & 7, ty = lane >> 3;
  const int64_t* Z = in.numbers + m * N;
  const double* R = in.positions + m * N * 3;
  const bool eeq = in.q_in == nullptr;
  double* s_at = s_dyn;
  double* s_f = s_dyn + (size_t)kChainAtom * n;
  for (int i = tid; i < n; i += NT) {
    const int64_t row = m * N + i;
    double* at = s_at + kChainAtom * i;
    at[0] = R[3 * i];
    at[1] = R[3 * i + 1];
    at[2] = R[3 * i + 2];
    at[3] = w.dldcn[row];
    at[4] = eeq ? w.beta[row] : 0.0;
    at[5] = eeq ? w.lam[row] : 0.0;
    at[6] = eeq ? w.q[row] : 0.0;
    at[7] = (double)Z[i];
  }
  for (int e = tid; e < kChainWaves * 3 * n; e += NT) s_f[e] = 0.0;
  __syncthreads();
  const int nb = (n + 7) >> 3;
  const int tiles = nb * (nb + 1) / 2;
  double* f = s_f + (size_t)wv * 3 * n;
  for (int tl = wv; tl < tiles; tl += kChainWaves) {
    int by = (int)((sqrt(8.0 * tl + 1.0) - 1.0) * 0.5);
    while (by * (by + 1) / 2 > tl) --by;
    while ((by + 1) * (by + 2) / 2 <= tl) ++by;
    const int bx = tl - by * (by + 1) / 2;
    const int x = 8 * bx + tx, y = 8 * by + ty;
    double fc[3] = {0.0, 0.0, 0.0};
    if (x < y && y < n) {
      const double* ax = s_at + kChainAtom * x;
      const double* ay = s_at + kChainAtom * y;
      const int zx = (int)ax[7], zy = (int)ay[7];
      if (zx != 0 && zy != 0) {
        const double dx = ax[0] - ay[0], dy = ax[1] - ay[1], dz = ax[2] - ay[2];
        const double r2 = dx * dx + dy * dy + dz * dz;
        const double ir = rsqrt_nt(r2);
        const double r = r2 * ir;
        const PairConst pc = pair_const(t, zx, zy);
        const double dc = erf_count_dr(r, pc.irc);
        double coef = 0.0;
        if (r <= p.cutcn) coef += (ax[3] + ay[3]) * pc.enf * dc;
        if (eeq) {
          if (r <= p.cuteeq) coef += (ax[4] + ay[4]) * dc;
          // d/dr erf(gamma r) / r
          const double gm = pc.gam;
          const double dA =
              ir * fma(2.0 * kInvSqrtPi * gm, exp(-gm * gm * r2), -erf(gm * r) * ir);
          coef -= (ax[5] * ay[6] + ay[5] * ax[6]) * dA;
        }
        coef *= ir;
        fc[0] = coef * dx;
        fc[1] = coef * dy;
        fc[2] = coef * dz;
      }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      // y atom 8 by + ty: minus the sum over its row's eight lanes
      const double rs = sum_lanes8(fc[c]);
      // x atom 8 bx + tx: the sum over the eight rows
      double cs = fc[c];
      cs += __shfl_xor(cs, 8);
      cs += __shfl_xor(cs, 16);
      cs += __shfl_xor(cs, 32);
      if (tx == 0 && y < n) f[c * n + y] -= rs;
      if (ty == 0 && x < n) f[c * n + x] += cs;
    }
  }
  __syncthreads();
  for (int e = tid; e < 3 * n; e += NT) {
    const int c = e / n, i = e - c * n;
    if (s_at[kChainAtom * i + 7] == 0.0) continue;
    double g = 0.0;
    for (int k = 0; k < kChainWaves; ++k) g += s_f[(size_t)k * 3 * n + e];
    grad_out[(m * N + i) * 3 + c] += g;
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

// Element-pair constants of the CN, chain and EEQ kernels (read through
// pair_const): 1 / (rcov_i + rcov_j), the electronegativity factor and
// 1 / sqrt(rad_i^2 + rad_j^2); zero where an element has no radius.
__global__ __launch_bounds__(256) void k_pair_consts(DevTables t, double* out) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= kMaxZ * kMaxZ) return;
  const int zi = k / kMaxZ, zj = k - zi * kMaxZ;
  const double rc = t.rcov[zi] + t.rcov[zj];
  const double rr = t.rad[zi] * t.rad[zi] + t.rad[zj] * t.rad[zj];
  out[4 * k] = rc > 0.0 ? 1.0 / rc : 0.0;
  out[4 * k + 1] = en_factor(t.en[zi], t.en[zj]);
  out[4 * k + 2] = rr > 0.0 ? 1.0 / sqrt(rr) : 0.0;
  out[4 * k + 3] = 0.0;
}

__global__ __launch_bounds__(256) void k_copy(const double* src, double* dst,
                                              int64_t count) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (k < count) dst[k] = src[k];
}

// Parameter gradients: fixed-order sums of the two-body row partials and the
// three-body work-group partials into dL/d(s6, s8, s10, s9, a1, a2, alp).
// mode: 0 pair-tile kernel (one partial per 16 x 16 pair tile), 1 one-pass
// kernel (one per split; tile8 picks the 8 x 8 variant), 2 screened one-pass
// kernel (one per work-group, all written).
__global__ __launch_bounds__(256) void k_pgrad_reduce(Inputs in, Work w, int mode,
                                                      int tile8, int64_t per_mol, double* out) {
  __shared__ double red[2 * kWaves];
  const int tid = threadIdx.x;
  const int64_t N = w.N;
  double a[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int64_t row = tid; row < w.B * N; row += 256) {
    const double* g = w.pg2 + row * 5;
    a[0] += g[0];
    a[1] += g[1];
    a[2] += g[2];
    a[4] += g[3];
    a[5] += g[4];
  }
  for (int64_t q = tid; q < w.B * per_mol; q += 256) {
    const int64_t m = q / per_mol;
    const int k = (int)(q - m * per_mol);
    const int n = w.natom[m];
    const int nb = (n + kTile - 1) / kTile;
    const int64_t valid = mode == 2   ? per_mol
                          : mode == 1 ? one_pass_splits(n, tile8 != 0)
                                      : nb * (nb + 1) / 2;
    if (n == 0 || k >= valid) continue;
    const double* g = w.pg3 + q * 4;
    a[3] += g[0];
    a[6] += g[1];
    a[4] += g[2];
    a[5] += g[3];
  }
  for (int k = 0; k < 7; ++k) {
    const double v = block_sum(a[k], red);
    if (tid == 0) out[k] = v;
  }
}

inline unsigned rows_grid_rl(const Work& w, int rl, int block = 256) {
  const int per = block / rl;
  return (unsigned)((w.B * w.N + per - 1) / per);
}

// Launch a row kernel instantiated for the row width of this batch (the
// three instantiations are passed in parentheses), BLK threads per work-group.
#define D4_LAUNCH_ROWS_B(BLK, K16, K32, K64, s, p, w, ...)          \
  do {                                                              \
    const int rl_ = (p).row_lanes ? (p).row_lanes : row_lanes((w).N); \
    const dim3 g_(rows_grid_rl((w), rl_, (BLK)));                   \
    if (rl_ == 16)                                                  \
      hipLaunchKernelGGL(K16, g_, dim3(BLK), 0, s, __VA_ARGS__);    \
    else if (rl_ == 32)                                             \
      hipLaunchKernelGGL(K32, g_, dim3(BLK), 0, s, __VA_ARGS__);    \
    else                                                            \
      hipLaunchKernelGGL(K64, g_, dim3(BLK), 0, s, __VA_ARGS__);    \
  } while (0)
#define D4_LAUNCH_ROWS(K16, K32, K64, s, p, w, ...) \
  D4_LAUNCH_ROWS_B(256, K16, K32, K64, s, p, w, __VA_ARGS__)

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
  if (w.N <= kAtm1MaxN) {
    if (w.B == 0) return hipSuccess;
    const size_t shmem = sizeof(double) * (size_t)(4 + 2 * kCnWaves) * (size_t)w.N;
    // (also the extents, the Z check and, for D4, the element slots)
    if (p.d4s)
      hipLaunchKernelGGL(k_cn_sym<false>, dim3((unsigned)w.B), dim3(64 * kCnWaves), shmem, s, t,
                         p, in, w);
    else
      hipLaunchKernelGGL(k_cn_sym<true>, dim3((unsigned)w.B), dim3(64 * kCnWaves), shmem, s, t,
                         p, in, w);
    return hipGetLastError();
  }
  D4_LAUNCH_ROWS((k_cn<16>), (k_cn<32>), (k_cn<64>), s, p, w, t, p, in, w);
  return hipGetLastError();
}

hipError_t launch_zslots(hipStream_t s, const Inputs& in, const Work& w) {
  hipLaunchKernelGGL(k_zslots, dim3((unsigned)w.B), dim3(64), 0, s, in, w);
  return hipGetLastError();
}

hipError_t launch_weights(hipStream_t s, const DevTables& t, const Params& p,
                          const Work& w, const Inputs& in, int part) {
  const unsigned grid = (unsigned)((w.B * w.N * 8 + 255) / 256);
  if (part != kPairsBoth && !p.d4s)  // with the contraction vectors of the part
    hipLaunchKernelGGL(k_weights<true>, dim3(grid), dim3(256), 0, s, t, p, w, in, part);
  else
    hipLaunchKernelGGL(k_weights<false>, dim3(grid), dim3(256), 0, s, t, p, w, in, part);
  return hipGetLastError();
}

template <bool S, int PART>
void launch_pairs_model(hipStream_t s, const DevTables& t, const Params& p, const Inputs& in,
                        const Work& w, bool grad, double* grad_out) {
  // the two-body pass runs beside the three-body kernel: one-wave
  // work-groups take a SIMD as soon as one of its three-body waves ends
  // (a four-wave group needs four such SIMDs of one CU at once)
  constexpr int blk = PART == kPartE2 ? 64 : 256;
  if (grad && p.pgrad)
    D4_LAUNCH_ROWS_B(blk, (k_pairs<true, true, S, 16, PART>), (k_pairs<true, true, S, 32, PART>),
                     (k_pairs<true, true, S, 64, PART>), s, p, w, t, p, in, w, grad_out);
  else if (grad)
    D4_LAUNCH_ROWS_B(blk, (k_pairs<true, false, S, 16, PART>),
                     (k_pairs<true, false, S, 32, PART>), (k_pairs<true, false, S, 64, PART>), s,
                     p, w, t, p, in, w, grad_out);
  else
    D4_LAUNCH_ROWS_B(blk, (k_pairs<false, false, S, 16, PART>),
                     (k_pairs<false, false, S, 32, PART>), (k_pairs<false, false, S, 64, PART>),
                     s, p, w, t, p, in, w, grad_out);
}

template <int PART>
void launch_pairs_part(hipStream_t s, const DevTables& t, const Params& p, const Inputs& in,
                       const Work& w, bool grad, double* grad_out) {
  if (p.d4s)
    launch_pairs_model<true, PART>(s, t, p, in, w, grad, grad_out);
  else
    launch_pairs_model<false, PART>(s, t, p, in, w, grad, grad_out);
}

hipError_t launch_pairs(hipStream_t s, const DevTables& t, const Params& p,
                        const Inputs& in, const Work& w, bool grad,
                        double* grad_out, int part) {
  if (part == kPairsAtm)
    launch_pairs_part<kPartAtm>(s, t, p, in, w, grad, grad_out);
  else if (part == kPairsTwoBody)
    launch_pairs_part<kPartE2>(s, t, p, in, w, grad, grad_out);
  else
    launch_pairs_part<kPartBoth>(s, t, p, in, w, grad, grad_out);
  return hipGetLastError();
}

hipError_t launch_atm8(hipStream_t s, const Params& p, const Inputs& in, const Work& w,
                       bool energy, bool grad) {
  const int S = atm8_splits((int)w.N);
  const int64_t nblocks = w.B;  // molecules; block b -> (m, s) as in k_atm8 (runs of 64
                                // molecules per round of the eight XCDs)
  const int64_t grid = (w.B + 63) / 64 * 64 * S;
  if (grid == 0) return hipSuccess;
  const size_t shmem = sizeof(double) * kAtm1Acc * (size_t)w.N;
  const dim3 g((unsigned)grid), blk(64);
  if (energy && grad && in.gE != nullptr) return hipErrorInvalidValue;  // unit dL/dE only
  if (energy && grad)
    hipLaunchKernelGGL((k_atm8<true, true, false>), g, blk, shmem, s, p, in, w, S, nblocks);
  else if (energy)
    hipLaunchKernelGGL((k_atm8<true, false, false>), g, blk, shmem, s, p, in, w, S, nblocks);
  else if (grad && p.pgrad)
    hipLaunchKernelGGL((k_atm8<false, true, true>), g, blk, shmem, s, p, in, w, S, nblocks);
  else if (grad)
    hipLaunchKernelGGL((k_atm8<false, true, false>), g, blk, shmem, s, p, in, w, S, nblocks);
  return hipGetLastError();
}

hipError_t launch_atm1(hipStream_t s, const Params& p, const Inputs& in, const Work& w,
                       bool energy, bool grad) {
  if (one_pass_tile8(p, grad)) return launch_atm8(s, p, in, w, energy, grad);
  const int S = atm1_splits((int)w.N);
  const int64_t nblocks = w.B * S;
  const int64_t grid = (nblocks + 7) / 8 * 8;  // multiple of 8 for the XCD remap
  if (energy && grad)
    hipLaunchKernelGGL((k_atm1<true, true, false>), dim3((unsigned)grid), dim3(256), 0, s, p,
                       in, w, S, nblocks);
  else if (energy)
    hipLaunchKernelGGL((k_atm1<true, false, false>), dim3((unsigned)grid), dim3(256), 0, s, p,
                       in, w, S, nblocks);
  else if (grad && p.pgrad)
    hipLaunchKernelGGL((k_atm1<false, true, true>), dim3((unsigned)grid), dim3(256), 0, s, p,
                       in, w, S, nblocks);
  else if (grad)
    hipLaunchKernelGGL((k_atm1<false, true, false>), dim3((unsigned)grid), dim3(256), 0, s, p,
                       in, w, S, nblocks);
  return hipGetLastError();
}

hipError_t launch_atm8s(hipStream_t s, const Params& p, const Inputs& in, const Work& w,
                        bool energy, bool grad) {
  const int64_t nblocks = w.B * atm8s_blocks_per_mol(w.N);
  if (nblocks == 0) return hipSuccess;
  hipError_t e = launch_groups8(s, in, w);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_atm8s_near, dim3((unsigned)(w.B * n_groups8(w.N) * near_words8(w.N))),
                     dim3(64), 0, s, p, in, w);
  hipLaunchKernelGGL(k_atm8s_index, dim3((unsigned)((w.B + 63) / 64)), dim3(64), 0, s, w);
  const dim3 g((unsigned)nblocks), blk(64);
  if (energy && grad)
    hipLaunchKernelGGL((k_atm8s<true, true, false>), g, blk, 0, s, p, in, w, nblocks);
  else if (energy)
    hipLaunchKernelGGL((k_atm8s<true, false, false>), g, blk, 0, s, p, in, w, nblocks);
  else if (grad && p.pgrad)
    hipLaunchKernelGGL((k_atm8s<false, true, true>), g, blk, 0, s, p, in, w, nblocks);
  else if (grad)
    hipLaunchKernelGGL((k_atm8s<false, true, false>), g, blk, 0, s, p, in, w, nblocks);
  return hipGetLastError();
}

hipError_t launch_atm(hipStream_t s, const Params& p, const Inputs& in,
                      const Work& w, bool energy, bool grad) {
  if (one_pass(p, w)) return launch_atm1(s, p, in, w, energy, grad);
  if (!p.pair_owner) return launch_atm8s(s, p, in, w, energy, grad);
  const int nb = (int)((w.N + kTile - 1) / kTile);
  const int tiles = nb * (nb + 1) / 2;
  const int64_t ntiles = w.B * tiles;
  const int64_t grid = (ntiles + 7) / 8 * 8;  // multiple of 8 for the XCD remap
  if (grid == 0) return hipSuccess;
  hipError_t e = launch_groups(s, in, w);
  if (e != hipSuccess) return e;
  if (energy && grad)
    hipLaunchKernelGGL((k_atm<true, true, false>), dim3((unsigned)grid), dim3(256), 0, s, p,
                       in, w, tiles, ntiles);
  else if (energy)
    hipLaunchKernelGGL((k_atm<true, false, false>), dim3((unsigned)grid), dim3(256), 0, s, p,
                       in, w, tiles, ntiles);
  else if (grad && p.pgrad)
    hipLaunchKernelGGL((k_atm<false, true, true>), dim3((unsigned)grid), dim3(256), 0, s, p,
                       in, w, tiles, ntiles);
  else if (grad)
    hipLaunchKernelGGL((k_atm<false, true, false>), dim3((unsigned)grid), dim3(256), 0, s, p,
                       in, w, tiles, ntiles);
  return hipGetLastError();
}

hipError_t launch_atm_rows(hipStream_t s, const Params& p, const Inputs& in, const Work& w,
                           bool energy, bool grad, double* energy_out,
                           double* grad_out) {
  if (one_pass(p, w)) {
    const unsigned grid = (unsigned)((w.B * w.N * 8 + 255) / 256);
    const int tile8 = one_pass_tile8(p, grad) ? 1 : 0;
    if (energy && grad)
      hipLaunchKernelGGL((k_atm1_rows<true, true>), dim3(grid), dim3(256), 0, s, in, w, tile8,
                         energy_out, grad_out);
    else if (energy)
      hipLaunchKernelGGL((k_atm1_rows<true, false>), dim3(grid), dim3(256), 0, s, in, w, tile8,
                         energy_out, grad_out);
    else if (grad)
      hipLaunchKernelGGL((k_atm1_rows<false, true>), dim3(grid), dim3(256), 0, s, in, w, tile8,
                         energy_out, grad_out);
    return hipGetLastError();
  }
  if (!p.pair_owner) {
    const dim3 grid((unsigned)(w.B * n_groups8(w.N))), blk(1024);
    if (energy && grad)
      hipLaunchKernelGGL((k_atm8s_rows<true, true>), grid, blk, 0, s, in, w, energy_out,
                         grad_out);
    else if (energy)
      hipLaunchKernelGGL((k_atm8s_rows<true, false>), grid, blk, 0, s, in, w, energy_out,
                         grad_out);
    else if (grad)
      hipLaunchKernelGGL((k_atm8s_rows<false, true>), grid, blk, 0, s, in, w, energy_out,
                         grad_out);
    return hipGetLastError();
  }
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

hipError_t launch_chain(hipStream_t s, const DevTables& t, const Params& p,
                        const Inputs& in, const Work& w, double* grad_out) {
  if (w.N <= kAtm1MaxN) {
    // at most (8 + 3 * 8) * 256 doubles = 64 KiB, the default dynamic LDS limit
    static_assert((kChainAtom + 3 * kChainWaves) * kAtm1MaxN * sizeof(double) <= 65536, "LDS");
    if (w.B == 0) return hipSuccess;
    const size_t shmem = sizeof(double) * (size_t)(kChainAtom + 3 * kChainWaves) * (size_t)w.N;
    hipLaunchKernelGGL(k_chain_sym, dim3((unsigned)w.B), dim3(64 * kChainWaves), shmem, s, t, p,
                       in, w, grad_out);
    return hipGetLastError();
  }
  D4_LAUNCH_ROWS((k_chain<16>), (k_chain<32>), (k_chain<64>), s, p, w, t, p, in, w, grad_out);
  return hipGetLastError();
}

hipError_t launch_properties(hipStream_t s, const DevTables& t, const Inputs& in,
                             const Work& w, double* c6_out, double* alpha_out) {
  hipLaunchKernelGGL(k_properties, dim3(rows_grid(w)), dim3(256), 0, s, t, in, w,
                     c6_out, alpha_out);
  return hipGetLastError();
}

hipError_t launch_pgrad_reduce(hipStream_t s, const Params& p, const Inputs& in,
                               const Work& w, double* out) {
  const bool op = one_pass(p, w);
  const bool tile8 = one_pass_tile8(p, true);
  const int64_t nb = n_groups(w.N);
  const int mode = op ? 1 : (p.pair_owner ? 0 : 2);
  const int64_t per_mol = mode == 1   ? one_pass_splits((int)w.N, tile8)
                          : mode == 2 ? atm8s_blocks_per_mol(w.N)
                                      : nb * (nb + 1) / 2;
  hipLaunchKernelGGL(k_pgrad_reduce, dim3(1), dim3(256), 0, s, in, w, mode, tile8 ? 1 : 0,
                     per_mol, out);
  return hipGetLastError();
}

hipError_t launch_pair_consts(hipStream_t s, const DevTables& t, double* out) {
  hipLaunchKernelGGL(k_pair_consts, dim3((kMaxZ * kMaxZ + 255) / 256), dim3(256), 0, s, t, out);
  return hipGetLastError();
}

hipError_t launch_copy(hipStream_t s, const double* src, double* dst, int64_t count) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_copy, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, src,
                     dst, count);
  return hipGetLastError();
}

}  // namespace d4
