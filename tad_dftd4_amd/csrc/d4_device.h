// Device helpers shared by the kernel translation units: constants of the
// model, wavefront and work-group reductions, counting functions and the
// EEQ matrix elements.
#pragma once

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
constexpr double kInvSqrtPi = 0.56418958354775628695;  // 1 / sqrt(pi)
constexpr double kInvK6 = 1.0 / kK6;
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

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
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

// 1 / sqrt(x), x > 0: hardware estimate and two Newton steps.
__device__ __forceinline__ double rsqrt_nt(double x) {
  double y = __builtin_amdgcn_rsq(x);
  const double h = 0.5 * x;
  y = y * fma(-h * y, y, 1.5);
  return y * fma(-h * y, y, 1.5);
}

// 1 / x: hardware estimate and two Newton steps.
__device__ __forceinline__ double rcp_nt(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

// 1/2 (1 + erf(-k (r/rc - 1)))  (tad-mctc erf_count), irc = 1 / rc
__device__ __forceinline__ double erf_count(double r, double irc) {
  return 0.5 * (1.0 + erf(-kKcn * fma(r, irc, -1.0)));
}

// d/dr of erf_count
__device__ __forceinline__ double erf_count_dr(double r, double irc) {
  const double x = fma(r, irc, -1.0);
  return -(kKcn * kInvSqrtPi) * irc * exp(-kKcn * kKcn * x * x);
}

// Electronegativity factor of the D4 coordination number.
__device__ __forceinline__ double en_factor(double eni, double enj) {
  const double d = fabs(eni - enj) + kK5;
  return kK4 * exp(-(d * d) * kInvK6);
}

// Element-pair constants (DevTables::pairk): 1 / (rcov_i + rcov_j), the
// electronegativity factor and the EEQ gamma = 1 / sqrt(rad_i^2 + rad_j^2).
struct PairConst {
  double irc, enf, gam;
};
__device__ __forceinline__ PairConst pair_const(const DevTables& t, int zi, int zj) {
  const double2* p = reinterpret_cast<const double2*>(t.pairk + (zi * kMaxZ + zj) * 4);
  const double2 a = p[0], b = p[1];
  return {a.x, a.y, b.x};
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

// zeta and its derivative together (the exponentials shared).
__device__ __forceinline__ void zeta_dzeta(double ga, double gam, double qref, double qmod,
                                           double& z, double& dz) {
  if (qmod > 0.0) {
    const double den = qmod - kEps;
    const double scale = exp(gam * (1.0 - qref / den));
    z = exp(ga * (1.0 - scale));
    dz = -ga * gam * scale * z * qref / (den * den);
  } else {
    z = exp(ga);
    dz = 0.0;
  }
}


__host__ __device__ __forceinline__ int tri(int i) { return (i * (i + 1)) >> 1; }

__device__ __forceinline__ double eeq_rhs(const DevTables& t, int z, double cnq) {
  return -t.chi[z] + t.kcn[z] * sqrt(fmax(cnq, kEps));
}

// Coulomb element erf(gamma r) / r from r^2 > 0.
__device__ __forceinline__ double eeq_offdiag_r2(double gamma, double r2) {
  const double ir = rsqrt_nt(r2);
  return erf(r2 * ir * gamma) * ir;
}

__device__ __forceinline__ double eeq_offdiag(const DevTables& t, int zi, int zj, double r2) {
  return eeq_offdiag_r2(t.pairk[(zi * kMaxZ + zj) * 4 + 2], r2);
}

__device__ __forceinline__ double eeq_diag(const DevTables& t, int z) {
  return t.eta[z] + sqrt(2.0 / kPi) / t.rad[z];
}

}  // namespace
}  // namespace d4
