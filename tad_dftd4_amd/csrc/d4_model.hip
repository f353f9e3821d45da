// HIP kernels (gfx950) behind the dispersion-model methods of the C ABI:
// reference weights from given coordination numbers and charges, C6 from
// given weights, weighted and static polarisabilities, and the backward
// (vector-Jacobian) products of each, for the D4 and D4S models.
//
// Inputs and outputs use the reference's layout: weights (..., n, 7) (D4S:
// (..., n, n, 7), entry [m][n][r] = atom n's weight of reference r in its
// pair with atom m) in the reference's slot order; the device tables keep
// each element's used reference systems compacted to the front and
// `refslot` maps compact slot c to the reference's slot.
//
// FP64 throughout; every output element has one writer and every sum a fixed
// order.

#include "d4_device.h"

namespace d4 {
namespace {

constexpr int kFreq = 23;  // Casimir-Polder frequencies of alpha(i w)

__device__ __forceinline__ int safe_z(int64_t z) {
  return (z > 0 && z < kMaxZ) ? (int)z : 0;
}

// Sum over each group of 8 lanes (lanes 8k .. 8k+7 of the wave).
__device__ __forceinline__ double sum8(double v) {
  v += __shfl_xor(v, 1, 64);
  v += __shfl_xor(v, 2, 64);
  v += __shfl_xor(v, 4, 64);
  return v;
}

// Gaussian part of the weights of an atom of element z (compact slot a of
// this lane, `used` if a < nref) at coordination number cn, for the
// weighting factor wf, and its derivative with respect to cn. D4 forms
// sum_k exp(-k wf d^2) (the reference's exp(-d^2)^(wf k), model/d4.py:
// 174-188, to ~1e-13 relative); D4S sum_k exp(-wf d^2)^k (model/d4s.py:
// 188-205). Exceptional case (every Gaussian underflows): the reference(s)
// with the largest reference CN take weight 1 (model/d4.py:209-217).
template <bool S>
__device__ __forceinline__ void gauss8(const DevTables& t, int z, int a, bool used, double cn,
                                       double wf, double& g, double& dg) {
  double e = 0.0, de = 0.0;
  double refcn = 0.0;
  if (used) {
    refcn = t.refcn[z * kMaxRef + a];
    const double d = cn - refcn;
    const int ng = t.ngw[z * kMaxRef + a];
    if (S) {
      const double tmp = exp(-(d * d) * wf);
      double sum = tmp, ksum = tmp;
      if (ng == 3) {
        const double t2 = tmp * tmp, t3 = t2 * tmp;
        sum = tmp + t2 + t3;
        ksum = fma(3.0, t3, fma(2.0, t2, tmp));
      }
      e = sum;
      de = -2.0 * wf * d * ksum;
    } else {
      for (int k = 1; k <= ng; ++k) {
        const double kw = k * wf;
        const double ek = exp(-kw * d * d);
        e += ek;
        de -= 2.0 * kw * d * ek;
      }
    }
  }
  const double norm = sum8(e);
  const double dnorm = sum8(de);
  g = 0.0;
  dg = 0.0;
  if (!used) return;
  if (!(norm > 0.0)) {
    g = refcn == t.maxcn[z] ? 1.0 : 0.0;
  } else {
    g = e / norm;
    dg = (de - g * dnorm) / norm;
  }
}

// D4 weights: 8 lanes per atom row (lane = compact reference slot). Forward
// writes gw (and -d gw / d cn, d gw / d q when asked: the reference returns
// the negated CN derivative, model/d4.py:231-250); backward (grad_gw given)
// writes dL/dcn and dL/dq of L = sum grad_gw * gw.
__global__ __launch_bounds__(256) void k_mw_d4(DevTables t, int64_t rows,
                                               const int64_t* __restrict__ numbers,
                                               const double* __restrict__ cn_in,
                                               const double* __restrict__ q_in, double ga,
                                               double gc, double wf, double* gw_out,
                                               double* dgwdcn_out, double* dgwdq_out,
                                               const double* __restrict__ grad_gw,
                                               double* grad_cn, double* grad_q) {
  const int64_t tid = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t row = tid >> 3;
  const int a = (int)(tid & 7);
  if (row >= rows) return;  // uniform over each group of 8 lanes
  const int z = safe_z(numbers[row]);
  const int nr = z != 0 ? t.nref[z] : 0;
  const bool used = a < nr;
  const double cn = cn_in ? cn_in[row] : 0.0;
  const double q = q_in ? q_in[row] : 0.0;
  double g, dg;
  gauss8<false>(t, z, a, used, cn, wf, g, dg);
  double zq = 0.0, dzq = 0.0;
  if (used) {
    const double zeff = t.zeff[z];
    zeta_dzeta(ga, t.gam[z] * gc, t.refq[z * kMaxRef + a] + zeff, q + zeff, zq, dzq);
  }
  const int slot = used ? t.refslot[z * kMaxRef + a] : -1;
  if (grad_gw) {
    const double G = used ? grad_gw[row * kMaxRef + slot] : 0.0;
    const double gc_ = sum8(G * zq * dg);
    const double gq_ = sum8(G * dzq * g);
    if (a == 0) {
      if (grad_cn) grad_cn[row] = gc_;
      if (grad_q) grad_q[row] = gq_;
    }
    return;
  }
  // lane a < 7 writes reference slot a: the value of the compact slot that
  // maps there (refcomp), zero for an unused slot
  const int c = a < kMaxRef ? t.refcomp[z * kMaxRef + a] : -1;
  const int src = (threadIdx.x & ~7) + (c < 0 ? 0 : c);
  const double vg = __shfl(zq * g, src, 64);
  const double vc = __shfl(-zq * dg, src, 64);
  const double vq = __shfl(dzq * g, src, 64);
  if (a < kMaxRef) {
    const int64_t o = row * kMaxRef + a;
    gw_out[o] = c < 0 ? 0.0 : vg;
    if (dgwdcn_out) dgwdcn_out[o] = c < 0 ? 0.0 : vc;
    if (dgwdq_out) dgwdq_out[o] = c < 0 ? 0.0 : vq;
  }
}

// D4S weights: 8 lanes per atom n, looping over the partner m; atom n's
// weights in its pair with m use wf(Z_n, Z_m) (model/d4s.py:61-67, 178-238),
// padding partners included as in the reference. Forward writes gw[m][n][:]
// (and d gw / d q); backward sums dL/dcn_n and dL/dq_n over m in order.
__global__ __launch_bounds__(256) void k_mw_d4s(DevTables t, int64_t B, int64_t N,
                                                const int64_t* __restrict__ numbers,
                                                const double* __restrict__ cn_in,
                                                const double* __restrict__ q_in, double ga,
                                                double gc, double* gw_out, double* dgwdq_out,
                                                const double* __restrict__ grad_gw,
                                                double* grad_cn, double* grad_q) {
  const int64_t tid = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t row = tid >> 3;  // (molecule, atom n)
  const int a = (int)(tid & 7);
  if (row >= B * N) return;
  const int64_t mol = row / N;
  const int64_t n = row - mol * N;
  const int z = safe_z(numbers[row]);
  const int nr = z != 0 ? t.nref[z] : 0;
  const bool used = a < nr;
  const double cn = cn_in ? cn_in[row] : 0.0;
  const double q = q_in ? q_in[row] : 0.0;
  double zq = 0.0, dzq = 0.0;
  if (used) {
    const double zeff = t.zeff[z];
    zeta_dzeta(ga, t.gam[z] * gc, t.refq[z * kMaxRef + a] + zeff, q + zeff, zq, dzq);
  }
  const int slot = used ? t.refslot[z * kMaxRef + a] : -1;
  const int c = a < kMaxRef ? t.refcomp[z * kMaxRef + a] : -1;
  const int src = (threadIdx.x & ~7) + (c < 0 ? 0 : c);
  double acc_cn = 0.0, acc_q = 0.0;
  for (int64_t m = 0; m < N; ++m) {
    const int zm = safe_z(numbers[mol * N + m]);
    const double wf = t.wfpair[z * kMaxZ + zm];
    double g, dg;
    gauss8<true>(t, z, a, used, cn, wf, g, dg);
    const int64_t base = ((mol * N + m) * N + n) * kMaxRef;
    if (grad_gw) {
      const double G = used ? grad_gw[base + slot] : 0.0;
      acc_cn += G * zq * dg;
      acc_q += G * dzq * g;
      continue;
    }
    const double vg = __shfl(zq * g, src, 64);
    const double vq = __shfl(dzq * g, src, 64);
    if (a < kMaxRef) {
      gw_out[base + a] = c < 0 ? 0.0 : vg;
      if (dgwdq_out) dgwdq_out[base + a] = c < 0 ? 0.0 : vq;
    }
  }
  if (grad_gw) {
    acc_cn = sum8(acc_cn);
    acc_q = sum8(acc_q);
    if (a == 0) {
      if (grad_cn) grad_cn[row] = acc_cn;
      if (grad_q) grad_q[row] = acc_q;
    }
  }
}

// Weights of atom i (reference slot order, stride `st` between atoms'
// vectors) gathered into compact slots.
__device__ __forceinline__ void gather_compact(const DevTables& t, int z, const double* gw,
                                               double* out) {
  const int nr = z != 0 ? t.nref[z] : 0;
#pragma unroll
  for (int c = 0; c < kMaxRef; ++c)
    out[c] = c < nr ? gw[t.refslot[z * kMaxRef + c]] : 0.0;
}

// sum_ab gi[a] RC6[zi][zj][a][b] gj[b] over the compact slots.
__device__ __forceinline__ double c6_contract(const DevTables& t, int zi, int zj,
                                              const double* gi, const double* gj) {
  if (zi == 0 || zj == 0) return 0.0;
  const double* blk = t.rc6 + ((size_t)zi * kMaxZ + zj) * kRc6Block;
  const int nri = t.nref[zi], nrj = t.nref[zj];
  double c6 = 0.0;
  for (int a = 0; a < nri; ++a) {
    double v = 0.0;
    for (int b = 0; b < nrj; ++b) v = fma(blk[a * kMaxRef + b], gj[b], v);
    c6 = fma(gi[a], v, c6);
  }
  return c6;
}

// C6 of every atom pair: one thread per (molecule, i, j).
// D4: c6_ij = sum_ab gw_ia RC6_ij,ab gw_jb (model/d4.py:268-289);
// D4S: c6_ij = sum_ab gw[j][i][a] RC6_ij,ab gw[i][j][b] (model/d4s.py:265-285).
template <bool S>
__global__ __launch_bounds__(256) void k_mc6(DevTables t, int64_t B, int64_t N,
                                             const int64_t* __restrict__ numbers,
                                             const double* __restrict__ gw, double* c6_out) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (k >= B * N * N) return;
  const int64_t mol = k / (N * N);
  const int64_t ij = k - mol * N * N;
  const int64_t i = ij / N, j = ij - (ij / N) * N;
  const int zi = safe_z(numbers[mol * N + i]), zj = safe_z(numbers[mol * N + j]);
  double gi[kMaxRef], gj[kMaxRef];
  if (S) {
    gather_compact(t, zi, gw + ((mol * N + j) * N + i) * kMaxRef, gi);
    gather_compact(t, zj, gw + ((mol * N + i) * N + j) * kMaxRef, gj);
  } else {
    gather_compact(t, zi, gw + (mol * N + i) * kMaxRef, gi);
    gather_compact(t, zj, gw + (mol * N + j) * kMaxRef, gj);
  }
  c6_out[k] = c6_contract(t, zi, zj, gi, gj);
}

// Backward of k_mc6 for L = sum_ij G_ij c6_ij. RC6 is symmetric under
// (i, a) <-> (j, b), so
//   D4:  dL/dgw_ia    = sum_j (G_ij + G_ji) sum_b RC6_ij,ab gw_jb;
//   D4S: dL/dgw[m][n][r] = (G_nm + G_mn) sum_b RC6_nm,rb gw[n][m][b].
// One thread per (molecule, atom, reference slot) (D4) or per (molecule,
// m, n, slot) (D4S); the D4 sum over j runs in order.
template <bool S>
__global__ __launch_bounds__(256) void k_mc6_bwd(DevTables t, int64_t B, int64_t N,
                                                 const int64_t* __restrict__ numbers,
                                                 const double* __restrict__ gw,
                                                 const double* __restrict__ G,
                                                 double* grad_gw) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t per = S ? N * N * kMaxRef : N * kMaxRef;
  if (k >= B * per) return;
  const int64_t mol = k / per;
  const int64_t rest = k - mol * per;
  const int s = (int)(rest % kMaxRef);  // reference slot of the output
  const int64_t* Z = numbers + mol * N;
  const double* Gm = G + mol * N * N;
  if (S) {
    const int64_t mn = rest / kMaxRef;
    const int64_t m = mn / N, n = mn - (mn / N) * N;
    const int zn = safe_z(Z[n]), zm = safe_z(Z[m]);
    double out = 0.0;
    const int c = zn != 0 ? t.refcomp[zn * kMaxRef + s] : -1;
    if (c >= 0 && zm != 0) {
      double gm[kMaxRef];
      gather_compact(t, zm, gw + ((mol * N + n) * N + m) * kMaxRef, gm);
      const double* blk = t.rc6 + ((size_t)zn * kMaxZ + zm) * kRc6Block;
      double v = 0.0;
      for (int b = 0; b < t.nref[zm]; ++b) v = fma(blk[c * kMaxRef + b], gm[b], v);
      out = (Gm[n * N + m] + Gm[m * N + n]) * v;
    }
    grad_gw[k] = out;
    return;
  }
  const int64_t i = rest / kMaxRef;
  const int zi = safe_z(Z[i]);
  const int c = zi != 0 ? t.refcomp[zi * kMaxRef + s] : -1;
  double out = 0.0;
  if (c >= 0) {
    for (int64_t j = 0; j < N; ++j) {
      const int zj = safe_z(Z[j]);
      if (zj == 0) continue;
      double gj[kMaxRef];
      gather_compact(t, zj, gw + (mol * N + j) * kMaxRef, gj);
      const double* blk = t.rc6 + ((size_t)zi * kMaxZ + zj) * kRc6Block;
      double v = 0.0;
      for (int b = 0; b < t.nref[zj]; ++b) v = fma(blk[c * kMaxRef + b], gj[b], v);
      out = fma(Gm[i * N + j] + Gm[j * N + i], v, out);
    }
  }
  grad_gw[k] = out;
}

// Polarisabilities from weights, alpha[z][c][w] the reference
// polarisabilities of compact slot c at the 23 frequencies.
//   mode 0 (D4 weighted, model/d4.py:291-306):   out[n][w] = sum_r gw[n][r] alpha_n[r][w]
//   mode 1 (D4 static, model/base.py:286-302):   out[n]    = sum_r gw[n][r] alpha_n[r][0]
//   mode 2 (D4S weighted, model/d4s.py:287-313): out[m][w] = sum_n sum_r gw[m][n][r]
//                                                alpha_n[r][w] / nat
//   mode 3 (D4S static, the reference's broadcast of model/base.py:302 over the
//           pair axis): out[m][n] = sum_r gw[m][n][r] alpha_n[r][0]
// One thread per output element; nat counts the molecule's atoms with Z != 0.
__global__ __launch_bounds__(256) void k_mpol(DevTables t, int64_t B, int64_t N, int mode,
                                              const int64_t* __restrict__ numbers,
                                              const double* __restrict__ gw, double* out) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t per = (mode == 0 || mode == 2) ? N * kFreq : (mode == 1 ? N : N * N);
  if (k >= B * per) return;
  const int64_t mol = k / per;
  const int64_t rest = k - mol * per;
  const int64_t* Z = numbers + mol * N;
  auto contract = [&](int64_t n, const double* g, int w) {
    const int z = safe_z(Z[n]);
    if (z == 0) return 0.0;
    const double* al = t.alpha + (size_t)z * kMaxRef * kFreq;
    double v = 0.0;
    for (int c = 0; c < t.nref[z]; ++c)
      v = fma(g[t.refslot[z * kMaxRef + c]], al[c * kFreq + w], v);
    return v;
  };
  if (mode == 0 || mode == 1) {
    const int64_t n = mode == 0 ? rest / kFreq : rest;
    const int w = mode == 0 ? (int)(rest - n * kFreq) : 0;
    out[k] = contract(n, gw + (mol * N + n) * kMaxRef, w);
  } else if (mode == 3) {
    const int64_t m = rest / N, n = rest - (rest / N) * N;
    out[k] = contract(n, gw + ((mol * N + m) * N + n) * kMaxRef, 0);
  } else {
    const int64_t m = rest / kFreq;
    const int w = (int)(rest - m * kFreq);
    double s = 0.0;
    int64_t nat = 0;
    for (int64_t n = 0; n < N; ++n) {
      nat += Z[n] != 0;
      s += contract(n, gw + ((mol * N + m) * N + n) * kMaxRef, w);
    }
    out[k] = s / (double)nat;
  }
}

// Backward of k_mpol: dL/dgw for L = sum G * out (same modes), one thread
// per weight element (molecule, [m,] n, reference slot).
__global__ __launch_bounds__(256) void k_mpol_bwd(DevTables t, int64_t B, int64_t N, int mode,
                                                  const int64_t* __restrict__ numbers,
                                                  const double* __restrict__ G,
                                                  double* grad_gw) {
  const bool pair = mode >= 2;
  const int64_t per = pair ? N * N * kMaxRef : N * kMaxRef;
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (k >= B * per) return;
  const int64_t mol = k / per;
  const int64_t rest = k - mol * per;
  const int s = (int)(rest % kMaxRef);
  const int64_t mn = rest / kMaxRef;
  const int64_t m = pair ? mn / N : mn, n = pair ? mn - (mn / N) * N : mn;
  const int64_t* Z = numbers + mol * N;
  const int z = safe_z(Z[n]);
  const int c = z != 0 ? t.refcomp[z * kMaxRef + s] : -1;
  double out = 0.0;
  if (c >= 0) {
    const double* al = t.alpha + ((size_t)z * kMaxRef + c) * kFreq;
    if (mode == 0) {
      const double* g = G + (mol * N + n) * kFreq;
      for (int w = 0; w < kFreq; ++w) out = fma(g[w], al[w], out);
    } else if (mode == 1) {
      out = G[mol * N + n] * al[0];
    } else if (mode == 3) {
      out = G[(mol * N + m) * N + n] * al[0];
    } else {
      int64_t nat = 0;
      for (int64_t j = 0; j < N; ++j) nat += Z[j] != 0;
      const double* g = G + (mol * N + m) * kFreq;
      for (int w = 0; w < kFreq; ++w) out = fma(g[w], al[w], out);
      out /= (double)nat;
    }
  }
  grad_gw[k] = out;
}

inline unsigned blocks(int64_t threads) { return (unsigned)((threads + 255) / 256); }

}  // namespace

hipError_t launch_model_weights(hipStream_t s, const DevTables& t, const ModelCall& mc,
                                const double* cn, const double* q, double* gw_out,
                                double* dgwdcn_out, double* dgwdq_out, const double* grad_gw,
                                double* grad_cn, double* grad_q) {
  const int64_t rows = mc.B * mc.N;
  if (rows == 0) return hipSuccess;
  if (mc.d4s)
    hipLaunchKernelGGL(k_mw_d4s, dim3(blocks(rows * 8)), dim3(256), 0, s, t, mc.B, mc.N,
                       mc.numbers, cn, q, mc.ga, mc.gc, gw_out, dgwdq_out, grad_gw, grad_cn,
                       grad_q);
  else
    hipLaunchKernelGGL(k_mw_d4, dim3(blocks(rows * 8)), dim3(256), 0, s, t, rows, mc.numbers,
                       cn, q, mc.ga, mc.gc, mc.wf, gw_out, dgwdcn_out, dgwdq_out, grad_gw,
                       grad_cn, grad_q);
  return hipGetLastError();
}

hipError_t launch_model_c6(hipStream_t s, const DevTables& t, const ModelCall& mc,
                           const double* gw, double* c6_out) {
  const int64_t n = mc.B * mc.N * mc.N;
  if (n == 0) return hipSuccess;
  if (mc.d4s)
    hipLaunchKernelGGL(k_mc6<true>, dim3(blocks(n)), dim3(256), 0, s, t, mc.B, mc.N,
                       mc.numbers, gw, c6_out);
  else
    hipLaunchKernelGGL(k_mc6<false>, dim3(blocks(n)), dim3(256), 0, s, t, mc.B, mc.N,
                       mc.numbers, gw, c6_out);
  return hipGetLastError();
}

hipError_t launch_model_c6_bwd(hipStream_t s, const DevTables& t, const ModelCall& mc,
                               const double* gw, const double* grad_c6, double* grad_gw) {
  const int64_t n = mc.B * (mc.d4s ? mc.N * mc.N : mc.N) * kMaxRef;
  if (n == 0) return hipSuccess;
  if (mc.d4s)
    hipLaunchKernelGGL(k_mc6_bwd<true>, dim3(blocks(n)), dim3(256), 0, s, t, mc.B, mc.N,
                       mc.numbers, gw, grad_c6, grad_gw);
  else
    hipLaunchKernelGGL(k_mc6_bwd<false>, dim3(blocks(n)), dim3(256), 0, s, t, mc.B, mc.N,
                       mc.numbers, gw, grad_c6, grad_gw);
  return hipGetLastError();
}

hipError_t launch_model_pols(hipStream_t s, const DevTables& t, const ModelCall& mc,
                             int static_only, const double* gw, double* out) {
  const int mode = (mc.d4s ? 2 : 0) + (static_only ? 1 : 0);
  const int64_t per = (mode == 0 || mode == 2) ? mc.N * kFreq : (mode == 1 ? mc.N : mc.N * mc.N);
  const int64_t n = mc.B * per;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_mpol, dim3(blocks(n)), dim3(256), 0, s, t, mc.B, mc.N, mode,
                     mc.numbers, gw, out);
  return hipGetLastError();
}

hipError_t launch_model_pols_bwd(hipStream_t s, const DevTables& t, const ModelCall& mc,
                                 int static_only, const double* grad_out, double* grad_gw) {
  const int mode = (mc.d4s ? 2 : 0) + (static_only ? 1 : 0);
  const int64_t n = mc.B * (mc.d4s ? mc.N * mc.N : mc.N) * kMaxRef;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_mpol_bwd, dim3(blocks(n)), dim3(256), 0, s, t, mc.B, mc.N, mode,
                     mc.numbers, grad_out, grad_gw);
  return hipGetLastError();
}

}  // namespace d4
