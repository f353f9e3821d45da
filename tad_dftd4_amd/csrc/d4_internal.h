// Internal declarations shared by the kernel and C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/d4amd.h"

namespace d4 {

constexpr int kMaxZ = D4_MAX_Z;
constexpr int kMaxRef = D4_MAX_REF;
constexpr int kRefStride = 8;  // per-atom weight vectors are padded to 8
constexpr int kRc6Block = kMaxRef * kMaxRef;
// Distinct elements of a molecule given C6 contraction vectors (more: the pair
// passes contract the reference C6 blocks per pair).
constexpr int kZSlots = 32;

// Molecule status bits (d4_molecule_status).
constexpr int kStatusLU = 1;
constexpr int kStatusSingular = 2;
constexpr int kStatusBadZ = 4;

// Device-resident element tables (pointers into one allocation).
struct DevTables {
  const double* rcov;
  const double* en;
  const double* r4r2;
  const double* chi;
  const double* eta;
  const double* kcn;
  const double* rad;
  const double* gam;
  const double* zeff;
  const double* maxcn;
  const double* refcn;
  const double* refq;
  const double* alpha0;
  const double* rc6;
  const double* wfpair;  // D4S pairwise weighting factors [zi][zj]
  const double* pairk;   // [zi][zj][4] element-pair constants (see pair_const)
  const int32_t* nref;
  const int32_t* ngw;
  // model methods (d4_model.hip); null when the context has no such tables
  const double* alpha;     // [z][c][23] reference polarisabilities alpha(i w), compact slots
  const int32_t* refslot;  // [z][c] reference slot of compact slot c (-1: unused)
  const int32_t* refcomp;  // [z][s] compact slot of reference slot s (-1: unused)
};

// One call of a dispersion-model method (d4_model.hip).
struct ModelCall {
  int64_t B, N;
  const int64_t* numbers;
  double ga, gc, wf;
  int d4s;
};

// Kernel-side copy of the parameters with derived constants.
struct Params {
  double s6, s8, s10, s9, a1, a2, alp;
  double ga, gc, wf;
  double cut2, cut3sq, cutcn, cuteeq;
  int s10on;
  int force_lu;
  int force_tiled;
  int force_sort;
  int force_pairtiles;
  int pgrad;  // accumulate parameter gradients (backward only)
  int d4s;    // D4S model: Gaussian weights with pairwise factors
  int atm16;  // one-pass three-body kernel with 16 x 16 pair tiles (testing)
  int pair_owner;  // large systems: pair-owner three-body kernel (testing)
  int row_lanes;   // lanes per atom row of the row passes (16, 32, 64; 0: by length)
  int atm8;        // energy-only calls: one-pass three-body kernel with 8 x 8 tiles
  int one_stream;  // no companion stream (testing, profiling)
};

// Workspace views for one batch (nbatch molecules padded to natoms).
struct Work {
  int32_t* status;  // [B]
  int32_t* natom;   // [B] extent: 1 + index of the last real atom
  double* cn;       // [B*N] D4 coordination number
  double* cnq;      // [B*N] capped EEQ coordination number
  double* dcnq;     // [B*N] d cap / d raw EEQ CN
  double* q;        // [B*N] charges used
  double* e2;       // [B*N] two-body atom energies
  double* dldcn;    // [B*N] dL/dCN
  double* dldq;     // [B*N] dL/dq
  double* beta;     // [B*N] lambda_i * d b_i / d raw EEQ CN
  double* lam;      // [B*N] EEQ adjoint
  double* gw;       // [B*N*8] weights with charges (D4S: zeta(q))
  double* gw0;      // [B*N*8] weights with q = 0 (ATM) (D4S: zeta(0))
  double* dgw;      // [B*N*8] d gw / d CN (D4S: unused)
  double* dgw0;     // [B*N*8] d gw0 / d CN (D4S: unused)
  double* dgwq;     // [B*N*8] d gw / d q (D4S: dzeta/dq)
  double* chol;     // [B*NT] packed lower Cholesky factor of the Coulomb block
  double* ez;       // [B*N] y_e = L^-1 e (forward-substituted border, e = 1 on real atoms)
  double* lu;       // [B*(N+1)^2] pivoted LU of the bordered matrix (fallback)
  int32_t* piv;     // [B*(N+1)]
  double* vec;      // [B*(N+1)] right-hand side / solution of the LU path
  double* pt;       // [B*N*N] ATM damping factor t_ij = (R0_ij / r_ij)^(alp/3)
  double* pp;       // [B*N*N] ATM factor p_ij = sqrt|C6(q=0)_ij| / r_ij^3
  double* pk;       // [B*N*N] d log sqrt|C6(q=0)_ij| / d CN_i
  double* pg;       // [B*N*N] ATM dL/d r_ij^2; pg, ph and pe together hold the window
                    // buffer of the screened one-pass kernel (atm8s_stride per molecule)
  double* ph;       // [B*N*N] ATM dL/d log sqrt|C6(q=0)_ij|
  double* pe;       // [B*N*N] ATM atom-energy partials (row atom)
  // spatial order of large molecules (see launch_sort)
  int32_t* perm;    // [B*N] sorted slot -> input index
  int64_t* snum;    // [B*N] atomic numbers in sorted order
  double* spos;     // [B*N*3] positions in sorted order
  double* sq;       // [B*N] given charges in sorted order
  double* sr4;      // [B*N] given r4r2 in sorted order
  double* sge;      // [B*N] dL/dE in sorted order
  double* sE;       // [B*N] per-atom outputs in sorted order
  double* sG;       // [B*N*3] gradient in sorted order
  double* gsph;     // [B*NG*4] bounding sphere (centre, radius) of each 16-atom group
  double* gsph8;    // [B*NG8*4] bounding sphere of each 8-atom group
  int64_t* wix;     // [B*(2*NG8+3)] slot tables of the screened one-pass kernel
  double* apart;    // [B*S*5*N] per-split atom sums of the one-pass three-body kernel
  double* rt4;      // [B*N] sqrt(r4r2) per atom (parameter gradients)
  int32_t* zslot;   // [B*kMaxZ] element -> slot in the molecule's element list (-1: none)
  int32_t* nzs;     // [B] distinct elements of the molecule, -1 if more than kZSlots
  int32_t* zlist;   // [B*kZSlots] slot -> element
  double* cv;       // [B*N*kZSlots*8] sum_b RC6[Z_e][Z_j][a][b] gw_j[b] per atom j, slot e
  double* cv0;      // the same with the weights at q = 0
  double* pg2;      // [B*N*5] two-body parameter-gradient partials per atom row
  double* pg3;      // [B*P*4] three-body parameter-gradient partials per work-group
  int32_t* eqord;   // [B] molecule order of the per-molecule EEQ kernels (largest first)
  double* bbox2;    // [B] squared diagonal of a molecule's bounding box (N <= kAtm1MaxN)
  uint64_t* gnear;  // [B*NG8*W8] near bits of 8-atom group pairs (screened kernel)
  int64_t B, N;
};

constexpr int kGroup = 16;  // atoms per group (the three-body tile edge)
__host__ __device__ inline int64_t n_groups(int64_t N) { return (N + kGroup - 1) / kGroup; }

// One-pass three-body kernel (molecules of up to kAtm1MaxN atoms): a
// molecule of n atoms is split over atm1_splits(n) work-groups, a number that
// depends on n alone, so a molecule's result does not depend on its batch.
constexpr int kAtm1MaxN = 256;
constexpr int kAtm1Acc = 5;  // energy share, gradient (3), dL/dCN
__host__ __device__ inline int atm1_splits(int n);
// The 8 x 8 variant (one wave per work-group, kAtm8-atom blocks; the
// gradient passes) splits a molecule over atm8_splits(n) waves, also a
// function of n alone: about one 40-atom molecule of work (n^3 = 65536) each,
// at most 64 (each split writes a partial sum per atom). On a 1024 x 20-120
// batch, 32768 per split gave the same kernel time alone but a 2.5 % slower
// two-stream step (twice the waves and partial sums); 49152 and 81920 or
// more were slower.
constexpr int kAtm8 = 8;
constexpr int kAtm8MaxSplits = 64;
__host__ __device__ inline int atm8_splits(int n) {
  const int nb = (n + kAtm8 - 1) / kAtm8;
  const int tiles = nb * (nb + 1) / 2;
  const long long work = (long long)n * n * n;
  int s = (int)((work + 65535) / 65536);
  s = s > kAtm8MaxSplits ? kAtm8MaxSplits : s;
  return tiles < 1 ? 1 : (s < 1 ? 1 : (s > tiles ? tiles : s));
}

// Screened one-pass three-body kernel (large systems, atoms in spatial
// order): 8-atom groups; x-group gx of a molecule with G groups is split over
// atm8s_splits(G, gx, smax) work items, about smax ((G - gx) / G)^2 (the
// work of gx falls with its forward neighbourhood), smax = atm8s_smax(n) at
// most kAtm8sMaxSplits. Item (gx, s) has one window slot (5 x 8 doubles) for
// each group g >= gx; the slots of a molecule fill at most atm8s_budget(n)
// doubles (12 n^2 from kAtm1MaxN to 16384 atoms, else 3 n^2: the pair-matrix
// space pg, ph, pe is allocated atm8s_stride(N) doubles per molecule), and
// smax is the largest power of two that fits. All of it depends on the atom
// count n alone.
constexpr int kAtm8sMaxSplits = 64;
__host__ __device__ inline int64_t n_groups8(int64_t N) { return (N + kAtm8 - 1) / kAtm8; }
// 64-bit words of one row of the group near bits
__host__ __device__ inline int64_t near_words8(int64_t N) { return (n_groups8(N) + 63) / 64; }
__host__ __device__ inline int atm8s_splits(int G, int gx, int smax) {
  const long long r = G - gx, GG = (long long)G * G;
  const long long s = (smax * r * r + GG - 1) / GG;
  return s < 1 ? 1 : (s > smax ? smax : (int)s);
}
__host__ __device__ inline long long atm8s_slots(int G, int smax) {
  long long t = 0;
  for (int gx = 0; gx < G; ++gx) t += (long long)atm8s_splits(G, gx, smax) * (G - gx);
  return t;
}
constexpr int64_t kAtm8sWideMax = 16384;
__host__ __device__ inline int64_t atm8s_budget(int64_t n) {
  const int64_t a = (n > kAtm1MaxN && n <= kAtm8sWideMax ? 12 : 3) * n * n;
  return a > 48 ? a : 48;
}
// the largest budget of a molecule of up to N atoms
__host__ __device__ inline int64_t atm8s_stride(int64_t N) {
  const int64_t a = atm8s_budget(N), b = atm8s_budget(N < kAtm8sWideMax ? N : kAtm8sWideMax);
  return a > b ? a : b;
}
__host__ __device__ inline int atm8s_smax(int n) {
  const int G = (n + kAtm8 - 1) / kAtm8;
  for (int s = kAtm8sMaxSplits; s > 1; s >>= 1)
    if (atm8s_slots(G, s) * kAtm1Acc * kAtm8 <= atm8s_budget(n)) return s;
  return 1;
}
// Work-groups per molecule (x-groups padded to a multiple of 8, times the
// largest split count) and the length of a molecule's slot table.
__host__ __device__ inline int64_t atm8s_blocks_per_mol(int64_t N) {
  return (n_groups8(N) + 7) / 8 * 8 * kAtm8sMaxSplits;
}
__host__ __device__ inline int64_t atm8s_table_len(int64_t N) { return 2 * n_groups8(N) + 3; }

// Three-body parameter-gradient partials per molecule: one per split of the
// one-pass kernels or one per pair tile of the pair-tile kernel (the largest).
__host__ __device__ inline int64_t pgrad_partials(int64_t N) {
  const int64_t nb = n_groups(N);
  const int64_t tiles = nb * (nb + 1) / 2;
  int64_t s = N <= kAtm1MaxN ? atm1_splits((int)N) : 0;
  const int64_t s8 = N <= kAtm1MaxN ? atm8_splits((int)N) : 0;
  s = s > s8 ? s : s8;
  s = s > tiles ? s : tiles;
  const int64_t s8s = atm8s_blocks_per_mol(N);
  return s > s8s ? s : s8s;
}
__host__ __device__ inline int atm1_splits(int n) {
  const int nb = (n + kGroup - 1) / kGroup;
  const int tiles = nb * (nb + 1) / 2;
  const long long work = (long long)n * n * n;
  const int s = (int)((work + 262143) / 262144);  // about one 64-atom molecule each
  return tiles < 1 ? 1 : (s < 1 ? 1 : (s > tiles ? tiles : s));
}
__host__ __device__ inline int one_pass_split_max(int n) {
  const int a = atm1_splits(n), b = atm8_splits(n);
  return a > b ? a : b;
}

// Small molecules take the one-pass three-body kernel.
inline bool one_pass(const Params& p, const Work& w) {
  return w.N <= kAtm1MaxN && !p.force_pairtiles;
}

// Splits of a molecule of n atoms in the one-pass kernel in use.
__host__ __device__ inline int one_pass_splits(int n, bool tile8) {
  return tile8 ? atm8_splits(n) : atm1_splits(n);
}

// The gradient passes use the 8 x 8 one-pass kernel, energies alone the
// 16 x 16 one (fewer, larger work-groups suit the lighter energy slots).
inline bool one_pass_tile8(const Params& p, bool grad) {
  return grad ? !p.atm16 : p.atm8 != 0;
}

inline size_t align_up(size_t x) { return (x + 255) & ~size_t(255); }

// Layout of the workspace; returns the total size. With base == nullptr only
// the size is computed.
inline size_t layout(int64_t B, int64_t N, char* base, Work* w) {
  size_t off = 0;
  auto take = [&](size_t bytes) -> char* {
    char* p = base ? base + off : nullptr;
    off = align_up(off + bytes);
    return p;
  };
  const size_t bn = size_t(B) * size_t(N);
  const size_t nt = size_t(N) * size_t(N + 1) / 2;
  const size_t n1 = size_t(N + 1);
  Work v{};
  v.B = B;
  v.N = N;
  v.status = (int32_t*)take(sizeof(int32_t) * B);
  v.natom = (int32_t*)take(sizeof(int32_t) * B);
  v.cn = (double*)take(8 * bn);
  v.cnq = (double*)take(8 * bn);
  v.dcnq = (double*)take(8 * bn);
  v.q = (double*)take(8 * bn);
  v.e2 = (double*)take(8 * bn);
  v.dldcn = (double*)take(8 * bn);
  v.dldq = (double*)take(8 * bn);
  v.beta = (double*)take(8 * bn);
  v.lam = (double*)take(8 * bn);
  v.gw = (double*)take(8 * bn * kRefStride);
  v.gw0 = (double*)take(8 * bn * kRefStride);
  v.dgw = (double*)take(8 * bn * kRefStride);
  v.dgw0 = (double*)take(8 * bn * kRefStride);
  v.dgwq = (double*)take(8 * bn * kRefStride);
  v.chol = (double*)take(8 * size_t(B) * nt);
  v.ez = (double*)take(8 * bn);
  v.lu = (double*)take(8 * size_t(B) * n1 * n1);
  v.piv = (int32_t*)take(sizeof(int32_t) * size_t(B) * n1);
  v.vec = (double*)take(8 * size_t(B) * n1);
  v.pt = (double*)take(8 * bn * size_t(N));
  v.pp = (double*)take(8 * bn * size_t(N));
  v.pk = (double*)take(8 * bn * size_t(N));
  {
    // pg, ph and pe also hold the window buffer of the screened one-pass
    // kernel: atm8s_stride(N) doubles per molecule from pg on
    const size_t nn = size_t(N) * size_t(N), st = size_t(atm8s_stride(N));
    v.pg = (double*)take(8 * size_t(B) * (st > 3 * nn ? st - 2 * nn : nn));
  }
  v.ph = (double*)take(8 * bn * size_t(N));
  v.pe = (double*)take(8 * bn * size_t(N));
  v.perm = (int32_t*)take(sizeof(int32_t) * bn);
  v.snum = (int64_t*)take(8 * bn);
  v.spos = (double*)take(24 * bn);
  v.sq = (double*)take(8 * bn);
  v.sr4 = (double*)take(8 * bn);
  v.sge = (double*)take(8 * bn);
  v.sE = (double*)take(8 * bn);
  v.sG = (double*)take(24 * bn);
  v.gsph = (double*)take(32 * size_t(B) * size_t(n_groups(N)));
  v.gsph8 = (double*)take(32 * size_t(B) * size_t(n_groups8(N)));
  v.wix = (int64_t*)take(8 * size_t(B) * size_t(atm8s_table_len(N)));
  v.apart = (double*)take(N <= kAtm1MaxN
                              ? 8 * size_t(B) * size_t(one_pass_split_max((int)N)) * kAtm1Acc *
                                    size_t(N)
                              : 0);
  v.rt4 = (double*)take(8 * bn);
  v.zslot = (int32_t*)take(4 * size_t(B) * kMaxZ);
  v.nzs = (int32_t*)take(4 * size_t(B));
  v.zlist = (int32_t*)take(4 * size_t(B) * kZSlots);
  v.cv = (double*)take(8 * bn * kZSlots * 8);
  v.cv0 = (double*)take(8 * bn * kZSlots * 8);
  v.pg2 = (double*)take(8 * bn * 5);
  v.pg3 = (double*)take(32 * size_t(B) * size_t(pgrad_partials(N)));
  v.eqord = (int32_t*)take(4 * size_t(B));
  v.bbox2 = (double*)take(8 * size_t(B));
  // (every N: the screened kernel also runs for small N when pair tiles are forced)
  v.gnear = (uint64_t*)take(8 * size_t(B) * size_t(n_groups8(N)) * size_t(near_words8(N)));
  if (w) *w = v;
  return off;
}

// Inputs of one call (device pointers).
struct Inputs {
  const int64_t* numbers;
  const double* positions;
  const double* charge;
  const double* q_in;     // optional
  const double* r4r2_in;  // optional
  const double* gE;       // optional dL/dE_i (nullptr: ones)
};

// Launchers (d4_kernels.hip). All enqueue on `s` and return hipError_t.
hipError_t launch_prep(hipStream_t s, const Inputs& in, const Work& w);
// Raise `kernel`'s dynamic-LDS limit to at least `bytes` on the current
// device (once per kernel, device and size; d4_eeq.hip).
hipError_t ensure_lds(const void* kernel, size_t bytes);
hipError_t launch_cn(hipStream_t s, const DevTables& t, const Params& p,
                     const Inputs& in, const Work& w);
// `paired`: where a molecule's factor fills a CU's LDS, two molecules per
// work-group (half the CUs), for evaluations whose charge chain is shorter
// than the three-body chain beside it (energies alone)
hipError_t launch_eeq(hipStream_t s, const DevTables& t, const Params& p,
                      const Inputs& in, const Work& w, bool paired);
// Parts of the weights and pair passes: the three-body side (weights and C6
// at q = 0, independent of the EEQ charges), the two-body side (at the EEQ
// charges) or both.
constexpr int kPairsBoth = 0;
constexpr int kPairsAtm = 1;
constexpr int kPairsTwoBody = 2;
hipError_t launch_weights(hipStream_t s, const DevTables& t, const Params& p,
                          const Work& w, const Inputs& in, int part);
// Element slots of every molecule (after launch_prep), and the C6 contraction
// vectors of one part (after its weights; D4 model only).
hipError_t launch_zslots(hipStream_t s, const Inputs& in, const Work& w);
hipError_t launch_pairs(hipStream_t s, const DevTables& t, const Params& p,
                        const Inputs& in, const Work& w, bool grad,
                        double* grad_out, int part);
hipError_t launch_atm(hipStream_t s, const Params& p, const Inputs& in,
                      const Work& w, bool energy, bool grad);
hipError_t launch_atm_rows(hipStream_t s, const Params& p, const Inputs& in, const Work& w,
                           bool energy, bool grad, double* energy_out,
                           double* grad_out);
hipError_t launch_adjoint(hipStream_t s, const DevTables& t, const Params& p,
                          const Inputs& in, const Work& w);
hipError_t launch_chain(hipStream_t s, const DevTables& t, const Params& p,
                        const Inputs& in, const Work& w, double* grad_out);
hipError_t launch_properties(hipStream_t s, const DevTables& t,
                             const Inputs& in, const Work& w, double* c6_out,
                             double* alpha_out);
hipError_t launch_copy(hipStream_t s, const double* src, double* dst, int64_t count);
// Fill t.pairk (kMaxZ^2 x 4 doubles at `out`) from the per-element tables.
hipError_t launch_pair_consts(hipStream_t s, const DevTables& t, double* out);
// Sum of the parameter-gradient partials into out[7] = dL/d(s6, s8, s10, s9,
// a1, a2, alp) (device pointer).
hipError_t launch_pgrad_reduce(hipStream_t s, const Params& p, const Inputs& in,
                               const Work& w, double* out);

// Spatial order (d4_sort.hip). launch_sort computes w.perm (only when
// `sort`); launch_gather fills the sorted copies of the inputs that are
// non-null; the scatters write sorted-order results back in input order.
hipError_t launch_sort(hipStream_t s, const Inputs& in, const Work& w);
hipError_t launch_gather(hipStream_t s, const Inputs& in, const Work& w);
hipError_t launch_scatter_atoms(hipStream_t s, const Work& w, const double* src,
                                double* dst, int comps);
hipError_t launch_scatter_pairs(hipStream_t s, const Work& w, const double* src,
                                double* dst);
hipError_t launch_groups(hipStream_t s, const Inputs& in, const Work& w);
hipError_t launch_groups8(hipStream_t s, const Inputs& in, const Work& w);

// Dispersion-model methods (d4_model.hip): weights from given CN and charges
// (forward, or with grad_gw the backward to dL/dcn, dL/dq), C6 from given
// weights and its backward, weighted (23 frequencies) or static
// polarisabilities and their backward.
hipError_t launch_model_weights(hipStream_t s, const DevTables& t, const ModelCall& mc,
                                const double* cn, const double* q, double* gw_out,
                                double* dgwdcn_out, double* dgwdq_out, const double* grad_gw,
                                double* grad_cn, double* grad_q);
hipError_t launch_model_c6(hipStream_t s, const DevTables& t, const ModelCall& mc,
                           const double* gw, double* c6_out);
hipError_t launch_model_c6_bwd(hipStream_t s, const DevTables& t, const ModelCall& mc,
                               const double* gw, const double* grad_c6, double* grad_gw);
hipError_t launch_model_pols(hipStream_t s, const DevTables& t, const ModelCall& mc,
                             int static_only, const double* gw, double* out);
hipError_t launch_model_pols_bwd(hipStream_t s, const DevTables& t, const ModelCall& mc,
                                 int static_only, const double* grad_out, double* grad_gw);

}  // namespace d4
