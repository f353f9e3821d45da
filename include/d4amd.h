/*
 * d4amd: C ABI of the MI355X (gfx950) DFT-D4 energy and gradient kernels.
 *
 * Plain pointers and sizes only. Every device pointer argument is memory
 * that the caller owns (allocated by the caller, e.g. by PyTorch); the
 * library owns only its per-device context (the element tables resident in
 * HBM). Streams are passed as `void*` holding a `hipStream_t`; every entry
 * point only enqueues work on that stream (no host synchronisation), so a
 * caller may capture the calls in a HIP graph.
 *
 * Batch convention (as in the reference): `numbers` is (nbatch, natoms)
 * int64 with Z = 0 marking padding atoms, `positions` is (nbatch, natoms, 3)
 * float64 in bohr, `charge` is (nbatch,) float64. Outputs for padding atoms
 * are zero.
 *
 * Errors: every function returns D4_OK (0) or a negative D4_ERR_* code and
 * never throws; d4_last_error() returns a thread-local message. Checks that
 * need the contents of device buffers are made on the device and do not fail
 * the call (the library never synchronises to read them): a molecule with an
 * atomic number outside 0..103 gets zero outputs and status bit 2, a
 * singular EEQ system status bit 1 (d4_molecule_status); callers that cannot
 * vouch for their inputs read the status words after the call.
 */
#ifndef D4AMD_H
#define D4AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define D4_OK 0
#define D4_ERR_INVALID (-1)   /* bad argument: null pointer, size, flags */
#define D4_ERR_DEVICE (-2)    /* HIP runtime error */
#define D4_ERR_NOMEM (-3)     /* device allocation failed */
#define D4_ERR_WORKSPACE (-4) /* workspace smaller than d4_workspace_bytes */

#define D4_MAX_Z 104   /* element slots 0..103 (0 = padding) */
#define D4_MAX_REF 7   /* reference systems per element */

/* d4_params.flags */
#define D4_FLAG_S10 1      /* "s10" key present: evaluate the order-10 term */
#define D4_FLAG_FORCE_LU 2    /* testing: solve every EEQ system by pivoted LU */
#define D4_FLAG_FORCE_TILED 4 /* testing: use the large-system (tiled) EEQ path at any size */
#define D4_FLAG_FORCE_SORT 8  /* testing: spatially sort the atoms of every molecule */
#define D4_FLAG_FORCE_PAIR_TILES 16 /* testing: the large-system three-body kernel at any size */
#define D4_FLAG_D4S 32     /* D4S model: Gaussian weights with pairwise factors
                              (reference src/tad_dftd4/model/d4s.py:56-285) */
#define D4_FLAG_ATM16 64   /* gradient calls: one-pass three-body kernel with 16 x 16 pair
                              tiles (default 8 x 8; energy-only calls use 16 x 16) */
#define D4_FLAG_PAIR_OWNER 128 /* testing: large systems take the pair-owner three-body
                                  kernel (each triple from its three pairs) */
/* Launch shapes (results agree to rounding; used by the shape tuner,
 * tad_dftd4_amd/tune.py). A two-bit field selects the lanes per atom row of
 * the CN, pair and chain passes; 0 chooses them from the padded length. */
#define D4_FLAG_ROWS_16 256
#define D4_FLAG_ROWS_32 512
#define D4_FLAG_ROWS_64 768
#define D4_FLAG_ROWS_MASK 768
#define D4_FLAG_ATM8 1024  /* energy-only calls: one-pass three-body kernel with 8 x 8
                              pair tiles (default 16 x 16; gradient calls use 8 x 8
                              unless D4_FLAG_ATM16) */
#define D4_FLAG_ONE_STREAM 2048 /* testing/profiling: run the three-body side on the
                                   caller's stream too (no companion stream), so a
                                   kernel trace shows each kernel's own time */

/* Element and reference-system tables, host memory, read once by
 * d4_context_create. Per-element arrays have D4_MAX_Z entries; per-reference
 * arrays have D4_MAX_Z * D4_MAX_REF entries with the used reference systems
 * of each element compacted to the front (nref[z] of them). */
typedef struct d4_tables {
  const double* rcov;  /* covalent radii (bohr) for the CN counting functions */
  const double* en;    /* Pauling electronegativities */
  const double* r4r2;  /* sqrt(1/2 <r4>/<r2> sqrt(Z)) */
  const double* eeq_chi;
  const double* eeq_eta;
  const double* eeq_kcn;
  const double* eeq_rad;
  const double* gam;   /* chemical hardness (multiplied by gc in the kernels) */
  const double* zeff;  /* effective nuclear charge */
  const double* maxcn; /* max reference CN over all 7 slots (exceptional weights) */
  const int32_t* nref; /* used reference systems per element */
  const int32_t* ngw;  /* Gaussians per reference (1 or 3) */
  const double* refcn; /* reference coordination numbers */
  const double* refq;  /* reference charges */
  const double* alpha0; /* static reference polarisabilities alpha(i*0) */
  const double* rc6;   /* [zi][zj][a][b] reference C6, D4_MAX_Z^2 * 49 */
  const double* wfpair; /* [zi][zj] D4S Gaussian weighting factors, D4_MAX_Z^2
                           (src/tad_dftd4/data/wfpair.py); NULL: no D4S */
  /* Needed by the d4_model_* functions only (NULL: those return
   * D4_ERR_INVALID): */
  const double* alpha;    /* reference polarisabilities alpha(i w) at the 23
                             Casimir-Polder frequencies, D4_MAX_Z * D4_MAX_REF * 23,
                             compacted like the other per-reference arrays */
  const int32_t* refslot; /* D4_MAX_Z * D4_MAX_REF: the reference's slot (0..6) of
                             each compacted reference system, -1 past nref[z] */
} d4_tables;

/* Damping parameters, model constants and cutoffs. Absent optional terms
 * are passed as zero (s8: order-8 term off; s10 with D4_FLAG_S10 clear). */
typedef struct d4_params {
  double s6, s8, s10, s9;
  double a1, a2, alp;
  double ga, gc, wf;
  double cutoff_disp2; /* two-body cutoff (bohr) */
  double cutoff_disp3; /* three-body cutoff (bohr) */
  double cutoff_cn;    /* D4 coordination number cutoff (bohr) */
  double cutoff_eeq;   /* EEQ coordination number cutoff (bohr) */
  int32_t flags;
  int32_t reserved;
} d4_params;

typedef struct d4_context d4_context;

/* Create a context on `device` and upload the tables to its HBM.
 * Replaces the per-call table build of D4Model (reference
 * src/tad_dftd4/model/base.py:107-151, model/base.py:367-431). */
int d4_context_create(int device, const d4_tables* tables, d4_context** out);

/* Release the context and its device tables. */
int d4_context_destroy(d4_context* ctx);

/* Bytes of device workspace needed for a batch of nbatch molecules padded to
 * natoms atoms (forward, backward and the fused energy+gradient path share
 * one layout; the forward leaves in it what the backward reads). */
int d4_workspace_bytes(int64_t nbatch, int64_t natoms, size_t* bytes);

/* Atom-resolved dispersion energy, energy_out (nbatch, natoms).
 * Replaces dftd4() (reference src/tad_dftd4/disp.py:44-146) and
 * Disp.calculate (src/tad_dftd4/dispersion/base.py:285-431).
 * q_in: optional (nbatch, natoms) charges (NULL: EEQ charges, as
 *   dispersion/base.py:401-407); r4r2_in: optional per-atom r4r2 values
 *   (NULL: element table, as dispersion/base.py:373-374).
 * q_out: optional, receives the charges used. */
int d4_forward(d4_context* ctx, void* stream, int64_t nbatch, int64_t natoms,
               const int64_t* numbers, const double* positions,
               const double* charge, const double* q_in, const double* r4r2_in,
               const d4_params* params, double* energy_out, double* q_out,
               void* workspace, size_t workspace_bytes);

/* Gradient of L = sum_i grad_energy[i] * E_i with respect to the positions,
 * written to grad_positions (nbatch, natoms, 3). Must follow d4_forward with
 * the same inputs and workspace. grad_energy NULL means all ones.
 * grad_params: optional device array of 7 doubles that receives dL/ds6,
 * dL/ds8, dL/ds10, dL/ds9, dL/da1, dL/da2, dL/dalp (summed over the batch;
 * the s8 and s10 entries are meaningful when those terms are on).
 * grad_q: optional (nbatch, natoms), receives dL/dq for caller-supplied
 * charges (requires q_in; zero on padding atoms).
 * Replaces the autograd backward of dftd4() with respect to the positions
 * (reference test/test_grad/test_pos.py:185-187), the damping parameters
 * (test/test_grad/test_param.py:40-86) and given charges (the reference
 * differentiates through q, src/tad_dftd4/dispersion/base.py:392-407 ->
 * model/d4.py:219-228). */
int d4_backward(d4_context* ctx, void* stream, int64_t nbatch, int64_t natoms,
                const int64_t* numbers, const double* positions,
                const double* charge, const double* q_in,
                const double* r4r2_in, const d4_params* params,
                const double* grad_energy, double* grad_positions,
                double* grad_params, double* grad_q, void* workspace,
                size_t workspace_bytes);

/* Fused atom energies and the gradient of the total energy in one pass.
 * Replaces dftd4() followed by torch.autograd.grad(energy.sum(), positions)
 * (reference examples/forces.py:47-50). */
int d4_energy_gradient(d4_context* ctx, void* stream, int64_t nbatch,
                       int64_t natoms, const int64_t* numbers,
                       const double* positions, const double* charge,
                       const double* q_in, const double* r4r2_in,
                       const d4_params* params, double* energy_out,
                       double* grad_positions, void* workspace,
                       size_t workspace_bytes);

/* d4_energy_gradient without a workspace argument and without any device
 * allocation: the workspace is a fixed scratch array of
 * D4_NOALLOC_SCRATCH_BYTES that the library holds on every device it runs on
 * until the process ends (its address is looked up at d4_context_create).
 * Batches whose d4_workspace_bytes exceeds it are rejected with
 * D4_ERR_INVALID. Calls on one GPU share the scratch: calls that overlap in
 * time on one GPU (other streams, other threads) race and must be
 * serialised by the caller. No counterpart in the reference: for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). */
#define D4_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int d4_energy_gradient_noalloc(d4_context* ctx, void* stream, int64_t nbatch,
                               int64_t natoms, const int64_t* numbers,
                               const double* positions, const double* charge,
                               const double* q_in, const double* r4r2_in,
                               const d4_params* params, double* energy_out,
                               double* grad_positions);

/* Per-atom properties: D4 coordination numbers, charges, the C6 matrix
 * (nbatch, natoms, natoms) and static polarisabilities (nbatch, natoms).
 * Replaces get_properties() (reference src/tad_dftd4/disp.py:149-197).
 * Any output pointer may be NULL; with only cn_out given, only the
 * coordination numbers are computed (tad-mctc's cn_d4, which dftd4() takes
 * as its default CN function, reference src/tad_dftd4/disp.py:30,56-57).
 * D4 model only (as the reference, which builds a D4Model there):
 * D4_FLAG_D4S is rejected. */
int d4_properties(d4_context* ctx, void* stream, int64_t nbatch,
                  int64_t natoms, const int64_t* numbers,
                  const double* positions, const double* charge,
                  const d4_params* params, double* cn_out, double* q_out,
                  double* c6_out, double* alpha_out, void* workspace,
                  size_t workspace_bytes);

/* Dispersion-model methods: the D4Model / D4SModel interface of the
 * reference on the device. `params` supplies ga, gc, wf and D4_FLAG_D4S (the
 * other fields are ignored); the context must hold `alpha` and `refslot`.
 * Weights are in the reference's slot order: (nbatch, natoms, 7), or for
 * D4S (nbatch, natoms, natoms, 7) with entry [m][n][r] the weight of atom
 * n's reference r in its pair with atom m. No workspace is needed.
 *
 * d4_model_weights: weights for coordination numbers `cn` and charges `q`
 * ((nbatch, natoms) each, NULL meaning zero, as the reference's defaults).
 * dgwdcn_out (optional, D4 only): -d gw / d cn, the reference's analytic
 * output, which carries the opposite sign of the derivative; dgwdq_out
 * (optional): d gw / d q. Replaces D4Model.weight_references
 * (src/tad_dftd4/model/d4.py:103-266) and D4SModel.weight_references
 * (model/d4s.py:109-263). */
int d4_model_weights(d4_context* ctx, void* stream, int64_t nbatch, int64_t natoms,
                     const int64_t* numbers, const double* cn, const double* q,
                     const d4_params* params, double* gw_out, double* dgwdcn_out,
                     double* dgwdq_out);

/* Vector-Jacobian product of d4_model_weights: dL/dcn and dL/dq
 * ((nbatch, natoms), either may be NULL) of L = sum grad_gw * gw. Replaces
 * autograd through weight_references (test/test_model/test_weights.py:
 * 157-282). */
int d4_model_weights_backward(d4_context* ctx, void* stream, int64_t nbatch, int64_t natoms,
                              const int64_t* numbers, const double* cn, const double* q,
                              const d4_params* params, const double* grad_gw,
                              double* grad_cn, double* grad_q);

/* C6 of every atom pair from weights, c6_out (nbatch, natoms, natoms).
 * Replaces D4Model.get_atomic_c6 (model/d4.py:268-289) and
 * D4SModel.get_atomic_c6 (model/d4s.py:265-285). */
int d4_model_atomic_c6(d4_context* ctx, void* stream, int64_t nbatch, int64_t natoms,
                       const int64_t* numbers, const double* gw, const d4_params* params,
                       double* c6_out);

/* Vector-Jacobian product of d4_model_atomic_c6: grad_gw_out (the weights'
 * shape) of L = sum grad_c6 * c6. */
int d4_model_atomic_c6_backward(d4_context* ctx, void* stream, int64_t nbatch,
                                int64_t natoms, const int64_t* numbers, const double* gw,
                                const d4_params* params, const double* grad_c6,
                                double* grad_gw_out);

/* Polarisabilities from weights. static_only = 0: weighted polarisabilities
 * at the 23 frequencies, out (nbatch, natoms, 23) (D4S: summed over the pair
 * axis and divided by the molecule's atom count, as the reference), replacing
 * get_weighted_pols (model/d4.py:291-306, model/d4s.py:287-313);
 * static_only = 1: static polarisabilities, out (nbatch, natoms) (D4S:
 * (nbatch, natoms, natoms), the reference's broadcast over the pair axis),
 * replacing BaseModel.get_polarizabilities (model/base.py:286-302). */
int d4_model_pols(d4_context* ctx, void* stream, int64_t nbatch, int64_t natoms,
                  const int64_t* numbers, const double* gw, const d4_params* params,
                  int static_only, double* out);

/* Vector-Jacobian product of d4_model_pols: grad_gw_out of L = sum grad_out * out. */
int d4_model_pols_backward(d4_context* ctx, void* stream, int64_t nbatch, int64_t natoms,
                           const int64_t* numbers, const d4_params* params, int static_only,
                           const double* grad_out, double* grad_gw_out);

/* Per-molecule status words of the last call that used `workspace`, copied
 * to host memory (synchronises the stream). Bit 0: EEQ solved by the
 * pivoted-LU path; bit 1: EEQ system singular; bit 2: Z out of range. */
int d4_molecule_status(void* stream, int64_t nbatch, const void* workspace,
                       int32_t* status_out);

/* Diagnostics (no counterpart in the reference): per-kernel HIP-event
 * timing on the launch stream. While enabled, every launch made through
 * `ctx` (on = 1), or every launch of the slots set in a mask (on =
 * D4_PROFILE_SLOTS | (1 << slot) | ...), is bracketed by events;
 * d4_profile_read synchronises the recorded events, adds the elapsed times
 * up per kernel slot (D4_SLOT_*) and resets the accumulators. */
#define D4_SLOT_PREP 0
#define D4_SLOT_CN 1
#define D4_SLOT_EEQ 2
#define D4_SLOT_WEIGHTS 3
#define D4_SLOT_PAIRS 4
#define D4_SLOT_ATM 5
#define D4_SLOT_ATM_ROWS 6
#define D4_SLOT_ADJOINT 7
#define D4_SLOT_CHAIN 8
#define D4_NUM_SLOTS 9
#define D4_PROFILE_SLOTS (1 << 16)
int d4_profile_enable(d4_context* ctx, int on);
int d4_profile_read(d4_context* ctx, double* ms_out, int64_t* count_out, int nslots);

/* Thread-local message describing the last error. */
const char* d4_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* D4AMD_H */
