// C ABI of the D4 kernels (include/d4amd.h). Validates arguments, owns the
// per-device table allocation and sequences the kernels on the caller's
// stream. No host synchronisation except in d4_molecule_status.

#include <string.h>

#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "d4_internal.h"

struct d4_event_pair {
  int slot;
  hipEvent_t start, stop;
};

// Companion stream of one caller stream: the three-body side (weights and
// pair factors at q = 0, the three-body kernel) runs on it while the caller's
// stream solves the EEQ system and evaluates the two-body term; two events
// fork and join it.
struct d4_aux {
  hipStream_t s = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
};

// Scratch of d4_energy_gradient_noalloc: the workspace of that entry point,
// a fixed array on every device the library runs on (held until the process
// ends). Host code never uses this name as an address: the device address of
// each GPU's copy comes from hipGetSymbolAddress at context creation.
constexpr size_t kScratchBytes = D4_NOALLOC_SCRATCH_BYTES;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

struct d4_context {
  int device;
  void* mem;
  void* scratch_dev = nullptr;  // this device's copy of `scratch`
  d4::DevTables t;
  bool has_wfpair = false;  // D4S factors supplied at creation
  bool has_model = false;   // alpha and refslot supplied (d4_model_* functions)
  std::mutex aux_mu;
  std::unordered_map<hipStream_t, d4_aux> aux;
  // diagnostics: per-kernel event timing
  unsigned profile = 0;  // bit k: bracket launches of slot k with events
  std::vector<d4_event_pair> pending;
  std::vector<hipEvent_t> pool;
  double prof_ms[D4_NUM_SLOTS] = {};
  int64_t prof_count[D4_NUM_SLOTS] = {};
};

namespace {

thread_local std::string g_error;

// device -> address of its copy of `scratch`, looked up once per device
std::mutex g_scratch_mu;
std::unordered_map<int, void*> g_scratch_addr;

int set_error(int code, const std::string& msg) {
  g_error = msg;
  return code;
}

int hip_error(hipError_t e, const char* where) {
  g_error = std::string(where) + ": " + hipGetErrorString(e);
  return D4_ERR_DEVICE;
}

#define D4_TRY(expr, where)                      \
  do {                                           \
    hipError_t _e = (expr);                      \
    if (_e != hipSuccess) return hip_error(_e, where); \
  } while (0)

d4::Params to_params(const d4_params* p) {
  d4::Params k{};
  k.s6 = p->s6;
  k.s8 = p->s8;
  k.s10 = p->s10;
  k.s9 = p->s9;
  k.a1 = p->a1;
  k.a2 = p->a2;
  k.alp = p->alp;
  k.ga = p->ga;
  k.gc = p->gc;
  k.wf = p->wf;
  k.cut2 = p->cutoff_disp2;
  k.cut3sq = p->cutoff_disp3 * p->cutoff_disp3;
  k.cutcn = p->cutoff_cn;
  k.cuteeq = p->cutoff_eeq;
  k.s10on = (p->flags & D4_FLAG_S10) ? 1 : 0;
  k.force_lu = (p->flags & D4_FLAG_FORCE_LU) ? 1 : 0;
  k.force_tiled = (p->flags & D4_FLAG_FORCE_TILED) ? 1 : 0;
  k.force_sort = (p->flags & D4_FLAG_FORCE_SORT) ? 1 : 0;
  k.force_pairtiles = (p->flags & D4_FLAG_FORCE_PAIR_TILES) ? 1 : 0;
  k.d4s = (p->flags & D4_FLAG_D4S) ? 1 : 0;
  k.atm16 = (p->flags & D4_FLAG_ATM16) ? 1 : 0;
  k.pair_owner = (p->flags & D4_FLAG_PAIR_OWNER) ? 1 : 0;
  const unsigned rows = ((unsigned)p->flags & D4_FLAG_ROWS_MASK) >> 8;
  k.row_lanes = rows == 0 ? 0 : 8 << rows;  // 1, 2, 3 -> 16, 32, 64
  k.atm8 = (p->flags & D4_FLAG_ATM8) ? 1 : 0;
  k.one_stream = (p->flags & D4_FLAG_ONE_STREAM) ? 1 : 0;
  return k;
}

int check_common(const d4_context* ctx, int64_t nbatch, int64_t natoms,
                 const int64_t* numbers, const double* positions,
                 const double* charge, const d4_params* params, void* ws,
                 size_t ws_bytes) {
  if (ctx == nullptr) return set_error(D4_ERR_INVALID, "context is NULL");
  if (params == nullptr) return set_error(D4_ERR_INVALID, "params is NULL");
  if (nbatch < 0 || natoms < 0)
    return set_error(D4_ERR_INVALID, "nbatch and natoms must be non-negative");
  if (natoms > 46000)
    return set_error(D4_ERR_INVALID, "natoms exceeds the supported maximum (46000)");
  if (nbatch == 0 || natoms == 0) return D4_OK;
  if (!numbers || !positions || !charge)
    return set_error(D4_ERR_INVALID, "numbers, positions and charge must be non-NULL");
  size_t need = 0;
  d4_workspace_bytes(nbatch, natoms, &need);
  if (ws == nullptr || ws_bytes < need)
    return set_error(D4_ERR_WORKSPACE, "workspace smaller than d4_workspace_bytes()");
  if (!(params->cutoff_disp2 >= 0.0) || !(params->cutoff_disp3 >= 0.0) ||
      !(params->cutoff_cn >= 0.0) || !(params->cutoff_eeq >= 0.0))
    return set_error(D4_ERR_INVALID, "cutoffs must be non-negative");
  if ((params->flags & D4_FLAG_D4S) && !ctx->has_wfpair)
    return set_error(D4_ERR_INVALID, "D4_FLAG_D4S needs the wfpair table at context creation");
  return D4_OK;
}

// Molecules of at least this padded length run in spatial (Morton) order.
constexpr int64_t kSortMinN = 256;

struct Call {
  d4::Work w;
  d4::Inputs in;
  d4::Params p;
  hipStream_t s;
  bool sorted;
  d4_aux aux;  // companion stream (null until fork)
};

Call make_call(void* stream, int64_t nbatch, int64_t natoms,
               const int64_t* numbers, const double* positions,
               const double* charge, const double* q_in, const double* r4r2_in,
               const d4_params* params, void* ws) {
  Call c;
  d4::layout(nbatch, natoms, (char*)ws, &c.w);
  c.in = d4::Inputs{numbers, positions, charge, q_in, r4r2_in, nullptr};
  c.p = to_params(params);
  c.s = (hipStream_t)stream;
  c.sorted = natoms >= kSortMinN || c.p.force_sort;
  return c;
}

// Sorted order: gather the inputs (sorting first unless `reuse` keeps the
// order a previous call left in the workspace) and point the call at the
// sorted copies.
hipError_t enter_sorted(Call& c, bool reuse) {
  if (!c.sorted) return hipSuccess;
  hipError_t e = hipSuccess;
  if (!reuse) e = d4::launch_sort(c.s, c.in, c.w);
  if (e == hipSuccess) e = d4::launch_gather(c.s, c.in, c.w);
  if (e != hipSuccess) return e;
  c.in.numbers = c.w.snum;
  c.in.positions = c.w.spos;
  if (c.in.q_in) c.in.q_in = c.w.sq;
  if (c.in.r4r2_in) c.in.r4r2_in = c.w.sr4;
  if (c.in.gE) c.in.gE = c.w.sge;
  return hipSuccess;
}

// Where a per-atom output is written by the kernels: the caller's buffer, or
// the sorted-order staging buffer that leave_sorted() scatters from.
double* out_atoms(const Call& c, double* user, double* staging) {
  return c.sorted ? staging : user;
}

hipError_t scatter_atoms(const Call& c, const double* src, double* dst, int comps) {
  if (!c.sorted || dst == nullptr) return hipSuccess;
  return d4::launch_scatter_atoms(c.s, c.w, src, dst, comps);
}

hipEvent_t take_event(d4_context* ctx) {
  if (!ctx->pool.empty()) {
    hipEvent_t e = ctx->pool.back();
    ctx->pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Launch through `fn`, bracketed by timing events when profiling is on.
template <class F>
hipError_t timed(d4_context* ctx, hipStream_t s, int slot, F&& fn) {
  if (!((ctx->profile >> slot) & 1u)) return fn();
  hipEvent_t a = take_event(ctx), b = take_event(ctx);
  if (!a || !b) return fn();
  hipError_t e = hipEventRecord(a, s);
  if (e != hipSuccess) return e;
  e = fn();
  if (e != hipSuccess) return e;
  e = hipEventRecord(b, s);
  if (e != hipSuccess) return e;
  ctx->pending.push_back({slot, a, b});
  return hipSuccess;
}

#define D4_RUN(slot, expr, where) \
  D4_TRY(timed(ctx, c.s, slot, [&]() { return (expr); }), where)
// the same on the companion stream (kernels launched on c.aux.s)
#define D4_RUN_AUX(slot, expr, where) \
  D4_TRY(timed(ctx, c.aux.s, slot, [&]() { return (expr); }), where)

// Companion stream of the call's stream, created on first use (on the
// context's device, default priority, not blocking on the null stream).
hipError_t companion(d4_context* ctx, hipStream_t s, d4_aux* out) {
  std::lock_guard<std::mutex> lock(ctx->aux_mu);
  auto it = ctx->aux.find(s);
  if (it != ctx->aux.end()) {
    *out = it->second;
    return hipSuccess;
  }
  int prev = 0;
  hipError_t e = hipGetDevice(&prev);
  if (e == hipSuccess) e = hipSetDevice(ctx->device);
  d4_aux a;
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&a.s, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&a.fork, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&a.join, hipEventDisableTiming);
  (void)hipSetDevice(prev);
  if (e != hipSuccess) return e;
  ctx->aux.emplace(s, a);
  *out = a;
  return hipSuccess;
}

// Start the companion stream after everything enqueued so far on c.s.
hipError_t fork_aux(d4_context* ctx, Call& c) {
  hipError_t e = companion(ctx, c.s, &c.aux);
  if (e == hipSuccess) e = hipEventRecord(c.aux.fork, c.s);
  if (e == hipSuccess) e = hipStreamWaitEvent(c.aux.s, c.aux.fork, 0);
  return e;
}

// Make c.s wait for the work enqueued on the companion stream.
hipError_t join_aux(Call& c) {
  hipError_t e = hipEventRecord(c.aux.join, c.aux.s);
  if (e == hipSuccess) e = hipStreamWaitEvent(c.s, c.aux.join, 0);
  return e;
}

// Sorted order, extents and coordination numbers on the call's stream.
int run_cn(d4_context* ctx, Call& c) {
  D4_RUN(D4_SLOT_PREP, enter_sorted(c, false), "spatial order");
  if (c.w.N > d4::kAtm1MaxN) {  // (smaller batches: one launch does all three)
    D4_RUN(D4_SLOT_PREP, d4::launch_prep(c.s, c.in, c.w), "prep");
    if (!c.p.d4s) D4_RUN(D4_SLOT_PREP, d4::launch_zslots(c.s, c.in, c.w), "element slots");
  }
  D4_RUN(D4_SLOT_CN, d4::launch_cn(c.s, ctx->t, c.p, c.in, c.w), "coordination numbers");
  return D4_OK;
}

// Everything up to the charges and all reference weights on one stream
// (properties).
int run_front(d4_context* ctx, Call& c) {
  int rc = run_cn(ctx, c);
  if (rc != D4_OK) return rc;
  D4_RUN(D4_SLOT_EEQ, d4::launch_eeq(c.s, ctx->t, c.p, c.in, c.w, false), "EEQ charges");
  D4_RUN(D4_SLOT_WEIGHTS, d4::launch_weights(c.s, ctx->t, c.p, c.w, c.in, d4::kPairsBoth),
         "reference weights");
  return D4_OK;
}

// The two halves of an evaluation after the coordination numbers:
//   companion: weights at q = 0 -> three-body pair factors -> three-body term
//   call's:    EEQ charges -> weights at q -> two-body term
// joined before the per-atom sums. `eeq` is false in the backward pass
// (the forward left charges and weights in the workspace). With `grad` the
// EEQ adjoint (which needs only dL/dq of the two-body term) also runs before
// the join.
int run_split(d4_context* ctx, Call& c, bool eeq, bool energy, bool grad, double* g_out) {
  // D4S forms the pair weights inside the pair pass. Up to 160 atoms one fused
  // pass on the call's stream, which forms them once, is faster (1024 x 20-120
  // atoms: 1.45 against 1.64 ms with gradients); for longer molecules the
  // three-body term dominates and the split below, which forms the q = 0 and
  // the q weights in separate passes but runs the three-body side beside the
  // charges, wins (256 x 8-200 atoms: 1.18 against 1.34 ms). Both give the
  // same bits.
  constexpr int64_t kD4sFusedMaxN = 160;
  if (c.p.d4s && c.w.N <= kD4sFusedMaxN) {
    if (eeq) {
      D4_RUN(D4_SLOT_EEQ, d4::launch_eeq(c.s, ctx->t, c.p, c.in, c.w, false), "EEQ charges");
      D4_RUN(D4_SLOT_WEIGHTS, d4::launch_weights(c.s, ctx->t, c.p, c.w, c.in, d4::kPairsBoth),
             "reference weights");
    }
    D4_RUN(D4_SLOT_PAIRS, d4::launch_pairs(c.s, ctx->t, c.p, c.in, c.w, grad, g_out, d4::kPairsBoth),
           "pair pass");
    // the three-body term and the EEQ adjoint both need only the pair pass
    if (grad && !c.p.one_stream) {
      D4_TRY(fork_aux(ctx, c), "fork companion stream");
      D4_RUN_AUX(D4_SLOT_ATM, d4::launch_atm(c.aux.s, c.p, c.in, c.w, energy, grad),
                 "three-body term");
      D4_RUN(D4_SLOT_ADJOINT, d4::launch_adjoint(c.s, ctx->t, c.p, c.in, c.w), "EEQ adjoint");
      D4_TRY(join_aux(c), "join companion stream");
      return D4_OK;
    }
    if (grad)
      D4_RUN(D4_SLOT_ADJOINT, d4::launch_adjoint(c.s, ctx->t, c.p, c.in, c.w), "EEQ adjoint");
    D4_RUN(D4_SLOT_ATM, d4::launch_atm(c.s, c.p, c.in, c.w, energy, grad), "three-body term");
    return D4_OK;
  }
  if (c.p.one_stream)
    c.aux.s = c.s;  // the three-body side in order on the caller's stream
  else
    D4_TRY(fork_aux(ctx, c), "fork companion stream");
  if (eeq) {
    D4_RUN_AUX(D4_SLOT_WEIGHTS, d4::launch_weights(c.aux.s, ctx->t, c.p, c.w, c.in, d4::kPairsAtm),
               "reference weights (q = 0)");
  }
  D4_RUN_AUX(D4_SLOT_PAIRS,
             d4::launch_pairs(c.aux.s, ctx->t, c.p, c.in, c.w, grad, nullptr, d4::kPairsAtm),
             "three-body pair factors");
  D4_RUN_AUX(D4_SLOT_ATM, d4::launch_atm(c.aux.s, c.p, c.in, c.w, energy, grad), "three-body term");
  // The three-body kernel's one-wave work-groups (two per SIMD) take every
  // slot they free, so a four-wave work-group launched on the other stream
  // behind them can wait for most of its run: the EEQ solve is one launch
  // (its pivoted-LU fallback runs inside it) and the two-body pass uses
  // one-wave work-groups.
  if (eeq) {
    // energies alone: the charge chain (EEQ, weights, two-body) is the shorter
    // one, so the solve may leave half the CUs to the three-body kernel
    D4_RUN(D4_SLOT_EEQ, d4::launch_eeq(c.s, ctx->t, c.p, c.in, c.w, !grad && !c.p.one_stream),
           "EEQ charges");
    D4_RUN(D4_SLOT_WEIGHTS, d4::launch_weights(c.s, ctx->t, c.p, c.w, c.in, d4::kPairsTwoBody),
           "reference weights");
  }
  D4_RUN(D4_SLOT_PAIRS, d4::launch_pairs(c.s, ctx->t, c.p, c.in, c.w, grad, g_out, d4::kPairsTwoBody),
         "two-body term");
  if (grad)
    D4_RUN(D4_SLOT_ADJOINT, d4::launch_adjoint(c.s, ctx->t, c.p, c.in, c.w), "EEQ adjoint");
  if (!c.p.one_stream) D4_TRY(join_aux(c), "join companion stream");
  return D4_OK;
}

}  // namespace

extern "C" {

int d4_context_create(int device, const d4_tables* tb, d4_context** out) {
  if (out == nullptr) return set_error(D4_ERR_INVALID, "out is NULL");
  *out = nullptr;
  if (tb == nullptr) return set_error(D4_ERR_INVALID, "tables is NULL");
  const double* dsrc[] = {tb->rcov,    tb->en,      tb->r4r2,   tb->eeq_chi,
                          tb->eeq_eta, tb->eeq_kcn, tb->eeq_rad, tb->gam,
                          tb->zeff,    tb->maxcn};
  for (const double* p : dsrc)
    if (p == nullptr) return set_error(D4_ERR_INVALID, "a per-element table is NULL");
  if (!tb->nref || !tb->ngw || !tb->refcn || !tb->refq || !tb->alpha0 || !tb->rc6)
    return set_error(D4_ERR_INVALID, "a reference table is NULL");

  const size_t ne = D4_MAX_Z, nr = size_t(D4_MAX_Z) * D4_MAX_REF;
  const size_t nrc6 = size_t(D4_MAX_Z) * D4_MAX_Z * D4_MAX_REF * D4_MAX_REF;
  const size_t nwf = size_t(D4_MAX_Z) * D4_MAX_Z;
  const size_t npk = 4 * nwf;  // element-pair constants, filled on the device
  const size_t nal = nr * 23;  // alpha(i w) of every compacted reference system
  const size_t ndbl = 10 * ne + 3 * nr + nrc6 + nwf + npk + nal;
  const size_t nint = ne + 3 * nr;  // nref, ngw, refslot, refcomp
  std::vector<double> hd;
  hd.reserve(ndbl);
  for (const double* p : dsrc) hd.insert(hd.end(), p, p + ne);
  hd.insert(hd.end(), tb->refcn, tb->refcn + nr);
  hd.insert(hd.end(), tb->refq, tb->refq + nr);
  hd.insert(hd.end(), tb->alpha0, tb->alpha0 + nr);
  hd.insert(hd.end(), tb->rc6, tb->rc6 + nrc6);
  if (tb->wfpair)
    hd.insert(hd.end(), tb->wfpair, tb->wfpair + nwf);
  else
    hd.insert(hd.end(), nwf, 0.0);
  hd.insert(hd.end(), npk, 0.0);
  const bool model = tb->alpha != nullptr && tb->refslot != nullptr;
  if (model)
    hd.insert(hd.end(), tb->alpha, tb->alpha + nal);
  else
    hd.insert(hd.end(), nal, 0.0);
  std::vector<int32_t> hi;
  hi.insert(hi.end(), tb->nref, tb->nref + ne);
  hi.insert(hi.end(), tb->ngw, tb->ngw + nr);
  for (size_t z = 0; z < ne; ++z)
    if (hi[z] < 0 || hi[z] > D4_MAX_REF)
      return set_error(D4_ERR_INVALID, "nref out of range");
  // refslot (compact -> reference slot) and its inverse refcomp
  std::vector<int32_t> slot(nr, -1), comp(nr, -1);
  if (model) {
    for (size_t z = 0; z < ne; ++z)
      for (int c = 0; c < D4_MAX_REF; ++c) {
        const int32_t r = tb->refslot[z * D4_MAX_REF + c];
        if (c < hi[z]) {
          if (r < 0 || r >= D4_MAX_REF || comp[z * D4_MAX_REF + r] != -1)
            return set_error(D4_ERR_INVALID, "refslot is not a map onto distinct slots");
          slot[z * D4_MAX_REF + c] = r;
          comp[z * D4_MAX_REF + r] = c;
        }
      }
  }
  hi.insert(hi.end(), slot.begin(), slot.end());
  hi.insert(hi.end(), comp.begin(), comp.end());

  int prev = 0;
  D4_TRY(hipGetDevice(&prev), "hipGetDevice");
  D4_TRY(hipSetDevice(device), "hipSetDevice");
  void* mem = nullptr;
  const size_t bytes = sizeof(double) * ndbl + sizeof(int32_t) * nint;
  hipError_t e = hipMalloc(&mem, bytes);
  if (e != hipSuccess) {
    (void)hipSetDevice(prev);
    return set_error(D4_ERR_NOMEM, std::string("hipMalloc tables: ") + hipGetErrorString(e));
  }
  e = hipMemcpy(mem, hd.data(), sizeof(double) * ndbl, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy((char*)mem + sizeof(double) * ndbl, hi.data(), sizeof(int32_t) * nint,
                  hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipSetDevice(prev);
    (void)hipFree(mem);
    return hip_error(e, "upload tables");
  }
  void* scratch_dev = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    auto it = g_scratch_addr.find(device);
    if (it != g_scratch_addr.end()) {
      scratch_dev = it->second;
    } else {
      // `device` is current here
      e = hipGetSymbolAddress(&scratch_dev, HIP_SYMBOL(scratch));
      if (e != hipSuccess || scratch_dev == nullptr) {
        (void)hipSetDevice(prev);
        (void)hipFree(mem);
        return hip_error(e != hipSuccess ? e : hipErrorInvalidSymbol,
                         "hipGetSymbolAddress(scratch)");
      }
      g_scratch_addr.emplace(device, scratch_dev);
    }
  }
  d4_context* ctx = new d4_context;
  ctx->device = device;
  ctx->mem = mem;
  ctx->scratch_dev = scratch_dev;
  ctx->has_wfpair = tb->wfpair != nullptr;
  ctx->has_model = model;
  const double* d = (const double*)mem;
  d4::DevTables& t = ctx->t;
  t.rcov = d + 0 * ne;
  t.en = d + 1 * ne;
  t.r4r2 = d + 2 * ne;
  t.chi = d + 3 * ne;
  t.eta = d + 4 * ne;
  t.kcn = d + 5 * ne;
  t.rad = d + 6 * ne;
  t.gam = d + 7 * ne;
  t.zeff = d + 8 * ne;
  t.maxcn = d + 9 * ne;
  t.refcn = d + 10 * ne;
  t.refq = t.refcn + nr;
  t.alpha0 = t.refq + nr;
  t.rc6 = t.alpha0 + nr;
  t.wfpair = t.rc6 + nrc6;
  t.pairk = t.wfpair + nwf;
  t.alpha = t.pairk + npk;
  const int32_t* ip = (const int32_t*)(d + ndbl);
  t.nref = ip;
  t.ngw = ip + ne;
  t.refslot = t.ngw + nr;
  t.refcomp = t.refslot + nr;
  e = d4::launch_pair_consts(nullptr, t, const_cast<double*>(t.pairk));
  if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
  (void)hipSetDevice(prev);
  if (e != hipSuccess) {
    (void)hipFree(mem);
    delete ctx;
    return hip_error(e, "element-pair constants");
  }
  *out = ctx;
  return D4_OK;
}

int d4_context_destroy(d4_context* ctx) {
  if (ctx == nullptr) return D4_OK;
  int prev = 0;
  (void)hipGetDevice(&prev);
  (void)hipSetDevice(ctx->device);
  for (auto& kv : ctx->aux) {
    (void)hipStreamSynchronize(kv.second.s);
    (void)hipEventDestroy(kv.second.fork);
    (void)hipEventDestroy(kv.second.join);
    (void)hipStreamDestroy(kv.second.s);
  }
  hipError_t e = hipFree(ctx->mem);
  (void)hipSetDevice(prev);
  delete ctx;
  if (e != hipSuccess) return hip_error(e, "hipFree tables");
  return D4_OK;
}

int d4_workspace_bytes(int64_t nbatch, int64_t natoms, size_t* bytes) {
  if (bytes == nullptr) return set_error(D4_ERR_INVALID, "bytes is NULL");
  if (nbatch < 0 || natoms < 0) return set_error(D4_ERR_INVALID, "negative size");
  *bytes = d4::layout(nbatch, natoms, nullptr, nullptr);
  return D4_OK;
}

int d4_forward(d4_context* ctx, void* stream, int64_t nbatch, int64_t natoms,
               const int64_t* numbers, const double* positions, const double* charge,
               const double* q_in, const double* r4r2_in, const d4_params* params,
               double* energy_out, double* q_out, void* workspace,
               size_t workspace_bytes) {
  int rc = check_common(ctx, nbatch, natoms, numbers, positions, charge, params,
                        workspace, workspace_bytes);
  if (rc != D4_OK || nbatch == 0 || natoms == 0) return rc;
  if (energy_out == nullptr) return set_error(D4_ERR_INVALID, "energy_out is NULL");
  Call c = make_call(stream, nbatch, natoms, numbers, positions, charge, q_in, r4r2_in,
                     params, workspace);
  if ((rc = run_cn(ctx, c)) != D4_OK) return rc;
  double* e_out = out_atoms(c, energy_out, c.w.sE);
  if ((rc = run_split(ctx, c, true, true, false, nullptr)) != D4_OK) return rc;
  D4_RUN(D4_SLOT_ATM_ROWS, d4::launch_atm_rows(c.s, c.p, c.in, c.w, true, false, e_out, nullptr),
         "energy rows");
  D4_TRY(scatter_atoms(c, c.w.sE, energy_out, 1), "scatter energies");
  if (q_out) {
    if (c.sorted)
      D4_TRY(scatter_atoms(c, c.w.q, q_out, 1), "scatter charges");
    else
      D4_TRY(d4::launch_copy(c.s, c.w.q, q_out, nbatch * natoms), "copy charges");
  }
  return D4_OK;
}

int d4_backward(d4_context* ctx, void* stream, int64_t nbatch, int64_t natoms,
                const int64_t* numbers, const double* positions, const double* charge,
                const double* q_in, const double* r4r2_in, const d4_params* params,
                const double* grad_energy, double* grad_positions, double* grad_params,
                double* grad_q, void* workspace, size_t workspace_bytes) {
  int rc = check_common(ctx, nbatch, natoms, numbers, positions, charge, params,
                        workspace, workspace_bytes);
  if (rc != D4_OK || nbatch == 0 || natoms == 0) return rc;
  if (grad_positions == nullptr)
    return set_error(D4_ERR_INVALID, "grad_positions is NULL");
  if (grad_q != nullptr && q_in == nullptr)
    return set_error(D4_ERR_INVALID, "grad_q needs caller-supplied charges (q_in)");
  Call c = make_call(stream, nbatch, natoms, numbers, positions, charge, q_in, r4r2_in,
                     params, workspace);
  c.in.gE = grad_energy;
  c.p.pgrad = grad_params != nullptr;
  D4_RUN(D4_SLOT_PREP, enter_sorted(c, true), "spatial order");  // the forward's order
  double* g_out = out_atoms(c, grad_positions, c.w.sG);
  if ((rc = run_split(ctx, c, false, false, true, g_out)) != D4_OK) return rc;
  D4_RUN(D4_SLOT_ATM_ROWS, d4::launch_atm_rows(c.s, c.p, c.in, c.w, false, true, nullptr, g_out),
         "three-body rows");
  D4_RUN(D4_SLOT_CHAIN, d4::launch_chain(c.s, ctx->t, c.p, c.in, c.w, g_out), "CN/EEQ chain");
  D4_TRY(scatter_atoms(c, c.w.sG, grad_positions, 3), "scatter gradient");
  if (grad_params)
    D4_TRY(d4::launch_pgrad_reduce(c.s, c.p, c.in, c.w, grad_params), "parameter gradients");
  if (grad_q) {
    // with given charges the two-body pass's dL/dq is the whole derivative
    // (the three-body term uses q = 0 and no EEQ adjoint runs)
    if (c.sorted)
      D4_TRY(scatter_atoms(c, c.w.dldq, grad_q, 1), "scatter charge gradient");
    else
      D4_TRY(d4::launch_copy(c.s, c.w.dldq, grad_q, nbatch * natoms), "copy charge gradient");
  }
  return D4_OK;
}

int d4_energy_gradient_noalloc(d4_context* ctx, void* stream, int64_t nbatch,
                               int64_t natoms, const int64_t* numbers,
                               const double* positions, const double* charge,
                               const double* q_in, const double* r4r2_in,
                               const d4_params* params, double* energy_out,
                               double* grad_positions) {
  if (ctx == nullptr) return set_error(D4_ERR_INVALID, "context is NULL");
  if (nbatch < 0 || natoms < 0)
    return set_error(D4_ERR_INVALID, "nbatch and natoms must be non-negative");
  size_t need = 0;
  d4_workspace_bytes(nbatch, natoms, &need);
  if (need > kScratchBytes)
    return set_error(D4_ERR_INVALID,
                     "d4_energy_gradient_noalloc: this batch needs " + std::to_string(need) +
                         " bytes of workspace, above its fixed scratch of " +
                         std::to_string(kScratchBytes) +
                         " bytes; call d4_energy_gradient with a workspace instead");
  // the scratch holds no state between calls: every array a call reads is
  // written earlier in the same call, as with a fresh workspace
  return d4_energy_gradient(ctx, stream, nbatch, natoms, numbers, positions, charge, q_in,
                            r4r2_in, params, energy_out, grad_positions, ctx->scratch_dev,
                            kScratchBytes);
}

int d4_energy_gradient(d4_context* ctx, void* stream, int64_t nbatch, int64_t natoms,
                       const int64_t* numbers, const double* positions,
                       const double* charge, const double* q_in, const double* r4r2_in,
                       const d4_params* params, double* energy_out,
                       double* grad_positions, void* workspace, size_t workspace_bytes) {
  int rc = check_common(ctx, nbatch, natoms, numbers, positions, charge, params,
                        workspace, workspace_bytes);
  if (rc != D4_OK || nbatch == 0 || natoms == 0) return rc;
  if (energy_out == nullptr || grad_positions == nullptr)
    return set_error(D4_ERR_INVALID, "energy_out and grad_positions must be non-NULL");
  Call c = make_call(stream, nbatch, natoms, numbers, positions, charge, q_in, r4r2_in,
                     params, workspace);
  if ((rc = run_cn(ctx, c)) != D4_OK) return rc;
  double* e_out = out_atoms(c, energy_out, c.w.sE);
  double* g_out = out_atoms(c, grad_positions, c.w.sG);
  if ((rc = run_split(ctx, c, true, true, true, g_out)) != D4_OK) return rc;
  D4_RUN(D4_SLOT_ATM_ROWS, d4::launch_atm_rows(c.s, c.p, c.in, c.w, true, true, e_out, g_out),
         "three-body rows");
  D4_RUN(D4_SLOT_CHAIN, d4::launch_chain(c.s, ctx->t, c.p, c.in, c.w, g_out), "CN/EEQ chain");
  D4_TRY(scatter_atoms(c, c.w.sE, energy_out, 1), "scatter energies");
  D4_TRY(scatter_atoms(c, c.w.sG, grad_positions, 3), "scatter gradient");
  return D4_OK;
}

int d4_properties(d4_context* ctx, void* stream, int64_t nbatch, int64_t natoms,
                  const int64_t* numbers, const double* positions, const double* charge,
                  const d4_params* params, double* cn_out, double* q_out, double* c6_out,
                  double* alpha_out, void* workspace, size_t workspace_bytes) {
  int rc = check_common(ctx, nbatch, natoms, numbers, positions, charge, params,
                        workspace, workspace_bytes);
  if (rc != D4_OK) return rc;
  if (params->flags & D4_FLAG_D4S)
    return set_error(D4_ERR_INVALID, "d4_properties evaluates the D4 model only");
  if (nbatch == 0 || natoms == 0) return rc;
  Call c = make_call(stream, nbatch, natoms, numbers, positions, charge, nullptr,
                     nullptr, params, workspace);
  if (q_out == nullptr && c6_out == nullptr && alpha_out == nullptr) {
    // coordination numbers alone (tad-mctc's cn_d4): no charges or weights
    if ((rc = run_cn(ctx, c)) != D4_OK) return rc;
    if (cn_out == nullptr) return D4_OK;
    if (c.sorted) {
      D4_TRY(scatter_atoms(c, c.w.cn, cn_out, 1), "scatter coordination numbers");
    } else {
      D4_TRY(d4::launch_copy(c.s, c.w.cn, cn_out, nbatch * natoms),
             "copy coordination numbers");
    }
    return D4_OK;
  }
  if ((rc = run_front(ctx, c)) != D4_OK) return rc;
  if (c.sorted) {
    // C6 is staged in the (unused here) three-body gradient pair array
    if (c6_out || alpha_out)
      D4_TRY(d4::launch_properties(c.s, ctx->t, c.in, c.w, c6_out ? c.w.pg : nullptr,
                                   alpha_out ? c.w.sE : nullptr),
             "properties");
    if (c6_out) D4_TRY(d4::launch_scatter_pairs(c.s, c.w, c.w.pg, c6_out), "scatter C6");
    D4_TRY(scatter_atoms(c, c.w.sE, alpha_out, 1), "scatter polarisabilities");
    D4_TRY(scatter_atoms(c, c.w.q, q_out, 1), "scatter charges");
    D4_TRY(scatter_atoms(c, c.w.cn, cn_out, 1), "scatter coordination numbers");
    return D4_OK;
  }
  if (c6_out || alpha_out)
    D4_TRY(d4::launch_properties(c.s, ctx->t, c.in, c.w, c6_out, alpha_out),
           "properties");
  if (q_out) D4_TRY(d4::launch_copy(c.s, c.w.q, q_out, nbatch * natoms), "copy charges");
  if (cn_out)
    D4_TRY(d4::launch_copy(c.s, c.w.cn, cn_out, nbatch * natoms), "copy coordination numbers");
  return D4_OK;
}

// ------------------------------------------------------- model methods

namespace {

int check_model(const d4_context* ctx, int64_t nbatch, int64_t natoms, const int64_t* numbers,
                const d4_params* params, d4::ModelCall* mc) {
  if (ctx == nullptr) return set_error(D4_ERR_INVALID, "context is NULL");
  if (params == nullptr) return set_error(D4_ERR_INVALID, "params is NULL");
  if (nbatch < 0 || natoms < 0)
    return set_error(D4_ERR_INVALID, "nbatch and natoms must be non-negative");
  if (!ctx->has_model)
    return set_error(D4_ERR_INVALID,
                     "the context was created without the alpha and refslot tables");
  if ((params->flags & D4_FLAG_D4S) && !ctx->has_wfpair)
    return set_error(D4_ERR_INVALID, "D4_FLAG_D4S needs the wfpair table at context creation");
  if (nbatch > 0 && natoms > 0 && numbers == nullptr)
    return set_error(D4_ERR_INVALID, "numbers is NULL");
  mc->B = nbatch;
  mc->N = natoms;
  mc->numbers = numbers;
  mc->ga = params->ga;
  mc->gc = params->gc;
  mc->wf = params->wf;
  mc->d4s = (params->flags & D4_FLAG_D4S) ? 1 : 0;
  return D4_OK;
}

}  // namespace

int d4_model_weights(d4_context* ctx, void* stream, int64_t nbatch, int64_t natoms,
                     const int64_t* numbers, const double* cn, const double* q,
                     const d4_params* params, double* gw_out, double* dgwdcn_out,
                     double* dgwdq_out) {
  d4::ModelCall mc;
  int rc = check_model(ctx, nbatch, natoms, numbers, params, &mc);
  if (rc != D4_OK || nbatch == 0 || natoms == 0) return rc;
  if (gw_out == nullptr) return set_error(D4_ERR_INVALID, "gw_out is NULL");
  if (mc.d4s && dgwdcn_out)
    return set_error(D4_ERR_INVALID, "the D4S model has no analytic CN derivative of its "
                                     "weights (as the reference, model/d4s.py:252-256)");
  D4_TRY(d4::launch_model_weights((hipStream_t)stream, ctx->t, mc, cn, q, gw_out, dgwdcn_out,
                                  dgwdq_out, nullptr, nullptr, nullptr),
         "model weights");
  return D4_OK;
}

int d4_model_weights_backward(d4_context* ctx, void* stream, int64_t nbatch, int64_t natoms,
                              const int64_t* numbers, const double* cn, const double* q,
                              const d4_params* params, const double* grad_gw,
                              double* grad_cn, double* grad_q) {
  d4::ModelCall mc;
  int rc = check_model(ctx, nbatch, natoms, numbers, params, &mc);
  if (rc != D4_OK || nbatch == 0 || natoms == 0) return rc;
  if (grad_gw == nullptr) return set_error(D4_ERR_INVALID, "grad_gw is NULL");
  if (grad_cn == nullptr && grad_q == nullptr) return D4_OK;
  D4_TRY(d4::launch_model_weights((hipStream_t)stream, ctx->t, mc, cn, q, nullptr, nullptr,
                                  nullptr, grad_gw, grad_cn, grad_q),
         "model weights backward");
  return D4_OK;
}

int d4_model_atomic_c6(d4_context* ctx, void* stream, int64_t nbatch, int64_t natoms,
                       const int64_t* numbers, const double* gw, const d4_params* params,
                       double* c6_out) {
  d4::ModelCall mc;
  int rc = check_model(ctx, nbatch, natoms, numbers, params, &mc);
  if (rc != D4_OK || nbatch == 0 || natoms == 0) return rc;
  if (gw == nullptr || c6_out == nullptr)
    return set_error(D4_ERR_INVALID, "gw and c6_out must be non-NULL");
  D4_TRY(d4::launch_model_c6((hipStream_t)stream, ctx->t, mc, gw, c6_out), "model C6");
  return D4_OK;
}

int d4_model_atomic_c6_backward(d4_context* ctx, void* stream, int64_t nbatch,
                                int64_t natoms, const int64_t* numbers, const double* gw,
                                const d4_params* params, const double* grad_c6,
                                double* grad_gw_out) {
  d4::ModelCall mc;
  int rc = check_model(ctx, nbatch, natoms, numbers, params, &mc);
  if (rc != D4_OK || nbatch == 0 || natoms == 0) return rc;
  if (gw == nullptr || grad_c6 == nullptr || grad_gw_out == nullptr)
    return set_error(D4_ERR_INVALID, "gw, grad_c6 and grad_gw_out must be non-NULL");
  D4_TRY(d4::launch_model_c6_bwd((hipStream_t)stream, ctx->t, mc, gw, grad_c6, grad_gw_out),
         "model C6 backward");
  return D4_OK;
}

int d4_model_pols(d4_context* ctx, void* stream, int64_t nbatch, int64_t natoms,
                  const int64_t* numbers, const double* gw, const d4_params* params,
                  int static_only, double* out) {
  d4::ModelCall mc;
  int rc = check_model(ctx, nbatch, natoms, numbers, params, &mc);
  if (rc != D4_OK || nbatch == 0 || natoms == 0) return rc;
  if (gw == nullptr || out == nullptr)
    return set_error(D4_ERR_INVALID, "gw and out must be non-NULL");
  D4_TRY(d4::launch_model_pols((hipStream_t)stream, ctx->t, mc, static_only ? 1 : 0, gw, out),
         "model polarisabilities");
  return D4_OK;
}

int d4_model_pols_backward(d4_context* ctx, void* stream, int64_t nbatch, int64_t natoms,
                           const int64_t* numbers, const d4_params* params, int static_only,
                           const double* grad_out, double* grad_gw_out) {
  d4::ModelCall mc;
  int rc = check_model(ctx, nbatch, natoms, numbers, params, &mc);
  if (rc != D4_OK || nbatch == 0 || natoms == 0) return rc;
  if (grad_out == nullptr || grad_gw_out == nullptr)
    return set_error(D4_ERR_INVALID, "grad_out and grad_gw_out must be non-NULL");
  D4_TRY(d4::launch_model_pols_bwd((hipStream_t)stream, ctx->t, mc, static_only ? 1 : 0,
                                   grad_out, grad_gw_out),
         "model polarisabilities backward");
  return D4_OK;
}

int d4_molecule_status(void* stream, int64_t nbatch, const void* workspace,
                       int32_t* status_out) {
  if (nbatch < 0 || status_out == nullptr || (nbatch > 0 && workspace == nullptr))
    return set_error(D4_ERR_INVALID, "bad arguments");
  if (nbatch == 0) return D4_OK;
  // status is the first array of the layout
  D4_TRY(hipMemcpyAsync(status_out, workspace, sizeof(int32_t) * nbatch,
                        hipMemcpyDeviceToHost, (hipStream_t)stream),
         "copy status");
  D4_TRY(hipStreamSynchronize((hipStream_t)stream), "synchronize");
  return D4_OK;
}

int d4_profile_enable(d4_context* ctx, int on) {
  if (ctx == nullptr) return set_error(D4_ERR_INVALID, "context is NULL");
  if (on == 0 || on == 1)
    ctx->profile = on ? (1u << D4_NUM_SLOTS) - 1u : 0u;
  else if (on & D4_PROFILE_SLOTS)
    ctx->profile = (unsigned)on & ((1u << D4_NUM_SLOTS) - 1u);
  else
    return set_error(D4_ERR_INVALID, "profile: 0, 1 or D4_PROFILE_SLOTS | slot bits");
  return D4_OK;
}

int d4_profile_read(d4_context* ctx, double* ms_out, int64_t* count_out, int nslots) {
  if (ctx == nullptr) return set_error(D4_ERR_INVALID, "context is NULL");
  for (const d4_event_pair& ev : ctx->pending) {
    D4_TRY(hipEventSynchronize(ev.stop), "hipEventSynchronize");
    float ms = 0.0f;
    D4_TRY(hipEventElapsedTime(&ms, ev.start, ev.stop), "hipEventElapsedTime");
    ctx->prof_ms[ev.slot] += ms;
    ctx->prof_count[ev.slot] += 1;
    ctx->pool.push_back(ev.start);
    ctx->pool.push_back(ev.stop);
  }
  ctx->pending.clear();
  for (int k = 0; k < nslots && k < D4_NUM_SLOTS; ++k) {
    if (ms_out) ms_out[k] = ctx->prof_ms[k];
    if (count_out) count_out[k] = ctx->prof_count[k];
    ctx->prof_ms[k] = 0.0;
    ctx->prof_count[k] = 0;
  }
  return D4_OK;
}

const char* d4_last_error(void) { return g_error.c_str(); }

}  // extern "C"
