// kg_capi.hip -- the C ABI of include/kg.h: handles, device buffers, the
// integrator stage sequences and their HIP-graph replay.
//
// Stage sequences (MoL/lines_method_lib.f90):
//   rk4  (:411-474): 4 x [stage kernel (RHS + combine), x/y/z ghost fill]
//   rk3  (:211-269): 3 x [stage kernel, fill]
//   pirk2(:271-332): psi part, Pi part (reads last step's u_stage(5)),
//                    final combine, fill of u
//   pirk3(:334-408): as pirk2 with the second stage state u_stage_2
// The fills that pirk2/pirk3 apply to u after stages 1 and 2 (:291, :353,
// :369) have no effect on the result: u is overwritten from u_p in the last
// stage and u_p was copied before them, so they are not issued.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <atomic>
#include <map>
#include <memory>
#include <mutex>
#include <utility>
#include <string>
#include <vector>

#include "../../include/kg.h"
#include "kg_internal.h"

using namespace kg;

namespace {

thread_local std::string g_last_error;

struct State5 {
  double* f[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  double* alloc = nullptr;  // slot 0 of the buffer the fields live in
};

}  // namespace

// One launch of a step, recorded instead of launched (batched stepping
// replays the recorded sequence of every member with one launch per entry).
struct RecOp {
  int kind;  // 0: stage kernel, 1: ghost fill (its three passes), 2: Sommerfeld outer column of a stage
  int mode;
  StageArgs a;
  FillArgs f;
};

// A recorded batched step led by a handle: the launches of both parity
// versions, their uploaded descriptors and (after the first replay) one
// captured graph per version.  Valid while the same members, in the same
// order, with the same tuning and parities step together, and while the
// lead's descriptor buffer still holds its upload (gen).
struct BatchPlan {
  std::vector<int64_t> sig;
  int n = 0;
  size_t nops = 0;
  std::vector<int> kind, mode;
  std::vector<int> total;  // per (version, launch): workgroups of the stage / column launch
  int nblocks[3] = {0, 0, 0}, wide[3] = {0, 0, 0};
  size_t o_fd = 0, o_ts = 0, o_bs = 0, o_ni = 0;
  char* dev = nullptr;
  uint64_t gen = 0;
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  int nwg = 0;  // workgroups of a persistent stage launch (0: one per tile)
  BatchPlan() = default;
  BatchPlan(BatchPlan&&) = default;
  BatchPlan& operator=(BatchPlan&&) = default;
  ~BatchPlan() {
    for (auto& e : exec)
      if (e) (void)hipGraphExecDestroy(e);
  }
};

struct kg_handle {
  kg_desc desc{};
  uint64_t serial = 0;  // unique per handle in the process (identifies members in cached batch plans)
  HostGrid hg{};
  Geom G{};
  int nsys = 1;
  int device = 0;
  hipStream_t stream = nullptr;
  double* d_tab = nullptr;
  State5 u;             // MoL%u (== u_p inside a step; updated in place)
  State5 S;             // MoL%u_stage; S.f[4] is one of the two Pi slots
  double* S5[2] = {nullptr, nullptr};
  State5 S2;            // MoL%u_stage_2
  double* S25[2] = {nullptr, nullptr};
  State5 Ua, Ub, A;     // RK3 stage states; RK4 k2, k3 and the buffer that
                        // alternates with u as (k1, next state)
  int p5 = 0, p25 = 0;  // slot holding the current u_stage(5), u_stage_2(5)
  int px = 0;           // RK4: which of the two state buffers holds u
  int tune_kc = 32;     // phi planes per workgroup of the stage kernel
  int tune_tj = 4;      // theta rows per workgroup
  int batch_wg_per_cu = 0;  // batched stage launches led by this handle: persistent, this many workgroups a CU
  int cus = 0;              // compute units of the device
  double t = 0.0;
  std::string err;
  std::vector<void*> allocs;
  std::map<int, hipGraphExec_t> graphs;
  bool prof = false;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  std::vector<std::pair<int, size_t>> ev_marks;  // (kind, index of start event)
  double prof_ms[3] = {0.0, 0.0, 0.0};  // stage kernels, ghost fills, Sommerfeld columns
  int64_t prof_n[3] = {0, 0, 0};
  void* batch_dev = nullptr;  // descriptors of batched launches led by this handle
  size_t batch_cap = 0;
  uint64_t batch_gen = 0;     // uploads into batch_dev so far
  std::unique_ptr<BatchPlan> plan;  // the last batched step this handle led
  mutable void* reduce_buf = nullptr;  // reductions of the diagnostics, kept between calls
  mutable size_t reduce_bytes = 0;
  hipEvent_t batch_done = nullptr;

  std::vector<RecOp>* rec = nullptr;  // when set, stage() and fill() record instead of launching
};

namespace {

kg_status fail(kg_handle* h, kg_status st, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (h) h->err = buf;
  g_last_error = buf;
  return st;
}

#define KG_HIP(h, expr)                                                                   \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return fail((h), KG_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));        \
  } while (0)

kg_status alloc_state(kg_handle* h, State5* s, int nfields_extra, double** extra) {
  const size_t per = (size_t)h->G.sys_stride * h->nsys;  // doubles per field buffer
  const size_t n = per * (5 + nfields_extra);
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, n * sizeof(double));
  if (e != hipSuccess) return fail(h, KG_ERR_NOMEM, "hipMalloc of %zu bytes failed: %s", n * 8, hipGetErrorString(e));
  h->allocs.push_back(p);
  KG_HIP(h, hipMemsetAsync(p, 0, n * sizeof(double), h->stream));
  double* d = (double*)p;
  s->alloc = d;
  for (int c = 0; c < 5; ++c) s->f[c] = d + c * per;
  for (int x = 0; x < nfields_extra; ++x) extra[x] = d + (5 + x) * per;
  return KG_OK;
}

kg_status ensure_pirk(kg_handle* h, bool need_s2) {
  if (!h->S.f[0]) {
    double* ex[1];
    kg_status st = alloc_state(h, &h->S, 1, ex);
    if (st) return st;
    h->S5[0] = h->S.f[4];
    h->S5[1] = ex[0];
  }
  if (need_s2 && !h->S2.f[0]) {
    double* ex[1];
    kg_status st = alloc_state(h, &h->S2, 1, ex);
    if (st) return st;
    h->S25[0] = h->S2.f[4];
    h->S25[1] = ex[0];
  }
  return KG_OK;
}

kg_status ensure_rk(kg_handle* h, bool need_acc) {
  if (!h->Ua.f[0]) {
    kg_status st = alloc_state(h, &h->Ua, 0, nullptr);
    if (st) return st;
    st = alloc_state(h, &h->Ub, 0, nullptr);
    if (st) return st;
  }
  if (need_acc && !h->A.f[0]) {
    kg_status st = alloc_state(h, &h->A, 0, nullptr);
    if (st) return st;
  }
  return KG_OK;
}

// ---------------------------------------------------------------------------
// launches
// ---------------------------------------------------------------------------
hipEvent_t next_event(kg_handle* h) {
  if (h->ev_used == h->ev_pool.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    h->ev_pool.push_back(e);
  }
  return h->ev_pool[h->ev_used++];
}

struct Box {
  int i_lo, i_hi, j_lo, j_hi, k_lo, k_hi;
};

Box full_box(const Geom& G) { return Box{0, G.ni, 0, G.nj, 0, G.nk}; }
// Points whose evolved value reaches the state after the ghost fill: the
// fill overwrites i = -g, i = Nx+g, the theta-ghost rows and the plane
// k = Nz from other points, so a stage followed by a fill only computes
// i in [-g+1, Nx+g-1], j in [1, Ny-1], k in [0, Nz-1].
Box filled_box(const Geom& G) { return Box{1, G.ni - 1, G.g + 1, G.g + G.Ny, 0, G.Nz}; }

struct Roles {
  const State5* r[R_NUM] = {nullptr, nullptr, nullptr, nullptr, nullptr};
};

#define KG_TRY(expr)              \
  do {                            \
    kg_status s_ = (expr);        \
    if (s_ != KG_OK) return s_;   \
  } while (0)

kg_status stage(kg_handle* h, int mode, const Box& b, const Roles& roles) {
  StageArgs a{};
  a.G = h->G;
  a.tab = h->d_tab;
  a.per = (int64_t)h->G.sys_stride * h->nsys;
  for (int r = 0; r < R_NUM; ++r) {
    const State5* st = roles.r[r];
    a.rb[r] = st ? st->alloc : nullptr;
    a.ri4[r] = 4;
    if (st && st->f[4]) {
      const int64_t slot = (st->f[4] - st->alloc) / a.per;
      if (slot != 4 && slot != 5) return fail(h, KG_ERR_INVALID, "internal: Pi slot %lld", (long long)slot);
      a.ri4[r] = (int)slot;
    }
  }
  a.i_lo = b.i_lo;
  a.i_hi = b.i_hi;
  a.j_lo = b.j_lo;
  a.j_hi = b.j_hi;
  a.k_lo = b.k_lo;
  a.k_hi = b.k_hi;
  a.kc = h->tune_kc;
  a.tj = h->tune_tj;
  a.nsys = h->nsys;
  const bool somm = h->G.bc_somm && somm_stage(mode);
  if (h->rec) {
    h->rec->push_back(RecOp{0, mode, a, FillArgs{}});
    if (somm) h->rec->push_back(RecOp{2, mode, a, FillArgs{}});
    return KG_OK;
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (h->prof) {
    e0 = next_event(h);
    e1 = next_event(h);
    if (!e0 || !e1) return fail(h, KG_ERR_HIP, "hipEventCreate failed");
    KG_HIP(h, hipEventRecord(e0, h->stream));
  }
  KG_HIP(h, launch_stage(mode, a, h->stream));
  if (h->prof) {
    KG_HIP(h, hipEventRecord(e1, h->stream));
    h->ev_marks.push_back({0, h->ev_used - 2});
  }
  // the Sommerfeld outer column after the stage kernel (it overwrites what a
  // full-box PIRK stage computed there); timed as kind 2
  if (somm) {
    if (h->prof) {
      e0 = next_event(h);
      e1 = next_event(h);
      if (!e0 || !e1) return fail(h, KG_ERR_HIP, "hipEventCreate failed");
      KG_HIP(h, hipEventRecord(e0, h->stream));
    }
    KG_HIP(h, launch_somm(mode, a, h->stream));
    if (h->prof) {
      KG_HIP(h, hipEventRecord(e1, h->stream));
      h->ev_marks.push_back({2, h->ev_used - 2});
    }
  }
  return KG_OK;
}

Roles roles(const State5* in, const State5* base, const State5* out, const State5* aux = nullptr,
            const State5* kin = nullptr) {
  Roles r;
  r.r[R_IN] = in;
  r.r[R_BASE] = base;
  r.r[R_OUT] = out;
  r.r[R_AUX] = aux;
  r.r[R_KIN] = kin;
  return r;
}

kg_status fill(kg_handle* h, const State5& u, const State5& u_p) {
  FillArgs f{};
  f.G = h->G;
  for (int c = 0; c < 5; ++c) {
    f.u[c] = u.f[c];
    f.u_p[c] = u_p.f[c];
  }
  if (h->rec) {
    h->rec->push_back(RecOp{1, -1, StageArgs{}, f});
    return KG_OK;
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (h->prof) {
    e0 = next_event(h);
    e1 = next_event(h);
    if (!e0 || !e1) return fail(h, KG_ERR_HIP, "hipEventCreate failed");
    KG_HIP(h, hipEventRecord(e0, h->stream));
  }
  for (int ph = 0; ph < 3; ++ph) KG_HIP(h, launch_fill(f, h->stream, ph, h->nsys));
  if (h->prof) {
    KG_HIP(h, hipEventRecord(e1, h->stream));
    h->ev_marks.push_back({1, h->ev_used - 2});
  }
  return KG_OK;
}


// rk4 (lines_method_lib.f90:424-446) on increments k_s = rhs*dt.  Stages 2
// and 4 read u_p + c*k (and its ghost fill, the index map of boundary())
// where they need it -- the same rounded values the reference stores in u.
// Stage 2 stores k2 as the two values later stages read: the stage-3 state
// Y = u_p + 0.5*k2, ghost-filled as the reference fills u (:434-435), which
// stage 3 reads as one stored array, and C = k1 + 2*k2, the first partial
// sum of the final combine (one array for stage 4 instead of k1 and k2).
// Stage 4 writes the new state over C, each point reading only its own C;
// u and that buffer trade places each step.
//   stage 1: u_p        -> k1 (Ua)
//   stage 2: u_p, k1    -> Y (Ub), C (X); fill Y
//   stage 3: Y          -> k3 (Ua)
//   stage 4: u_p, k3, C -> X (over C); fill X
kg_status enqueue_rk4(kg_handle* h, const State5& P, const State5& X) {
  const Box fb = filled_box(h->G);
  const State5& K = h->Ua;
  const State5& Y = h->Ub;
  KG_TRY(stage(h, M_RK4_S1, fb, roles(&P, nullptr, &K)));
  KG_TRY(stage(h, M_RK4_S2, fb, roles(&P, nullptr, &Y, &X, &K)));
  KG_TRY(fill(h, Y, P));
  KG_TRY(stage(h, M_RK4_S3, fb, roles(&Y, nullptr, &K)));
  KG_TRY(stage(h, M_RK4_S4, fb, roles(&P, &P, &X, &X, &K)));
  KG_TRY(fill(h, X, P));
  return KG_OK;
}

kg_status enqueue_rk3(kg_handle* h) {
  const Box fb = filled_box(h->G);
  KG_TRY(stage(h, M_RK3_S1, fb, roles(&h->u, &h->u, &h->Ua)));
  KG_TRY(fill(h, h->Ua, h->u));
  KG_TRY(stage(h, M_RK3_S2, fb, roles(&h->Ua, &h->u, &h->Ub)));
  KG_TRY(fill(h, h->Ub, h->u));
  KG_TRY(stage(h, M_RK3_S3, fb, roles(&h->Ub, &h->u, &h->u)));
  KG_TRY(fill(h, h->u, h->u));
  return KG_OK;
}

// PIRK stage 1 (:278-289): S(1..4) from u, then S(5) from S(1..4) and the
// previous step's S(5) held in slot p5.  Leaves the new S(5) in slot 1-p5.
kg_status enqueue_pirk_stage1(kg_handle* h, int p5) {
  const Box all = full_box(h->G);
  KG_TRY(stage(h, M_PIRK_1A, all, roles(&h->u, &h->u, &h->S)));
  State5 in = h->S;
  in.f[4] = h->S5[p5];
  State5 out = h->S;
  out.f[4] = h->S5[1 - p5];
  KG_TRY(stage(h, M_PIRK_1B, all, roles(&in, &h->u, &out)));
  return KG_OK;
}

kg_status enqueue_pirk2(kg_handle* h, int p5) {
  KG_TRY(enqueue_pirk_stage1(h, p5));
  State5 s = h->S;
  s.f[4] = h->S5[1 - p5];
  // :294-304 u = 0.5*(u_p + u_stage + dt*rhs(u_stage)), then boundary (:306)
  KG_TRY(stage(h, M_PIRK2_2, filled_box(h->G), roles(&s, &h->u, &h->u)));
  KG_TRY(fill(h, h->u, h->u));
  return KG_OK;
}

kg_status enqueue_pirk3(kg_handle* h, int p5, int p25) {
  KG_TRY(enqueue_pirk_stage1(h, p5));
  State5 s = h->S;
  s.f[4] = h->S5[1 - p5];
  const Box all = full_box(h->G);
  // :356-364 u_stage_2(1..4)
  KG_TRY(stage(h, M_PIRK3_2A, all, roles(&s, &h->u, &h->S2, &s)));
  // :366-367 u_stage_2(5) from rhs_pi(u_stage_2) with the previous u_stage_2(5)
  State5 in2 = h->S2;
  in2.f[4] = h->S25[p25];
  State5 out2 = h->S2;
  out2.f[4] = h->S25[1 - p25];
  KG_TRY(stage(h, M_PIRK3_2B, all, roles(&in2, &h->u, &out2, &s)));
  // :372-384
  State5 s2 = h->S2;
  s2.f[4] = h->S25[1 - p25];
  KG_TRY(stage(h, M_PIRK3_3, filled_box(h->G), roles(&s2, &h->u, &h->u)));
  KG_TRY(fill(h, h->u, h->u));
  return KG_OK;
}

kg_status enqueue_step(kg_handle* h, int scheme) {
  switch (scheme) {
    case KG_RK4: return enqueue_rk4(h, h->u, h->A);
    case KG_RK3: return enqueue_rk3(h);
    case KG_PIRK2: return enqueue_pirk2(h, h->p5);
    case KG_PIRK3: return enqueue_pirk3(h, h->p5, h->p25);
  }
  return fail(h, KG_ERR_INVALID, "unknown scheme %d", scheme);
}

void advance_parity(kg_handle* h, int scheme) {
  if (scheme == KG_PIRK2 || scheme == KG_PIRK3) h->p5 ^= 1;
  if (scheme == KG_PIRK3) h->p25 ^= 1;
  if (scheme == KG_RK4) {
    std::swap(h->u, h->A);  // the new state was written into the other buffer
    h->px ^= 1;
  }
}

kg_status ensure_buffers(kg_handle* h, int scheme) {
  if (scheme == KG_RK4) return ensure_rk(h, true);
  if (scheme == KG_RK3) return ensure_rk(h, false);
  if (scheme == KG_PIRK2) return ensure_pirk(h, false);
  if (scheme == KG_PIRK3) return ensure_pirk(h, true);
  return fail(h, KG_ERR_INVALID, "unknown scheme %d", scheme);
}

kg_status do_steps(kg_handle* h, int scheme, int nsteps) {
  if (nsteps < 0) return fail(h, KG_ERR_INVALID, "nsteps < 0");
  if (h->G.fd4 && (scheme == KG_PIRK2 || scheme == KG_PIRK3))
    return fail(h, KG_ERR_UNSUPPORTED, "the fd4 operator set runs with rk3 and rk4");
  KG_TRY(ensure_buffers(h, scheme));
  for (int it = 0; it < nsteps; ++it) {
    if (h->prof) {
      KG_TRY(enqueue_step(h, scheme));
    } else {
      const int key = scheme * 8 + h->px * 4 + h->p5 * 2 + h->p25;
      auto itg = h->graphs.find(key);
      if (itg == h->graphs.end()) {
        hipGraph_t graph;
        KG_HIP(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
        kg_status st = enqueue_step(h, scheme);
        hipError_t ec = hipStreamEndCapture(h->stream, &graph);
        if (st != KG_OK) return st;
        if (ec != hipSuccess) return fail(h, KG_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(ec));
        hipGraphExec_t exec;
        hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ei != hipSuccess) return fail(h, KG_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(ei));
        itg = h->graphs.emplace(key, exec).first;
      }
      KG_HIP(h, hipGraphLaunch(itg->second, h->stream));
    }
    advance_parity(h, scheme);
    h->t = h->t + h->desc.dt;  // this%t = this%t + this%dt
  }
  return KG_OK;
}

kg_status xfer(kg_handle* h, const State5& st, double* host, int32_t layout, int to_dev) {
  if (!host) return fail(h, KG_ERR_INVALID, "null host pointer");
  if (layout != KG_LAYOUT_FORTRAN && layout != KG_LAYOUT_SOA) return fail(h, KG_ERR_INVALID, "bad layout %d", layout);
  const Geom& G = h->G;
  const size_t n = (size_t)G.ni * G.nj * G.nk * 5 * h->nsys;
  // the host array crosses PCIe as it is; the device reorders it (for the
  // Fortran layout, u(5,i,j,k) with the parts fastest, in the pack kernel)
  const int fortran = layout == KG_LAYOUT_FORTRAN;
  void* dbuf = nullptr;
  KG_HIP(h, hipMalloc(&dbuf, n * sizeof(double)));
  kg_status rc = KG_OK;
  hipError_t e = hipSuccess;
  if (to_dev) {
    e = hipMemcpyAsync(dbuf, host, n * sizeof(double), hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess) e = launch_pack(G, (double*)dbuf, st.f, h->nsys, 1, h->stream, fortran);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  } else {
    e = launch_pack(G, (double*)dbuf, st.f, h->nsys, 0, h->stream, fortran);
    if (e == hipSuccess) e = hipMemcpyAsync(host, dbuf, n * sizeof(double), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  }
  if (e != hipSuccess) rc = fail(h, KG_ERR_HIP, "state transfer failed: %s", hipGetErrorString(e));
  (void)hipFree(dbuf);
  return rc;
}

}  // namespace

// ---------------------------------------------------------------------------
// Batched stepping: every handle of a batch advances with one launch per
// stage kernel and one per fill pass.  A step of each member is recorded
// (the launches enqueue_step would make, with the member's own buffers) for
// both values of its buffer and stale-Pi parities, then replayed for all
// members together: member q's part of launch o is exactly its own launch o.
// ---------------------------------------------------------------------------
namespace {
// Descriptor block of kg_batch_step_noalloc: device memory of the library
// itself (one copy per GPU, held until the process ends), so that call makes
// no allocation.  Host code reaches it only through the address resolved per
// device at set-up (kg_create, with that device current).
constexpr size_t kBatchScratchBytes = size_t(4) << 20;
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxDevices] = {};
}  // namespace

static __device__ __align__(256) char scratch[kBatchScratchBytes];

namespace {
hipError_t resolve_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_dev[device]) return hipSuccess;
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_scratch_dev[device] = p;
  return hipSuccess;
}

void* scratch_of(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return g_scratch_dev[device];
}

// the stage buffers a step of `scheme` uses already exist
bool buffers_ready(const kg_handle* h, int scheme) {
  switch (scheme) {
    case KG_RK4: return h->Ua.f[0] && h->Ub.f[0] && h->A.f[0];
    case KG_RK3: return h->Ua.f[0] && h->Ub.f[0];
    case KG_PIRK2: return h->S.f[0] != nullptr;
    case KG_PIRK3: return h->S.f[0] && h->S2.f[0];
  }
  return false;
}
}  // namespace

// noalloc: the descriptors go to the library's own device array instead of
// the lead handle's growable buffer, no buffer or event is created and the
// kernel-timing events are not recorded.
kg_status batch_step(kg_handle* const* hs, int n, int nsteps, bool noalloc) {
  kg_handle* lead = hs[0];
  const int ef = lead->G.metric_ef;
  const int scheme = lead->desc.scheme;
  for (int q = 0; q < n; ++q) {
    kg_handle* h = hs[q];
    if (!h) return fail(lead, KG_ERR_INVALID, "batch: null handle %d", q);
    if (h->desc.scheme != scheme) return fail(lead, KG_ERR_UNSUPPORTED, "batch: handle %d runs another scheme", q);
    if (h->G.fd4 != lead->G.fd4) return fail(lead, KG_ERR_UNSUPPORTED, "batch: mixed operator sets");
    if (h->G.metric_ef != ef) return fail(lead, KG_ERR_UNSUPPORTED, "batch: mixed metric forms");
    if (h->G.bc_somm != lead->G.bc_somm)
      return fail(lead, KG_ERR_UNSUPPORTED, "batch: mixed outer boundaries (Sommerfeld and filled)");
    if (h->device != lead->device) return fail(lead, KG_ERR_INVALID, "batch: handles on different devices");
    for (int p = 0; p < q; ++p)
      if (hs[p] == h) return fail(lead, KG_ERR_INVALID, "batch: handle %d repeated", q);
    if (noalloc) {
      if (!buffers_ready(h, scheme))
        return fail(lead, KG_ERR_INVALID, "batch without allocation: handle %d has no stage buffers for its scheme", q);
    } else {
      KG_TRY(ensure_buffers(h, scheme));
    }
    if (h != lead) KG_HIP(lead, hipStreamSynchronize(h->stream));
  }
  // A plan (recorded launches and uploaded descriptors) is kept on the lead
  // handle and reused, with its two step graphs, while the same members step
  // together with the same tuning and parities.
  std::vector<int64_t> sig;
  sig.reserve(4 + 6 * (size_t)n);
  sig.push_back(n);
  sig.push_back(lead->tune_tj);
  sig.push_back(lead->batch_wg_per_cu);
  for (int q = 0; q < n; ++q) {
    const kg_handle* h = hs[q];
    sig.push_back((int64_t)h->serial);
    sig.push_back(h->tune_kc);
    sig.push_back(h->px);
    sig.push_back(h->p5);
    sig.push_back(h->p25);
  }
  const bool use_graphs = !noalloc && !lead->prof;
  BatchPlan local;
  BatchPlan* plan = nullptr;
  if (!noalloc && lead->plan && lead->plan->sig == sig && lead->plan->gen == lead->batch_gen) plan = lead->plan.get();
  if (!plan) {
    plan = &local;
    plan->sig = sig;
    plan->n = n;
    plan->nwg = lead->batch_wg_per_cu > 0 ? (lead->cus * lead->batch_wg_per_cu + 7) / 8 * 8 : 0;
    // two consecutive steps of every member: version v is the step taken after
    // v parity advances (each parity has period 2 under advance_parity)
    std::vector<std::vector<RecOp>> rec(2 * (size_t)n);
    for (int q = 0; q < n; ++q) {
      kg_handle* h = hs[q];
      const State5 u0 = h->u, a0 = h->A;
      const int px0 = h->px, p50 = h->p5, p250 = h->p25;
      kg_status st = KG_OK;
      for (int v = 0; v < 2 && st == KG_OK; ++v) {
        h->rec = &rec[(size_t)v * n + q];
        st = enqueue_step(h, scheme);
        h->rec = nullptr;
        advance_parity(h, scheme);
      }
      h->u = u0;
      h->A = a0;
      h->px = px0;
      h->p5 = p50;
      h->p25 = p250;
      if (st != KG_OK) return st;
    }
    const size_t nops = rec[0].size();
    for (size_t r = 0; r < rec.size(); ++r) {
      if (rec[r].size() != nops) return fail(lead, KG_ERR_INVALID, "internal: batch step shapes differ");
      for (size_t o = 0; o < nops; ++o)
        if (rec[r][o].kind != rec[0][o].kind || rec[r][o].mode != rec[0][o].mode)
          return fail(lead, KG_ERR_INVALID, "internal: batch step launches differ");
    }
    plan->nops = nops;
    for (size_t o = 0; o < nops; ++o) {
      plan->kind.push_back(rec[0][o].kind);
      plan->mode.push_back(rec[0][o].mode);
    }
    // descriptors: stage args and fill args per (version, launch, member), the
    // workgroup prefix of each stage launch, the fill passes' block prefixes
    std::vector<StageArgs> sd(2 * nops * n);
    std::vector<FillArgs> fd(2 * nops * n);
    std::vector<int> tstart(2 * nops * (n + 1), 0), bstart(3 * (n + 1));
    std::vector<int64_t> nitems(3 * n);
    plan->total.assign(2 * nops, 0);
    for (int v = 0; v < 2; ++v)
      for (size_t o = 0; o < nops; ++o) {
        const size_t vo = (size_t)v * nops + o;
        for (int q = 0; q < n; ++q) {
          const RecOp& op = rec[(size_t)v * n + q][o];
          if (op.kind == 0) {
            StageArgs a = op.a;  // the member's own phi chunk; theta rows are the lead's
            a.tj = lead->G.fd4 ? 8 : stage_tj(op.mode, lead->tune_tj);
            finalize_tiles(a);
            sd[vo * n + q] = a;
            tstart[vo * (n + 1) + q + 1] = tstart[vo * (n + 1) + q] + a.nti * a.ntj * a.ntk * a.nsys;
          } else if (op.kind == 2) {
            // Sommerfeld column: blocks of somm_block() work items
            sd[vo * n + q] = op.a;
            const int64_t nb = (somm_items(op.a) + somm_block() - 1) / somm_block();
            tstart[vo * (n + 1) + q + 1] = tstart[vo * (n + 1) + q] + (int)nb;
          } else {
            fd[vo * n + q] = op.f;
          }
        }
        plan->total[vo] = tstart[vo * (n + 1) + n];
      }
    for (int ph = 0; ph < 3; ++ph) {  // a pass with 2^32 items or more in some grid
      plan->wide[ph] = 0;
      bstart[ph * (n + 1)] = 0;
      for (int q = 0; q < n; ++q) {
        const int64_t it = fill_work_items(hs[q]->G, hs[q]->nsys, ph);
        nitems[ph * n + q] = it;
        if (it >= (int64_t{1} << 32)) plan->wide[ph] = 1;
        bstart[ph * (n + 1) + q + 1] = bstart[ph * (n + 1) + q] + (int)((it + 255) / 256);
      }
      plan->nblocks[ph] = bstart[ph * (n + 1) + n];
    }
    // one device block: [sd][fd][tstart][bstart][nitems]
    plan->o_fd = sd.size() * sizeof(StageArgs);
    plan->o_ts = plan->o_fd + fd.size() * sizeof(FillArgs);
    plan->o_bs = plan->o_ts + tstart.size() * sizeof(int);
    plan->o_ni = (plan->o_bs + bstart.size() * sizeof(int) + 7) / 8 * 8;
    const size_t bytes = plan->o_ni + nitems.size() * sizeof(int64_t);
    std::vector<char> host(bytes);
    memcpy(host.data(), sd.data(), plan->o_fd);
    memcpy(host.data() + plan->o_fd, fd.data(), fd.size() * sizeof(FillArgs));
    memcpy(host.data() + plan->o_ts, tstart.data(), tstart.size() * sizeof(int));
    memcpy(host.data() + plan->o_bs, bstart.data(), bstart.size() * sizeof(int));
    memcpy(host.data() + plan->o_ni, nitems.data(), nitems.size() * sizeof(int64_t));
    if (noalloc) {
      if (bytes > kBatchScratchBytes)
        return fail(lead, KG_ERR_INVALID,
                    "batch without allocation: %zu bytes of launch descriptors exceed its %zu-byte device array "
                    "(fewer members per call)", bytes, kBatchScratchBytes);
      plan->dev = (char*)scratch_of(lead->device);
      if (!plan->dev) return fail(lead, KG_ERR_INVALID, "batch without allocation: device %d not set up", lead->device);
    } else {
      if (lead->batch_cap < bytes) {
        KG_HIP(lead, hipStreamSynchronize(lead->stream));
        if (lead->batch_dev) KG_HIP(lead, hipFree(lead->batch_dev));
        lead->batch_dev = nullptr;
        lead->batch_cap = 0;
        KG_HIP(lead, hipMalloc(&lead->batch_dev, bytes));
        lead->batch_cap = bytes;
      }
      plan->dev = (char*)lead->batch_dev;
      plan->gen = ++lead->batch_gen;  // an older plan's descriptors are overwritten now
    }
    KG_HIP(lead, hipMemcpyAsync(plan->dev, host.data(), bytes, hipMemcpyHostToDevice, lead->stream));
  }
  // the launches of one step of version v (0: the step the members take now)
  const BatchPlan& P = *plan;
  const StageArgs* dsd = (const StageArgs*)P.dev;
  const FillArgs* dfd = (const FillArgs*)(P.dev + P.o_fd);
  const int* dts = (const int*)(P.dev + P.o_ts);
  const int* dbs = (const int*)(P.dev + P.o_bs);
  const int64_t* dni = (const int64_t*)(P.dev + P.o_ni);
  auto launch_version = [&](int v, bool prof) -> kg_status {
    for (size_t o = 0; o < P.nops; ++o) {
      const size_t vo = (size_t)v * P.nops + o;
      const int kind = P.kind[o], mode = P.mode[o];
      hipEvent_t e0 = nullptr, e1 = nullptr;
      if (prof) {
        e0 = next_event(lead);
        e1 = next_event(lead);
        if (!e0 || !e1) return fail(lead, KG_ERR_HIP, "hipEventCreate failed");
        KG_HIP(lead, hipEventRecord(e0, lead->stream));
      }
      if (kind == 0) {
        KG_HIP(lead, launch_stage_batch(mode, dsd + vo * n, dts + vo * (n + 1), n, P.total[vo], P.nwg,
                                        stage_tj(mode, lead->tune_tj), ef, lead->G.fd4, lead->stream));
      } else if (kind == 2) {
        KG_HIP(lead, launch_somm_batch(mode, dsd + vo * n, dts + vo * (n + 1), n, P.total[vo], lead->G.fd4,
                                       lead->stream));
      } else {
        for (int ph = 0; ph < 3; ++ph)
          KG_HIP(lead, launch_fill_batch(dfd + vo * n, dbs + ph * (n + 1), n, dni + ph * n, P.nblocks[ph], ph,
                                         P.wide[ph], lead->stream));
      }
      if (prof) {
        KG_HIP(lead, hipEventRecord(e1, lead->stream));
        lead->ev_marks.push_back({kind, lead->ev_used - 2});
      }
    }
    return KG_OK;
  };
  if (use_graphs) {
    for (int v = 0; v < 2 && v < nsteps; ++v) {
      if (plan->exec[v]) continue;
      hipGraph_t graph;
      KG_HIP(lead, hipStreamBeginCapture(lead->stream, hipStreamCaptureModeThreadLocal));
      const kg_status st = launch_version(v, false);
      const hipError_t ec = hipStreamEndCapture(lead->stream, &graph);
      if (st != KG_OK) return st;
      if (ec != hipSuccess) return fail(lead, KG_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(ec));
      const hipError_t ei = hipGraphInstantiate(&plan->exec[v], graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (ei != hipSuccess) return fail(lead, KG_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(ei));
    }
    for (int it = 0; it < nsteps; ++it) KG_HIP(lead, hipGraphLaunch(plan->exec[it & 1], lead->stream));
  } else {
    const bool prof = lead->prof && !noalloc;
    for (int it = 0; it < nsteps; ++it) KG_TRY(launch_version(it & 1, prof));
  }
  if (plan == &local && !noalloc) {
    // keep the plan (and its graphs) for the next call with the same members
    lead->plan.reset(new BatchPlan(std::move(local)));
    local.exec[0] = local.exec[1] = nullptr;
  }
  // bookkeeping and ordering: later work on any member waits for the batch
  // (the event is made at kg_create)
  if (!lead->batch_done) return fail(lead, KG_ERR_INVALID, "internal: batch event missing");
  KG_HIP(lead, hipEventRecord(lead->batch_done, lead->stream));
  for (int q = 0; q < n; ++q) {
    kg_handle* h = hs[q];
    if (h != lead) KG_HIP(lead, hipStreamWaitEvent(h->stream, lead->batch_done, 0));
    for (int it = 0; it < nsteps; ++it) {
      advance_parity(h, scheme);
      h->t = h->t + h->desc.dt;
    }
  }
  return KG_OK;
}

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char* kg_last_error(const kg_handle* h) { return h ? h->err.c_str() : g_last_error.c_str(); }

kg_status kg_create(const kg_desc* d, kg_handle** out) {
  if (!d || !out) return fail(nullptr, KG_ERR_INVALID, "null argument");
  *out = nullptr;
  if (d->nvars != 5) return fail(nullptr, KG_ERR_INVALID, "nvars must be 5 (got %d)", d->nvars);
  if (d->g_pts < 1 || d->g_pts > 8) return fail(nullptr, KG_ERR_INVALID, "g_pts out of range");
  if (d->Nx < 4) return fail(nullptr, KG_ERR_INVALID, "Nx must be >= 4");
  if (d->Ny < d->g_pts + 2) return fail(nullptr, KG_ERR_INVALID, "Ny must be >= g_pts+2");
  if (d->Nz < 4) return fail(nullptr, KG_ERR_INVALID, "Nz must be >= 4");
  if (d->scheme < KG_PIRK2 || d->scheme > KG_RK4) return fail(nullptr, KG_ERR_INVALID, "bad scheme");
  if (d->metric != KG_METRIC_KERR_SCHILD && d->metric != KG_METRIC_SCHW_EF)
    return fail(nullptr, KG_ERR_UNSUPPORTED, "Metric %d not implemented yet (rhs_pi supports Kerr_Schild, Schw_EF)",
                d->metric);
  if (d->bc_rmax != KG_BC_OUTFLOW && d->bc_rmax != KG_BC_INFLOW && d->bc_rmax != KG_BC_SOMMERFELD)
    return fail(nullptr, KG_ERR_UNSUPPORTED, "This boundary condition type for r_max is not implemented yet");
  if (!(d->xmax > d->xmin)) return fail(nullptr, KG_ERR_INVALID, "xmax must exceed xmin");
  if (!(d->dt > 0.0) && !(d->CFL > 0.0)) return fail(nullptr, KG_ERR_INVALID, "need dt > 0 or CFL > 0");
  if (d->op_set != KG_OPS_REF2 && d->op_set != KG_OPS_FD4) return fail(nullptr, KG_ERR_INVALID, "bad op_set");
  if (d->op_set == KG_OPS_FD4 && (d->scheme == KG_PIRK2 || d->scheme == KG_PIRK3))
    return fail(nullptr, KG_ERR_UNSUPPORTED, "the fd4 operator set runs with rk3 and rk4");
  if (d->op_set == KG_OPS_FD4 && d->g_pts < 2)
    return fail(nullptr, KG_ERR_INVALID, "the fd4 operator set needs g_pts >= 2");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(nullptr, KG_ERR_NO_DEVICE, "no GPU visible");

  static std::atomic<uint64_t> next_serial{1};
  kg_handle* h = new kg_handle();
  h->serial = next_serial++;
  h->desc = *d;
  h->nsys = d->is_complex ? 2 : 1;
  build_grid(d->Nx, d->Ny, d->Nz, d->g_pts, d->xmin, d->xmax, d->CFL, d->dt, &h->hg);
  h->desc.dt = h->hg.dt;
  h->t = d->tmin;
  Geom& G = h->G;
  G.ni = h->hg.ni;
  G.nj = h->hg.nj;
  G.nk = h->hg.nk;
  G.ld = h->hg.ld;
  G.Nx = d->Nx;
  G.Ny = d->Ny;
  G.Nz = d->Nz;
  G.g = d->g_pts;
  G.plane = (int64_t)G.ld * G.nj;
  G.fstride = (G.plane * G.nk + 31) / 32 * 32;
  G.sys_stride = G.fstride;
  if (G.fstride * 8 * h->nsys >= (int64_t(1) << 31)) {
    fail(nullptr, KG_ERR_INVALID, "grid too large: one field must stay below 2 GiB");
    delete h;
    return KG_ERR_INVALID;
  }
  G.rr = 0.50 / h->hg.dx;
  G.rt = 0.50 / h->hg.dy;
  G.rp = 0.50 / h->hg.dz;
  G.r4r = 1.0 / (12.0 * h->hg.dx);
  G.r4t = 1.0 / (12.0 * h->hg.dy);
  G.r4p = 1.0 / (12.0 * h->hg.dz);
  G.rar = 1.0 / h->hg.dx;
  G.fd4 = d->op_set == KG_OPS_FD4;
  G.dt = h->hg.dt;
  G.metric_ef = d->metric == KG_METRIC_SCHW_EF;
  G.bc_inflow = d->bc_rmax == KG_BC_INFLOW;
  G.bc_somm = d->bc_rmax == KG_BC_SOMMERFELD;
  G.r_out = h->hg.xmin + (double)(G.ni - 1) * h->hg.dx;  // x(Nx+g) (mesh_refinement_lib.f90:130)

  hipError_t e = hipGetDevice(&h->device);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&h->cus, hipDeviceAttributeMultiprocessorCount, h->device);
  if (e == hipSuccess) e = hipStreamCreate(&h->stream);
  if (e != hipSuccess) {
    fail(nullptr, KG_ERR_HIP, "stream creation failed: %s", hipGetErrorString(e));
    delete h;
    return KG_ERR_HIP;
  }
  // set-up of the batched paths on this device: the ordering event, and the
  // address of the no-allocation descriptor array (once per device)
  e = hipEventCreateWithFlags(&h->batch_done, hipEventDisableTiming);
  if (e == hipSuccess) e = resolve_scratch(h->device);

  if (e != hipSuccess) {
    fail(nullptr, KG_ERR_HIP, "batch set-up failed: %s", hipGetErrorString(e));
    kg_destroy(h);
    return KG_ERR_HIP;
  }
  std::vector<double> tab((size_t)kNumTables * G.nj * G.ld);
  build_tables(h->hg, d->metric, d->a, d->M_bh, d->m, G.fd4, tab.data());
  e = hipMalloc(&h->d_tab, tab.size() * sizeof(double));
  if (e == hipSuccess) {
    h->allocs.push_back(h->d_tab);
    e = hipMemcpy(h->d_tab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice);
  }
  if (e != hipSuccess) {
    fail(nullptr, KG_ERR_NOMEM, "metric table upload failed: %s", hipGetErrorString(e));
    kg_destroy(h);
    return KG_ERR_NOMEM;
  }
  kg_status st = alloc_state(h, &h->u, 0, nullptr);
  if (st == KG_OK) st = ensure_buffers(h, d->scheme);
  if (st == KG_OK && hipStreamSynchronize(h->stream) != hipSuccess) st = KG_ERR_HIP;
  if (st != KG_OK) {
    g_last_error = h->err;
    kg_destroy(h);
    return st;
  }
  *out = h;
  return KG_OK;
}

kg_status kg_destroy(kg_handle* h) {
  if (!h) return KG_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  h->plan.reset();
  for (auto& kv : h->graphs) (void)hipGraphExecDestroy(kv.second);
  for (auto e : h->ev_pool) (void)hipEventDestroy(e);
  for (void* p : h->allocs) (void)hipFree(p);
  if (h->batch_dev) (void)hipFree(h->batch_dev);
  if (h->reduce_buf) (void)hipFree(h->reduce_buf);
  if (h->batch_done) (void)hipEventDestroy(h->batch_done);

  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return KG_OK;
}

kg_status kg_get_dims(const kg_handle* h, int64_t dims[8]) {
  if (!h || !dims) return fail(nullptr, KG_ERR_INVALID, "null argument");
  const Geom& G = h->G;
  dims[0] = G.Nx;
  dims[1] = G.Ny;
  dims[2] = G.Nz;
  dims[3] = G.g;
  dims[4] = G.ni;
  dims[5] = G.nj;
  dims[6] = G.nk;
  dims[7] = (int64_t)G.ni * G.nj * G.nk;
  return KG_OK;
}

kg_status kg_get_time(const kg_handle* h, double* t, double* dt) {
  if (!h) return fail(nullptr, KG_ERR_INVALID, "null handle");
  if (t) *t = h->t;
  if (dt) *dt = h->desc.dt;
  return KG_OK;
}

kg_status kg_set_time(kg_handle* h, double t) {
  if (!h) return fail(nullptr, KG_ERR_INVALID, "null handle");
  h->t = t;
  return KG_OK;
}

kg_status kg_upload_state(kg_handle* h, const double* u, int32_t layout) {
  if (!h) return fail(nullptr, KG_ERR_INVALID, "null handle");
  return xfer(h, h->u, const_cast<double*>(u), layout, 1);
}

kg_status kg_download_state(kg_handle* h, double* u, int32_t layout) {
  if (!h) return fail(nullptr, KG_ERR_INVALID, "null handle");
  return xfer(h, h->u, u, layout, 0);
}

static kg_status aux_state(kg_handle* h, int32_t which, State5* s) {
  if (which == 0) {
    KG_TRY(ensure_pirk(h, false));
    *s = h->S;
    s->f[4] = h->S5[h->p5];
  } else if (which == 1) {
    KG_TRY(ensure_pirk(h, true));
    *s = h->S2;
    s->f[4] = h->S25[h->p25];
  } else {
    return fail(h, KG_ERR_INVALID, "aux index must be 0 (u_stage) or 1 (u_stage_2)");
  }
  return KG_OK;
}

kg_status kg_upload_aux(kg_handle* h, int32_t which, const double* u, int32_t layout) {
  if (!h) return fail(nullptr, KG_ERR_INVALID, "null handle");
  State5 s;
  KG_TRY(aux_state(h, which, &s));
  return xfer(h, s, const_cast<double*>(u), layout, 1);
}

kg_status kg_download_aux(kg_handle* h, int32_t which, double* u, int32_t layout) {
  if (!h) return fail(nullptr, KG_ERR_INVALID, "null handle");
  State5 s;
  KG_TRY(aux_state(h, which, &s));
  return xfer(h, s, u, layout, 0);
}

kg_status kg_copy_state_device(kg_handle* h, void* dev_ptr, int32_t dir) {
  if (!h || !dev_ptr) return fail(h, KG_ERR_INVALID, "null argument");
  KG_HIP(h, launch_pack(h->G, (double*)dev_ptr, h->u.f, h->nsys, dir == 0 ? 1 : 0, h->stream));
  KG_HIP(h, hipStreamSynchronize(h->stream));
  return KG_OK;
}

// ---------------------------------------------------------------------------
// Checkpoint / restart (the role of check_point / read_from_check_point,
// lines_method_lib.f90:599-667, in a binary format of this engine; the
// reference's reader deletes its file before reading it).  Layout:
//   "KGCKPT01", kg_desc, double t, int32 nsys, int32 narrays, int64 count,
//   then narrays x count doubles in KG_LAYOUT_SOA order: u, and for PIRK2 /
//   PIRK3 the stage arrays u_stage (/ u_stage_2), whose Pi carries over
//   between steps.
// ---------------------------------------------------------------------------
namespace {
constexpr char kCkptMagic[8] = {'K', 'G', 'C', 'K', 'P', 'T', '0', '1'};

int aux_count(int scheme) { return scheme == KG_PIRK2 ? 1 : scheme == KG_PIRK3 ? 2 : 0; }

bool same_grid(const kg_desc& a, const kg_desc& b) {
  return a.nvars == b.nvars && a.g_pts == b.g_pts && a.Nx == b.Nx && a.Ny == b.Ny && a.Nz == b.Nz &&
         a.scheme == b.scheme && a.metric == b.metric && a.bc_rmax == b.bc_rmax && a.is_complex == b.is_complex &&
         a.op_set == b.op_set &&
         a.xmin == b.xmin && a.xmax == b.xmax && a.dt == b.dt && a.a == b.a && a.M_bh == b.M_bh && a.m == b.m;
}
}  // namespace

kg_status kg_checkpoint_save(kg_handle* h, const char* path) {
  if (!h || !path) return fail(h, KG_ERR_INVALID, "null argument");
  const int naux = aux_count(h->desc.scheme);
  const int64_t count = (int64_t)h->G.ni * h->G.nj * h->G.nk * 5 * h->nsys;
  std::vector<double> buf((size_t)count);
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(h, KG_ERR_INVALID, "cannot open %s for writing", path);
  const int32_t meta[2] = {h->nsys, 1 + naux};
  bool ok = std::fwrite(kCkptMagic, 1, 8, f) == 8 && std::fwrite(&h->desc, sizeof(kg_desc), 1, f) == 1 &&
            std::fwrite(&h->t, sizeof(double), 1, f) == 1 && std::fwrite(meta, sizeof(int32_t), 2, f) == 2 &&
            std::fwrite(&count, sizeof(int64_t), 1, f) == 1;
  kg_status st = KG_OK;
  for (int a = 0; ok && st == KG_OK && a <= naux; ++a) {
    if (a == 0) {
      st = xfer(h, h->u, buf.data(), KG_LAYOUT_SOA, 0);
    } else {
      State5 s;
      st = aux_state(h, a - 1, &s);
      if (st == KG_OK) st = xfer(h, s, buf.data(), KG_LAYOUT_SOA, 0);
    }
    if (st == KG_OK) ok = std::fwrite(buf.data(), sizeof(double), (size_t)count, f) == (size_t)count;
  }
  ok = (std::fclose(f) == 0) && ok;
  if (st != KG_OK) return st;
  if (!ok) return fail(h, KG_ERR_INVALID, "write to %s failed", path);
  return KG_OK;
}

kg_status kg_checkpoint_load(kg_handle* h, const char* path) {
  if (!h || !path) return fail(h, KG_ERR_INVALID, "null argument");
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(h, KG_ERR_INVALID, "cannot open %s", path);
  char magic[8];
  kg_desc d{};
  double t = 0.0;
  int32_t meta[2] = {0, 0};
  int64_t count = 0;
  bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, kCkptMagic, 8) == 0 &&
            std::fread(&d, sizeof(kg_desc), 1, f) == 1 && std::fread(&t, sizeof(double), 1, f) == 1 &&
            std::fread(meta, sizeof(int32_t), 2, f) == 2 && std::fread(&count, sizeof(int64_t), 1, f) == 1;
  const int naux = aux_count(h->desc.scheme);
  const int64_t want = (int64_t)h->G.ni * h->G.nj * h->G.nk * 5 * h->nsys;
  if (!ok) {
    std::fclose(f);
    return fail(h, KG_ERR_INVALID, "%s is not a checkpoint of this engine", path);
  }
  if (!same_grid(d, h->desc) || meta[0] != h->nsys || meta[1] != 1 + naux || count != want) {
    std::fclose(f);
    return fail(h, KG_ERR_INVALID, "%s holds a different grid, scheme or physics", path);
  }
  std::vector<double> buf((size_t)count);
  kg_status st = KG_OK;
  for (int a = 0; st == KG_OK && a <= naux; ++a) {
    if (std::fread(buf.data(), sizeof(double), (size_t)count, f) != (size_t)count) {
      st = fail(h, KG_ERR_INVALID, "%s is truncated", path);
      break;
    }
    if (a == 0) {
      st = xfer(h, h->u, buf.data(), KG_LAYOUT_SOA, 1);
    } else {
      State5 s;
      st = aux_state(h, a - 1, &s);
      if (st == KG_OK) st = xfer(h, s, buf.data(), KG_LAYOUT_SOA, 1);
    }
  }
  std::fclose(f);
  if (st == KG_OK) h->t = t;
  return st;
}

kg_status kg_fill_ghosts(kg_handle* h) {
  if (!h) return fail(nullptr, KG_ERR_INVALID, "null handle");
  return fill(h, h->u, h->u);
}

kg_status kg_pirk2(kg_handle* h) { return h ? do_steps(h, KG_PIRK2, 1) : fail(nullptr, KG_ERR_INVALID, "null handle"); }
kg_status kg_pirk3(kg_handle* h) { return h ? do_steps(h, KG_PIRK3, 1) : fail(nullptr, KG_ERR_INVALID, "null handle"); }
kg_status kg_rk3(kg_handle* h) { return h ? do_steps(h, KG_RK3, 1) : fail(nullptr, KG_ERR_INVALID, "null handle"); }
kg_status kg_rk4(kg_handle* h) { return h ? do_steps(h, KG_RK4, 1) : fail(nullptr, KG_ERR_INVALID, "null handle"); }

kg_status kg_step(kg_handle* h, int32_t nsteps) {
  if (!h) return fail(nullptr, KG_ERR_INVALID, "null handle");
  return do_steps(h, h->desc.scheme, nsteps);
}

kg_status kg_batch_step(kg_handle* const* hs, int32_t n, int32_t nsteps) {
  if (!hs || n < 1 || !hs[0]) return fail(nullptr, KG_ERR_INVALID, "batch: need at least one handle");
  if (nsteps < 0) return fail(hs[0], KG_ERR_INVALID, "nsteps < 0");
  return batch_step(hs, n, nsteps, false);
}

kg_status kg_batch_step_noalloc(kg_handle* const* hs, int32_t n, int32_t nsteps) {
  if (!hs || n < 1 || !hs[0]) return fail(nullptr, KG_ERR_INVALID, "batch: need at least one handle");
  if (nsteps < 0) return fail(hs[0], KG_ERR_INVALID, "nsteps < 0");
  return batch_step(hs, n, nsteps, true);
}

kg_status kg_sync(kg_handle* h) {
  if (!h) return fail(nullptr, KG_ERR_INVALID, "null handle");
  KG_HIP(h, hipStreamSynchronize(h->stream));
  return KG_OK;
}

kg_status kg_set_tuning(kg_handle* h, int32_t kc, int32_t tj) {
  if (!h) return fail(nullptr, KG_ERR_INVALID, "null handle");
  if (kc < 1 || kc > 4096 || (tj != 4 && tj != 8)) return fail(h, KG_ERR_INVALID, "tuning: kc in [1,4096], tj in {4,8}");
  if (kc == h->tune_kc && tj == h->tune_tj) return KG_OK;
  KG_HIP(h, hipStreamSynchronize(h->stream));
  h->tune_kc = kc;
  h->tune_tj = tj;
  for (auto& kv : h->graphs) (void)hipGraphExecDestroy(kv.second);
  h->graphs.clear();
  return KG_OK;
}

kg_status kg_set_batch_persistent(kg_handle* h, int32_t wg_per_cu) {
  if (!h) return fail(nullptr, KG_ERR_INVALID, "null handle");
  if (wg_per_cu < 0 || wg_per_cu > 16) return fail(h, KG_ERR_INVALID, "batch workgroups per CU must be in [0, 16]");
  h->batch_wg_per_cu = wg_per_cu;
  return KG_OK;
}

kg_status kg_profile_enable(kg_handle* h, int32_t on) {
  if (!h) return fail(nullptr, KG_ERR_INVALID, "null handle");
  KG_HIP(h, hipStreamSynchronize(h->stream));
  h->prof = on != 0;
  h->ev_used = 0;
  h->ev_marks.clear();
  for (int k = 0; k < 3; ++k) {
    h->prof_ms[k] = 0.0;
    h->prof_n[k] = 0;
  }
  return KG_OK;
}

kg_status kg_profile_read(kg_handle* h, int32_t kind, double* total_ms, int64_t* launches) {
  if (!h) return fail(nullptr, KG_ERR_INVALID, "null handle");
  if (kind < 0 || kind > 2) return fail(h, KG_ERR_INVALID, "kind must be 0, 1 or 2");
  KG_HIP(h, hipStreamSynchronize(h->stream));
  for (auto& m : h->ev_marks) {
    float ms = 0.f;
    KG_HIP(h, hipEventElapsedTime(&ms, h->ev_pool[m.second], h->ev_pool[m.second + 1]));
    h->prof_ms[m.first] += ms;
    h->prof_n[m.first] += 1;
  }
  h->ev_marks.clear();
  h->ev_used = 0;
  if (total_ms) *total_ms = h->prof_ms[kind];
  if (launches) *launches = h->prof_n[kind];
  return KG_OK;
}

}  // extern "C"

namespace kg {
hipError_t handle_scratch(const kg_handle* h, size_t bytes, void** out) {
  if (h->reduce_bytes < bytes) {
    if (h->reduce_buf) {
      (void)hipStreamSynchronize(h->stream);
      (void)hipFree(h->reduce_buf);
      h->reduce_buf = nullptr;
      h->reduce_bytes = 0;
    }
    hipError_t e = hipMalloc(&h->reduce_buf, bytes);
    if (e != hipSuccess) return e;
    h->reduce_bytes = bytes;
  }
  *out = h->reduce_buf;
  return hipSuccess;
}

int handle_view(const kg_handle* h, HandleView* v) {
  if (!h || !v) return 1;
  v->handle = h;
  v->device = h->device;
  v->stream = h->stream;
  v->G = h->G;
  v->hg = h->hg;
  v->nsys = h->nsys;
  for (int f = 0; f < kNumFields; ++f) v->u[f] = h->u.f[f];
  v->tab = h->d_tab;
  v->m = h->desc.m;
  v->metric = h->desc.metric;
  return 0;
}
}  // namespace kg
