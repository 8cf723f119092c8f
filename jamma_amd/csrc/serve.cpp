// Native continuous-batching runtime for association requests on resident cohorts
// (host code; no reference counterpart: the reference runs one association per
// call, run_lmm_association_streaming in src/jamma/lmm/runner_streaming.py:94-628).
//
// A request is one prepared problem (jm_lmm_problem + its partner rows V) against a
// contiguous SNP range of a cohort, with a caller-owned device result block. The
// runtime owns one scheduler thread and a first-fit allocator over ONE caller-owned
// device arena: each cohort's step workspace is carved from it when the cohort is
// added, and each request's moments slice when it is admitted, so admission is a
// real HBM budget and the serving path itself calls no hipMalloc. Per cohort:
//   * one HIP stream (default priority) and `max_inflight` pinned descriptor slots,
//     each with an event; a slot is reused only once its event has completed;
//   * a step takes the next slice (<= slice_snps) of every active request, up to
//     the cohort's max_batch SNPs and max_reqs requests, writes the step's
//     descriptors into the slot (gather index, moments column descriptors, per-SNP
//     problem index and moments address, the problems and their modes), copies them
//     up with one hipMemcpyAsync, and launches pack + rotation (or the FP64
//     rotation), ONE jm_spectrum_moments_multi and ONE jm_lmm_assoc_multi over the
//     whole batch, then copies each request's rows into its result block (device or
//     pinned host memory) on the stream;
//   * a request completes when the event of the step that took its last slice has
//     completed; its arena space is released then.
// Admission walks the queue in submission order (first fit, with aging: a request
// that later ones overtook max_skips times stops further overtaking). A step whose
// launch fails fails that cohort's active requests alone (the stream is drained
// first); the loop goes on. jm_serve_poll / jm_serve_wait hand back completion
// records. With JM_SERVE_STUB the runtime makes no HIP call at all (the arena is a
// range of offsets, steps complete as they are formed): the admission, batching and
// completion logic then runs on a machine without a GPU.

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdarg>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "jm_common.h"

namespace jm {
namespace {

constexpr int64_t kAlign = 256;
inline int64_t align_up(int64_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

// First-fit allocator over [0, size): offsets only; coalesces on release.
class Arena {
 public:
  explicit Arena(int64_t size) : size_(size) {
    if (size > 0) free_[0] = size;
  }
  bool alloc(int64_t bytes, int64_t& off) {
    bytes = align_up(std::max<int64_t>(bytes, 1));
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= bytes) {
        off = it->first;
        const int64_t rest = it->second - bytes;
        free_.erase(it);
        if (rest > 0) free_[off + bytes] = rest;
        used_ += bytes;
        return true;
      }
    }
    return false;
  }
  void release(int64_t off, int64_t bytes) {
    bytes = align_up(std::max<int64_t>(bytes, 1));
    used_ -= bytes;
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  int64_t used() const { return used_; }
  int64_t size() const { return size_; }

 private:
  int64_t size_, used_ = 0;
  std::map<int64_t, int64_t> free_;
};

struct Request {
  int64_t id = 0;
  int cohort = 0;
  jm_lmm_problem prob{};
  const double* V = nullptr;
  int64_t ldv = 0;
  int nv = 0;
  int mode = 4;
  int64_t start = 0, end = 0, next = 0;
  double* out = nullptr;
  int64_t ws_off = -1, ws_bytes = 0;
  int skips = 0;
  int64_t done_seq = -1;  // step sequence number that completes it
  int64_t remaining() const { return end - next; }
};

struct Slot {
  uint8_t* host = nullptr;  // pinned descriptor staging
  hipEvent_t ev = nullptr;
  int64_t seq = -1;
  bool busy = false;
};

struct Cohort {
  jm_serve_cohort_desc d{};
  jm_spectrum_plan plan{};
  hipStream_t stream = nullptr;
  std::vector<Slot> slots;
  int64_t ws_off = -1, ws_bytes = 0;
  // byte offsets of the step workspace parts inside the arena
  int64_t o_desc = 0, desc_bytes = 0, o_gp = 0, o_mm = 0, o_z = 0, o_work = 0, o_sout = 0;
  // descriptor layout (offsets inside the descriptor block)
  int64_t d_col = 0, d_cols = 0, d_prob = 0, d_xmom = 0, d_probs = 0, d_modes = 0;
  int64_t seq_launched = 0, seq_done = 0;
  std::vector<Request*> active;
};

}  // namespace
}  // namespace jm

struct jm_serve {
  jm_serve_config cfg{};
  bool stub = false;
  int device = 0;
  uint8_t* arena_base = nullptr;
  std::unique_ptr<jm::Arena> arena;
  std::vector<std::unique_ptr<jm::Cohort>> cohorts;
  std::deque<jm::Request*> queue;
  std::deque<jm_serve_done> done;
  std::mutex mu;
  std::condition_variable cv;       // work for the scheduler
  std::condition_variable cv_done;  // completions for waiters
  std::thread thread;
  bool stop = false;
  int64_t next_id = 1;
  int64_t steps = 0, snps = 0, max_active = 0;
  bool manual = false;  // no thread: jm_serve_run_once drives the loop
  bool paused = false;  // no admission and no new step (completions still polled)
  int inject_fail = -1;  // cohort whose next step fails (failure-handling tests)
  int64_t reserved = 0;  // arena bytes held by the cohorts' step workspaces
};

namespace jm {
namespace {

void finish(jm_serve* s, Request* r, int status, const char* msg) {
  jm_serve_done rec{};
  rec.id = r->id;
  rec.status = status;
  if (msg) std::snprintf(rec.message, sizeof(rec.message), "%s", msg);
  if (r->ws_off >= 0) s->arena->release(r->ws_off, r->ws_bytes);
  s->done.push_back(rec);
  delete r;
}

// SNPs per step of a cohort: its own cap if it has one, else the server's.
int64_t cohort_batch(const jm_serve_config& cfg, const jm_serve_cohort_desc& d) {
  return d.max_batch > 0 ? std::min<int64_t>(d.max_batch, cfg.max_batch) : cfg.max_batch;
}

int64_t request_bytes(const jm_serve* s, const Cohort& c, int nv) {
  return align_up(s->cfg.slice_snps * (int64_t)(nv + 1) * c.d.n_terms * (int64_t)sizeof(double));
}

// Step workspace layout of a cohort (device, in the arena) and its descriptor block.
int64_t layout(const jm_serve_config& cfg, const jm_serve_cohort_desc& d,
               const jm_spectrum_plan* plan, Cohort* c) {
  const int64_t MB = cohort_batch(cfg, d), MR = cfg.max_reqs;
  int64_t dd = 0;
  auto put = [&](int64_t bytes) {
    const int64_t o = dd;
    dd += align_up(bytes);
    return o;
  };
  const int64_t d_col = put(4 * MB), d_cols = put((int64_t)sizeof(jm_mom_col) * MB),
                d_prob = put(4 * MB), d_xmom = put(8 * MB),
                d_probs = put((int64_t)sizeof(jm_lmm_problem) * MR), d_modes = put(4 * MR);
  int64_t o = 0;
  auto take = [&](int64_t bytes) {
    const int64_t x = o;
    o += align_up(bytes);
    return x;
  };
  const int64_t o_desc = take(dd);
  const bool mixed = d.precision == JM_SERVE_MIXED;
  const int64_t o_gp = take(mixed ? jm_packed_geno_bytes(d.n, MB) : 0);
  const int64_t o_mm = take(mixed ? 4 * MB : 0);
  const int64_t o_z = take((int64_t)d.n * MB * (mixed ? 4 : 8));
  const int64_t o_work = take(jm_moments_work_bytes(plan, MB, cfg.max_cvt + 1));
  const int64_t o_sout = take(MB * (int64_t)JM_NOUT * 8);
  if (c) {
    c->o_desc = o_desc;
    c->desc_bytes = dd;
    c->o_gp = o_gp;
    c->o_mm = o_mm;
    c->o_z = o_z;
    c->o_work = o_work;
    c->o_sout = o_sout;
    c->d_col = d_col;
    c->d_cols = d_cols;
    c->d_prob = d_prob;
    c->d_xmom = d_xmom;
    c->d_probs = d_probs;
    c->d_modes = d_modes;
  }
  return o;
}

// Admission in submission order, first fit with aging. Caller holds s->mu.
void admit(jm_serve* s) {
  std::deque<Request*> keep;
  bool blocked = false;
  std::vector<Request*> passed;  // queued items an admitted one overtook
  for (Request* r : s->queue) {
    const Cohort& c = *s->cohorts[r->cohort];
    int64_t off = -1;
    if (!blocked && s->arena->alloc(request_bytes(s, c, r->nv), off)) {
      r->ws_off = off;
      r->ws_bytes = request_bytes(s, c, r->nv);
      for (Request* w : passed) ++w->skips;
      s->cohorts[r->cohort]->active.push_back(r);
    } else {
      keep.push_back(r);
      passed.push_back(r);
      if (r->skips >= s->cfg.max_skips) blocked = true;  // aged: no more overtaking
    }
  }
  s->queue.swap(keep);
}

void fail_cohort(jm_serve* s, Cohort& c, int status, const char* msg) {
  if (!s->stub && c.stream) hipStreamSynchronize(c.stream);  // queued work done first
  for (Slot& sl : c.slots) sl.busy = false;
  c.seq_done = c.seq_launched;
  for (Request* r : c.active) finish(s, r, status, msg);
  c.active.clear();
}

// Completed steps -> completed requests. Caller holds s->mu.
int poll_cohort(jm_serve* s, Cohort& c) {
  if (!s->stub) {
    for (;;) {  // slots complete in launch order
      Slot* oldest = nullptr;
      for (Slot& sl : c.slots)
        if (sl.busy && (!oldest || sl.seq < oldest->seq)) oldest = &sl;
      if (!oldest) break;
      const hipError_t e = hipEventQuery(oldest->ev);
      if (e == hipErrorNotReady) break;
      oldest->busy = false;
      if (e != hipSuccess) {
        fail_cohort(s, c, JM_ERR_HIP, hipGetErrorString(e));
        return 1;
      }
      c.seq_done = oldest->seq;
    }
  }
  int n = 0;
  std::vector<Request*> keep;
  for (Request* r : c.active) {
    if (r->done_seq >= 0 && r->done_seq <= c.seq_done) {
      finish(s, r, JM_OK, nullptr);
      ++n;
    } else {
      keep.push_back(r);
    }
  }
  c.active.swap(keep);
  return n;
}

// One step of a cohort: returns 1 if launched, 0 if nothing to do or no free slot,
// < 0 on a launch error (message via set_error). Caller holds s->mu.
int launch_step(jm_serve* s, int ci) {
  Cohort& c = *s->cohorts[(size_t)ci];
  Slot* slot = nullptr;
  for (Slot& sl : c.slots)
    if (!sl.busy) {
      slot = &sl;
      break;
    }
  if (!slot) return 0;
  struct Take {
    Request* r;
    int64_t off, take;
  };
  std::vector<Take> takes;
  int64_t B = 0;
  int max_cvt = 1;
  const int64_t MB = cohort_batch(s->cfg, c.d);
  for (Request* r : c.active) {
    if (r->remaining() <= 0) continue;
    if (B >= MB || (int64_t)takes.size() >= s->cfg.max_reqs) break;
    const int64_t t = std::min({s->cfg.slice_snps, r->remaining(), MB - B});
    takes.push_back({r, r->next - r->start, t});
    B += t;
    max_cvt = std::max(max_cvt, r->prob.n_cvt);
  }
  if (takes.empty()) return 0;
  if (s->inject_fail == ci) {
    // failure-handling hook: this cohort's step fails on the host before any launch
    s->inject_fail = -1;
    set_error("injected step error");
    for (const Take& t : takes) t.r->next += t.take;
    return JM_ERR_UNSUPPORTED;
  }
  const int T = c.d.n_terms;
  // device addresses as integers: in stub mode the arena is a range of offsets
  const uintptr_t abase = (uintptr_t)s->arena_base;
  uint8_t* dev = (uint8_t*)(abase + (uintptr_t)c.ws_off);
  uint8_t* ddesc = dev + c.o_desc;
  // descriptors into the pinned slot
  uint8_t* h = slot->host;
  std::vector<uint8_t> stub_buf;
  if (s->stub) {
    stub_buf.resize((size_t)c.desc_bytes);
    h = stub_buf.data();
  }
  int32_t* col = (int32_t*)(h + c.d_col);
  jm_mom_col* cols = (jm_mom_col*)(h + c.d_cols);
  int32_t* sprob = (int32_t*)(h + c.d_prob);
  uint64_t* sxmom = (uint64_t*)(h + c.d_xmom);
  jm_lmm_problem* probs = (jm_lmm_problem*)(h + c.d_probs);
  int32_t* modes = (int32_t*)(h + c.d_modes);
  int64_t o = 0;
  for (size_t k = 0; k < takes.size(); ++k) {
    Request* r = takes[k].r;
    const int rows = r->prob.n_cvt + 2;
    double* xm = (double*)(abase + (uintptr_t)r->ws_off);
    for (int64_t j = 0; j < takes[k].take; ++j) {
      col[o + j] = (int32_t)(r->next + j);
      cols[o + j].V = r->V;
      cols[o + j].ldv = r->ldv;
      cols[o + j].nv = r->nv;
      cols[o + j].reserved = 0;
      cols[o + j].out = xm + j * (int64_t)rows * T;
      cols[o + j].ldm = T;
      sprob[o + j] = (int32_t)k;
      sxmom[o + j] = (uint64_t)(uintptr_t)(xm + j * (int64_t)rows * T);
    }
    probs[k] = r->prob;
    modes[k] = r->mode;
    o += takes[k].take;
  }
  for (const Take& t : takes) t.r->next += t.take;
  const int64_t seq = ++c.seq_launched;
  for (const Take& t : takes)
    if (t.r->remaining() == 0) t.r->done_seq = seq;
  ++s->steps;
  s->snps += B;
  if (s->stub) {  // no device: the step completes as it is formed
    c.seq_done = seq;
    return 1;
  }
  hipStream_t st = c.stream;
  JM_CHECK_HIP(hipMemcpyAsync(ddesc, h, (size_t)c.desc_bytes, hipMemcpyHostToDevice, st));
  const int32_t* dcol = (const int32_t*)(ddesc + c.d_col);
  const bool mixed = c.d.precision == JM_SERVE_MIXED;
  void* Z = dev + c.o_z;
  int rc;
  if (mixed) {
    rc = jm_pack_genotypes(c.d.g, c.d.ld_g, dcol, c.d.means, c.d.n, B, dev + c.o_gp,
                           (uint32_t*)(dev + c.o_mm), st);
    if (rc) return rc;
    rc = jm_rotate_bf16x3(c.d.planes, c.d.nplanes, dev + c.o_gp, (const uint32_t*)(dev + c.o_mm),
                          c.d.n, B, (float*)Z, B, st);
  } else {
    rc = jm_rotate_f64(c.d.UT, c.d.ld_u, c.d.g, c.d.ld_g, c.d.means, dcol, c.d.n, B, (double*)Z, B,
                       st);
  }
  if (rc) return rc;
  rc = jm_spectrum_moments_multi(Z, mixed ? 1 : 0, B, B, (const jm_mom_col*)(ddesc + c.d_cols),
                                 max_cvt + 1, &c.plan, (double*)(dev + c.o_work), st);
  if (rc) return rc;
  double* sout = (double*)(dev + c.o_sout);
  // the search kernel is chosen for the configured max_cvt, not for this step's
  // widest problem: its variant, and with it every result bit, is then the same
  // whatever else shares the step
  rc = jm_lmm_assoc_multi((const jm_lmm_problem*)(ddesc + c.d_probs),
                          (const int32_t*)(ddesc + c.d_modes), (int)takes.size(), max_cvt,
                          (int)s->cfg.max_cvt, T,
                          c.d.n_clusters, (const int32_t*)(ddesc + c.d_prob),
                          (const double* const*)(ddesc + c.d_xmom), T, B, sout, st);
  if (rc) return rc;
  o = 0;
  for (const Take& t : takes) {
    JM_CHECK_HIP(hipMemcpyAsync(t.r->out + t.off * JM_NOUT, sout + o * JM_NOUT,
                                (size_t)t.take * JM_NOUT * sizeof(double), hipMemcpyDefault,
                                st));
    o += t.take;
  }
  JM_CHECK_HIP(hipEventRecord(slot->ev, st));
  slot->seq = seq;
  slot->busy = true;
  return 1;
}

bool busy(const jm_serve* s) {
  if (!s->queue.empty()) return true;
  for (const auto& c : s->cohorts)
    if (!c->active.empty()) return true;
  return false;
}

bool in_flight(const jm_serve* s) {
  for (const auto& c : s->cohorts)
    if (c->seq_done < c->seq_launched) return true;
  return false;
}

// One pass: admit, then per cohort poll / launch a step / poll. Caller holds s->mu.
// -> launched steps * 65536 + completed requests.
int64_t iterate(jm_serve* s) {
  if (!s->paused) admit(s);
  int64_t active = 0;
  for (const auto& c : s->cohorts) active += (int64_t)c->active.size();
  s->max_active = std::max(s->max_active, active);
  int64_t launched = 0, completed = 0;
  for (size_t ci = 0; ci < s->cohorts.size(); ++ci) {
    Cohort& c = *s->cohorts[ci];
    completed += poll_cohort(s, c);
    const int rc = s->paused ? 0 : launch_step(s, (int)ci);
    if (rc > 0) {
      ++launched;
    } else if (rc < 0) {
      char msg[256];
      jm_last_error(msg, sizeof(msg));
      const int64_t n_fail = (int64_t)c.active.size();
      fail_cohort(s, c, rc, msg);
      completed += n_fail;
    }
    completed += poll_cohort(s, c);
  }
  if (completed) s->cv_done.notify_all();
  return launched * 65536 + completed;
}

void loop(jm_serve* s) {
  if (!s->stub) hipSetDevice(s->device);
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    if (s->stop && !busy(s)) break;
    const int64_t r = iterate(s);
    const bool launched = r >= 65536, completed = (r % 65536) != 0;
    if (!launched && !completed) {
      if (!busy(s) || (s->paused && !in_flight(s))) {
        if (s->stop && !busy(s)) break;
        s->cv.wait(lk);  // submit, resume and destroy notify
      } else {
        // device work in flight (or requests waiting for budget): wait a little
        s->cv.wait_for(lk, std::chrono::microseconds(200));
      }
    }
  }
  s->cv_done.notify_all();
}

}  // namespace
}  // namespace jm

using namespace jm;

extern "C" int64_t jm_serve_cohort_ws_bytes(const jm_serve_config* cfg,
                                             const jm_serve_cohort_desc* d,
                                             const jm_spectrum_plan* plan) {
  if (!cfg || !d || !plan || cfg->max_batch < 1 || cfg->max_reqs < 1 || d->n < 1) return 0;
  return layout(*cfg, *d, plan, nullptr);
}

extern "C" int64_t jm_serve_request_ws_bytes(const jm_serve_config* cfg, int n_terms, int nv) {
  if (!cfg || n_terms < 1 || nv < 1) return 0;
  return align_up(cfg->slice_snps * (int64_t)(nv + 1) * n_terms * (int64_t)sizeof(double));
}

extern "C" int jm_serve_create(const jm_serve_config* cfg, void* arena, int64_t arena_bytes,
                               jm_serve** out) {
  JM_CHECK_ARG(cfg && out, "null argument");
  JM_CHECK_ARG(cfg->slice_snps >= 1 && cfg->max_batch >= cfg->slice_snps && cfg->max_reqs >= 1 &&
                   cfg->max_inflight >= 1 && cfg->max_inflight <= 64 && cfg->max_skips >= 0 &&
                   cfg->max_cvt >= 1 && cfg->max_cvt <= 20,
               "bad serving configuration");
  JM_CHECK_ARG(cfg->max_batch < (1LL << 31), "max_batch exceeds int32 SNP indices");
  const bool stub = (cfg->flags & JM_SERVE_STUB) != 0;
  JM_CHECK_ARG(!(cfg->flags & JM_SERVE_MANUAL) || stub, "JM_SERVE_MANUAL requires JM_SERVE_STUB");
  JM_CHECK_ARG(arena_bytes > 0 && (stub || (arena && ((uintptr_t)arena % kAlign) == 0)),
               "a %lld-byte aligned device arena is required", (long long)kAlign);
  auto s = std::make_unique<jm_serve>();
  s->cfg = *cfg;
  s->stub = stub;
  if (!stub) JM_CHECK_HIP(hipGetDevice(&s->device));
  s->arena_base = (uint8_t*)arena;
  s->arena = std::make_unique<Arena>(arena_bytes);
  s->manual = (cfg->flags & JM_SERVE_MANUAL) != 0;
  jm_serve* raw = s.release();
  if (!raw->manual) raw->thread = std::thread(loop, raw);
  *out = raw;
  return JM_OK;
}

extern "C" int jm_serve_add_cohort(jm_serve* s, const jm_serve_cohort_desc* d,
                                   const jm_spectrum_plan* plan, int* cohort_id) {
  JM_CHECK_ARG(s && d && plan && cohort_id, "null argument");
  JM_CHECK_ARG(d->n >= 1 && d->n_snps >= 1 && d->ld_g >= d->n && d->n_terms >= 1 &&
                   d->n_clusters >= 1 && d->n_terms == plan->n_terms,
               "bad cohort description");
  JM_CHECK_ARG(d->precision == JM_SERVE_MIXED || d->precision == JM_SERVE_F64,
               "precision must be JM_SERVE_MIXED or JM_SERVE_F64");
  JM_CHECK_ARG(s->stub || (d->g && d->means &&
                           (d->precision == JM_SERVE_MIXED ? (d->planes && (d->nplanes == 2 ||
                                                                              d->nplanes == 3))
                                                           : (d->UT && d->ld_u >= d->n))),
               "missing cohort arrays");
  JM_CHECK_ARG(d->n_snps < (1LL << 31), "n_snps exceeds int32 SNP indices");
  JM_CHECK_ARG(d->max_batch == 0 || d->max_batch >= s->cfg.slice_snps,
               "a cohort's max_batch (%d) must be 0 or at least slice_snps (%lld)", d->max_batch,
               (long long)s->cfg.slice_snps);
  auto c = std::make_unique<Cohort>();
  c->d = *d;
  c->plan = *plan;
  c->ws_bytes = layout(s->cfg, *d, plan, c.get());
  {
    std::lock_guard<std::mutex> lk(s->mu);
    JM_CHECK_ARG(s->arena->alloc(c->ws_bytes, c->ws_off),
                 "the arena cannot hold the cohort's step workspace (%lld bytes)",
                 (long long)c->ws_bytes);
    s->reserved += align_up(c->ws_bytes);
  }
  if (!s->stub) {
    int rc = JM_OK;
    auto fail = [&](hipError_t e) {
      set_error("jm_serve_add_cohort: %s", hipGetErrorString(e));
      rc = JM_ERR_HIP;
    };
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) fail(e);
    c->slots.resize((size_t)s->cfg.max_inflight);
    for (Slot& sl : c->slots) {
      if (rc) break;
      e = hipHostMalloc((void**)&sl.host, (size_t)c->desc_bytes, hipHostMallocDefault);
      if (e != hipSuccess) {
        fail(e);
        break;
      }
      e = hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming);
      if (e != hipSuccess) fail(e);
    }
    if (rc) {
      for (Slot& sl : c->slots) {
        if (sl.host) hipHostFree(sl.host);
        if (sl.ev) hipEventDestroy(sl.ev);
      }
      if (c->stream) hipStreamDestroy(c->stream);
      std::lock_guard<std::mutex> lk(s->mu);
      s->arena->release(c->ws_off, c->ws_bytes);
      s->reserved -= align_up(c->ws_bytes);
      return rc;
    }
  } else {
    c->slots.resize((size_t)s->cfg.max_inflight);
  }
  std::lock_guard<std::mutex> lk(s->mu);
  *cohort_id = (int)s->cohorts.size();
  s->cohorts.push_back(std::move(c));
  return JM_OK;
}

extern "C" int jm_serve_submit(jm_serve* s, int cohort, const jm_lmm_problem* prob,
                               const double* V, int64_t ldv, int mode, int64_t snp_start,
                               int64_t snp_end, double* out, int64_t* request_id) {
  JM_CHECK_ARG(s && prob && request_id, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  JM_CHECK_ARG(cohort >= 0 && cohort < (int)s->cohorts.size(), "unknown cohort %d", cohort);
  const Cohort& c = *s->cohorts[cohort];
  JM_CHECK_ARG(mode >= 1 && mode <= 4, "mode must be 1..4");
  JM_CHECK_ARG(0 <= snp_start && snp_start < snp_end && snp_end <= c.d.n_snps,
               "SNP range [%lld, %lld) outside [0, %lld)", (long long)snp_start,
               (long long)snp_end, (long long)c.d.n_snps);
  JM_CHECK_ARG(prob->n == c.d.n && prob->n_terms == c.d.n_terms &&
                   prob->n_clusters == c.d.n_clusters,
               "the problem was prepared on another cohort");
  JM_CHECK_ARG(prob->n_cvt >= 1 && prob->n_cvt <= s->cfg.max_cvt,
               "n_cvt %d outside [1, max_cvt = %d]", prob->n_cvt, s->cfg.max_cvt);
  JM_CHECK_ARG(s->stub || (V && out && ldv >= c.d.n), "missing V / out");
  JM_CHECK_ARG(request_bytes(s, c, prob->n_cvt + 1) <= s->arena->size() - s->reserved,
               "the request can never fit the arena (%lld bytes needed, %lld beyond the cohorts' "
               "step workspaces)",
               (long long)request_bytes(s, c, prob->n_cvt + 1),
               (long long)(s->arena->size() - s->reserved));
  JM_CHECK_ARG(!s->stop, "the runtime is stopping");
  auto* r = new Request();
  r->id = s->next_id++;
  r->cohort = cohort;
  r->prob = *prob;
  r->V = V;
  r->ldv = ldv;
  r->nv = prob->n_cvt + 1;
  r->mode = mode;
  r->start = r->next = snp_start;
  r->end = snp_end;
  r->out = out;
  s->queue.push_back(r);
  *request_id = r->id;
  s->cv.notify_all();
  return JM_OK;
}

extern "C" int jm_serve_poll(jm_serve* s, jm_serve_done* recs, int max_recs, int* n_recs) {
  JM_CHECK_ARG(s && n_recs && (max_recs == 0 || recs) && max_recs >= 0, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  int n = 0;
  while (n < max_recs && !s->done.empty()) {
    recs[n++] = s->done.front();
    s->done.pop_front();
  }
  *n_recs = n;
  return JM_OK;
}

extern "C" int jm_serve_wait(jm_serve* s, int timeout_ms) {
  JM_CHECK_ARG(s && timeout_ms >= 0, "bad argument");
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms),
                      [&] { return !s->done.empty() || (s->stop && !busy(s)); });
  return (int)std::min<size_t>(s->done.size(), 1 << 30);
}

extern "C" int64_t jm_serve_run_once(jm_serve* s) {
  if (!s || !s->manual) return JM_ERR_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  return iterate(s);
}

extern "C" int jm_serve_stats(jm_serve* s, jm_serve_stats_t* st) {
  JM_CHECK_ARG(s && st, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  st->queued = (int64_t)s->queue.size();
  int64_t active = 0;
  for (const auto& c : s->cohorts) active += (int64_t)c->active.size();
  st->active = active;
  st->steps = s->steps;
  st->snps = s->snps;
  st->arena_used = s->arena->used();
  st->arena_bytes = s->arena->size();
  st->max_active = s->max_active;
  st->reserved = s->reserved;
  return JM_OK;
}

extern "C" int jm_serve_pause(jm_serve* s, int paused) {
  JM_CHECK_ARG(s, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  s->paused = paused != 0;
  s->cv.notify_all();
  return JM_OK;
}

extern "C" int jm_serve_cohort_stream(jm_serve* s, int cohort, void** stream) {
  JM_CHECK_ARG(s && stream, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  JM_CHECK_ARG(cohort >= 0 && cohort < (int)s->cohorts.size(), "unknown cohort %d", cohort);
  *stream = (void*)s->cohorts[(size_t)cohort]->stream;
  return JM_OK;
}

extern "C" int jm_serve_inject_step_error(jm_serve* s, int cohort) {
  JM_CHECK_ARG(s, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  JM_CHECK_ARG(cohort >= 0 && cohort < (int)s->cohorts.size(), "unknown cohort %d", cohort);
  s->inject_fail = cohort;
  return JM_OK;
}

extern "C" int jm_serve_destroy(jm_serve* s) {
  if (!s) return JM_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    s->paused = false;  // destroy drains everything
    s->cv.notify_all();
  }
  if (s->thread.joinable()) s->thread.join();  // drains every admitted and queued request
  if (s->manual) {
    std::lock_guard<std::mutex> lk(s->mu);
    for (int guard = 0; busy(s) && guard < (1 << 20); ++guard) iterate(s);
  }
  for (auto& c : s->cohorts) {
    if (!s->stub) {
      if (c->stream) hipStreamSynchronize(c->stream);
      for (Slot& sl : c->slots) {
        if (sl.host) hipHostFree(sl.host);
        if (sl.ev) hipEventDestroy(sl.ev);
      }
      if (c->stream) hipStreamDestroy(c->stream);
    }
  }
  for (Request* r : s->queue) delete r;
  delete s;
  return JM_OK;
}
