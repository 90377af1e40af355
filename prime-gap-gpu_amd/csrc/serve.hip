// Serving engine: per-m next-prime / prev-prime searches with continuous
// batching on one GPU.
//
// A request carries the sieved candidates of both sides of m*K, in ascending
// X (N = m*K + X and N = m*K - X).  The answer per side is the first
// candidate that passes the reference's verdict (base-2 strong test, as the
// reference's GPU stage with ROUNDS = 1, gpu_testing.cu:62), i.e. the next /
// previous probable prime among the sieved offsets; the reference finds the
// same with mpz_nextprime / mpz_prevprime on the CPU (overflow.cpp:148, 158).
//
// Device-resident state: every width class (kernel limb count) keeps its
// requests' candidates in one limb-interleaved pool in HBM, with a column
// allocator; admission needs a free column range and room under the byte
// budget (in-use candidate and verdict bytes of admitted requests).
//
// Iteration-level scheduling: each class runs iterations in up to
// `slots_per_class` slots, each slot with its own stream.  An iteration takes
// the next `chunk` untested candidates of every unfinished side that no other
// iteration holds, of every admitted request of the class, tests them with
// one indexed launch, and gathers their verdicts to the host.  pgx_serve_poll
// applies every finished iteration: it retires sides that found a PRP (early
// exit) or ran out of candidates, completes requests, frees their memory,
// admits waiting requests into the freed room and launches an iteration in a
// free slot.  Requests that arrive while an iteration runs join the next one,
// which with two slots starts at once beside it; classes overlap on the
// device.  A side is in at most one iteration at a time, so a request's
// admission copy (on the stream of the slot whose iteration first reads it)
// is ordered before every read of its columns.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <new>
#include <vector>

#include "launch.hpp"
#include "runtime.hpp"

using pgx::fail;

namespace {

using Clock = std::chrono::steady_clock;

uint64_t now_ns() {
  return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
             Clock::now().time_since_epoch())
      .count();
}

// First-fit allocator of column ranges in a pool.
class Columns {
 public:
  explicit Columns(uint32_t cap) { free_[0] = cap; }
  // returns the first column of a free range of n columns, or UINT32_MAX
  uint32_t take(uint32_t n) {
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < n) continue;
      const uint32_t at = it->first, len = it->second;
      free_.erase(it);
      if (len > n) free_[at + n] = len - n;
      return at;
    }
    return UINT32_MAX;
  }
  void give(uint32_t at, uint32_t n) {
    auto it = free_.emplace(at, n).first;
    auto next = std::next(it);
    if (next != free_.end() && it->first + it->second == next->first) {
      it->second += next->second;
      free_.erase(next);
    }
    if (it != free_.begin()) {
      auto prev = std::prev(it);
      if (prev->first + prev->second == it->first) {
        prev->second += it->second;
        free_.erase(it);
      }
    }
  }

 private:
  std::map<uint32_t, uint32_t> free_;
};

struct Request {
  uint64_t id = 0;
  uint64_t arrival_ns = 0;
  uint64_t bytes = 0;
  uint32_t n_side[2] = {0, 0};  // next, prev
  uint32_t cursor[2] = {0, 0};  // candidates tested so far per side
  int32_t first[2] = {-1, -1};  // index of the first PRP per side
  bool done[2] = {false, false};
  bool busy[2] = {false, false};  // a chunk of the side is in an unfinished iteration
  uint32_t col0 = 0;            // pool columns [col0, col0 + n_next + n_prev)
  uint32_t tests = 0;
  std::vector<uint32_t> cand;   // host copy until admitted: W rows x (n_next + n_prev)
};

// Work of one side in an iteration: items [off, off + count) of the list.
struct Segment {
  Request* req;
  int side;
  uint32_t off, count;
};

// One iteration in flight at a time: its stream, buffers and segments.
struct Slot {
  uint32_t* d_index = nullptr;
  uint32_t* d_out = nullptr;
  uint32_t* h_index = nullptr;  // pinned
  uint32_t* h_out = nullptr;    // pinned
  uint32_t* h_stage = nullptr;  // pinned admission staging, W rows x cap_stage
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  // timing of the iterations that tested candidates: `start` before the
  // iteration's first copy, `done` after its last; prev_done is the previous
  // iteration's end on this stream, so prev_done -> start is the time the
  // device waited for the host to schedule (what a device-side loop removes)
  hipEvent_t start = nullptr, prev_done = nullptr;
  bool timed = false;     // the iteration in flight has a start event
  bool has_prev = false;  // prev_done ends the iteration before it on this stream
  bool in_flight = false;
  std::vector<Segment> segs;
};

struct Class {
  int S = 0;
  uint32_t W = 0;      // 32-bit words per candidate stored in the pool
  uint32_t chunk = 0;  // candidates per side per iteration
  uint32_t cap = 0;
  std::unique_ptr<Columns> cols;
  uint32_t* d_pool = nullptr;
  uint32_t* d_verdict = nullptr;
  uint32_t cap_stage = 0;
  uint32_t max_items = 0;
  std::vector<Slot> slots;
  std::vector<std::unique_ptr<Request>> active;
  std::deque<std::unique_ptr<Request>> waiting;
};

}  // namespace

struct pgx_server {
  pgx_runner* h = nullptr;  // owns the device, base list {2}
  pgx_serve_config_t cfg{};
  uint64_t in_use = 0;
  std::map<int, std::unique_ptr<Class>> classes;
  std::vector<pgx_serve_result_t> finished;
  uint64_t tests_total = 0;
  uint64_t iterations = 0, gaps = 0;  // timed iterations, and gaps measured between them
  double iter_ms = 0, gap_ms = 0;
};

namespace {

void release_class(Class& c) {
  for (Slot& q : c.slots) {
    if (q.done) (void)hipEventDestroy(q.done);
    if (q.start) (void)hipEventDestroy(q.start);
    if (q.prev_done) (void)hipEventDestroy(q.prev_done);
    if (q.stream) (void)hipStreamDestroy(q.stream);
    for (uint32_t* p : {q.d_index, q.d_out})
      if (p) (void)hipFree(p);
    for (uint32_t* p : {q.h_index, q.h_out, q.h_stage})
      if (p) (void)hipHostFree(p);
  }
  for (uint32_t* p : {c.d_pool, c.d_verdict})
    if (p) (void)hipFree(p);
}

int make_class(pgx_server* s, int S, Class** out) {
  auto it = s->classes.find(S);
  if (it != s->classes.end()) {
    *out = it->second.get();
    return PGX_OK;
  }
  auto c = std::make_unique<Class>();
  c->S = S;
  c->W = (uint32_t)((S <= 31 ? 29 : 28) * S + 31) / 32;
  // PRPs thin out as 1/bits (about 35/bits of sieved candidates are prime),
  // so the chunk grows with the class width to decide most sides in one
  // iteration: cfg.chunk at 512 bits, rounded up to a multiple of 16.
  const uint32_t bits = (uint32_t)pgx::kernel_capacity_bits(S);
  c->chunk = std::max<uint32_t>(16, (s->cfg.chunk * bits / 512 + 15) / 16 * 16);
  c->cap = s->cfg.pool_columns;
  c->cols = std::make_unique<Columns>(c->cap);
  c->max_items = s->cfg.max_requests * 2u * c->chunk;
  c->cap_stage = std::max(s->cfg.stage_columns, s->cfg.max_side * 2u);
  c->slots.resize(s->cfg.slots_per_class ? s->cfg.slots_per_class : 1);
  hipError_t e = hipMalloc((void**)&c->d_pool, (size_t)c->W * c->cap * 4);
  if (e == hipSuccess) e = hipMalloc((void**)&c->d_verdict, (size_t)c->cap * 4);
  for (Slot& q : c->slots) {
    if (e == hipSuccess) e = hipMalloc((void**)&q.d_index, (size_t)c->max_items * 4);
    if (e == hipSuccess) e = hipMalloc((void**)&q.d_out, (size_t)c->max_items * 4);
    if (e == hipSuccess) e = hipHostMalloc((void**)&q.h_index, (size_t)c->max_items * 4, 0);
    if (e == hipSuccess) e = hipHostMalloc((void**)&q.h_out, (size_t)c->max_items * 4, 0);
    if (e == hipSuccess)
      e = hipHostMalloc((void**)&q.h_stage, (size_t)c->W * c->cap_stage * 4, 0);
  }
  if (e == hipSuccess) {
    // A stream created with a CU mask gets a hardware queue of its own, so a
    // narrow class never waits behind a wide class's kernels in a queue that
    // the runtime shares between streams.  The mask holds every CU, or with
    // wide_cu_percent the wide classes' share [0, n_wide) or the narrow
    // classes' [n_wide, n_cu): a wide class's wavefronts then never take the
    // SIMDs that a narrow class's short iterations run on.
    const int n_cu = s->h->n_cu;
    int lo = 0, hi = n_cu;
    if (s->cfg.wide_cu_percent) {
      const int n_wide = std::min(n_cu - 1, std::max(1, n_cu * (int)s->cfg.wide_cu_percent / 100));
      const bool wide = s->cfg.group_min_bits && bits >= s->cfg.group_min_bits;
      lo = wide ? 0 : n_wide;
      hi = wide ? n_wide : n_cu;
    }
    std::vector<uint32_t> mask((n_cu + 31) / 32, 0u);
    for (int cu = lo; cu < hi; ++cu) mask[cu / 32] |= 1u << (cu % 32);
    for (Slot& q : c->slots) {
      if (e == hipSuccess)
        e = hipExtStreamCreateWithCUMask(&q.stream, (uint32_t)mask.size(), mask.data());
      if (e == hipSuccess) e = hipEventCreate(&q.done);
      if (e == hipSuccess) e = hipEventCreate(&q.start);
      if (e == hipSuccess) e = hipEventCreate(&q.prev_done);
    }
  }
  if (e != hipSuccess) {
    release_class(*c);
    return fail(s->h, PGX_E_HIP, "serving class allocation: %s", hipGetErrorString(e));
  }
  *out = c.get();
  s->classes[S] = std::move(c);
  return PGX_OK;
}

// Moves waiting requests of the class into the pool while room allows
// (FIFO: a request that does not fit blocks the ones behind it), through the
// staging and on the stream of slot q, whose next iteration reads them first.
int admit(pgx_server* s, Class& c, Slot& q, uint32_t* staged_out) {
  uint32_t staged = 0;
  while (!c.waiting.empty() && c.active.size() < s->cfg.max_requests) {
    Request& r = *c.waiting.front();
    const uint32_t n = r.n_side[0] + r.n_side[1];
    if (s->in_use + r.bytes > s->cfg.hbm_budget) break;
    if (staged + n > c.cap_stage) break;  // staging of this iteration is full
    const uint32_t at = c.cols->take(n);
    if (at == UINT32_MAX) break;
    r.col0 = at;
    // pack into this iteration's staging slice, then one 2D copy per request
    uint32_t* st = q.h_stage + (size_t)staged * c.W;
    std::memcpy(st, r.cand.data(), (size_t)c.W * n * 4);
    PGX_HIP(s->h, hipMemcpy2DAsync(c.d_pool + at, (size_t)c.cap * 4, st, (size_t)n * 4,
                                   (size_t)n * 4, c.W, hipMemcpyHostToDevice, q.stream));
    staged += n;
    s->in_use += r.bytes;
    r.cand.clear();
    r.cand.shrink_to_fit();
    c.active.push_back(std::move(c.waiting.front()));
    c.waiting.pop_front();
  }
  *staged_out = staged;
  return PGX_OK;
}

// Builds and launches an iteration of the class in slot q over every side
// that is neither decided nor held by another iteration.  The slot's event is
// recorded whenever anything was enqueued (admission copies included), so its
// pinned staging is never refilled while a copy from it may be pending.
int launch_iteration(pgx_server* s, Class& c, Slot& q, bool copies_pending) {
  q.segs.clear();
  uint32_t items = 0;
  for (auto& rp : c.active) {
    Request& r = *rp;
    for (int side = 0; side < 2; ++side) {
      if (r.done[side] || r.busy[side]) continue;
      const uint32_t k = std::min(c.chunk, r.n_side[side] - r.cursor[side]);
      const uint32_t base = r.col0 + (side ? r.n_side[0] : 0) + r.cursor[side];
      for (uint32_t j = 0; j < k; ++j) q.h_index[items + j] = base + j;
      q.segs.push_back(Segment{&r, side, items, k});
      r.busy[side] = true;
      items += k;
    }
  }
  if (items == 0) {
    if (copies_pending) {
      PGX_HIP(s->h, hipEventRecord(q.done, q.stream));
      q.in_flight = true;
      q.timed = false;
      q.has_prev = false;  // copies only: no iteration to measure a gap after
    }
    return PGX_OK;
  }
  if (q.timed) std::swap(q.done, q.prev_done);  // the last iteration's end
  q.has_prev = q.timed;
  q.timed = true;
  PGX_HIP(s->h, hipEventRecord(q.start, q.stream));
  PGX_HIP(s->h, hipMemcpyAsync(q.d_index, q.h_index, (size_t)items * 4, hipMemcpyHostToDevice,
                               q.stream));
  // a small iteration of a wide class: a group of lanes per candidate
  const uint32_t cap_bits = (uint32_t)pgx::kernel_capacity_bits(c.S);
  const int G = (int)s->cfg.group_lanes;
  const int Sp = (s->cfg.group_max_items && items <= s->cfg.group_max_items &&
                  cap_bits >= s->cfg.group_min_bits)
                     ? pgx::group_width_for_bits(cap_bits, G)
                     : 0;
  pgx::PrpArgs a{};
  a.cand = c.d_pool;
  a.ld = c.cap;
  a.n = items;
  a.buf_limbs = c.W;
  a.mode = PGX_MODE_REFERENCE;
  a.bases = s->h->bases;  // {2}
  a.verdict = c.d_verdict;
  a.index = q.d_index;
  PGX_HIP(s->h, (Sp ? pgx::group_table(G)[Sp] : pgx::prp_table()[c.S])(a, q.stream));
  PGX_HIP(s->h, pgx::launch_gather(c.d_verdict, q.d_index, items, q.d_out, q.stream));
  PGX_HIP(s->h, hipMemcpyAsync(q.h_out, q.d_out, (size_t)items * 4, hipMemcpyDeviceToHost,
                               q.stream));
  PGX_HIP(s->h, hipEventRecord(q.done, q.stream));
  q.in_flight = true;
  return PGX_OK;
}

// Applies a finished iteration's verdicts (segments of requests still active:
// a request leaves only when neither side is held by an iteration).
void apply(pgx_server* s, Slot& q) {
  for (const Segment& g : q.segs) {
    Request& r = *g.req;
    r.tests += g.count;
    s->tests_total += g.count;
    for (uint32_t j = 0; j < g.count; ++j)
      if (q.h_out[g.off + j] == 1u) {  // the single base (2) passed
        r.first[g.side] = (int32_t)(r.cursor[g.side] + j);
        r.done[g.side] = true;
        break;
      }
    r.cursor[g.side] += g.count;
    if (r.cursor[g.side] >= r.n_side[g.side]) r.done[g.side] = true;
    r.busy[g.side] = false;
  }
  q.segs.clear();
}

// Requests with both sides decided leave the class and free their memory.
void complete(pgx_server* s, Class& c) {
  const uint64_t t = now_ns();
  auto keep = c.active.begin();
  for (auto it = c.active.begin(); it != c.active.end(); ++it) {
    Request& r = **it;
    if (r.done[0] && r.done[1]) {
      pgx_serve_result_t res{};
      res.id = r.id;
      res.next_first = r.first[0];
      res.prev_first = r.first[1];
      res.tests = r.tests;
      res.latency_ns = t - r.arrival_ns;
      s->finished.push_back(res);
      c.cols->give(r.col0, r.n_side[0] + r.n_side[1]);
      s->in_use -= r.bytes;
    } else {
      *keep++ = std::move(*it);
    }
  }
  c.active.erase(keep, c.active.end());
}

int advance(pgx_server* s, Class& c) {
  Slot* free_slot = nullptr;
  for (Slot& q : c.slots) {
    if (q.in_flight) {
      const hipError_t e = hipEventQuery(q.done);
      if (e == hipErrorNotReady) continue;
      if (e != hipSuccess)
        return fail(s->h, PGX_E_HIP, "serving iteration: %s", hipGetErrorString(e));
      q.in_flight = false;
      if (q.timed) {
        float busy = 0, gap = 0;
        if (hipEventElapsedTime(&busy, q.start, q.done) == hipSuccess) {
          s->iterations += 1;
          s->iter_ms += busy;
        }
        if (q.has_prev && hipEventElapsedTime(&gap, q.prev_done, q.start) == hipSuccess) {
          s->gaps += 1;
          s->gap_ms += gap;
        }
      }
      apply(s, q);
    }
    if (!free_slot) free_slot = &q;
  }
  complete(s, c);
  if (!free_slot) return PGX_OK;
  uint32_t staged = 0;
  int rc = admit(s, c, *free_slot, &staged);
  if (rc != PGX_OK) return rc;
  return launch_iteration(s, c, *free_slot, staged > 0);
}

}  // namespace

extern "C" {

int pgx_serve_create(pgx_server_t** out, int device, const pgx_serve_config_t* cfg) {
  if (!out || !cfg) return fail(nullptr, PGX_E_ARG, "null argument");
  *out = nullptr;
  if (cfg->chunk == 0 || cfg->max_requests == 0 || cfg->pool_columns == 0 ||
      cfg->max_side == 0 || cfg->max_side * 2ull > cfg->pool_columns ||
      (cfg->group_max_items && !pgx::valid_group((int)cfg->group_lanes)) ||
      cfg->wide_cu_percent > 99 || (cfg->wide_cu_percent && !cfg->group_min_bits) ||
      cfg->slots_per_class > 4)
    return fail(nullptr, PGX_E_ARG, "bad serving config");
  const uint32_t base2 = 2;
  pgx_runner_t* h = nullptr;
  int rc = pgx_runner_create(&h, device, 1, 1, &base2, 1);
  if (rc != PGX_OK) return rc;
  pgx_server* s = new (std::nothrow) pgx_server();
  if (!s) {
    pgx_runner_destroy(h);
    return fail(nullptr, PGX_E_ARG, "out of host memory");
  }
  s->h = h;
  s->cfg = *cfg;
  *out = s;
  return PGX_OK;
}

int pgx_serve_submit(pgx_server_t* s, const pgx_serve_request_t* req) {
  if (!s || !req) return PGX_E_ARG;
  if ((req->n_next && !req->next) || (req->n_prev && !req->prev) || req->limbs == 0)
    return fail(s->h, PGX_E_ARG, "request %llu: null candidates or zero limbs",
                (unsigned long long)req->id);
  if (req->n_next > s->cfg.max_side || req->n_prev > s->cfg.max_side)
    return fail(s->h, PGX_E_CAPACITY, "request %llu: more than %u candidates on a side",
                (unsigned long long)req->id, s->cfg.max_side);
  const int S = pgx::kernel_for_bits(req->max_bits);
  if (S == 0 || req->max_bits < 2)
    return fail(s->h, PGX_E_INPUT, "request %llu: %u-bit candidates are not supported",
                (unsigned long long)req->id, req->max_bits);
  Class* c = nullptr;
  int rc = make_class(s, S, &c);
  if (rc != PGX_OK) return rc;
  auto r = std::make_unique<Request>();
  r->id = req->id;
  r->arrival_ns = req->arrival_ns ? req->arrival_ns : now_ns();
  r->n_side[0] = req->n_next;
  r->n_side[1] = req->n_prev;
  const uint32_t n = req->n_next + req->n_prev;
  if (n == 0) {  // nothing to search: answered at once
    pgx_serve_result_t res{};
    res.id = req->id;
    res.next_first = res.prev_first = -1;
    res.latency_ns = now_ns() - r->arrival_ns;
    s->finished.push_back(res);
    return PGX_OK;
  }
  r->bytes = (uint64_t)n * (c->W + 1) * 4;
  if (r->bytes > s->cfg.hbm_budget)
    return fail(s->h, PGX_E_CAPACITY, "request %llu needs %llu bytes, budget %llu",
                (unsigned long long)req->id, (unsigned long long)r->bytes,
                (unsigned long long)s->cfg.hbm_budget);
  // limb-interleaved copy (W rows x n columns); reject what the class cannot hold
  r->cand.assign((size_t)c->W * n, 0u);
  for (int side = 0; side < 2; ++side) {
    const uint32_t* src = side ? req->prev : req->next;
    const uint32_t cnt = r->n_side[side], col = side ? req->n_next : 0;
    for (uint32_t i = 0; i < cnt; ++i) {
      const uint32_t* x = src + (size_t)i * req->limbs;
      for (uint32_t j = c->W; j < req->limbs; ++j)
        if (x[j]) return fail(s->h, PGX_E_INPUT, "request %llu: candidate wider than max_bits",
                              (unsigned long long)req->id);
      for (uint32_t j = 0; j < std::min(c->W, req->limbs); ++j)
        r->cand[(size_t)j * n + col + i] = x[j];
    }
  }
  for (int side = 0; side < 2; ++side) r->done[side] = (r->n_side[side] == 0);
  c->waiting.push_back(std::move(r));
  return PGX_OK;
}

int pgx_serve_poll(pgx_server_t* s, pgx_serve_result_t* out, uint32_t max_out, uint32_t* n_out) {
  if (!s || !n_out || (max_out && !out)) return PGX_E_ARG;
  PGX_HIP(s->h, hipSetDevice(s->h->device));
  for (auto& kv : s->classes) {
    int rc = advance(s, *kv.second);
    if (rc != PGX_OK) return rc;
  }
  const uint32_t k = (uint32_t)std::min<size_t>(max_out, s->finished.size());
  for (uint32_t i = 0; i < k; ++i) out[i] = s->finished[i];
  s->finished.erase(s->finished.begin(), s->finished.begin() + k);
  *n_out = k;
  return PGX_OK;
}

int pgx_serve_stats(const pgx_server_t* s, pgx_serve_stats_t* st) {
  if (!s || !st) return PGX_E_ARG;
  std::memset(st, 0, sizeof *st);
  st->bytes_in_use = s->in_use;
  st->tests = s->tests_total;
  st->finished_unread = (uint32_t)s->finished.size();
  st->iterations = s->iterations;
  st->iter_ms = s->iter_ms;
  st->gaps = s->gaps;
  st->gap_ms = s->gap_ms;
  for (auto& kv : s->classes) {
    st->active += (uint32_t)kv.second->active.size();
    st->waiting += (uint32_t)kv.second->waiting.size();
    for (const Slot& q : kv.second->slots) st->in_flight += q.in_flight ? 1u : 0u;
  }
  return PGX_OK;
}

int pgx_serve_destroy(pgx_server_t* s) {
  if (!s) return PGX_E_ARG;
  (void)hipSetDevice(s->h->device);
  for (auto& kv : s->classes) {
    for (Slot& q : kv.second->slots)
      if (q.stream) (void)hipStreamSynchronize(q.stream);
    release_class(*kv.second);
  }
  pgx_runner_destroy(s->h);
  delete s;
  return PGX_OK;
}

const char* pgx_serve_last_error(const pgx_server_t* s) {
  return s ? pgx_last_error(s->h) : pgx_last_error(nullptr);
}

}  // extern "C"
