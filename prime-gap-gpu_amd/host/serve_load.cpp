// serve_load.cpp — drives the serving engine (pgx_serve_*) with a Poisson
// stream of per-m next/prev-prime requests from one C++ thread, as a server
// process would, so that the engine's capacity is measured without an
// interpreter's cost per request.  The requests are the overflow searches the
// reference runs with mpz_nextprime / mpz_prevprime (overflow.cpp:142-160);
// the reference has no serving path to compare with.
//
// usage: serve_load TEMPLATES OUT RATE DURATION_S SEED [key=value ...]
//   TEMPLATES (little-endian): "PGXS", n_templates, then per template:
//     limbs, max_bits, n_next, n_prev, next[n_next * limbs], prev[n_prev * limbs]
//     (candidate-major words, ascending X)
//   RATE requests/s offered, DURATION_S seconds of arrivals (at most 2^20
//   requests), SEED of the arrival times and template picks (mt19937_64).
//   keys (pgx_serve_config_t): hbm_gb chunk max_requests pool_columns max_side
//     stage_columns group_min_bits group_max_items group_lanes wide_cu_percent
//     slots_per_class
//   OUT: n_req, then per request in id order: template (u32), next_first (i32),
//     prev_first (i32), tests (u32), latency_ns (u64; arrival to completion)
//   stdout: one summary line.
// Each request is submitted at its scheduled arrival time and its latency runs
// from that time, so the generator's own lateness counts against the engine.
// Before the stream, one request of each candidate width runs alone (the width
// class's allocation and first launch).  Exit status 0 when every request
// completed, 1 on an engine error or 30 s without progress, 2 on bad arguments.
#include <time.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "pgx.h"

namespace {

struct Template {
  uint32_t limbs = 0, max_bits = 0, n_next = 0, n_prev = 0;
  std::vector<uint32_t> next, prev;
};

uint64_t now_ns() {
  timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (uint64_t)ts.tv_sec * 1000000000ull + (uint64_t)ts.tv_nsec;
}

void read_exact(FILE* f, void* p, size_t bytes) {
  if (fread(p, 1, bytes, f) != bytes) throw std::runtime_error("short template file");
}

std::vector<Template> read_templates(const char* path) {
  FILE* f = fopen(path, "rb");
  if (!f) throw std::runtime_error(std::string("cannot open ") + path);
  char magic[4];
  uint32_t n = 0;
  read_exact(f, magic, 4);
  if (std::memcmp(magic, "PGXS", 4) != 0) throw std::runtime_error("bad magic");
  read_exact(f, &n, 4);
  std::vector<Template> out(n);
  for (Template& t : out) {
    uint32_t hdr[4];
    read_exact(f, hdr, sizeof hdr);
    t.limbs = hdr[0];
    t.max_bits = hdr[1];
    t.n_next = hdr[2];
    t.n_prev = hdr[3];
    if (t.limbs == 0 || t.limbs > 4096 || t.n_next > (1u << 20) || t.n_prev > (1u << 20))
      throw std::runtime_error("bad template header");
    t.next.resize((size_t)t.n_next * t.limbs);
    t.prev.resize((size_t)t.n_prev * t.limbs);
    read_exact(f, t.next.data(), t.next.size() * 4);
    read_exact(f, t.prev.data(), t.prev.size() * 4);
  }
  fclose(f);
  if (out.empty()) throw std::runtime_error("no templates");
  return out;
}

pgx_serve_request_t request_of(const Template& t, uint64_t id, uint64_t arrival) {
  pgx_serve_request_t r{};
  r.id = id;
  r.arrival_ns = arrival;
  r.limbs = t.limbs;
  r.max_bits = t.max_bits;
  r.n_next = t.n_next;
  r.n_prev = t.n_prev;
  r.next = t.n_next ? t.next.data() : nullptr;
  r.prev = t.n_prev ? t.prev.data() : nullptr;
  return r;
}

struct Record {
  uint32_t tmpl = 0;
  int32_t next_first = -2, prev_first = -2;
  uint32_t tests = 0;
  uint64_t latency_ns = 0;
};

}  // namespace

int main(int argc, char** argv) {
  if (argc < 6) {
    fprintf(stderr, "usage: %s TEMPLATES OUT RATE DURATION_S SEED [key=value ...]\n", argv[0]);
    return 2;
  }
  std::vector<Template> tmpl;
  try {
    tmpl = read_templates(argv[1]);
  } catch (const std::exception& e) {
    fprintf(stderr, "serve_load: %s\n", e.what());
    return 2;
  }
  const double rate = atof(argv[3]), duration = atof(argv[4]);
  const uint64_t seed = strtoull(argv[5], nullptr, 0);
  const double want_req = rate * duration;
  if (!(rate > 0) || !(duration > 0) || want_req < 1 || want_req > (double)(1u << 20)) {
    fprintf(stderr, "serve_load: RATE * DURATION_S must give 1 .. 2^20 requests\n");
    return 2;
  }
  pgx_serve_config_t cfg{};
  cfg.hbm_budget = 64ull << 30;
  cfg.chunk = 32;
  cfg.max_requests = 8192;
  cfg.pool_columns = 1u << 21;
  cfg.max_side = 512;
  cfg.stage_columns = 1u << 17;
  for (int k = 6; k < argc; ++k) {
    const char* eq = strchr(argv[k], '=');
    if (!eq) {
      fprintf(stderr, "serve_load: expected key=value, got %s\n", argv[k]);
      return 2;
    }
    const std::string key(argv[k], eq - argv[k]);
    const unsigned long long v = strtoull(eq + 1, nullptr, 0);
    if (key == "hbm_gb") cfg.hbm_budget = v << 30;
    else if (key == "chunk") cfg.chunk = (uint32_t)v;
    else if (key == "max_requests") cfg.max_requests = (uint32_t)v;
    else if (key == "pool_columns") cfg.pool_columns = (uint32_t)v;
    else if (key == "max_side") cfg.max_side = (uint32_t)v;
    else if (key == "stage_columns") cfg.stage_columns = (uint32_t)v;
    else if (key == "group_min_bits") cfg.group_min_bits = (uint32_t)v;
    else if (key == "group_max_items") cfg.group_max_items = (uint32_t)v;
    else if (key == "group_lanes") cfg.group_lanes = (uint32_t)v;
    else if (key == "wide_cu_percent") cfg.wide_cu_percent = (uint32_t)v;
    else if (key == "slots_per_class") cfg.slots_per_class = (uint32_t)v;
    else {
      fprintf(stderr, "serve_load: unknown key %s\n", key.c_str());
      return 2;
    }
  }

  pgx_server_t* srv = nullptr;
  if (pgx_serve_create(&srv, 0, &cfg) != PGX_OK) {
    fprintf(stderr, "serve_load: pgx_serve_create: %s\n", pgx_last_error(nullptr));
    return 1;
  }
  std::vector<pgx_serve_result_t> out(4096);
  int status = 0;
  auto fail = [&](const char* what) {
    fprintf(stderr, "serve_load: %s: %s\n", what, pgx_serve_last_error(srv));
    status = 1;
  };

  // warm-up: one request of each candidate width, alone
  std::map<uint32_t, size_t> first_of_width;
  for (size_t k = 0; k < tmpl.size(); ++k) first_of_width.emplace(tmpl[k].limbs, k);
  const uint64_t warm_id = 1ull << 62;
  size_t warm_left = 0;
  for (const auto& [limbs, k] : first_of_width) {
    const pgx_serve_request_t r = request_of(tmpl[k], warm_id + k, 0);
    if (pgx_serve_submit(srv, &r) != PGX_OK) {
      fail("warm-up submit");
      break;
    }
    ++warm_left;
  }
  const uint64_t warm_start = now_ns();
  while (status == 0 && warm_left > 0) {
    uint32_t got = 0;
    if (pgx_serve_poll(srv, out.data(), (uint32_t)out.size(), &got) != PGX_OK) {
      fail("warm-up poll");
      break;
    }
    warm_left -= got;
    if (now_ns() - warm_start > 60ull * 1000000000ull) {
      fprintf(stderr, "serve_load: warm-up did not finish in 60 s\n");
      status = 1;
    }
  }

  // the stream: arrival times and template picks from the seed
  const size_t n_req = (size_t)want_req;
  std::vector<Record> rec(n_req);
  std::vector<uint64_t> arrival(n_req);
  {
    std::mt19937_64 gen(seed);
    std::exponential_distribution<double> gap(rate);
    std::uniform_int_distribution<uint32_t> pick(0, (uint32_t)tmpl.size() - 1);
    double t = 0;
    for (size_t i = 0; i < n_req; ++i) {
      t += gap(gen);
      arrival[i] = (uint64_t)(t * 1e9);
      rec[i].tmpl = pick(gen);
    }
  }
  size_t next_req = 0, done = 0;
  const uint64_t start = now_ns();
  for (uint64_t& a : arrival) a += start;
  uint64_t last_progress = start;
  while (status == 0 && done < n_req) {
    const uint64_t now = now_ns();
    while (next_req < n_req && arrival[next_req] <= now) {
      const pgx_serve_request_t r =
          request_of(tmpl[rec[next_req].tmpl], next_req, arrival[next_req]);
      if (pgx_serve_submit(srv, &r) != PGX_OK) {
        fail("submit");
        break;
      }
      ++next_req;
    }
    uint32_t got = 0;
    if (status == 0 && pgx_serve_poll(srv, out.data(), (uint32_t)out.size(), &got) != PGX_OK)
      fail("poll");
    for (uint32_t j = 0; j < got; ++j) {
      const pgx_serve_result_t& r = out[j];
      if (r.id >= n_req) continue;  // a warm-up request
      Record& x = rec[r.id];
      x.next_first = r.next_first;
      x.prev_first = r.prev_first;
      x.tests = r.tests;
      x.latency_ns = r.latency_ns;
      ++done;
      last_progress = now;
    }
    if (status == 0 && now - last_progress > 30ull * 1000000000ull) {
      fprintf(stderr, "serve_load: no request completed for 30 s (%zu of %zu done)\n", done,
              n_req);
      status = 1;
    }
  }
  const double wall = (double)(now_ns() - start) * 1e-9;
  pgx_serve_stats_t st{};
  (void)pgx_serve_stats(srv, &st);
  pgx_serve_destroy(srv);
  if (status != 0) return status;

  FILE* f = fopen(argv[2], "wb");
  if (!f) {
    fprintf(stderr, "serve_load: cannot write %s\n", argv[2]);
    return 2;
  }
  const uint32_t n32 = (uint32_t)n_req;
  fwrite(&n32, 4, 1, f);
  uint64_t tests = 0;
  for (const Record& x : rec) {
    fwrite(&x.tmpl, 4, 1, f);
    fwrite(&x.next_first, 4, 1, f);
    fwrite(&x.prev_first, 4, 1, f);
    fwrite(&x.tests, 4, 1, f);
    fwrite(&x.latency_ns, 8, 1, f);
    tests += x.tests;
  }
  fclose(f);
  printf("serve_load: %zu requests offered at %.0f/s, served in %.6f s (%.0f/s), %llu tests; "
         "iterations %llu, iteration_ms %.3f, gaps %llu, gap_ms %.3f\n",
         n_req, rate, wall, n_req / wall, (unsigned long long)tests,
         (unsigned long long)st.iterations, st.iter_ms, (unsigned long long)st.gaps, st.gap_ms);
  return 0;
}
