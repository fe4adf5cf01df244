// Open-loop load generator for the serving benchmark (serve_bench.py): replays
// one arrival trace of NTT / coset-LDE jobs through the plain path (one call
// per job, synchronous) or the server (sfn_server_*), natively, so the
// driver's own per-job cost stays far below the jobs' cost.
//
//   serve_load <trace> <plain|server> <device|host> [inflight] [max_group] [busy]
//
// <trace>: little-endian u32 job count J and u32 rate count R, then J records
// of 6 u32 (kind 0 = NTT / 1 = LDE, field, log_n, batch, added_bits, 0), then
// R x J f64 arrival times in seconds (one arrival sequence per offered rate).
// Inputs: one shared buffer of valid field values (BabyBear values are below
// both moduli); outputs: one device buffer per job (device) or one shared
// pinned output (host: uploads and downloads are timed, results are checked by
// the tests instead).  Every shape runs once untimed first (tables, streams,
// staging).  Prints one JSON line per arrival sequence: per-job latency
// (arrival -> result in place) in microseconds, the span from the first
// arrival to the last completion, and the mean fused group size (server).
// With "busy" the plain path also times every call's launch sequence on the
// device (hipEvent pair, sfn_call_timing) and reports their sum over the span
// as device_busy: the share of the time the GPU had work.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <string>
#include <thread>
#include <vector>

#include "sfntt.h"

namespace {

struct Job {
  uint32_t kind, field, log_n, batch, added_bits, pad;
};

uint64_t now_ns() { return sfn_now(); }

void wait_until(uint64_t t_ns) {
  for (;;) {
    const uint64_t n = now_ns();
    if (n >= t_ns) return;
    if (t_ns - n > 200000) std::this_thread::sleep_for(std::chrono::nanoseconds(t_ns - n - 100000));
  }
}

#define CHECK(x)                                                                       \
  do {                                                                                 \
    sfn_status s_ = (x);                                                               \
    if (s_ != SFN_OK) {                                                                \
      fprintf(stderr, "serve_load: %s failed: %s\n", #x, sfn_status_string(s_)); \
      exit(2);                                                                         \
    }                                                                                  \
  } while (0)

}  // namespace

int main(int argc, char **argv) {
  if (argc < 4) {
    fprintf(stderr, "usage: serve_load <trace> <plain|server> <device|host> [inflight] [max_group]\n");
    return 2;
  }
  const bool server = std::string(argv[2]) == "server";
  const bool device = std::string(argv[3]) == "device";
  const uint32_t inflight = argc > 4 ? (uint32_t)atoi(argv[4]) : 4;
  const uint32_t max_group = argc > 5 ? (uint32_t)atoi(argv[5]) : 0;
  const bool busy = argc > 6 && std::string(argv[6]) == "busy";
  std::vector<Job> jobs;
  std::vector<std::vector<double>> arrivals;
  {
    FILE *f = fopen(argv[1], "rb");
    if (!f) return 2;
    uint32_t nj = 0, nr = 0;
    if (fread(&nj, 4, 1, f) != 1 || fread(&nr, 4, 1, f) != 1) return 2;
    jobs.resize(nj);
    if (nj && fread(jobs.data(), sizeof(Job), nj, f) != nj) return 2;
    arrivals.assign(nr, std::vector<double>(nj));
    for (auto &a : arrivals)
      if (nj && fread(a.data(), 8, nj, f) != nj) return 2;
    fclose(f);
  }
  if (jobs.empty() || arrivals.empty()) return 2;
  sfn_ctx *ctx = nullptr;
  CHECK(sfn_ctx_create(0, &ctx));
  size_t max_in = 0, max_out = 0;
  for (const Job &j : jobs) {
    max_in = std::max(max_in, (size_t)j.batch << j.log_n);
    max_out = std::max(max_out, (size_t)j.batch << (j.log_n + j.added_bits));
  }
  // valid inputs for every field: small values (below every modulus)
  std::vector<uint32_t> host_in(max_in);
  for (size_t i = 0; i < max_in; ++i) host_in[i] = (uint32_t)((i * 2654435761ull) % 0x78000001u);
  void *d_in = nullptr;
  CHECK(sfn_buf_alloc(ctx, max_in * 4, &d_in));
  CHECK(sfn_upload(ctx, d_in, host_in.data(), max_in * 4, nullptr));
  std::vector<void *> d_out;
  void *h_in = nullptr, *h_out = nullptr;
  if (device) {
    for (const Job &j : jobs) {
      void *p = nullptr;
      CHECK(sfn_buf_alloc(ctx, ((size_t)j.batch << (j.log_n + j.added_bits)) * 4, &p));
      d_out.push_back(p);
    }
  } else {
    CHECK(sfn_host_alloc(ctx, max_in * 4, &h_in));
    CHECK(sfn_host_alloc(ctx, max_out * 4, &h_out));
    memcpy(h_in, host_in.data(), max_in * 4);
  }
  CHECK(sfn_synchronize(ctx, nullptr));
  double device_ns = 0;  // "busy": the launch sequences' device time
  auto plain_call = [&](size_t k) {
    const Job &j = jobs[k];
    uint64_t ns = 0;
    if (device) {
      if (j.kind == 1)
        CHECK(sfn_coset_lde_batch_device(ctx, j.log_n, j.batch, j.added_bits, 0, SFN_LAYOUT_SOA, (const uint32_t *)d_in,
                                         (uint32_t *)d_out[k], nullptr));
      else
        CHECK(sfn_ntt_device(ctx, (sfn_field)j.field, SFN_VARIANT_AUTO, SFN_FORWARD, j.log_n, j.batch, SFN_LAYOUT_SOA,
                             (const uint32_t *)d_in, (uint32_t *)d_out[k], nullptr));
      CHECK(sfn_synchronize(ctx, nullptr));
      if (busy) CHECK(sfn_last_device_ns(ctx, &ns));
    } else if (j.kind == 1) {
      CHECK(sfn_coset_lde_batch(ctx, j.log_n, j.batch, j.added_bits, 0, SFN_LAYOUT_SOA, (const uint32_t *)h_in,
                                (uint32_t *)h_out, &ns));
    } else {
      CHECK(sfn_ntt(ctx, (sfn_field)j.field, SFN_VARIANT_AUTO, SFN_FORWARD, j.log_n, j.batch, SFN_LAYOUT_SOA,
                    (const uint32_t *)h_in, (uint32_t *)h_out, &ns));
    }
    device_ns += (double)ns;
  };
  if (busy) CHECK(sfn_call_timing(ctx, 1));
  auto request = [&](size_t k) {
    const Job &j = jobs[k];
    sfn_request r{};
    r.kind = j.kind == 1 ? SFN_JOB_LDE : SFN_JOB_NTT;
    r.field = (sfn_field)j.field;
    r.dir = SFN_FORWARD;
    r.log_n = j.log_n;
    r.batch = j.batch;
    r.added_bits = j.added_bits;
    r.device_ptrs = device ? 1 : 0;
    r.in = device ? (const uint32_t *)d_in : (const uint32_t *)h_in;
    r.out = device ? (uint32_t *)d_out[k] : (uint32_t *)h_out;
    return r;
  };
  // untimed pass over every job: tables, kernels, streams
  for (size_t k = 0; k < jobs.size(); ++k) plain_call(k);
  sfn_server *srv = nullptr;
  std::vector<uint64_t> tickets(jobs.size());
  if (server) {
    sfn_server_config cfg{};
    cfg.max_inflight = inflight;
    cfg.max_group = max_group;
    CHECK(sfn_server_create(ctx, &cfg, &srv));
    // warm the workers' streams and staging with the whole trace, untimed
    for (size_t k = 0; k < jobs.size(); ++k) {
      sfn_request r = request(k);
      CHECK(sfn_server_submit(srv, &r, &tickets[k]));
    }
    for (size_t k = 0; k < jobs.size(); ++k) CHECK(sfn_server_wait(srv, tickets[k], nullptr));
  }
  for (const std::vector<double> &arr : arrivals) {
    std::vector<double> lat(jobs.size());
    double span = 0, mean_group = 1;
    const uint64_t t0 = now_ns() + 1000000;
    uint64_t last = t0;
    device_ns = 0;
    if (!server) {
      for (size_t k = 0; k < jobs.size(); ++k) {
        const uint64_t at = t0 + (uint64_t)(arr[k] * 1e9);
        wait_until(at);
        plain_call(k);
        last = now_ns();
        lat[k] = (double)(last - at) / 1e3;
      }
    } else {
      std::thread submitter([&] {
        for (size_t k = 0; k < jobs.size(); ++k) {
          wait_until(t0 + (uint64_t)(arr[k] * 1e9));
          sfn_request r = request(k);
          CHECK(sfn_server_submit(srv, &r, &tickets[k]));
        }
      });
      submitter.join();
      double groups = 0;
      for (size_t k = 0; k < jobs.size(); ++k) {
        sfn_job_timing tm{};
        CHECK(sfn_server_wait(srv, tickets[k], &tm));
        const uint64_t at = t0 + (uint64_t)(arr[k] * 1e9);
        lat[k] = (double)((int64_t)tm.end_ns - (int64_t)at) / 1e3;
        last = std::max<uint64_t>(last, tm.end_ns);
        groups += tm.group_size;
      }
      mean_group = groups / (double)jobs.size();
    }
    span = (double)(last - t0) / 1e9;
    printf("{\"span_s\": %.6f, \"mean_group\": %.3f, ", span, mean_group);
    if (busy && !server) printf("\"device_busy\": %.4f, ", device_ns / 1e9 / span);
    printf("\"lat_us\": [");
    for (size_t k = 0; k < lat.size(); ++k) printf("%s%.1f", k ? ", " : "", lat[k]);
    printf("]}\n");
    fflush(stdout);
  }
  if (srv) sfn_server_destroy(srv);
  for (void *p : d_out) sfn_buf_free(ctx, p);
  sfn_buf_free(ctx, d_in);
  if (h_in) sfn_host_free(ctx, h_in);
  if (h_out) sfn_host_free(ctx, h_out);
  sfn_ctx_destroy(ctx);
  return 0;
}
