// tga_server.cpp -- continuous-batching serving loop over HBM-resident columns.
//
// Requests are queued by tga_server_submit.  One scheduler thread
//   * admits queued requests, oldest first, while their device footprint fits
//     the memory budget,
//   * hands each admitted request of >= 2^24 rows to the solo lanes: worker
//     threads, each with its own lane context (stream and arena), that run
//     such a request alone through the plain operator (device-wide scans,
//     direct emission) while the scheduler goes on,
//   * turns each admitted group of smaller requests into one batched
//     persistent launch (the work queue of tga_batch.hip) on its own stream,
//     so flights overlap on the device with each other and with the solo
//     lanes while new requests keep arriving,
//   * polls the flights' events, finishes completed flights (overflowed
//     queries rerun alone) and hands each request its own result,
// and records latency = completion - submission for p50/p99.
//
// The reference has no serving mode (one blocking call per dataset,
// test/test.cu:137-161); this follows the batched path's design.
#include <algorithm>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <thread>
#include <unordered_set>

#include "tga_batch_internal.hpp"

using namespace tga;
using clk = std::chrono::steady_clock;

namespace {

struct Req {
  int64_t ticket;
  tga_query q;
  std::vector<tga_column> vals;
  std::vector<tga_agg> aggs;
  clk::time_point t_submit;
  uint64_t bytes;
  uint64_t est;
};

struct Flight {
  std::unique_ptr<BatchExec> ex;
  std::vector<Req> reqs;
  int slot;
  hipEvent_t ev;
  uint64_t bytes;
};

struct Done {
  tga_result *r;
  tga_status st;
};

double pct(std::vector<double> v, double p) {
  if (v.empty()) return 0;
  std::sort(v.begin(), v.end());
  const double idx = p * (double)(v.size() - 1);
  const size_t lo = (size_t)idx;
  const size_t hi = std::min(v.size() - 1, lo + 1);
  return v[lo] + (v[hi] - v[lo]) * (idx - (double)lo);
}

}  // namespace

// rows per flight before a new flight is started
constexpr uint64_t kFlightRows = 1ULL << 26;
// A request this large fills the GPU by itself: it runs alone through the
// plain operator (device-wide scans, direct emission) on a solo lane, while
// the flights of small requests keep running beside it.
constexpr uint64_t kSoloRows = 1ULL << 24;
// Solo lanes: while one waits for its request's group count, the next
// request's kernels run on another.
constexpr int kSoloLanes = 2;

struct tga_server {
  tga_context *ctx = nullptr;
  uint64_t budget = 0;
  int max_batch = 256;
  int nslots = 3;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Req> pending;
  std::map<int64_t, Done> done;
  std::unordered_set<int64_t> live;  // tickets submitted and not yet taken
  std::vector<Flight> flights;
  std::vector<hipStream_t> streams;
  std::vector<bool> slot_busy;
  std::vector<double> lat_us;
  uint64_t reserved = 0, peak = 0;
  int64_t next_ticket = 1, submitted = 0, completed = 0, rejected = 0, launches = 0, solo_runs = 0;
  double build_s = 0, finish_s = 0;  // scheduler thread: building and finishing flights
  clk::time_point t0;
  bool started = false, stop = false;
  std::thread worker;
  // solo lanes: admitted large requests wait in `solo`; lane i runs them on
  // its own context
  std::deque<Req> solo;
  std::condition_variable cv_solo;
  std::vector<tga_context *> solo_ctx;
  std::vector<std::thread> solo_workers;

  void loop();
  bool finish_ready();
  bool launch_one();
  bool admit_solo();
  void solo_loop(tga_context *lc);
  void run_solo(tga_context *lc, Req &r);
};

// Moves the oldest request to the solo queue when it is a large one and its
// footprint fits the budget (oldest first: a large request that does not fit
// yet holds back the requests behind it).
bool tga_server::admit_solo() {
  std::lock_guard<std::mutex> g(mu);
  if (pending.empty()) return false;
  const Req &f = pending.front();
  if ((uint64_t)f.q.n_rows < kSoloRows || reserved + f.bytes > budget) return false;
  reserved += f.bytes;
  peak = std::max(peak, reserved);
  solo.push_back(std::move(pending.front()));
  pending.pop_front();
  cv_solo.notify_one();
  return true;
}

void tga_server::solo_loop(tga_context *lc) {
  hipSetDevice(lc->device);
  for (;;) {
    Req r;
    {
      std::unique_lock<std::mutex> g(mu);
      cv_solo.wait(g, [&] { return stop || !solo.empty(); });
      if (solo.empty()) return;  // stop, and nothing left
      r = std::move(solo.front());
      solo.pop_front();
    }
    run_solo(lc, r);
  }
}

void tga_server::run_solo(tga_context *lc, Req &r) {
  Done d{nullptr, TGA_OK};
  {
    tga_context *ctx = lc;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    hipSetDevice(ctx->device);
    QuerySpec spec;
    build_spec(&r.q.key, r.vals.data(), (int32_t)r.vals.size(), r.aggs.data(), (int32_t)r.aggs.size(), spec);
    // the caller's hint, else the estimator (r.est is only the admission bound)
    uint64_t est = r.q.est_groups > 0 ? (uint64_t)r.q.est_groups : 0;
    if (!est) d.st = estimate_groups(ctx, r.q.key.data, spec.key_bytes, (uint64_t)r.q.n_rows, &est);
    PlainOut po;
    uint64_t G = 0;
    if (!d.st) {
      const Plan p = autotune(ctx, spec, (uint64_t)r.q.n_rows, est, nullptr);
      d.st = run_plain(ctx, spec, r.q.key.data, (uint64_t)r.q.n_rows, p, &po, nullptr, nullptr, 0, &G);
    }
    if (!d.st) {
      auto *res = new tga_result();
      res->ctx = ctx;
      res->key_dtype = spec.key_dtype;
      res->naggs = spec.schema.nacc;
      for (int a = 0; a < spec.schema.nacc; ++a) res->dtype[a] = spec.out_dtype[a];
      res->out = std::move(po);
      d.r = res;
    }
  }
  const auto now = clk::now();
  std::lock_guard<std::mutex> g(mu);
  done[r.ticket] = d;
  lat_us.push_back(std::chrono::duration<double, std::micro>(now - r.t_submit).count());
  ++completed;
  ++launches;
  ++solo_runs;
  reserved -= r.bytes;
  cv_done.notify_all();
}

bool tga_server::finish_ready() {
  bool any = false;
  for (size_t i = 0; i < flights.size();) {
    Flight &f = flights[i];
    if (hipEventQuery(f.ev) != hipSuccess) { ++i; continue; }
    any = true;
    const auto tf = clk::now();
    std::vector<uint64_t> groups;
    tga_status st;
    {
      std::lock_guard<std::recursive_mutex> lk(ctx->mu);
      hipSetDevice(ctx->device);
      st = batch_finalize(*f.ex, groups);
    }
    const auto now = clk::now();
    std::lock_guard<std::mutex> g(mu);
    for (size_t j = 0; j < f.reqs.size(); ++j) {
      Done d{nullptr, st};
      if (!st) {
        auto *r = new tga_result();
        r->ctx = ctx;
        const QuerySpec &q = f.ex->spec[j];
        r->key_dtype = q.key_dtype;
        r->naggs = q.schema.nacc;
        for (int a = 0; a < q.schema.nacc; ++a) r->dtype[a] = q.out_dtype[a];
        r->out.keys = std::move(f.ex->okeys[j]);
        for (int a = 0; a < q.schema.nacc; ++a) r->out.cols[a] = std::move(f.ex->ocols[j][a]);
        r->out.groups = groups[j];
        r->out.capacity = f.ex->cap[j];
        d.r = r;
      }
      done[f.reqs[j].ticket] = d;
      lat_us.push_back(std::chrono::duration<double, std::micro>(now - f.reqs[j].t_submit).count());
      ++completed;
    }
    reserved -= f.bytes;
    slot_busy[f.slot] = false;
    hipEventDestroy(f.ev);
    {
      std::lock_guard<std::recursive_mutex> lk(ctx->mu);
      f.ex.reset();  // returns tables and buffers to the arena
    }
    flights.erase(flights.begin() + (long)i);
    finish_s += std::chrono::duration<double>(clk::now() - tf).count();
    cv_done.notify_all();
  }
  return any;
}

bool tga_server::launch_one() {
  std::vector<Req> take;
  int slot = -1;
  uint64_t bytes = 0;
  {
    std::lock_guard<std::mutex> g(mu);
    if (pending.empty()) return false;
    for (int s = 0; s < nslots; ++s)
      if (!slot_busy[s]) { slot = s; break; }
    if (slot < 0) return false;
    // A flight holds at most max_batch requests and about kFlightRows rows,
    // so a large request travels (nearly) alone and the small requests
    // admitted after it do not wait for it: they go out in the next flight,
    // on another stream, and overlap with it on the device.
    uint64_t rows = 0;
    while (!pending.empty() && (int)take.size() < max_batch) {
      const Req &r = pending.front();
      if (reserved + bytes + r.bytes > budget) break;  // oldest first: wait for memory
      if ((uint64_t)r.q.n_rows >= kSoloRows) break;      // a solo lane's (admit_solo)
      if (!take.empty() && rows + (uint64_t)r.q.n_rows > kFlightRows) break;
      bytes += r.bytes;
      rows += (uint64_t)r.q.n_rows;
      take.push_back(std::move(pending.front()));
      pending.pop_front();
    }
    if (take.empty()) return false;
    reserved += bytes;
    peak = std::max(peak, reserved);
    slot_busy[slot] = true;
  }
  std::vector<tga_query> qs(take.size());
  for (size_t j = 0; j < take.size(); ++j) {
    qs[j] = take[j].q;
    qs[j].values = take[j].vals.data();
    qs[j].aggs = take[j].aggs.data();
    // the caller's hint (0: batch_prepare estimates it); take[j].est, every
    // row its own group when there is no hint, only bounds the admission
    qs[j].est_groups = take[j].q.est_groups;
  }
  const auto tb = clk::now();
  Flight f;
  f.ex = std::make_unique<BatchExec>();
  f.slot = slot;
  f.bytes = bytes;
  tga_status st;
  {
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    hipSetDevice(ctx->device);
    st = batch_prepare(ctx, qs.data(), (int32_t)qs.size(), *f.ex);
    if (!st) st = batch_build(*f.ex, qs.data(), streams[slot]);
    if (!st) st = batch_enqueue(*f.ex, streams[slot]);
    if (!st && hipEventCreateWithFlags(&f.ev, hipEventDisableTiming) != hipSuccess) st = TGA_ERR_HIP;
    if (!st && hipEventRecord(f.ev, streams[slot]) != hipSuccess) st = TGA_ERR_HIP;
    if (st) { hipStreamSynchronize(streams[slot]); f.ex.reset(); }
  }
  std::lock_guard<std::mutex> g(mu);
  ++launches;
  build_s += std::chrono::duration<double>(clk::now() - tb).count();
  if (st) {
    for (auto &r : take) { done[r.ticket] = Done{nullptr, st}; ++completed; }
    reserved -= bytes;
    slot_busy[slot] = false;
    cv_done.notify_all();
    return true;
  }
  f.reqs = std::move(take);
  flights.push_back(std::move(f));
  return true;
}

void tga_server::loop() {
  hipSetDevice(ctx->device);
  for (;;) {
    bool did = false;
    if (!flights.empty()) did |= finish_ready();
    while (admit_solo()) did = true;
    did |= launch_one();
    std::unique_lock<std::mutex> g(mu);
    if (stop && pending.empty() && flights.empty()) break;
    if (!did) {
      if (flights.empty() && pending.empty()) cv_work.wait_for(g, std::chrono::milliseconds(2));
      else {
        g.unlock();
        std::this_thread::sleep_for(std::chrono::microseconds(10));
      }
    }
  }
}

#define TGA_API extern "C" __attribute__((visibility("default")))

TGA_API tga_status tga_server_create(tga_context *ctx, uint64_t budget_bytes, int32_t max_batch, tga_server **out) {
  if (!ctx || !out || max_batch < 0) return TGA_ERR_INVALID_ARG;
  *out = nullptr;
  auto *s = new tga_server();
  s->ctx = ctx;
  s->budget = budget_bytes ? budget_bytes : ctx->arena->budget() / 2;
  s->max_batch = max_batch ? max_batch : 256;
  {
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    hipSetDevice(ctx->device);
    for (int i = 0; i < s->nslots; ++i) {
      hipStream_t st = nullptr;
      if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
        for (auto x : s->streams) hipStreamDestroy(x);
        delete s;
        return fail(ctx, TGA_ERR_HIP, "server stream");
      }
      s->streams.push_back(st);
    }
  }
  s->slot_busy.assign(s->nslots, false);
  {
    // the solo lanes' contexts belong to ctx (its lane list): results of
    // solo requests hold their memory and may outlive the server
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    tga_status lst = TGA_OK;
    const std::vector<tga_context *> lanes = lanes_of(ctx, 1 + kSoloLanes, &lst);
    if (lst) {
      for (auto x : s->streams) hipStreamDestroy(x);
      delete s;
      return lst;
    }
    s->solo_ctx.assign(lanes.begin() + 1, lanes.end());
  }
  s->t0 = clk::now();
  for (tga_context *lc : s->solo_ctx) s->solo_workers.emplace_back([s, lc] { s->solo_loop(lc); });
  s->worker = std::thread([s] { s->loop(); });
  *out = s;
  return TGA_OK;
}

TGA_API tga_status tga_server_submit(tga_server *s, const tga_query *q, int64_t *ticket) {
  if (!s || !q || !ticket) return TGA_ERR_INVALID_ARG;
  Req r;
  QuerySpec spec;
  std::string e = build_spec(&q->key, q->values, q->n_values, q->aggs, q->n_aggs, spec);
  if (e.empty() && q->key.location != TGA_DEVICE) e = "served queries need device-resident columns";
  if (e.empty() && q->n_rows < 0) e = "negative row count";
  if (!e.empty()) return fail(s->ctx, TGA_ERR_INVALID_ARG, "%s", e.c_str());
  r.q = *q;
  r.vals.assign(q->values, q->values + q->n_values);
  r.aggs.assign(q->aggs, q->aggs + q->n_aggs);
  // without a hint, admission assumes every row is its own group
  r.est = q->est_groups > 0 ? (uint64_t)q->est_groups : (uint64_t)std::max<int64_t>(1, q->n_rows);
  r.bytes = batch_query_bytes(s->ctx, spec, (uint64_t)q->n_rows, r.est);
  r.t_submit = clk::now();
  std::lock_guard<std::mutex> g(s->mu);
  if (r.bytes > s->budget) {
    ++s->rejected;
    return fail(s->ctx, TGA_ERR_BUDGET, "request needs %llu B > budget %llu B", (unsigned long long)r.bytes,
                (unsigned long long)s->budget);
  }
  if (!s->started) { s->started = true; s->t0 = r.t_submit; }
  r.ticket = s->next_ticket++;
  *ticket = r.ticket;
  s->live.insert(r.ticket);
  ++s->submitted;
  s->pending.push_back(std::move(r));
  s->cv_work.notify_one();
  return TGA_OK;
}

static tga_status take_result(tga_server *s, int64_t ticket, tga_result **out, bool wait) {
  std::unique_lock<std::mutex> g(s->mu);
  for (;;) {
    auto it = s->done.find(ticket);
    if (it != s->done.end()) {
      Done d = it->second;
      s->done.erase(it);
      s->live.erase(ticket);
      if (out) *out = d.r;
      else if (d.r) tga_result_free(d.r);
      return d.st;
    }
    if (!s->live.count(ticket)) return TGA_ERR_INVALID_ARG;  // never issued, or already taken
    if (!wait) return TGA_ERR_NOT_READY;
    s->cv_done.wait_for(g, std::chrono::milliseconds(5));
  }
}

TGA_API tga_status tga_server_poll(tga_server *s, int64_t ticket, tga_result **out) {
  if (!s) return TGA_ERR_INVALID_ARG;
  return take_result(s, ticket, out, false);
}

TGA_API tga_status tga_server_wait(tga_server *s, int64_t ticket, tga_result **out) {
  if (!s) return TGA_ERR_INVALID_ARG;
  return take_result(s, ticket, out, true);
}

TGA_API tga_status tga_server_stats_get(tga_server *s, tga_server_stats *st) {
  if (!s || !st) return TGA_ERR_INVALID_ARG;
  std::lock_guard<std::mutex> g(s->mu);
  *st = tga_server_stats{};
  st->submitted = s->submitted;
  st->completed = s->completed;
  st->rejected = s->rejected;
  st->inflight = s->submitted - s->completed;
  st->p50_us = pct(s->lat_us, 0.50);
  st->p99_us = pct(s->lat_us, 0.99);
  double sum = 0, mx = 0;
  for (double x : s->lat_us) { sum += x; mx = std::max(mx, x); }
  st->mean_us = s->lat_us.empty() ? 0 : sum / (double)s->lat_us.size();
  st->max_us = mx;
  st->elapsed_s = std::chrono::duration<double>(clk::now() - s->t0).count();
  st->qps = st->elapsed_s > 0 ? (double)s->completed / st->elapsed_s : 0;
  st->bytes_in_use = s->reserved;
  st->peak_bytes = s->peak;
  st->budget_bytes = s->budget;
  st->launches = s->launches;
  st->solo_runs = s->solo_runs;
  st->build_s = s->build_s;
  st->finish_s = s->finish_s;
  return TGA_OK;
}

TGA_API tga_status tga_server_destroy(tga_server *s) {
  if (!s) return TGA_ERR_INVALID_ARG;
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;
    s->cv_work.notify_all();
  }
  if (s->worker.joinable()) s->worker.join();
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->cv_solo.notify_all();
  }
  for (auto &w : s->solo_workers) w.join();
  for (auto &kv : s->done)
    if (kv.second.r) tga_result_free(kv.second.r);
  {
    std::lock_guard<std::recursive_mutex> lk(s->ctx->mu);
    for (auto x : s->streams) { hipStreamSynchronize(x); hipStreamDestroy(x); }
  }
  delete s;
  return TGA_OK;
}

// Replays an arrival trace: query i is submitted at t0 + arrival_us[i] (busy
// wait on the calling thread), then every result is awaited and freed.
// The stats of the run are returned; wall_s receives first submit -> last done.
TGA_API tga_status tga_server_replay(tga_server *s, const tga_query *qs, const double *arrival_us, int64_t n,
                                     tga_server_stats *out, double *wall_s, int64_t *groups) {
  if (!s || (n > 0 && (!qs || !arrival_us)) || n < 0) return TGA_ERR_INVALID_ARG;
  std::vector<int64_t> tickets((size_t)n, 0);
  const auto t0 = clk::now();
  tga_status first = TGA_OK;
  for (int64_t i = 0; i < n; ++i) {
    const auto due = t0 + std::chrono::duration_cast<clk::duration>(std::chrono::duration<double, std::micro>(arrival_us[i]));
    while (clk::now() < due) {}
    tga_status st = tga_server_submit(s, &qs[i], &tickets[(size_t)i]);
    if (st && !first) first = st;
  }
  for (int64_t i = 0; i < n; ++i) {
    if (groups) groups[i] = -1;
    if (!tickets[(size_t)i]) continue;
    tga_result *r = nullptr;
    tga_status st = tga_server_wait(s, tickets[(size_t)i], &r);
    if (st && !first) first = st;
    if (r) {
      if (groups) groups[i] = (int64_t)r->out.groups;
      tga_result_free(r);
    }
  }
  if (wall_s) *wall_s = std::chrono::duration<double>(clk::now() - t0).count();
  if (out) tga_server_stats_get(s, out);
  return first;
}

// The same trace through the plain operator, one query at a time in arrival
// order on the calling thread (latency includes the wait for earlier queries).
TGA_API tga_status tga_sequential_replay(tga_context *ctx, const tga_query *qs, const double *arrival_us, int64_t n,
                                         tga_server_stats *out, double *wall_s, int64_t *groups) {
  if (!ctx || (n > 0 && (!qs || !arrival_us)) || n < 0 || !out) return TGA_ERR_INVALID_ARG;
  std::vector<double> lat;
  const auto t0 = clk::now();
  tga_status first = TGA_OK;
  for (int64_t i = 0; i < n; ++i) {
    const auto due = t0 + std::chrono::duration_cast<clk::duration>(std::chrono::duration<double, std::micro>(arrival_us[i]));
    while (clk::now() < due) {}
    tga_options o{};
    o.est_groups = qs[i].est_groups;
    tga_result *r = nullptr;
    tga_status st = tga_groupby(ctx, &qs[i].key, qs[i].values, qs[i].n_values, qs[i].n_rows, qs[i].aggs, qs[i].n_aggs,
                                &o, &r);
    if (st && !first) first = st;
    if (groups) groups[i] = r ? (int64_t)r->out.groups : -1;
    if (r) tga_result_free(r);
    lat.push_back(std::chrono::duration<double, std::micro>(clk::now() - due).count());
  }
  const double w = std::chrono::duration<double>(clk::now() - t0).count();
  if (wall_s) *wall_s = w;
  *out = tga_server_stats{};
  out->submitted = out->completed = n;
  out->p50_us = pct(lat, 0.5);
  out->p99_us = pct(lat, 0.99);
  double sum = 0, mx = 0;
  for (double x : lat) { sum += x; mx = std::max(mx, x); }
  out->mean_us = n ? sum / (double)n : 0;
  out->max_us = mx;
  out->elapsed_s = w;
  out->qps = w > 0 ? (double)n / w : 0;
  out->launches = n;
  return first;
}
