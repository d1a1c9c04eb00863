// serving.cpp — the request loop and the batched solve entry.
//
// Server: admitted requests wait in one FIFO and are taken by two kinds of device lanes, each
// a thread owning one HIP stream (default priority) and its own cache of resident plans:
//   * plan workers take requests one at a time and solve them with a resident plan;
//   * batch lanes take the requests the lock-step batched engine handles: the first one in
//     the queue and every later queued request with the same options (up to batch_max), and
//     solve them as one batch, one launch per operation over every member.  A lane that is busy
//     leaves new arrivals to the next free lane, so requests that arrive while others run start
//     as soon as a lane and memory free up, and the lanes' kernels overlap on the device.
// A request is admitted while the bytes reserved by running requests plus its own fit the HBM
// budget; its bytes are the measured size of an earlier plan of the same shape and options
// (an estimate before the first).  Operators (hierarchy, coefficient fields, boundary data)
// stay resident per lane in an LRU cache keyed by a hash of the problem minus its source; a
// hit only uploads the new source and rebuilds the right-hand side.  Deadlines end running
// solves on the device: a watchdog raises a plan solve's cancel flag; a batch lane marks the
// member broken down between polling chunks.  Every solve runs the same kernels with the same
// launch shapes (batch members with fixed work-item shapes and per-member reductions) whatever
// else is in flight, so each result is bitwise the same alone and in a mix.
#include "serving.hpp"
#include "batch.hpp"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <functional>
#include <list>
#include <map>
#include <thread>
#include <unordered_map>

namespace pdehip {

namespace {

uint64_t mix(uint64_t h, uint64_t v) {
  h ^= v + 0x9e3779b97f4a7c15ULL + (h << 6) + (h >> 2);
  return h;
}
uint64_t hash_bytes(uint64_t h, const void* p, size_t n) {
  const unsigned char* c = static_cast<const unsigned char*>(p);
  uint64_t x = 1469598103934665603ULL;
  for (size_t i = 0; i < n; ++i) x = (x ^ c[i]) * 1099511628211ULL;
  return mix(h, x);
}
uint64_t hash_d(uint64_t h, double v) { return hash_bytes(h, &v, sizeof v); }

// Owned deep copy of a problem (the caller's arrays may go away after submit).
struct OwnedProblem {
  pdehip_problem p{};
  std::vector<double> coef[PDEHIP_NCOEF], kappa, f, initial, vel;
  std::vector<double> samples[6][4];
  void take(const pdehip_problem& src) {
    p = src;
    const size_t n = (size_t)src.nx * src.ny * std::max(1, src.nz);
    auto cp = [&](const double* s, std::vector<double>& d, const double** out) {
      if (s) { d.assign(s, s + n); *out = d.data(); }
    };
    for (int m = 0; m < PDEHIP_NCOEF; ++m) cp(src.coef[m], coef[m], &p.coef[m]);
    cp(src.kappa, kappa, &p.kappa);
    cp(src.f_field, f, &p.f_field);
    cp(src.initial, initial, &p.initial);
    cp(src.initial_velocity, vel, &p.initial_velocity);
    const Geo g = make_geo(src);
    for (int fc = 0; fc < 6; ++fc) {
      const size_t m = fc < 2 ? (size_t)g.ny * g.nz : (fc < 4 ? (size_t)g.nx * g.nz : (size_t)g.sxy());
      pdehip_expr* ex[4] = {&p.bc[fc].value, &p.bc[fc].alpha, &p.bc[fc].beta, &p.bc[fc].gamma};
      for (int t = 0; t < 4; ++t)
        if (ex[t]->samples) {
          samples[fc][t].assign(ex[t]->samples, ex[t]->samples + m);
          ex[t]->samples = samples[fc][t].data();
        }
    }
  }
  // Operator identity: everything except the source term.
  uint64_t op_key(const pdehip_options& o) const {
    uint64_t h = 0;
    const size_t n = (size_t)p.nx * p.ny * std::max(1, p.nz);
    h = hash_bytes(h, &p.nx, 3 * sizeof(int));
    for (double v : {p.xmin, p.xmax, p.ymin, p.ymax, p.zmin, p.zmax, p.a, p.b, p.az, p.c, p.d, p.dz, p.e})
      h = hash_d(h, v);
    h = mix(h, (uint64_t)p.op_kind * 131 + p.convection * 7 + p.bc_convention);
    for (int m = 0; m < PDEHIP_NCOEF; ++m) if (p.coef[m]) h = hash_bytes(h, p.coef[m], n * 8);
    if (p.kappa) h = hash_bytes(h, p.kappa, n * 8);
    for (int fc = 0; fc < 6; ++fc) {
      const pdehip_face_bc& b = p.bc[fc];
      h = mix(h, (uint64_t)b.kind);
      const pdehip_expr* ex[4] = {&b.value, &b.alpha, &b.beta, &b.gamma};
      for (int t = 0; t < 4; ++t) {
        h = hash_d(h, ex[t]->c); h = hash_d(h, ex[t]->cx); h = hash_d(h, ex[t]->cy); h = hash_d(h, ex[t]->cz);
        if (!samples[fc][t].empty()) h = hash_bytes(h, samples[fc][t].data(), samples[fc][t].size() * 8);
      }
    }
    h = mix(h, (uint64_t)std::max(0, std::min(p.n_nonlinear, PDEHIP_MAX_NONLINEAR)));
    for (int t = 0; t < std::min(p.n_nonlinear, PDEHIP_MAX_NONLINEAR); ++t) {
      h = mix(h, (uint64_t)p.nonlinear[t].kind * 1000003ULL + (uint64_t)(int64_t)p.nonlinear[t].power);
      h = hash_d(h, p.nonlinear[t].coeff);
    }
    h = hash_d(h, p.ut);
    h = hash_d(h, p.utt);
    // named option fields (the struct's padding bytes are not part of the key)
    for (int v : {o.method, o.max_iter, o.tol_mode, o.gmres_restart, o.mg_pre, o.mg_post,
                  o.mg_coarse_iters, o.mg_max_levels, o.mg_single_cycle, o.precond, o.precision,
                  o.bc_rows, o.check_every})
      h = mix(h, (uint64_t)(int64_t)v);
    for (double v : {o.tol, o.sor_omega, o.inner_tol}) h = hash_d(h, v);
    return h;
  }
};

struct Request {
  uint64_t ticket = 0;
  OwnedProblem prob;
  pdehip_options opt{};
  double* u_out = nullptr;
  size_t bytes = 0;          // reserved device bytes (measured size of an earlier plan, or estimate)
  std::string shape;         // key of the measured-bytes table
  bool lane = false;         // the batched engine takes it
  uint64_t op_key = 0;       // resident-operator identity
  uint64_t group = 0;        // options identity (batch lanes group by it)
  int result = PDEHIP_PENDING;  // written by the lane that ran it
  int status = PDEHIP_PENDING;  // published under the server lock when the request ends
  pdehip_report rep{};
  std::chrono::steady_clock::time_point t_submit, deadline;
  bool timed = false;        // has a deadline
  volatile int cancel = 0;   // raised by the watchdog when the deadline passes
  bool expired = false;      // ended by the deadline
};

using Clock = std::chrono::steady_clock;

uint64_t options_key(const pdehip_options& o) {
  uint64_t h = 0;
  for (int v : {o.method, o.max_iter, o.tol_mode, o.gmres_restart, o.mg_pre, o.mg_post,
                o.mg_coarse_iters, o.mg_max_levels, o.mg_single_cycle, o.precond, o.precision,
                o.bc_rows, o.check_every})
    h = mix(h, (uint64_t)(int64_t)v);
  for (double v : {o.tol, o.sor_omega, o.inner_tol}) h = hash_d(h, v);
  return h;
}

std::string shape_key(const pdehip_problem& p, const pdehip_options& o) {
  return std::to_string(p.nx) + "x" + std::to_string(p.ny) + "x" + std::to_string(std::max(1, p.nz)) +
         ":" + std::to_string(p.op_kind) + ":" + std::to_string(o.method) + ":" +
         std::to_string(o.precision) + ":" + std::to_string(o.precond) + ":" +
         std::to_string(o.gmres_restart) + ":" + std::to_string(o.mg_max_levels);
}

// The batch lanes take 2D CONST problems the lock-step engine solves: MG-PCG (any precision),
// FP64 BiCGStab or FP64 GMRES.  (Mixed BiCGStab batches add an FP64 completion pass that
// restarts unconverged members; those requests run on plan workers.)
bool lane_eligible(const pdehip_problem& p, const pdehip_options& o) {
  if (o.method == PDEHIP_BICGSTAB && o.precision == PDEHIP_MIXED) return false;
  return batch_eligible(1, &p, o, nullptr);
}

}  // namespace

size_t footprint_bytes(const pdehip_problem& p, const pdehip_options& o) {
  const double n = (double)p.nx * p.ny * std::max(1, p.nz);
  double vec = 8.0;   // FP64 vectors: x, b, x0, r, p, q, z (+ GMRES basis)
  double nvec = 9.0;
  if (o.method == PDEHIP_GMRES) nvec += std::max(1, o.gmres_restart) + 1;
  if (o.method == PDEHIP_BICGSTAB) nvec += 4;
  if (o.precision == PDEHIP_MIXED) nvec += 3.5;   // FP32 inner vectors (in FP64 units)
  double coef = p.op_kind == PDEHIP_OP_FLUX ? 1.5 : (p.op_kind == PDEHIP_OP_POINTWISE ? 7.0 : 0.0);
  const double mg = (o.method == PDEHIP_MGPCG || o.method == PDEHIP_MG) ? (p.nz > 1 ? 8.0 / 7.0 : 4.0 / 3.0) : 1.0;
  return static_cast<size_t>(n * vec * (nvec + coef) * mg + (8u << 20));
}

}  // namespace pdehip

using namespace pdehip;

struct pdehip_server {
  Ctx* parent = nullptr;
  pdehip_server_config cfg{};
  size_t budget = 0;
  int max_inflight = 256;
  std::mutex mu;
  std::condition_variable cv_work, cv_done, cv_watch;
  std::deque<std::shared_ptr<Request>> queue;
  std::unordered_map<uint64_t, std::shared_ptr<Request>> all;
  std::vector<std::shared_ptr<Request>> running;   // on plan workers (the watchdog's list)
  std::map<std::string, size_t> measured;          // plan bytes by shape key
  size_t reserved = 0;
  int inflight = 0;
  bool stop = false;
  uint64_t next_ticket = 1;
  pdehip_server_stats stats{};
  std::vector<std::thread> threads;
  std::thread watchdog;
  // per lane: own context (stream) and LRU cache of resident plans
  struct Lane {
    Ctx ctx;
    std::list<std::pair<uint64_t, std::unique_ptr<Plan>>> lru;
  };
  std::vector<std::unique_ptr<Lane>> workers, lanes;

  void run_worker(int wi);
  void run_lane(int li);
  void run_watchdog();
  // under mu: take the next request(s) for a worker (lane == false) or a batch lane
  std::vector<std::shared_ptr<Request>> take(bool lane);
  void finish(const std::vector<std::shared_ptr<Request>>& reqs);
  std::unique_ptr<Plan> cached_plan(Lane& w, Request& r);
  void keep_plan(Lane& w, uint64_t key, std::unique_ptr<Plan> P);
};

// Requests whose deadline passed while they waited end at once, unsolved.
std::vector<std::shared_ptr<Request>> pdehip_server::take(bool lane) {
  std::vector<std::shared_ptr<Request>> got;
  const bool has_lanes = !lanes.empty();
  const int cap = lane ? std::max(1, cfg.batch_max > 0 ? cfg.batch_max : 64) : 1;
  uint64_t group = 0;
  bool first = true;
  for (auto it = queue.begin(); it != queue.end() && static_cast<int>(got.size()) < cap;) {
    Request& r = **it;
    const bool mine = lane ? r.lane : (!r.lane || !has_lanes);
    if (!mine || (!first && r.group != group)) { ++it; continue; }
    if (inflight >= max_inflight || reserved + r.bytes > budget) break;   // FIFO: no overtaking
    first = false;
    group = r.group;
    reserved += r.bytes;
    ++inflight;
    got.push_back(*it);
    it = queue.erase(it);
  }
  stats.bytes_in_use = reserved;
  stats.peak_bytes = std::max(stats.peak_bytes, reserved);
  return got;
}

void pdehip_server::finish(const std::vector<std::shared_ptr<Request>>& reqs) {
  {
    std::lock_guard<std::mutex> lk(mu);
    for (const auto& r : reqs) {
      if (r->expired && r->result != PDEHIP_OK) r->result = PDEHIP_ERR_TIMEOUT;
      r->rep.status = r->result;
      r->status = r->result;
      reserved -= r->bytes;
      --inflight;
      ++stats.completed;
      if (r->status == PDEHIP_ERR_TIMEOUT) ++stats.timed_out;
      running.erase(std::remove(running.begin(), running.end(), r), running.end());
    }
    stats.bytes_in_use = reserved;
  }
  cv_done.notify_all();
  cv_work.notify_all();
}

std::unique_ptr<Plan> pdehip_server::cached_plan(Lane& w, Request& r) {
  for (auto it = w.lru.begin(); it != w.lru.end(); ++it)
    if (it->first == r.op_key) {
      std::unique_ptr<Plan> P = std::move(it->second);
      w.lru.erase(it);
      {
        std::lock_guard<std::mutex> g(mu);
        ++stats.cache_hits;
      }
      plan_set_source(*P, r.prob.p.f_field, r.prob.p.f);
      return P;
    }
  auto P = plan_create(&w.ctx, r.prob.p, r.opt);
  const size_t b = plan_bytes(*P);
  std::lock_guard<std::mutex> g(mu);
  measured[r.shape] = b;                 // later requests of this shape are admitted by it
  return P;
}

void pdehip_server::keep_plan(Lane& w, uint64_t key, std::unique_ptr<Plan> P) {
  if (!cfg.cache_operators || !P) return;
  P->opt.cancel = nullptr;
  w.lru.emplace_front(key, std::move(P));
  // keep this lane's resident cache within a share of the budget
  const size_t nl = workers.size() + lanes.size();
  const size_t cap = budget / (4 * std::max<size_t>(1, nl));
  size_t held = 0;
  for (auto& e : w.lru) held += plan_bytes(*e.second);
  while (w.lru.size() > 1 && held > cap) {
    held -= plan_bytes(*w.lru.back().second);
    w.lru.pop_back();
  }
}

void pdehip_server::run_worker(int wi) {
  Lane& w = *workers[wi];
  (void)hipSetDevice(w.ctx.device);
  for (;;) {
    std::vector<std::shared_ptr<Request>> got;
    {
      std::unique_lock<std::mutex> lk(mu);
      cv_work.wait(lk, [&] {
        got = take(false);
        return !got.empty() || stop;   // on stop, drain what is still queued first
      });
      if (got.empty()) return;
      for (auto& r : got) running.push_back(r);
    }
    cv_watch.notify_all();
    Request& r = *got[0];
    if (r.timed && Clock::now() > r.deadline) {
      r.expired = true;
      r.result = PDEHIP_ERR_TIMEOUT;
    } else {
      std::unique_ptr<Plan> P;
      try {
        P = cached_plan(w, r);
        P->opt.cancel = &r.cancel;
        plan_solve(*P, &r.rep);
        r.result = r.rep.status;
        if (r.u_out) plan_download(*P, r.u_out);
        keep_plan(w, r.op_key, std::move(P));
      } catch (const Error& e) {
        r.result = e.code;
        if (e.code == PDEHIP_ERR_CANCELLED && r.cancel) r.expired = true;
        else w.lru.clear();   // a device or allocation error: drop the resident plans
      } catch (const std::exception&) {
        r.result = PDEHIP_ERR_DEVICE;
        w.lru.clear();
      }
    }
    finish(got);
  }
}

void pdehip_server::run_lane(int li) {
  Lane& w = *lanes[li];
  (void)hipSetDevice(w.ctx.device);
  pdehip_batch_config bcfg{};
  bcfg.tile_y = 16;   // fixed work-item shapes: a member's arithmetic does not depend on the others
  for (;;) {
    std::vector<std::shared_ptr<Request>> got;
    {
      std::unique_lock<std::mutex> lk(mu);
      cv_work.wait(lk, [&] {
        got = take(true);
        return !got.empty() || stop;
      });
      if (got.empty()) return;
    }
    // members: plans (resident or new); expired or failing requests end here
    std::vector<std::shared_ptr<Request>> members;
    std::vector<std::unique_ptr<Plan>> plans;
    std::vector<pdehip_problem> probs;
    for (auto& rp : got) {
      Request& r = *rp;
      if (r.timed && Clock::now() > r.deadline) { r.expired = true; r.result = PDEHIP_ERR_TIMEOUT; continue; }
      try {
        auto P = cached_plan(w, r);
        if (!batch_plan_ok(*P, r.opt)) {
          // solved alone on this lane (a property of the problem, so the same alone or mixed)
          plan_solve(*P, &r.rep);
          r.result = r.rep.status;
          if (r.u_out) plan_download(*P, r.u_out);
          keep_plan(w, r.op_key, std::move(P));
          continue;
        }
        members.push_back(rp);
        plans.push_back(std::move(P));
        probs.push_back(r.prob.p);
      } catch (const Error& e) {
        r.result = e.code;
      } catch (const std::exception&) {
        r.result = PDEHIP_ERR_DEVICE;
      }
    }
    if (!members.empty()) {
      const int k = static_cast<int>(members.size());
      try {
        bcfg.precision = members[0]->opt.precision;
        BatchPtr B = batch_create(&w.ctx, k, probs.data(), members[0]->opt, &bcfg, &plans);
        std::vector<char> cancelled(k, 0);
        batch_set_chunk_hook(*B, [&] {
          const auto now = Clock::now();
          for (int i = 0; i < k; ++i)
            if (!cancelled[i] && members[i]->timed && now > members[i]->deadline) {
              batch_cancel(*B, i);
              cancelled[i] = 1;
            }
        });
        std::vector<pdehip_report> reps(k);
        batch_solve(*B, reps.data());
        for (int i = 0; i < k; ++i) {
          Request& r = *members[i];
          r.rep = reps[i];
          r.result = reps[i].status;
          if (cancelled[i] && r.result != PDEHIP_OK) r.expired = true;
          if (r.u_out) batch_download(*B, i, r.u_out);
        }
        plans = batch_release_plans(*B);
        batch_destroy(B);
        for (int i = 0; i < k; ++i)
          if (i < static_cast<int>(plans.size())) keep_plan(w, members[i]->op_key, std::move(plans[i]));
        std::lock_guard<std::mutex> g(mu);
        stats.batched += k;
        stats.batches += 1;
      } catch (const Error& e) {
        for (auto& r : members) r->result = e.code;
        w.lru.clear();
      } catch (const std::exception&) {
        for (auto& r : members) r->result = PDEHIP_ERR_DEVICE;
        w.lru.clear();
      }
    }
    finish(got);
  }
}

// Raises the cancel flag of plan-worker solves whose deadline has passed.
void pdehip_server::run_watchdog() {
  std::unique_lock<std::mutex> lk(mu);
  while (!stop) {
    const auto now = Clock::now();
    for (auto& r : running)
      if (r->timed && !r->cancel && now > r->deadline) r->cancel = 1;
    cv_watch.wait_for(lk, std::chrono::microseconds(500));
  }
}

namespace pdehip {

pdehip_server* server_create(Ctx* ctx, const pdehip_server_config* cfg) {
  auto* s = new pdehip_server();
  s->parent = ctx;
  if (cfg) s->cfg = *cfg;
  else { s->cfg.streams = 4; s->cfg.cache_operators = 1; }
  s->budget = s->cfg.hbm_budget_bytes ? s->cfg.hbm_budget_bytes : ctx->budget;
  s->max_inflight = s->cfg.max_inflight > 0 ? s->cfg.max_inflight : 256;
  const int nw = s->cfg.streams > 0 ? s->cfg.streams : 4;
  const int nl = s->cfg.batch_lanes < 0 ? 0 : (s->cfg.batch_lanes > 0 ? s->cfg.batch_lanes : 2);
  auto make = [&](int count, std::vector<std::unique_ptr<pdehip_server::Lane>>& out) {
    for (int i = 0; i < count; ++i) {
      auto w = std::make_unique<pdehip_server::Lane>();
      w->ctx.device = ctx->device;
      w->ctx.cus = ctx->cus;
      w->ctx.budget = 0;   // admission is the gate
      // released plan buffers are reused by the lane's next plans: a hipFree would synchronise
      // the device, stalling every other lane's stream
      w->ctx.pool_cap = s->budget / (8 * static_cast<size_t>(nw + nl));
      (void)hipSetDevice(ctx->device);
      HIPCHK(hipStreamCreateWithFlags(&w->ctx.stream, hipStreamNonBlocking));
      out.push_back(std::move(w));
    }
  };
  make(nw, s->workers);
  make(nl, s->lanes);
  for (int i = 0; i < nw; ++i) s->threads.emplace_back([s, i] { s->run_worker(i); });
  for (int i = 0; i < nl; ++i) s->threads.emplace_back([s, i] { s->run_lane(i); });
  if (s->cfg.timeout_ms > 0) s->watchdog = std::thread([s] { s->run_watchdog(); });
  return s;
}

int server_submit(pdehip_server* s, const pdehip_problem& p, const pdehip_options& o, double* u_out,
                  uint64_t* ticket) {
  auto r = std::make_shared<Request>();
  r->prob.take(p);
  r->opt = o;
  r->opt.cancel = nullptr;
  r->u_out = u_out;
  r->shape = shape_key(p, o);
  r->lane = !s->lanes.empty() && lane_eligible(p, o);
  r->op_key = r->prob.op_key(r->opt);
  r->group = options_key(r->opt);
  r->t_submit = Clock::now();
  r->timed = s->cfg.timeout_ms > 0;
  if (r->timed)
    r->deadline = r->t_submit + std::chrono::microseconds(static_cast<long long>(s->cfg.timeout_ms * 1e3));
  std::lock_guard<std::mutex> lk(s->mu);
  auto m = s->measured.find(r->shape);
  r->bytes = m != s->measured.end() ? m->second : footprint_bytes(p, o);
  ++s->stats.submitted;
  if (r->bytes > s->budget) {
    ++s->stats.rejected;
    throw Error(PDEHIP_ERR_ADMISSION, "request footprint exceeds the HBM budget");
  }
  r->ticket = s->next_ticket++;
  *ticket = r->ticket;
  s->all[r->ticket] = r;
  s->queue.push_back(r);
  s->cv_work.notify_all();
  return PDEHIP_OK;
}

int server_poll(pdehip_server* s, uint64_t ticket, pdehip_report* rep, bool wait) {
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->all.find(ticket);
  if (it == s->all.end()) throw Error(PDEHIP_ERR_INVALID, "unknown ticket");
  auto r = it->second;
  if (wait) s->cv_done.wait(lk, [&] { return r->status != PDEHIP_PENDING; });
  if (r->status == PDEHIP_PENDING) return PDEHIP_PENDING;
  if (rep) *rep = r->rep;
  s->all.erase(it);
  return r->status;
}

void server_stats(pdehip_server* s, pdehip_server_stats* st) {
  std::lock_guard<std::mutex> lk(s->mu);
  *st = s->stats;
}

void server_destroy(pdehip_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  s->cv_watch.notify_all();
  for (auto& t : s->threads) t.join();
  if (s->watchdog.joinable()) s->watchdog.join();
  for (auto* group : {&s->workers, &s->lanes})
    for (auto& w : *group) {
      w->lru.clear();
      (void)hipStreamSynchronize(w->ctx.stream);
      (void)hipStreamDestroy(w->ctx.stream);
    }
  delete s;
}

// Batched solve, first form: every problem through a resident plan on the context stream.
int solve_batch(Ctx* ctx, int count, const pdehip_problem* probs, const pdehip_options& opt,
                const pdehip_batch_config* cfg, double* const* u_out, pdehip_report* reports) {
  int worst = PDEHIP_OK;
  if (batch_eligible(count, probs, opt, nullptr)) {
    // one lock-step batched solve over resident per-problem plans
    auto B = batch_create(ctx, count, probs, opt, cfg);
    std::vector<pdehip_report> reps(count);
    batch_solve(*B, reps.data());
    for (int i = 0; i < count; ++i) {
      if (u_out[i]) batch_download(*B, i, u_out[i]);
      if (reports) reports[i] = reps[i];
      if (reps[i].status != PDEHIP_OK) worst = reps[i].status;
    }
    batch_destroy(B);
    return worst;
  }
  // other problem classes: one resident plan after the other
  for (int i = 0; i < count; ++i) {
    pdehip_report rep{};
    auto P = plan_create(ctx, probs[i], opt);
    plan_solve(*P, &rep);
    if (u_out[i]) plan_download(*P, u_out[i]);
    if (reports) reports[i] = rep;
    if (rep.status != PDEHIP_OK) worst = rep.status;
  }
  return worst;
}

}  // namespace pdehip
