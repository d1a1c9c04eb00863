// serving.cpp — the request loop and the batched solve entry.
//
// Server: worker threads, each owning one HIP stream (default priority), pull admitted
// requests from a FIFO.  A request is admitted only while the bytes reserved by running
// requests plus its own footprint fit the HBM budget, so requests that arrive while others
// run start as soon as slots and memory free up; their kernels overlap on the device through
// the workers' streams.  Operators (level hierarchy, coefficient fields, boundary data) stay
// resident per worker in an LRU cache keyed by a hash of the problem minus its source term;
// a cache hit only uploads the new source and rebuilds the right-hand side.  Every solve
// runs the same kernels with the same launch shapes whatever else is in flight, so its
// result is bitwise the same alone and in the mix.
#include "serving.hpp"
#include "batch.hpp"

#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
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
  size_t bytes = 0;
  int status = PDEHIP_PENDING;
  pdehip_report rep{};
  std::chrono::steady_clock::time_point t_submit;
};

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
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<Request>> queue;
  std::unordered_map<uint64_t, std::shared_ptr<Request>> all;
  size_t reserved = 0;
  int inflight = 0;
  bool stop = false;
  uint64_t next_ticket = 1;
  pdehip_server_stats stats{};
  std::vector<std::thread> workers;
  // per worker: own context (stream) and LRU cache of resident plans
  struct Worker {
    Ctx ctx;
    std::list<std::pair<uint64_t, std::unique_ptr<Plan>>> lru;
    size_t cached_bytes = 0;
  };
  std::vector<std::unique_ptr<Worker>> ws;

  void run(int wi);
};

void pdehip_server::run(int wi) {
  Worker& w = *ws[wi];
  (void)hipSetDevice(w.ctx.device);
  for (;;) {
    std::shared_ptr<Request> req;
    {
      std::unique_lock<std::mutex> lk(mu);
      cv_work.wait(lk, [&] {
        if (stop) return true;
        if (queue.empty()) return false;
        const int max_inflight = cfg.max_inflight > 0 ? cfg.max_inflight : 8;
        return inflight < max_inflight && reserved + queue.front()->bytes <= budget;
      });
      if (stop && queue.empty()) return;
      if (queue.empty()) continue;
      req = queue.front();
      queue.pop_front();
      reserved += req->bytes;
      ++inflight;
      stats.bytes_in_use = reserved;
      stats.peak_bytes = std::max(stats.peak_bytes, reserved);
    }
    int st = PDEHIP_OK;
    try {
      const uint64_t key = req->prob.op_key(req->opt);
      Plan* plan = nullptr;
      for (auto it = w.lru.begin(); it != w.lru.end(); ++it)
        if (it->first == key) {
          w.lru.splice(w.lru.begin(), w.lru, it);
          plan = w.lru.front().second.get();
          std::lock_guard<std::mutex> g(mu);
          ++stats.cache_hits;
          break;
        }
      if (plan) {
        plan_set_source(*plan, req->prob.p.f_field, req->prob.p.f);
      } else {
        auto P = plan_create(&w.ctx, req->prob.p, req->opt);
        plan = P.get();
        if (cfg.cache_operators) {
          w.lru.emplace_front(key, std::move(P));
        } else {
          w.lru.emplace_front(0, std::move(P));
        }
      }
      plan_solve(*plan, &req->rep);
      st = req->rep.status;
      if (req->u_out) plan_download(*plan, req->u_out);
      if (!cfg.cache_operators) w.lru.clear();
      // keep this worker's resident cache within a quarter of the budget
      auto held = [&] {
        size_t t = 0;
        for (auto& e : w.lru) t += plan_bytes(*e.second);
        return t;
      };
      w.cached_bytes = held();
      while (cfg.cache_operators && w.lru.size() > 1 && w.cached_bytes > budget / 4) {
        w.lru.pop_back();
        w.cached_bytes = held();
      }
    } catch (const Error& e) {
      st = e.code;
      w.lru.clear();
      w.cached_bytes = 0;
    } catch (const std::exception&) {
      st = PDEHIP_ERR_DEVICE;
      w.lru.clear();
      w.cached_bytes = 0;
    }
    {
      std::lock_guard<std::mutex> lk(mu);
      if (cfg.timeout_ms > 0) {
        const double el = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - req->t_submit).count();
        if (el > cfg.timeout_ms && st == PDEHIP_OK) st = PDEHIP_ERR_TIMEOUT;
      }
      req->status = st;
      req->rep.status = st;
      reserved -= req->bytes;
      --inflight;
      ++stats.completed;
      stats.bytes_in_use = reserved;
    }
    cv_done.notify_all();
    cv_work.notify_all();
  }
}

namespace pdehip {

pdehip_server* server_create(Ctx* ctx, const pdehip_server_config* cfg) {
  auto* s = new pdehip_server();
  s->parent = ctx;
  if (cfg) s->cfg = *cfg;
  else { s->cfg.max_inflight = 8; s->cfg.streams = 4; s->cfg.cache_operators = 1; }
  s->budget = s->cfg.hbm_budget_bytes ? s->cfg.hbm_budget_bytes : ctx->budget;
  const int nw = s->cfg.streams > 0 ? s->cfg.streams : 4;
  for (int i = 0; i < nw; ++i) {
    auto w = std::make_unique<pdehip_server::Worker>();
    w->ctx.device = ctx->device;
    w->ctx.cus = ctx->cus;
    w->ctx.budget = 0;   // admission is the gate
    // a worker's released plan buffers are reused by its next plans: a hipFree would
    // synchronise the device, stalling every other worker's stream
    w->ctx.pool_cap = s->budget / (8 * static_cast<size_t>(nw));
    (void)hipSetDevice(ctx->device);
    HIPCHK(hipStreamCreateWithFlags(&w->ctx.stream, hipStreamNonBlocking));
    s->ws.push_back(std::move(w));
  }
  for (int i = 0; i < nw; ++i) s->workers.emplace_back([s, i] { s->run(i); });
  return s;
}

int server_submit(pdehip_server* s, const pdehip_problem& p, const pdehip_options& o, double* u_out,
                  uint64_t* ticket) {
  auto r = std::make_shared<Request>();
  r->prob.take(p);
  r->opt = o;
  r->opt.cancel = nullptr;
  r->u_out = u_out;
  r->bytes = footprint_bytes(p, o);
  r->t_submit = std::chrono::steady_clock::now();
  std::lock_guard<std::mutex> lk(s->mu);
  ++s->stats.submitted;
  if (r->bytes > s->budget) {
    ++s->stats.rejected;
    throw Error(PDEHIP_ERR_ADMISSION, "request footprint exceeds the HBM budget");
  }
  r->ticket = s->next_ticket++;
  *ticket = r->ticket;
  s->all[r->ticket] = r;
  s->queue.push_back(r);
  s->cv_work.notify_one();
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
  for (auto& t : s->workers) t.join();
  for (auto& w : s->ws) {
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
