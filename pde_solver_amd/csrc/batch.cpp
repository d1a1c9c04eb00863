// batch.cpp — host side of the batched MG-PCG engine (see batch.hpp).
//
// Set-up builds one resident Plan per problem (the same operators, folded boundary rows,
// hierarchy and dense coarsest inverse a single solve would use), then the device tables and
// the work-item lists:
//   * per level l: smoother strips (problem, 124-column strip, row range), residual/prolongation
//     row blocks, restriction coarse-row blocks (only problems whose hierarchy continues below
//     l) and dense-solve rows (problems whose coarsest level is l);
//   * level 0: element blocks for the vector kernels and row blocks for the operator apply,
//     each with per-problem [first, count) ranges for the segmented reductions.
// The solve runs the PCG (FP64) or the FP64 refinement around FP32/FP16 MG-PCG (mixed) of
// every problem in lock step; converged problems drop out item by item on the device.  The
// host reads one counter of unfinished problems every few iterations.
#include "batch.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <map>
#include <string>

namespace pdehip {

namespace {

struct List {
  DBuf items;
  int n = 0;
  DBuf first, cnt;   // per problem (reduction lists only)
  const BItem* d() const { return items.as<BItem>(); }
};

void upload_list(Ctx* ctx, hipStream_t s, List& L, const std::vector<BItem>& v,
                 const std::vector<int>* first = nullptr, const std::vector<int>* cnt = nullptr) {
  L.n = static_cast<int>(v.size());
  if (!v.empty()) {
    L.items.alloc(ctx, v.size() * sizeof(BItem));
    HIPCHK(hipMemcpyAsync(L.items.get(), v.data(), v.size() * sizeof(BItem), hipMemcpyHostToDevice, s));
  }
  if (first) {
    L.first.alloc(ctx, first->size() * sizeof(int));
    L.cnt.alloc(ctx, cnt->size() * sizeof(int));
    HIPCHK(hipMemcpyAsync(L.first.get(), first->data(), first->size() * sizeof(int), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(L.cnt.get(), cnt->data(), cnt->size() * sizeof(int), hipMemcpyHostToDevice, s));
  }
  HIPCHK(hipStreamSynchronize(s));   // the host vectors may be temporaries
}

}  // namespace

struct BatchPlan {
  Ctx* ctx = nullptr;
  int n = 0;
  pdehip_options opt{};
  pdehip_batch_config cfg{};
  bool mixed = false;
  bool krylov = false;                     // BiCGStab or GMRES batch (no hierarchy)
  bool gmres = false;                      // GMRES(mm) batch
  int mm = 0;                              // GMRES restart length
  DBuf gvecs;                              // GMRES: BGVec [n]
  int L = 0;                               // deepest hierarchy
  std::vector<std::unique_ptr<Plan>> plans;
  std::vector<int> levels;                 // per problem
  DBuf tab;                                // BEnt [L][n]
  DBuf vecs;                               // BVec [n]
  DBuf dense;                              // BDense [n]
  std::vector<List> sweep, rows, crows, drows;   // per level
  List el0, rows0;                         // level-0 element / strip items (with ranges)
  List rb0;                                // BiCGStab: level-0 row-block items (with ranges)
  List gel;                                // GMRES: level-0 element items (with ranges)
  DBuf kvecs;                              // BiCGStab: BKVec [n] (the batch's precision)
  DBuf kvecs64;                            // mixed BiCGStab: FP64 views for the completion pass
  std::vector<LevelShape> shape;           // per level work-item shapes
  DBuf part, maxp;
  DBuf active;
  DBuf faces, psum, pmax, rep;             // finish: per-problem boundary data, partials, results
  BReport* hrep = nullptr;                 // pinned copy of rep
  int* hactive = nullptr;                  // two poll slots
  hipEvent_t evpoll[2] = {nullptr, nullptr};
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipGraphExec_t graph[2] = {nullptr, nullptr};   // captured iterations (FP64 / inner) or GMRES cycle
  std::function<void()> on_chunk;          // host hook after each polling chunk is enqueued
  size_t bytes = 0;                        // device memory held by the batch and its plans
  ~BatchPlan() {
    for (hipGraphExec_t g : graph)
      if (g) (void)hipGraphExecDestroy(g);
    if (hactive) (void)hipHostFree(hactive);
    for (hipEvent_t e : evpoll)
      if (e) (void)hipEventDestroy(e);
    if (hrep) (void)hipHostFree(hrep);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
  }
  const BEnt* T(int l) const { return tab.as<BEnt>() + (size_t)l * n; }
  const BVec* V() const { return vecs.as<BVec>(); }
  const BKVec* K() const { return kvecs.as<BKVec>(); }
  const BKVec* K64() const { return kvecs64.as<BKVec>(); }
  const BGVec* G() const { return gvecs.as<BGVec>(); }
  Prec coarse = P16;                       // mixed: storage of levels >= 1 (from the plans)
  Prec prec(int l) const { return mixed ? (l == 0 ? P32 : coarse) : P64; }
};

namespace {

// Work-item shapes.  Smoother items are 124-column strips of `ychunk` rows (a wavefront marches
// the rows, so short chunks trade ring re-loads for more items in flight); row items cover
// whole rows, about `row_nodes` nodes each (at least one row); restriction items cover about
// `crow_nodes` coarse nodes.
LevelShape default_shape(const BatchPlan& B, int l) {
  int maxny = 0;
  for (int p = 0; p < B.n; ++p)
    if (l < B.levels[p]) maxny = std::max(maxny, B.plans[p]->L[l].g.ny);
  LevelShape sh;
  sh.ychunk = maxny <= 65 ? 8 : (maxny <= 257 ? 16 : 32);
  sh.row_nodes = 256;
  sh.crow_nodes = 128;
  return sh;
}

std::vector<BItem> sweep_items(const BatchPlan& B, int l, int ychunk) {
  std::vector<BItem> v;
  const int strip = 124;   // output columns of one wavefront strip (k_batch.hip)
  for (int p = 0; p < B.n; ++p) {
    if (l >= B.levels[p] - 1) continue;            // coarsest level: dense solve
    const Geo& g = B.plans[p]->L[l].g;
    for (int sx = 0; sx * strip < g.nx; ++sx)
      for (int j0 = 0; j0 < g.ny; j0 += ychunk) v.push_back(BItem{p, sx, j0, std::min(g.ny, j0 + ychunk)});
  }
  return v;
}

// Level-0 strips of every problem (the operator applies' items), whatever its hierarchy.
std::vector<BItem> strip_items0(const BatchPlan& B, int ychunk) {
  std::vector<BItem> v;
  const int strip = 124;
  for (int p = 0; p < B.n; ++p) {
    const Geo& g = B.plans[p]->g;
    for (int sx = 0; sx * strip < g.nx; ++sx)
      for (int j0 = 0; j0 < g.ny; j0 += ychunk) v.push_back(BItem{p, sx, j0, std::min(g.ny, j0 + ychunk)});
  }
  return v;
}

// Row blocks of level l.  coarse_side: the level is the coarse side of a restriction (only
// problems whose hierarchy reaches it from above); otherwise problems that smooth on level l.
// first/cnt: per-problem item ranges for the segmented reductions.
std::vector<BItem> row_items(const BatchPlan& B, int l, int nodes, bool coarse_side,
                             std::vector<int>* first = nullptr, std::vector<int>* cnt = nullptr) {
  std::vector<BItem> v;
  if (first) { first->assign(B.n, 0); cnt->assign(B.n, 0); }
  for (int p = 0; p < B.n; ++p) {
    if (first) (*first)[p] = static_cast<int>(v.size());
    const bool has = coarse_side ? (l >= 1 && l < B.levels[p]) : (l < B.levels[p] - 1 || (first && l == 0));
    if (has) {
      const Geo& g = B.plans[p]->L[l].g;
      const int rc = std::max(1, nodes / g.nx);
      for (int j0 = 0; j0 < g.ny; j0 += rc) v.push_back(BItem{p, j0, std::min(g.ny, j0 + rc), 0});
    }
    if (first) (*cnt)[p] = static_cast<int>(v.size()) - (*first)[p];
  }
  return v;
}

std::vector<BItem> dense_items(const BatchPlan& B, int l, const std::vector<bk::BDense>& dense) {
  std::vector<BItem> v;
  for (int p = 0; p < B.n; ++p)
    if (l == B.levels[p] - 1)
      for (int r = 0; r < dense[p].m; ++r) v.push_back(BItem{p, r, 0, 0});
  return v;
}

// Mean time of `reps` launches of f (after one warm-up) on the batch stream.
template <typename F>
float time_ms(BatchPlan& B, F&& f, int reps) {
  const hipStream_t s = B.ctx->stream;
  f();
  HIPCHK(hipEventRecord(B.ev0, s));
  for (int r = 0; r < reps; ++r) f();
  HIPCHK(hipEventRecord(B.ev1, s));
  HIPCHK(hipEventSynchronize(B.ev1));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, B.ev0, B.ev1));
  return ms / reps;
}

std::string shape_key(const BatchPlan& B) {
  std::map<std::pair<int, int>, int> hist;
  for (const auto& P : B.plans) hist[{P->g.nx, P->g.ny}] += 1;
  std::string key = B.mixed ? "mixed" : "fp64";
  for (const auto& kv : hist)
    key += ":" + std::to_string(kv.first.first) + "x" + std::to_string(kv.first.second) + "*" +
           std::to_string(kv.second);
  return key;
}

void upload_shapes(BatchPlan& B, const std::vector<bk::BDense>* dense) {
  Ctx* ctx = B.ctx;
  const hipStream_t s = ctx->stream;
  for (int l = 0; l < B.L; ++l) {
    upload_list(ctx, s, B.sweep[l], sweep_items(B, l, B.shape[l].ychunk));
    upload_list(ctx, s, B.rows[l], row_items(B, l, B.shape[l].row_nodes, false));
    if (l + 1 < B.L)
      upload_list(ctx, s, B.crows[l + 1], row_items(B, l + 1, B.shape[l + 1].crow_nodes, true));
    if (dense) upload_list(ctx, s, B.drows[l], dense_items(B, l, *dense));
  }
}

// Autotuner: for every level, time the smoother sweep pair over each candidate strip height,
// the residual + prolongation over each row-block size and the restriction over each coarse
// block size, on this batch's own operators and buffers with every problem active, and keep
// the fastest.  Results are cached in the context under the batch's size histogram.  The
// level buffers hold zeros here, and every timed kernel maps zeros to zeros.
void autotune(BatchPlan& B) {
  Ctx* ctx = B.ctx;
  const hipStream_t s = ctx->stream;
  const std::string key = shape_key(B);
  {
    std::lock_guard<std::mutex> g(ctx->tune_mu);
    auto it = ctx->batch_shapes.find(key);
    if (it != ctx->batch_shapes.end() && static_cast<int>(it->second.size()) == B.L * 3) {
      for (int l = 0; l < B.L; ++l)
        B.shape[l] = LevelShape{it->second[3 * l], it->second[3 * l + 1], it->second[3 * l + 2]};
      upload_shapes(B, nullptr);
      return;
    }
  }
  bk::init_scalars(s, B.V(), B.n, 1e-300);
  const int reps = 3;
  for (int l = 0; l < B.L; ++l) {
    const Prec pl = B.prec(l);
    if (B.sweep[l].n > 0) {
      float best = 1e30f;
      int pick = B.shape[l].ychunk;
      for (int yc : {8, 16, 32, 64}) {
        List tmp;
        upload_list(ctx, s, tmp, sweep_items(B, l, yc));
        const float t = time_ms(B, [&] {
          bk::sweep(s, pl, B.T(l), tmp.d(), tmp.n, 0, 0, false);
          bk::sweep(s, pl, B.T(l), tmp.d(), tmp.n, 0, 1, false);
          bk::residual(s, pl, B.T(l), tmp.d(), tmp.n);
        }, reps);
        if (t < best) { best = t; pick = yc; }
      }
      B.shape[l].ychunk = pick;
    }
    if (l + 1 < B.L && B.rows[l].n > 0) {
      const Prec pc = B.prec(l + 1);
      const bool s0 = B.mixed && l == 0;
      float best = 1e30f;
      int pick = B.shape[l].row_nodes;
      for (int rn : {64, 256, 1024, 4096}) {
        List tmp;
        upload_list(ctx, s, tmp, row_items(B, l, rn, false));
        const float t = time_ms(B, [&] {
          bk::prolong(s, pc, pl, B.T(l + 1), B.T(l), tmp.d(), tmp.n, s0);
        }, reps);
        if (t < best) { best = t; pick = rn; }
      }
      B.shape[l].row_nodes = pick;
      best = 1e30f;
      pick = B.shape[l + 1].crow_nodes;
      for (int cn : {32, 128, 512, 2048}) {
        List tmp;
        upload_list(ctx, s, tmp, row_items(B, l + 1, cn, true));
        const float t = time_ms(B, [&] {
          bk::restrict_(s, pl, pc, B.T(l), B.T(l + 1), tmp.d(), tmp.n, s0);
        }, reps);
        if (t < best) { best = t; pick = cn; }
      }
      B.shape[l + 1].crow_nodes = pick;
    }
  }
  upload_shapes(B, nullptr);
  std::vector<int> flat;
  for (const LevelShape& sh : B.shape) { flat.push_back(sh.ychunk); flat.push_back(sh.row_nodes); flat.push_back(sh.crow_nodes); }
  std::lock_guard<std::mutex> g(ctx->tune_mu);
  ctx->batch_shapes[key] = flat;
}

}  // namespace

bool batch_eligible(int count, const pdehip_problem* probs, const pdehip_options& opt,
                    std::string* why) {
  auto no = [&](const char* r) { if (why) *why = r; return false; };
  if (count < 1) return no("empty batch");
  const bool gm = opt.method == PDEHIP_GMRES;
  const bool bicg = opt.method == PDEHIP_BICGSTAB || gm;   // the Krylov engines
  if (opt.method != PDEHIP_MGPCG && !bicg) return no("batched engine runs MG-PCG, BiCGStab or GMRES");
  if (bicg && opt.precond != PDEHIP_PC_NONE && opt.precond != PDEHIP_PC_JACOBI)
    return no("batched BiCGStab / GMRES take no preconditioner or Jacobi");
  if (gm && opt.precision == PDEHIP_MIXED) return no("batched GMRES runs in FP64");
  if (!bicg && (opt.mg_pre < 2 || opt.mg_pre % 2 || opt.mg_post < 2 || opt.mg_post % 2))
    return no("batched engine needs even pre/post smoothing counts");
  for (int i = 0; i < count; ++i) {
    const pdehip_problem& p = probs[i];
    if (p.nz > 1) return no("batched engine is 2D");
    if (p.op_kind != PDEHIP_OP_CONST) return no("batched engine needs constant coefficients");
    if (!bicg && (p.c != 0.0 || p.d != 0.0)) return no("batched MG-PCG needs no convection");
    if (p.n_nonlinear > 0) return no("batched engine needs linear problems");
  }
  return true;
}

bool batch_plan_ok(const Plan& P, const pdehip_options& opt) {
  if (opt.method == PDEHIP_BICGSTAB || opt.method == PDEHIP_GMRES) return true;
  const int nl = static_cast<int>(P.L.size());
  return nl >= 2 && P.L[nl - 1].m > 0;
}

namespace {

// Finish buffers shared by both engines: per-problem boundary data, residual-norm partials,
// the device report array and its pinned copy; polling slots and events.
void setup_finish(BatchPlan& B) {
  Ctx* ctx = B.ctx;
  const hipStream_t s = ctx->stream;
  const int n = B.n;
  B.active.alloc(ctx, 2 * sizeof(int));
  HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&B.hactive), 2 * sizeof(int)));
  for (hipEvent_t& e : B.evpoll) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  std::vector<BFace> fc(n);
  for (int p = 0; p < n; ++p) {
    const Plan& P = *B.plans[p];
    fc[p] = BFace{P.faces.fd, P.f ? P.f.as<double>() : nullptr, P.f_const};
  }
  B.faces.alloc(ctx, n * sizeof(BFace));
  HIPCHK(hipMemcpyAsync(B.faces.get(), fc.data(), n * sizeof(BFace), hipMemcpyHostToDevice, s));
  B.psum.alloc(ctx, (size_t)std::max(1, B.el0.n) * sizeof(double));
  B.pmax.alloc(ctx, (size_t)std::max(1, B.el0.n) * sizeof(double));
  B.rep.alloc(ctx, n * sizeof(BReport));
  HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&B.hrep), n * sizeof(BReport)));
}

// Element items (4096 nodes) and level-0 strip items of every problem, with per-problem ranges.
void level0_lists(BatchPlan& B, int ychunk) {
  Ctx* ctx = B.ctx;
  const hipStream_t s = ctx->stream;
  const int n = B.n;
  std::vector<BItem> el;
  std::vector<int> ef(n), ec(n);
  for (int p = 0; p < n; ++p) {
    const Geo& g = B.plans[p]->g;
    ef[p] = static_cast<int>(el.size());
    for (long long e0 = 0; e0 < g.n; e0 += 4096)
      el.push_back(BItem{p, static_cast<int>(e0), static_cast<int>(std::min(g.n, e0 + 4096)), 0});
    ec[p] = static_cast<int>(el.size()) - ef[p];
  }
  upload_list(ctx, s, B.el0, el, &ef, &ec);
  const std::vector<BItem> rw = strip_items0(B, ychunk);
  std::vector<int> rf(n, 0), rcn(n, 0);
  for (int t = static_cast<int>(rw.size()) - 1; t >= 0; --t) { rf[rw[t].p] = t; rcn[rw[t].p] += 1; }
  upload_list(ctx, s, B.rows0, rw, &rf, &rcn);
}

// Problem i's plan: the caller's resident one when given (moved into the batch), else new.
std::unique_ptr<Plan> take_or_create(Ctx* ctx, const pdehip_problem* probs, const pdehip_options& opt,
                                     std::vector<std::unique_ptr<Plan>>* given, int i) {
  if (given && i < static_cast<int>(given->size()) && (*given)[i]) {
    if ((*given)[i]->ctx != ctx) throw Error(PDEHIP_ERR_INVALID, "a batch's plans must share its context");
    return std::move((*given)[i]);
  }
  return plan_create(ctx, probs[i], opt);
}

// Per-problem BiCGStab vectors: the plan's FP64 Krylov vectors, or (fp32) its FP32 ones; the
// solution, right-hand side and initial guess are FP64 either way.  A mixed-precision plan
// allocates its FP64 Krylov vectors only when a completion pass needs them
// (ensure_fp64_views).
BKVec krylov_view(Plan& P, bool jac, bool fp32) {
  BKVec k;
  std::memset(&k, 0, sizeof(k));
  if (fp32) {
    k.r = P.v32[0].get(); k.rh = P.v32[1].get(); k.p = P.v32[2].get(); k.v = P.v32[3].get();
    k.s = P.v32[4].get(); k.t = P.v32[5].get();
    k.ph = jac ? P.v32[6].get() : k.p;
    k.sh = jac ? P.v32[7].get() : k.s;
    k.op = P.L[0].op32;
  } else {
    k.r = P.r.get(); k.rh = P.z.get(); k.p = P.p.get(); k.v = P.q.get();
    k.s = P.w1.get(); k.t = P.w2.get();
    k.ph = jac ? P.w4.get() : k.p;
    k.sh = jac ? P.w3.get() : k.s;
    k.op = P.L[0].op64;
  }
  k.x = P.x.as<double>();
  k.b = P.bvec.as<double>();
  k.x0 = P.x0.as<double>();
  k.sc = P.sc(0);
  k.op64 = P.L[0].op64;
  k.nx = P.g.nx;
  k.ny = P.g.ny;
  k.jacobi = jac ? 1 : 0;
  const void* need[] = {k.r, k.rh, k.p, k.v, k.s, k.t, k.ph, k.sh, k.x, k.b, k.x0, k.sc};
  for (const void* q : need)
    if (!q) throw Error(PDEHIP_ERR_DEVICE, "batched BiCGStab: a Krylov vector of a plan is missing");
  return k;
}

// Per-problem GMRES data: the plan's FP64 basis, Hessenberg system and work vectors.
BGVec gmres_view(Plan& P, bool jac) {
  BGVec g;
  std::memset(&g, 0, sizeof(g));
  g.V = P.V.as<double>();
  g.H = P.H.as<double>();
  g.cs = P.cs.as<double>();
  g.sn = P.sn.as<double>();
  g.g = P.gv.as<double>();
  g.y = P.yv.as<double>();
  g.hc = P.hcb.as<double>();
  g.x = P.x.as<double>();
  g.r = P.r.as<double>();
  g.b = P.bvec.as<double>();
  g.x0 = P.x0.as<double>();
  g.w = P.w1.as<double>();
  g.sc = P.sc(0);
  g.op = P.L[0].op64;
  g.n = P.g.n;
  g.nx = P.g.nx;
  g.ny = P.g.ny;
  g.jacobi = jac ? 1 : 0;
  const void* need[] = {g.V, g.H, g.cs, g.sn, g.g, g.y, g.hc, g.x, g.r, g.b, g.x0, g.w, g.sc};
  for (const void* q : need)
    if (!q) throw Error(PDEHIP_ERR_DEVICE, "batched GMRES: a vector of a plan is missing");
  return g;
}

// BiCGStab batch: one resident plan per problem (operator, folded boundary rows, right-hand
// side, Krylov vectors), the BKVec table, row-block items for the vector kernels, level-0
// strips for the applies, and the finish buffers of the MG-PCG engine.
BatchPtr create_krylov(BatchPtr B, Ctx* ctx, int count, const pdehip_problem* probs,
                       const pdehip_options& opt, size_t used0,
                       std::vector<std::unique_ptr<Plan>>* given) {
  const hipStream_t s = ctx->stream;
  B->krylov = true;
  B->gmres = opt.method == PDEHIP_GMRES;
  B->mm = std::max(1, std::min(opt.gmres_restart, std::max(1, opt.max_iter)));
  B->L = 1;
  const bool jac = opt.precond == PDEHIP_PC_JACOBI;
  std::vector<BVec> vecs(count);
  std::vector<BKVec> kv(B->gmres ? 0 : count);
  std::vector<BGVec> gv(B->gmres ? count : 0);
  for (int i = 0; i < count; ++i) {
    B->plans.push_back(take_or_create(ctx, probs, opt, given, i));
    Plan& P = *B->plans.back();
    B->levels.push_back(1);
    BVec& v = vecs[i];
    std::memset(&v, 0, sizeof(v));
    v.x64 = P.x.as<double>();
    v.r64 = P.r.as<double>();
    v.b64 = P.bvec.as<double>();
    v.x0 = P.x0.as<double>();
    v.so = P.sc(0);
    v.si = P.sc(1);
    v.n = P.g.n;
    v.nx = P.g.nx;
    v.ny = P.g.ny;
    v.op64 = P.L[0].op64;
    if (B->gmres) gv[i] = gmres_view(P, jac);
    else kv[i] = krylov_view(P, jac, B->mixed);
  }
  B->vecs.alloc(ctx, count * sizeof(BVec));
  HIPCHK(hipMemcpyAsync(B->vecs.get(), vecs.data(), count * sizeof(BVec), hipMemcpyHostToDevice, s));
  if (B->gmres) {
    B->gvecs.alloc(ctx, count * sizeof(BGVec));
    HIPCHK(hipMemcpyAsync(B->gvecs.get(), gv.data(), count * sizeof(BGVec), hipMemcpyHostToDevice, s));
  } else {
    B->kvecs.alloc(ctx, count * sizeof(BKVec));
    HIPCHK(hipMemcpyAsync(B->kvecs.get(), kv.data(), count * sizeof(BKVec), hipMemcpyHostToDevice, s));
  }
  HIPCHK(hipEventCreate(&B->ev0));
  HIPCHK(hipEventCreate(&B->ev1));
  const int ychunk = B->cfg.tile_y > 0 ? B->cfg.tile_y : 16;
  level0_lists(*B, ychunk);
  {
    const int nodes = B->cfg.tile_x > 0 ? B->cfg.tile_x : 4096;
    std::vector<BItem> rb;
    std::vector<int> f(count), c(count);
    for (int p = 0; p < count; ++p) {
      const Geo& g = B->plans[p]->g;
      f[p] = static_cast<int>(rb.size());
      const int rc = std::max(1, nodes / g.nx);
      for (int j0 = 0; j0 < g.ny; j0 += rc) rb.push_back(BItem{p, j0, std::min(g.ny, j0 + rc), 0});
      c[p] = static_cast<int>(rb.size()) - f[p];
    }
    upload_list(ctx, s, B->rb0, rb, &f, &c);
  }
  if (B->gmres) {
    // element items of the GMRES vector kernels (bk::kGmresItemNodes nodes at most)
    std::vector<BItem> el;
    std::vector<int> f(count), c(count);
    for (int p = 0; p < count; ++p) {
      const long long nn = B->plans[p]->g.n;
      f[p] = static_cast<int>(el.size());
      for (long long e0 = 0; e0 < nn; e0 += bk::kGmresItemNodes)
        el.push_back(BItem{p, static_cast<int>(e0), static_cast<int>(std::min(nn, e0 + bk::kGmresItemNodes)), 0});
      c[p] = static_cast<int>(el.size()) - f[p];
    }
    upload_list(ctx, s, B->gel, el, &f, &c);
  }
  // GMRES: the CGS partials hold (mm + 1) sums per element item
  const int np = std::max({B->el0.n, B->rows0.n, B->rb0.n, B->gel.n * (B->gmres ? B->mm + 1 : 1)});
  B->part.alloc(ctx, (size_t)np * sizeof(double));
  B->maxp.alloc(ctx, (size_t)np * sizeof(double));
  setup_finish(*B);
  HIPCHK(hipStreamSynchronize(s));
  B->bytes = ctx->used.load() - used0;
  return B;
}

}  // namespace

BatchPtr batch_create(Ctx* ctx, int count, const pdehip_problem* probs, const pdehip_options& opt,
                      const pdehip_batch_config* cfg, std::vector<std::unique_ptr<Plan>>* plans) {
  std::string why;
  if (!batch_eligible(count, probs, opt, &why)) throw Error(PDEHIP_ERR_UNSUPPORTED, why);
  BatchPtr B(new BatchPlan());
  B->ctx = ctx;
  B->n = count;
  B->opt = opt;
  if (cfg) B->cfg = *cfg;
  B->mixed = opt.precision == PDEHIP_MIXED;
  const hipStream_t s = ctx->stream;
  const size_t used0 = ctx->used.load();
  B->plans.reserve(count);
  if (opt.method == PDEHIP_BICGSTAB || opt.method == PDEHIP_GMRES)
    return create_krylov(std::move(B), ctx, count, probs, opt, used0, plans);
  for (int i = 0; i < count; ++i) {
    B->plans.push_back(take_or_create(ctx, probs, opt, plans, i));
    Plan& P = *B->plans.back();
    const int nl = static_cast<int>(P.L.size());
    if (nl < 2 || P.L[nl - 1].m <= 0)
      throw Error(PDEHIP_ERR_UNSUPPORTED, "batched engine needs a dense coarsest level");
    B->levels.push_back(nl);
    B->L = std::max(B->L, nl);
    B->coarse = P.L[1].prec;
  }
  const int n = count, L = B->L;

  // ---- device tables
  std::vector<BEnt> tab((size_t)L * n);
  std::vector<BVec> vecs(n);
  std::vector<bk::BDense> dense(n);
  for (int p = 0; p < n; ++p) {
    Plan& P = *B->plans[p];
    const int nl = B->levels[p];
    KScalars* sc_inner = B->mixed ? P.sc(1) : P.sc(0);
    for (int l = 0; l < nl; ++l) {
      Level& lv = P.L[l];
      BEnt& e = tab[(size_t)l * n + p];
      std::memset(&e, 0, sizeof(e));
      e.op = B->mixed ? lv.op32 : lv.op64;
      if (l == 0) {
        e.u = B->mixed ? P.z32.get() : P.z.get();
        e.b = B->mixed ? P.r32.get() : P.r.get();
      } else {
        e.u = lv.u.get();
        e.b = lv.b.get();
      }
      e.r = lv.r.get();
      e.alt = lv.alt.get();
      e.sc = sc_inner;
      e.pfac = (l + 1 < nl) ? P.L[l + 1].sigma / lv.sigma : 1.0;
    }
    BVec& v = vecs[p];
    std::memset(&v, 0, sizeof(v));
    v.x = B->mixed ? P.x32.get() : P.x.get();
    v.r = B->mixed ? P.r32.get() : P.r.get();
    v.p = B->mixed ? P.p32.get() : P.p.get();
    v.q = B->mixed ? P.q32.get() : P.q.get();
    v.z = B->mixed ? P.z32.get() : P.z.get();
    v.x64 = P.x.as<double>();
    v.r64 = P.r.as<double>();
    v.b64 = P.bvec.as<double>();
    v.x0 = P.x0.as<double>();
    v.so = P.sc(0);
    v.si = P.sc(1);
    v.n = P.g.n;
    v.nx = P.g.nx;
    v.ny = P.g.ny;
    v.op64 = P.L[0].op64;
    Level& cl = P.L[nl - 1];
    dense[p] = bk::BDense{cl.ainv.as<double>(), cl.map.as<int>(), cl.m, 0};
  }
  B->tab.alloc(ctx, tab.size() * sizeof(BEnt));
  B->vecs.alloc(ctx, vecs.size() * sizeof(BVec));
  B->dense.alloc(ctx, dense.size() * sizeof(bk::BDense));
  HIPCHK(hipMemcpyAsync(B->tab.get(), tab.data(), tab.size() * sizeof(BEnt), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(B->vecs.get(), vecs.data(), vecs.size() * sizeof(BVec), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(B->dense.get(), dense.data(), dense.size() * sizeof(bk::BDense), hipMemcpyHostToDevice, s));

  // ---- work items (shapes per level: defaults, explicit tile, or timed candidates)
  B->shape.assign(L, LevelShape{});
  for (int l = 0; l < L; ++l) B->shape[l] = default_shape(*B, l);
  if (B->cfg.tile_y > 0 || B->cfg.tile_x > 0)
    for (LevelShape& sh : B->shape) {
      if (B->cfg.tile_y > 0) sh.ychunk = B->cfg.tile_y;
      if (B->cfg.tile_x > 0) sh.row_nodes = sh.crow_nodes = B->cfg.tile_x;
    }
  B->sweep.resize(L); B->rows.resize(L); B->crows.resize(L); B->drows.resize(L);
  HIPCHK(hipEventCreate(&B->ev0));
  HIPCHK(hipEventCreate(&B->ev1));
  upload_shapes(*B, &dense);
  if (B->cfg.autotune) autotune(*B);
  // level-0 element items and the smoother's level-0 strips, with per-problem ranges
  level0_lists(*B, B->shape[0].ychunk);
  const int np = std::max(B->el0.n, B->rows0.n);
  B->part.alloc(ctx, (size_t)np * sizeof(double));
  B->maxp.alloc(ctx, (size_t)np * sizeof(double));
  setup_finish(*B);
  HIPCHK(hipStreamSynchronize(s));
  B->bytes = ctx->used.load() - used0;
  return B;
}

namespace {

// One V-cycle of every unfinished problem: level-0 input tab[0].b, output tab[0].u.
void vcycle(BatchPlan& B) {
  const hipStream_t s = B.ctx->stream;
  const int L = B.L;
  const int pre = std::max(2, B.opt.mg_pre), post = std::max(2, B.opt.mg_post);
  for (int l = 0; l < L; ++l) {
    bk::coarse_dense(s, B.prec(l), B.T(l), B.dense.as<bk::BDense>(), B.drows[l].d(), B.drows[l].n);
    const List& sw = B.sweep[l];
    if (sw.n == 0) continue;
    for (int it = 0; it < pre; it += 2) {
      bk::sweep(s, B.prec(l), B.T(l), sw.d(), sw.n, 0, 0, it == 0);
      bk::sweep(s, B.prec(l), B.T(l), sw.d(), sw.n, 0, 1, false);
    }
    bk::residual(s, B.prec(l), B.T(l), sw.d(), sw.n);
    bk::restrict_(s, B.prec(l), B.prec(l + 1), B.T(l), B.T(l + 1), B.crows[l + 1].d(),
                  B.crows[l + 1].n, B.mixed && l == 0);
  }
  for (int l = L - 2; l >= 0; --l) {
    const List& sw = B.sweep[l];
    if (sw.n == 0) continue;
    bk::prolong(s, B.prec(l + 1), B.prec(l), B.T(l + 1), B.T(l), B.rows[l].d(), B.rows[l].n,
                B.mixed && l == 0);
    for (int it = 0; it < post; it += 2) {
      bk::sweep(s, B.prec(l), B.T(l), sw.d(), sw.n, 1, 0, false);
      bk::sweep(s, B.prec(l), B.T(l), sw.d(), sw.n, 1, 1, false);
    }
  }
}

// Count of unfinished problems, enqueued into poll slot `slot` (device counter -> pinned host
// word, then an event); poll_read waits for that event only.
void poll_enqueue(BatchPlan& B, int inner, int slot) {
  const hipStream_t s = B.ctx->stream;
  bk::active_count(s, B.V(), B.n, inner, B.active.as<int>() + slot);
  HIPCHK(hipMemcpyAsync(B.hactive + slot, B.active.as<int>() + slot, sizeof(int),
                        hipMemcpyDeviceToHost, s));
  HIPCHK(hipEventRecord(B.evpoll[slot], s));
}

int poll_read(BatchPlan& B, int slot) {
  HIPCHK(hipEventSynchronize(B.evpoll[slot]));
  return B.hactive[slot];
}

// PCG on every problem in lock step.  Krylov precision kp (P64: the FP64 path on the outer
// scalars; P32: mixed precision on the inner scalars, with the correction accumulated in the
// FP64 solution and the FP32 residual replaced by the FP64 true residual whenever it has
// dropped by `delta` since the last replacement or claims convergence; only a true residual
// stops a problem).
void pcg(BatchPlan& B, Prec kp, int max_iter, double delta) {
  const hipStream_t s = B.ctx->stream;
  const int inner = kp == P32 ? 1 : 0;
  double* part = B.part.as<double>();
  double* maxp = B.mixed && inner ? B.maxp.as<double>() : nullptr;
  const List& el = B.el0;
  const List& rw = B.rows0;
  auto fin = [&](const List& Lst, int op, double* mp) {
    bk::finalize(s, B.V(), Lst.first.as<int>(), Lst.cnt.as<int>(), B.n, part, mp, op, inner, 0.0, 0);
  };
  vcycle(B);
  bk::vec(s, kp, bk::V_DOT_RZ, B.V(), el.d(), el.n, part, nullptr);
  fin(el, FIN_CG_INIT, nullptr);
  bk::vec(s, kp, bk::V_COPY_ZP, B.V(), el.d(), el.n, nullptr, nullptr);
  const int every = 2;
  auto iterations = [&] {
    for (int c = 0; c < every; ++c) {
      bk::apply_dot_tab(s, kp, B.V(), B.T(0), rw.d(), rw.n, part, false);
      fin(rw, FIN_CG_ALPHA, nullptr);
      bk::vec(s, kp, bk::V_CG_UPDATE, B.V(), el.d(), el.n, part, maxp);
      if (inner) {
        bk::finalize(s, B.V(), el.first.as<int>(), el.cnt.as<int>(), B.n, part, maxp,
                     FIN_RU_CHECK, 1, delta, 0);
        bk::replace_residual(s, B.V(), rw.d(), rw.n, part);
        fin(rw, FIN_RU_APPLY, nullptr);
      } else {
        fin(el, FIN_CG_RR, maxp);
      }
      vcycle(B);
      bk::vec(s, kp, bk::V_DOT_RZ, B.V(), el.d(), el.n, part, nullptr);
      fin(el, FIN_CG_BETA, nullptr);
      bk::vec(s, kp, bk::V_XPAY, B.V(), el.d(), el.n, nullptr, nullptr);
    }
  };
  // Polling is pipelined: the count after chunk c is read while chunk c + 1 is queued behind
  // it, so the device never waits for the host to enqueue; a chunk that runs after every
  // problem has finished only launches items that return at once.
  int outstanding = -1;
  int chunk = 0;
  for (int it = 0; it < max_iter; it += every, ++chunk) {
    if (B.cfg.use_graph) {
      // the launch sequence of `every` iterations, captured once per Krylov precision (every
      // pointer and item list is resident, so the graph stays valid for the batch's life)
      hipGraphExec_t& ge = B.graph[inner];
      if (!ge) {
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try {
          iterations();
        } catch (...) {
          (void)hipStreamEndCapture(s, &g);
          if (g) (void)hipGraphDestroy(g);
          throw;
        }
        HIPCHK(hipStreamEndCapture(s, &g));
        const hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HIPCHK(e);
      }
      HIPCHK(hipGraphLaunch(ge, s));
    } else {
      iterations();
    }
    const int slot = chunk & 1;
    if (B.on_chunk) B.on_chunk();
    poll_enqueue(B, inner, slot);
    if (outstanding >= 0 && poll_read(B, outstanding) == 0) { outstanding = -1; break; }
    outstanding = slot;
  }
  if (outstanding >= 0) (void)poll_read(B, outstanding);
}

// BiCGStab on every problem in lock step (reference BiCGStabSolve, krylov.cpp:82-185: the
// residual recomputed from the solution every iteration; Jacobi preconditioning optional), the
// launch sequence of the single-problem path (runtime.cpp solve_bicgstab) with every launch
// covering every unfinished problem.  Unlike the reference, a breakdown restarts the problem
// from its true residual (at most 8 times) instead of ending its solve.  Mixed precision: FP32
// Krylov vectors and applies, FP64 solution and residual.
// Mixed BiCGStab batch: the FP64 Krylov vectors of every plan (allocated once, zeroed) and the
// FP64 view table the completion pass runs on.
void ensure_fp64_views(BatchPlan& B) {
  if (B.kvecs64) return;
  Ctx* ctx = B.ctx;
  const hipStream_t s = ctx->stream;
  const size_t used0 = ctx->used.load();
  const bool jac = B.opt.precond == PDEHIP_PC_JACOBI;
  std::vector<BKVec> kv(B.n);
  for (int i = 0; i < B.n; ++i) {
    Plan& P = *B.plans[i];
    const size_t n8 = static_cast<size_t>(P.g.n) * sizeof(double);
    for (DBuf* b : {&P.p, &P.q, &P.z, &P.w1, &P.w2, &P.w3, &P.w4}) {
      if (b == &P.w4 && !jac) continue;
      if (!*b) { b->alloc(ctx, n8); fill(s, P64, b->get(), P.g.n, 0.0); }
    }
    kv[i] = krylov_view(P, jac, false);
  }
  B.kvecs64.alloc(ctx, B.n * sizeof(BKVec));
  HIPCHK(hipMemcpyAsync(B.kvecs64.get(), kv.data(), B.n * sizeof(BKVec), hipMemcpyHostToDevice, s));
  HIPCHK(hipStreamSynchronize(s));
  B.bytes += ctx->used.load() - used0;
}

void bicgstab(BatchPlan& B, int max_iter, Prec kp, const BKVec* K, int graph_slot, int stall = 0) {
  const hipStream_t s = B.ctx->stream;
  double* p1 = B.part.as<double>();
  double* p2 = B.maxp.as<double>();
  const List& rb = B.rb0;
  const List& rw = B.rows0;
  auto fin = [&](const List& Lst, int op, const double* q2) {
    bk::kfinalize(s, K, Lst.first.as<int>(), Lst.cnt.as<int>(), B.n, p1, q2, op, 0.0, 0);
  };
  bk::kvec(s, kp, bk::KK_RH, K, rb.d(), rb.n, nullptr);
  const int every = 8;
  auto iterations = [&] {
    for (int c = 0; c < every; ++c) {
      bk::kvec(s, kp, bk::KK_RESTART, K, rb.d(), rb.n, nullptr);
      bk::krestart(s, K, B.n);
      bk::kvec(s, kp, bk::KK_RHO, K, rb.d(), rb.n, p1);
      fin(rb, bk::KK_FRHO, nullptr);
      bk::kvec(s, kp, bk::KK_P, K, rb.d(), rb.n, nullptr);
      bk::kapply(s, kp, 0, K, rw.d(), rw.n, p1, nullptr);
      fin(rw, bk::KK_FALPHA, nullptr);
      bk::kvec(s, kp, bk::KK_S, K, rb.d(), rb.n, p1);
      fin(rb, bk::KK_FS, nullptr);
      bk::kvec(s, kp, bk::KK_XS, K, rb.d(), rb.n, nullptr);
      bk::kapply(s, kp, 1, K, rw.d(), rw.n, p1, p2);
      fin(rw, bk::KK_FOMEGA, p2);
      bk::kvec(s, kp, bk::KK_X, K, rb.d(), rb.n, nullptr);
      bk::kapply(s, kp, 2, K, rw.d(), rw.n, p1, nullptr);
      bk::kfinalize(s, K, rw.first.as<int>(), rw.cnt.as<int>(), B.n, p1, nullptr, bk::KK_FR,
                    static_cast<double>(stall), 0);
    }
  };
  int outstanding = -1;
  int chunk = 0;
  for (int it = 0; it < max_iter; it += every, ++chunk) {
    if (B.cfg.use_graph) {
      hipGraphExec_t& ge = B.graph[graph_slot];
      if (!ge) {
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try {
          iterations();
        } catch (...) {
          (void)hipStreamEndCapture(s, &g);
          if (g) (void)hipGraphDestroy(g);
          throw;
        }
        HIPCHK(hipStreamEndCapture(s, &g));
        const hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HIPCHK(e);
      }
      HIPCHK(hipGraphLaunch(ge, s));
    } else {
      iterations();
    }
    const int slot = chunk & 1;
    if (B.on_chunk) B.on_chunk();
    poll_enqueue(B, 0, slot);
    if (outstanding >= 0 && poll_read(B, outstanding) == 0) { outstanding = -1; break; }
    outstanding = slot;
  }
  if (outstanding >= 0) (void)poll_read(B, outstanding);
}

// GMRES(mm) on every problem in lock step: the single-problem path's cycle (runtime.cpp
// solve_gmres: V_0 = r / |r|; per step w = A M^-1 V_j, two classical Gram-Schmidt passes, the
// norm, the Givens step; then back substitution, x += M^-1 V y and the true residual) with every
// launch covering every problem.  A problem whose residual estimate meets the target leaves the
// cycle; the true residual at the cycle's end ends it.  One cycle is captured as a graph and
// replayed; the host reads the count of unfinished problems after each cycle (pipelined).
void gmres(BatchPlan& B, int max_iter) {
  const hipStream_t s = B.ctx->stream;
  const int mm = B.mm;
  const BGVec* G = B.G();
  double* p1 = B.part.as<double>();
  double* p2 = B.maxp.as<double>();
  const List& rb = B.gel;
  const List& rw = B.rows0;
  const int* rbf = rb.first.as<int>();
  const int* rbc = rb.cnt.as<int>();
  auto cycle = [&](int m) {
    bk::gcycle(s, G, B.n, bk::GC_START, mm);
    bk::gvec(s, bk::KG_START, G, rb.d(), rb.n, 0, mm, nullptr, nullptr);
    for (int j = 0; j < m; ++j) {
      bk::gapply(s, 0, G, rw.d(), rw.n, j, nullptr);
      bk::gvec(s, bk::KG_DOT, G, rb.d(), rb.n, j, mm, p1, nullptr);
      bk::gfinalize(s, G, rbf, rbc, rb.n, B.n, p1, bk::GF_H1, j, mm, 0.0, 0);
      bk::gvec(s, bk::KG_AXPY1, G, rb.d(), rb.n, j, mm, nullptr, nullptr);
      bk::gvec(s, bk::KG_DOT, G, rb.d(), rb.n, j, mm, p1, nullptr);
      bk::gfinalize(s, G, rbf, rbc, rb.n, B.n, p1, bk::GF_H2, j, mm, 0.0, 0);
      bk::gvec(s, bk::KG_AXPY2, G, rb.d(), rb.n, j, mm, nullptr, p2);
      bk::gfinalize(s, G, rbf, rbc, rb.n, B.n, p2, bk::GF_NORM, j, mm, 0.0, 0);
      bk::gvec(s, bk::KG_NORMALIZE, G, rb.d(), rb.n, j, mm, nullptr, nullptr);
    }
    bk::gcycle(s, G, B.n, bk::GC_BACKSUB, mm);
    bk::gvec(s, bk::KG_UPDATE, G, rb.d(), rb.n, 0, mm, nullptr, nullptr);
    bk::gapply(s, 1, G, rw.d(), rw.n, 0, p1);
    bk::gfinalize(s, G, rw.first.as<int>(), rw.cnt.as<int>(), rw.n, B.n, p1, bk::GF_R, 0, mm, 0.0, 0);
  };
  int outstanding = -1;
  int chunk = 0;
  for (int it = 0; it < max_iter; it += mm, ++chunk) {
    const int m = std::min(mm, max_iter - it);
    if (B.cfg.use_graph && m == mm) {
      hipGraphExec_t& ge = B.graph[0];
      if (!ge) {
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try {
          cycle(m);
        } catch (...) {
          (void)hipStreamEndCapture(s, &g);
          if (g) (void)hipGraphDestroy(g);
          throw;
        }
        HIPCHK(hipStreamEndCapture(s, &g));
        const hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HIPCHK(e);
      }
      HIPCHK(hipGraphLaunch(ge, s));
    } else {
      cycle(m);
    }
    const int slot = chunk & 1;
    if (B.on_chunk) B.on_chunk();
    poll_enqueue(B, 0, slot);
    if (outstanding >= 0 && poll_read(B, outstanding) == 0) { outstanding = -1; break; }
    outstanding = slot;
  }
  if (outstanding >= 0) (void)poll_read(B, outstanding);
}

}  // namespace

int batch_solve(BatchPlan& B, pdehip_report* reports) {
  const hipStream_t s = B.ctx->stream;
  const int n = B.n;
  double* part = B.part.as<double>();
  const List& rw = B.rows0;
  const List& el = B.el0;
  const int max_iter = std::max(1, B.opt.max_iter);
  bk::init_scalars(s, B.V(), n, B.krylov ? 1e-30 : 1e-300);
  HIPCHK(hipEventRecord(B.ev0, s));
  bk::vec(s, P64, bk::V_COPY_X0, B.V(), el.d(), el.n, nullptr, nullptr);
  if (B.gmres) {
    // r = b - A x0, |r0| and the target, then the restart cycles
    bk::gapply(s, 1, B.G(), rw.d(), rw.n, 0, part);
    bk::gfinalize(s, B.G(), rw.first.as<int>(), rw.cnt.as<int>(), rw.n, n, part, bk::GF_R0, 0,
                  B.mm, B.opt.tol, B.opt.tol_mode);
    gmres(B, max_iter);
  } else if (B.krylov) {
    // r = b - A x0 in FP64 (stored in the Krylov precision), |r0| and the target
    const Prec kp = B.mixed ? P32 : P64;
    bk::kapply(s, kp, 2, B.K(), rw.d(), rw.n, part, nullptr);
    bk::kfinalize(s, B.K(), rw.first.as<int>(), rw.cnt.as<int>(), n, part, nullptr, bk::KK_R0,
                  B.opt.tol, B.opt.tol_mode);
    bicgstab(B, max_iter, kp, B.K(), 0, B.mixed ? bk::kMixedStallWindow : 0);
    if (B.mixed) {
      // FP64 completion: problems the FP32 recurrence did not bring to the target (diverged,
      // stalled for the window, or out of iterations) continue in FP64 with the first pass's
      // target, from their current solution when the pass cut the true residual by 1e3, else
      // from x0 (refinements = -1 in the report)
      bk::kfailed(s, B.K(), n, B.active.as<int>());
      HIPCHK(hipMemcpyAsync(B.hactive, B.active.get(), sizeof(int), hipMemcpyDeviceToHost, s));
      HIPCHK(hipStreamSynchronize(s));
      if (B.hactive[0] > 0) {
        ensure_fp64_views(B);
        bk::kreset_failed(s, B.K64(), n, 1e-30);
        bk::kvec(s, P64, bk::KK_X0, B.K64(), B.rb0.d(), B.rb0.n, nullptr);
        bk::kapply(s, P64, 2, B.K64(), rw.d(), rw.n, part, nullptr);
        bk::kfinalize(s, B.K64(), rw.first.as<int>(), rw.cnt.as<int>(), n, part, nullptr,
                      bk::KK_R0K, 0.0, 0);
        bicgstab(B, max_iter, P64, B.K64(), 1);
      }
    }
  } else {
    bk::apply_dot_tab(s, P64, B.V(), B.T(0), rw.d(), rw.n, part, true);
    bk::finalize(s, B.V(), rw.first.as<int>(), rw.cnt.as<int>(), n, part, nullptr, FIN_R0, 0,
                 B.opt.tol, B.opt.tol_mode);
    if (!B.mixed) {
      pcg(B, P64, max_iter, 0.0);
    } else {
      // inner scale fixed at |r0|; inner stop only through replacements (target 0 here)
      const double delta = B.opt.inner_tol > 0 ? B.opt.inner_tol : 1e-2;
      bk::ru_start(s, B.V(), n);
      bk::vec(s, P32, bk::V_TO32, B.V(), el.d(), el.n, nullptr, nullptr);
      bk::vec(s, P32, bk::V_RR, B.V(), el.d(), el.n, part, B.maxp.as<double>());
      bk::finalize(s, B.V(), el.first.as<int>(), el.cnt.as<int>(), n, part, B.maxp.as<double>(),
                   FIN_R0, 1, 0.0, PDEHIP_TOL_ABS);
      pcg(B, P32, max_iter, delta);
    }
  }
  HIPCHK(hipEventRecord(B.ev1, s));
  // finish, batched: residual of the solved system, boundary closures, reference residual
  // norms, one gather and one copy of every problem's results
  const BVec* V = B.V();
  bk::apply_dot_tab(s, P64, V, B.T(0), rw.d(), rw.n, part, true, true);
  for (int phase = 0; phase < 4; ++phase) bk::closures(s, V, B.faces.as<BFace>(), n, phase);
  bk::resnorm(s, V, B.faces.as<BFace>(), el.d(), el.n, B.psum.as<double>(), B.pmax.as<double>());
  bk::finish_reduce(s, V, rw.first.as<int>(), rw.cnt.as<int>(), part, el.first.as<int>(),
                    el.cnt.as<int>(), B.psum.as<double>(), B.pmax.as<double>(), n,
                    B.rep.as<BReport>());
  HIPCHK(hipMemcpyAsync(B.hrep, B.rep.get(), n * sizeof(BReport), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, B.ev0, B.ev1));
  int status = PDEHIP_OK;
  for (int p = 0; p < n; ++p) {
    const BReport& k = B.hrep[p];
    pdehip_report rep{};
    rep.solve_ms = ms;
    rep.rel_residual = k.r0 > 0 ? std::sqrt(k.rr) / k.r0 : std::sqrt(k.rr);
    rep.residual_l2 = std::sqrt(k.l2sq);
    rep.residual_linf = k.linf;
    rep.converged = k.done && !k.breakdown;
    if (B.krylov) {
      rep.iterations = k.iter;
      if (B.mixed && k.steps < 0) {   // completed by the FP64 pass
        rep.refinements = -1;
      } else if (B.mixed) {           // every iteration's residual is the FP64 b - A x
        rep.inner_iterations = k.iter;
        rep.refinements = k.iter;
      }
    } else if (B.mixed) {
      rep.refinements = k.steps;
      rep.iterations = k.steps;
      rep.inner_iterations = k.inner;
    } else {
      rep.iterations = k.iter;
    }
    rep.final_metric = k.metric;
    rep.status = k.breakdown ? PDEHIP_ERR_BREAKDOWN : (k.done ? PDEHIP_OK : PDEHIP_ERR_NOT_CONVERGED);
    if (rep.status == PDEHIP_OK && !std::isfinite(rep.rel_residual)) rep.status = PDEHIP_ERR_BREAKDOWN;
    if (rep.status != PDEHIP_OK && status == PDEHIP_OK) status = rep.status;
    if (reports) reports[p] = rep;
  }
  return status;
}

double batch_time_kernel(BatchPlan& B, int kernel, int reps, double* algorithmic_bytes) {
  const hipStream_t s = B.ctx->stream;
  const Prec p0 = B.prec(0);
  const Prec kp = B.mixed ? P32 : P64;
  const double es = p0 == P64 ? 8.0 : 4.0;
  long long n0 = 0;
  for (const auto& P : B.plans) n0 += P->g.n;
  bk::init_scalars(s, B.V(), B.n, 1e-300);   // every problem active (the next solve resets them)
  double bytes = 0.0;
  float ms = 0.f;
  double* part = B.part.as<double>();
  if (B.gmres) throw Error(PDEHIP_ERR_UNSUPPORTED, "batch_time_kernel: not for GMRES batches");
  if (B.krylov) {
    const double ek = kp == P64 ? 8.0 : 4.0;
    if (kernel == 2) {          // v = A ph with rh.v partials (read ph, rh; write v)
      bytes = 3.0 * ek * static_cast<double>(n0);
      ms = time_ms(B, [&] { bk::kapply(s, kp, 0, B.K(), B.rows0.d(), B.rows0.n, part, nullptr); },
                   std::max(1, reps));
    } else if (kernel == 3) {   // r = b - A x in FP64 (read x, b; write r)
      bytes = (16.0 + ek) * static_cast<double>(n0);
      ms = time_ms(B, [&] { bk::kapply(s, kp, 2, B.K(), B.rows0.d(), B.rows0.n, part, nullptr); },
                   std::max(1, reps));
    } else {
      throw Error(PDEHIP_ERR_INVALID, "batch_time_kernel: a BiCGStab batch times kernels 2 and 3");
    }
    if (algorithmic_bytes) *algorithmic_bytes = bytes;
    return ms;
  }
  switch (kernel) {
    case 0:   // level-0 smoother: one red + black sweep, out of place (read u, b; write alt)
      bytes = 3.0 * es * static_cast<double>(n0);
      ms = time_ms(B, [&] { bk::sweep(s, p0, B.T(0), B.sweep[0].d(), B.sweep[0].n, 0, 0, false); },
                   std::max(1, reps));
      break;
    case 1:   // level-0 residual r = b - A u
      bytes = 3.0 * es * static_cast<double>(n0);
      ms = time_ms(B, [&] { bk::residual(s, p0, B.T(0), B.sweep[0].d(), B.sweep[0].n); }, std::max(1, reps));
      break;
    case 2:   // Krylov apply q = A p with p.q partials
      bytes = 2.0 * (kp == P64 ? 8.0 : 4.0) * static_cast<double>(n0);
      ms = time_ms(B, [&] { bk::apply_dot_tab(s, kp, B.V(), B.T(0), B.rows0.d(), B.rows0.n, part, false); },
                   std::max(1, reps));
      break;
    case 3:   // FP64 outer residual r64 = b64 - A64 x64
      bytes = 24.0 * static_cast<double>(n0);
      ms = time_ms(B, [&] { bk::apply_dot_tab(s, P64, B.V(), B.T(0), B.rows0.d(), B.rows0.n, part, true); },
                   std::max(1, reps));
      break;
    default:
      throw Error(PDEHIP_ERR_INVALID, "batch_time_kernel: kernel 0..3");
  }
  if (algorithmic_bytes) *algorithmic_bytes = bytes;
  return ms;
}

void batch_download(BatchPlan& B, int i, double* u_out) { plan_download(*B.plans.at(i), u_out); }

size_t batch_bytes(const BatchPlan& B) { return B.bytes; }

std::vector<std::unique_ptr<Plan>> batch_release_plans(BatchPlan& B) {
  (void)hipStreamSynchronize(B.ctx->stream);
  std::vector<std::unique_ptr<Plan>> out = std::move(B.plans);
  B.plans.clear();
  return out;
}

void batch_set_chunk_hook(BatchPlan& B, std::function<void()> hook) { B.on_chunk = std::move(hook); }

void batch_cancel(BatchPlan& B, int i) {
  if (i < 0 || i >= B.n) throw Error(PDEHIP_ERR_INVALID, "batch_cancel: problem index out of range");
  bk::cancel(B.ctx->stream, B.V(), i);
}

void BatchDeleter::operator()(BatchPlan* b) const { delete b; }

void batch_destroy(BatchPtr& B) {
  if (B) (void)hipStreamSynchronize(B->ctx->stream);
  B.reset();
}

}  // namespace pdehip
