// capi.cpp — the extern "C" boundary of libpdehip (declarations and the reference interface
// each entry point replaces: include/pdehip.h).
#include <cstring>
#include <memory>
#include <mutex>
#include <string>

#include "runtime.hpp"
#include "batch.hpp"
#include "serving.hpp"

using namespace pdehip;

struct pdehip_ctx {
  Ctx c;
};

struct pdehip_plan {
  std::unique_ptr<Plan> p;
  pdehip_ctx* owner;
};

namespace {

thread_local std::string g_tls_error;

int fail(pdehip_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->c.set_error(msg);
  g_tls_error = msg;
  return code;
}

template <typename F>
int guarded(pdehip_ctx* ctx, F&& f) {
  try {
    return f();
  } catch (const Error& e) {
    return fail(ctx, e.code, e.what());
  } catch (const std::bad_alloc&) {
    return fail(ctx, PDEHIP_ERR_OOM, "host allocation failed");
  } catch (const std::exception& e) {
    return fail(ctx, PDEHIP_ERR_DEVICE, e.what());
  }
}

// SolveOutput.error strings of the reference for the statuses it reports.
const char* status_message(int st) {
  switch (st) {
    case PDEHIP_ERR_NOT_CONVERGED: return "HIP Krylov solver failed to converge";
    case PDEHIP_ERR_BREAKDOWN: return "HIP Krylov solver broke down (denominator below threshold)";
    default: return "";
  }
}

}  // namespace

extern "C" {

int pdehip_abi_version(void) { return PDEHIP_ABI_VERSION; }

void pdehip_default_options(pdehip_options* o) {
  std::memset(o, 0, sizeof(*o));
  o->method = PDEHIP_JACOBI;
  o->max_iter = 10000;
  o->tol = 1e-6;
  o->tol_mode = PDEHIP_TOL_ABS;
  o->sor_omega = 1.5;
  o->gmres_restart = 30;
  o->mg_pre = 2;
  o->mg_post = 2;
  o->mg_coarse_iters = 50;
  o->mg_max_levels = 10;
  o->precond = PDEHIP_PC_NONE;
  o->precision = PDEHIP_FP64;
  o->bc_rows = PDEHIP_ROWS_REFERENCE;
}

void pdehip_default_problem(pdehip_problem* p) {
  std::memset(p, 0, sizeof(*p));
  p->nx = 64;
  p->ny = 64;
  p->nz = 1;
  p->xmax = p->ymax = p->zmax = 1.0;
}

int pdehip_is_available(char* note, int note_len) {
  int count = 0;
  const hipError_t e = hipGetDeviceCount(&count);
  std::string msg;
  int ok = 0;
  if (e != hipSuccess || count == 0) {
    msg = e != hipSuccess ? std::string("hipGetDeviceCount failed: ") + hipGetErrorString(e)
                          : "no HIP devices found";
  } else {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) == hipSuccess) {
      msg = std::string("device: ") + prop.name + " (" + prop.gcnArchName + ", " +
            std::to_string(prop.multiProcessorCount) + " CUs)";
      ok = std::strncmp(prop.gcnArchName, "gfx9", 4) == 0;
      if (!ok) msg += " is not a gfx9 (CDNA) device";
    }
  }
  if (note && note_len > 0) {
    std::strncpy(note, msg.c_str(), note_len - 1);
    note[note_len - 1] = 0;
  }
  return ok;
}

int pdehip_create(int device, size_t hbm_budget_bytes, pdehip_ctx** out) {
  if (!out) return PDEHIP_ERR_INVALID;
  *out = nullptr;
  auto ctx = std::make_unique<pdehip_ctx>();
  const int st = guarded(nullptr, [&]() {
    HIPCHK(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device));
    ctx->c.device = device;
    ctx->c.cus = prop.multiProcessorCount;
    ctx->c.arch = prop.gcnArchName;
    size_t free_b = 0, total_b = 0;
    HIPCHK(hipMemGetInfo(&free_b, &total_b));
    ctx->c.budget = hbm_budget_bytes ? hbm_budget_bytes : static_cast<size_t>(0.8 * free_b);
    ctx->c.pool_cap = ctx->c.budget / 8;   // released blocks kept for reuse (no hipFree syncs)
    HIPCHK(hipStreamCreateWithFlags(&ctx->c.stream, hipStreamNonBlocking));
    noalloc_init(device);   // the static scratch array's address on this device
    return PDEHIP_OK;
  });
  if (st == PDEHIP_OK) *out = ctx.release();
  return st;
}

void pdehip_destroy(pdehip_ctx* ctx) {
  if (!ctx) return;
  if (ctx->c.stream) {
    (void)hipStreamSynchronize(ctx->c.stream);
    (void)hipStreamDestroy(ctx->c.stream);
  }
  delete ctx;
}

const char* pdehip_last_error(const pdehip_ctx* ctx) {
  return ctx ? ctx->c.last_error.c_str() : g_tls_error.c_str();
}

int pdehip_solve(pdehip_ctx* ctx, const pdehip_problem* prob, const pdehip_options* opt,
                 double* u_out, pdehip_report* report) {
  if (!ctx || !prob || !opt) return PDEHIP_ERR_INVALID;
  pdehip_report local{};
  pdehip_report* rep = report ? report : &local;
  std::memset(rep, 0, sizeof(*rep));
  return guarded(ctx, [&]() {
    std::lock_guard<std::mutex> g(ctx->c.mu);
    HIPCHK(hipSetDevice(ctx->c.device));
    auto plan = plan_create(&ctx->c, *prob, *opt);
    plan_solve(*plan, rep);
    if (u_out) plan_download(*plan, u_out);
    if (rep->status != PDEHIP_OK) ctx->c.set_error(status_message(rep->status));
    return rep->status;
  });
}

int pdehip_apply(pdehip_ctx* ctx, const pdehip_problem* prob, const double* x, double* y) {
  if (!ctx || !prob || !x || !y) return PDEHIP_ERR_INVALID;
  return guarded(ctx, [&]() {
    std::lock_guard<std::mutex> g(ctx->c.mu);
    HIPCHK(hipSetDevice(ctx->c.device));
    op_apply(&ctx->c, *prob, x, y);
    return PDEHIP_OK;
  });
}

int pdehip_residual_norms(pdehip_ctx* ctx, const pdehip_problem* prob, const double* u,
                          double* l2, double* linf) {
  if (!ctx || !prob || !u || !l2 || !linf) return PDEHIP_ERR_INVALID;
  return guarded(ctx, [&]() {
    std::lock_guard<std::mutex> g(ctx->c.mu);
    HIPCHK(hipSetDevice(ctx->c.device));
    pdehip_options o;
    pdehip_default_options(&o);
    o.max_iter = 1;
    auto plan = plan_create(&ctx->c, *prob, o);
    DBuf du(&ctx->c, plan->g.n * sizeof(double));
    HIPCHK(hipMemcpyAsync(du.get(), u, plan->g.n * sizeof(double), hipMemcpyHostToDevice, ctx->c.stream));
    plan_residual_norms(*plan, du.as<double>(), l2, linf);
    return PDEHIP_OK;
  });
}

int pdehip_residual_norms_noalloc(pdehip_ctx* ctx, const pdehip_problem* prob, const double* u_dev,
                                  double* l2, double* linf) {
  if (!ctx || !prob || !u_dev || !l2 || !linf) return PDEHIP_ERR_INVALID;
  return guarded(ctx, [&]() {
    std::lock_guard<std::mutex> g(ctx->c.mu);
    HIPCHK(hipSetDevice(ctx->c.device));
    residual_norms_noalloc(&ctx->c, *prob, u_dev, l2, linf);
    return PDEHIP_OK;
  });
}

int pdehip_plan_create(pdehip_ctx* ctx, const pdehip_problem* prob, const pdehip_options* opt,
                       pdehip_plan** out) {
  if (!ctx || !prob || !opt || !out) return PDEHIP_ERR_INVALID;
  *out = nullptr;
  return guarded(ctx, [&]() {
    std::lock_guard<std::mutex> g(ctx->c.mu);
    HIPCHK(hipSetDevice(ctx->c.device));
    auto h = std::make_unique<pdehip_plan>();
    h->p = plan_create(&ctx->c, *prob, *opt);
    h->owner = ctx;
    *out = h.release();
    return PDEHIP_OK;
  });
}

int pdehip_plan_set_source(pdehip_plan* plan, const double* f_field) {
  if (!plan) return PDEHIP_ERR_INVALID;
  return guarded(plan->owner, [&]() {
    std::lock_guard<std::mutex> g(plan->owner->c.mu);
    plan_set_source(*plan->p, f_field, plan->p->f_const);
    return PDEHIP_OK;
  });
}

int pdehip_plan_solve(pdehip_plan* plan, pdehip_report* report) {
  if (!plan) return PDEHIP_ERR_INVALID;
  pdehip_report local{};
  pdehip_report* rep = report ? report : &local;
  return guarded(plan->owner, [&]() {
    std::lock_guard<std::mutex> g(plan->owner->c.mu);
    plan_solve(*plan->p, rep);
    if (rep->status != PDEHIP_OK) plan->owner->c.set_error(status_message(rep->status));
    return rep->status;
  });
}

int pdehip_plan_download(pdehip_plan* plan, double* u_out) {
  if (!plan || !u_out) return PDEHIP_ERR_INVALID;
  return guarded(plan->owner, [&]() {
    std::lock_guard<std::mutex> g(plan->owner->c.mu);
    plan_download(*plan->p, u_out);
    return PDEHIP_OK;
  });
}

size_t pdehip_plan_bytes(const pdehip_plan* plan) { return plan ? plan_bytes(*plan->p) : 0; }

int pdehip_plan_time_kernel(pdehip_plan* plan, int kernel, int precision, int reps, double* avg_ms,
                            double* algorithmic_bytes) {
  if (!plan || !avg_ms || !algorithmic_bytes) return PDEHIP_ERR_INVALID;
  return guarded(plan->owner, [&]() {
    std::lock_guard<std::mutex> g(plan->owner->c.mu);
    *avg_ms = time_kernel(*plan->p, kernel, precision, reps, algorithmic_bytes);
    return PDEHIP_OK;
  });
}

int pdehip_tune_set(const char* key, int value) {
  if (!key) return PDEHIP_ERR_INVALID;
  SmoothTuning t = smooth_tuning();
  const std::string k(key);
  if (k == "smooth.fused" && (value == 0 || value == 1)) t.fused = value;
  else if (k == "smooth.rows" && (value == 2 || value == 4)) t.rows = value;
  else if (k == "smooth.zchunk" && (value == 0 || (value >= 4 && value <= 512))) t.zchunk = value;
  else if (k == "smooth.ychunk" && value >= 16 && value <= 4096) t.ychunk = value;
  else if (k == "smooth.pair" && (value == 0 || value == 1)) t.pair = value;
  else if (k == "smooth.ring" && (value == 0 || value == 1)) t.ring = value;
  else if (k == "smooth.reg" && (value == 0 || value == 1)) t.reg = value;
  else if (k == "smooth.dot" && (value == 0 || value == 1)) t.dot = value;
  else if (k == "smooth.box" && value >= 0 && value <= 1025) t.box = value;
  else if (k == "smooth.box2" && value >= 0 && value <= 1025) t.box2 = value;
  else if (k == "smooth.box_z" && (value == 4 || value == 8)) t.box_z = value;
  else if (k == "smooth.box_rr" && value >= 0 && value <= 4097) t.box_rr = value;
  else if (k == "smooth.box_small" && value >= 0 && value <= 1025) t.box_small = value;
  else if (k == "smooth.box_sxy" && (value == 4 || value == 8)) t.box_sxy = value;
  else if (k == "stencil.plane" && (value == 0 || value == 1)) { set_plane_stencils(value == 1); return PDEHIP_OK; }
  else if (k == "stencil.ring" && (value == 0 || value == 1)) { set_plane_ring(value == 1); return PDEHIP_OK; }
  else if (k == "mixed.fp16" && (value == 0 || value == 1)) { set_mixed_fp16(value == 1); return PDEHIP_OK; }
  else if (k == "mixed.replace" && (value == 0 || value == 1)) { set_fused_replace(value == 1); return PDEHIP_OK; }
  else if (k == "mixed.xpay" && (value == 0 || value == 1)) { set_fused_xpay(value == 1); return PDEHIP_OK; }
  else if (k == "restrict.march" && (value == 0 || value == 1)) { set_restrict_march(value == 1); return PDEHIP_OK; }
  else if (k == "time.fused" && (value == 0 || value == 1)) { set_fused_steps(value == 1); return PDEHIP_OK; }
  else return PDEHIP_ERR_INVALID;
  set_smooth_tuning(t);
  return PDEHIP_OK;
}

int pdehip_tune_get(const char* key, int* value) {
  if (!key || !value) return PDEHIP_ERR_INVALID;
  const SmoothTuning t = smooth_tuning();
  const std::string k(key);
  if (k == "smooth.fused") *value = t.fused;
  else if (k == "smooth.rows") *value = t.rows;
  else if (k == "smooth.zchunk") *value = t.zchunk;
  else if (k == "smooth.ychunk") *value = t.ychunk;
  else if (k == "smooth.pair") *value = t.pair;
  else if (k == "smooth.ring") *value = t.ring;
  else if (k == "smooth.reg") *value = t.reg;
  else if (k == "smooth.dot") *value = t.dot;
  else if (k == "smooth.box") *value = t.box;
  else if (k == "smooth.box2") *value = t.box2;
  else if (k == "smooth.box_z") *value = t.box_z;
  else if (k == "smooth.box_rr") *value = t.box_rr;
  else if (k == "smooth.box_small") *value = t.box_small;
  else if (k == "smooth.box_sxy") *value = t.box_sxy;
  else if (k == "stencil.plane") *value = plane_stencils() ? 1 : 0;
  else if (k == "stencil.ring") *value = plane_ring() ? 1 : 0;
  else if (k == "mixed.fp16") *value = mixed_fp16() ? 1 : 0;
  else if (k == "mixed.replace") *value = fused_replace() ? 1 : 0;
  else if (k == "mixed.xpay") *value = fused_xpay() ? 1 : 0;
  else if (k == "restrict.march") *value = restrict_march() ? 1 : 0;
  else if (k == "time.fused") *value = fused_steps() ? 1 : 0;
  else return PDEHIP_ERR_INVALID;
  return PDEHIP_OK;
}

void pdehip_plan_destroy(pdehip_plan* plan) {
  if (!plan) return;
  {
    std::lock_guard<std::mutex> g(plan->owner->c.mu);
    (void)hipStreamSynchronize(plan->owner->c.stream);
    plan->p.reset();
  }
  delete plan;
}

int pdehip_time_series(pdehip_ctx* ctx, const pdehip_problem* prob, const pdehip_time_config* tc,
                       pdehip_frame_fn on_frame, void* user, double* u_out, pdehip_report* report) {
  if (!ctx || !prob || !tc) return PDEHIP_ERR_INVALID;
  return guarded(ctx, [&]() {
    std::lock_guard<std::mutex> g(ctx->c.mu);
    HIPCHK(hipSetDevice(ctx->c.device));
    time_series(&ctx->c, *prob, *tc, on_frame, user, u_out, report);
    return PDEHIP_OK;
  });
}

int pdehip_solve_batch(pdehip_ctx* ctx, int count, const pdehip_problem* probs,
                       const pdehip_options* opt, const pdehip_batch_config* cfg,
                       double* const* u_out, pdehip_report* reports) {
  if (!ctx || count < 0 || (count && (!probs || !opt || !u_out))) return PDEHIP_ERR_INVALID;
  return guarded(ctx, [&]() {
    std::lock_guard<std::mutex> g(ctx->c.mu);
    HIPCHK(hipSetDevice(ctx->c.device));
    return solve_batch(&ctx->c, count, probs, *opt, cfg, u_out, reports);
  });
}

struct pdehip_batch {
  pdehip_ctx* owner;
  pdehip::BatchPtr b;
};

int pdehip_batch_create(pdehip_ctx* ctx, int count, const pdehip_problem* probs,
                        const pdehip_options* opt, const pdehip_batch_config* cfg,
                        pdehip_batch** out) {
  if (!ctx || !out || count < 1 || !probs || !opt) return PDEHIP_ERR_INVALID;
  *out = nullptr;
  return guarded(ctx, [&]() {
    std::lock_guard<std::mutex> g(ctx->c.mu);
    HIPCHK(hipSetDevice(ctx->c.device));
    auto h = std::make_unique<pdehip_batch>();
    h->owner = ctx;
    h->b = batch_create(&ctx->c, count, probs, *opt, cfg);
    *out = h.release();
    return PDEHIP_OK;
  });
}

int pdehip_batch_solve(pdehip_batch* b, pdehip_report* reports) {
  if (!b) return PDEHIP_ERR_INVALID;
  return guarded(b->owner, [&]() {
    std::lock_guard<std::mutex> g(b->owner->c.mu);
    HIPCHK(hipSetDevice(b->owner->c.device));
    return batch_solve(*b->b, reports);
  });
}

int pdehip_batch_download(pdehip_batch* b, int index, double* u_out) {
  if (!b || !u_out) return PDEHIP_ERR_INVALID;
  return guarded(b->owner, [&]() {
    std::lock_guard<std::mutex> g(b->owner->c.mu);
    batch_download(*b->b, index, u_out);
    return PDEHIP_OK;
  });
}

size_t pdehip_batch_bytes(const pdehip_batch* b) { return b ? batch_bytes(*b->b) : 0; }

int pdehip_batch_time_kernel(pdehip_batch* b, int kernel, int reps, double* avg_ms,
                             double* algorithmic_bytes) {
  if (!b || !avg_ms) return PDEHIP_ERR_INVALID;
  return guarded(b->owner, [&]() {
    std::lock_guard<std::mutex> g(b->owner->c.mu);
    HIPCHK(hipSetDevice(b->owner->c.device));
    *avg_ms = batch_time_kernel(*b->b, kernel, reps, algorithmic_bytes);
    return PDEHIP_OK;
  });
}

void pdehip_batch_destroy(pdehip_batch* b) {
  if (!b) return;
  {
    std::lock_guard<std::mutex> g(b->owner->c.mu);
    batch_destroy(b->b);
  }
  delete b;
}

int pdehip_server_create(pdehip_ctx* ctx, const pdehip_server_config* cfg, pdehip_server** out) {
  if (!ctx || !out) return PDEHIP_ERR_INVALID;
  *out = nullptr;
  return guarded(ctx, [&]() {
    *out = server_create(&ctx->c, cfg);
    return PDEHIP_OK;
  });
}

int pdehip_server_submit(pdehip_server* srv, const pdehip_problem* prob, const pdehip_options* opt,
                         double* u_out, uint64_t* ticket) {
  if (!srv || !prob || !opt || !ticket) return PDEHIP_ERR_INVALID;
  return guarded(nullptr, [&]() { return server_submit(srv, *prob, *opt, u_out, ticket); });
}

int pdehip_server_poll(pdehip_server* srv, uint64_t ticket, pdehip_report* report) {
  if (!srv) return PDEHIP_ERR_INVALID;
  return guarded(nullptr, [&]() { return server_poll(srv, ticket, report, false); });
}

int pdehip_server_wait(pdehip_server* srv, uint64_t ticket, pdehip_report* report) {
  if (!srv) return PDEHIP_ERR_INVALID;
  return guarded(nullptr, [&]() { return server_poll(srv, ticket, report, true); });
}

int pdehip_server_stats_get(pdehip_server* srv, pdehip_server_stats* stats) {
  if (!srv || !stats) return PDEHIP_ERR_INVALID;
  return guarded(nullptr, [&]() { server_stats(srv, stats); return PDEHIP_OK; });
}

void pdehip_server_destroy(pdehip_server* srv) { server_destroy(srv); }

}  // extern "C"
