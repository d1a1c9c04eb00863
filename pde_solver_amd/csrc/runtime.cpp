// runtime.cpp — host runtime of libpdehip (see runtime.hpp).
#include "runtime.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

namespace pdehip {

int device_cus() {
  static const int n = [] {
    int dev = 0;
    hipDeviceProp_t prop{};
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return 256;
    return prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
  }();
  return n;
}

void hip_check(hipError_t e, const char* what) {
  if (e == hipSuccess) return;
  if (e == hipErrorOutOfMemory) throw Error(PDEHIP_ERR_OOM, std::string(what) + ": out of device memory");
  throw Error(PDEHIP_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

// ------------------------------------------------------------------ buffers
namespace {
std::atomic<bool> g_mixed_fp16{true};
std::atomic<bool> g_fused_replace{true};
std::atomic<bool> g_fused_xpay{true};
}  // namespace
void set_mixed_fp16(bool on) { g_mixed_fp16.store(on); }
bool mixed_fp16() { return g_mixed_fp16.load(); }
void set_fused_replace(bool on) { g_fused_replace.store(on); }
bool fused_replace() { return g_fused_replace.load(); }
void set_fused_xpay(bool on) { g_fused_xpay.store(on); }
bool fused_xpay() { return g_fused_xpay.load(); }

void Ctx::trim_pool() {
  std::lock_guard<std::mutex> g(pool_mu);
  for (auto& kv : pool)
    for (void* p : kv.second) (void)hipFree(p);
  pool.clear();
  pooled = 0;
}

void DBuf::alloc(Ctx* c, size_t bytes) {
  release();
  if (bytes == 0) return;
  const size_t now = c->used.fetch_add(bytes) + bytes;
  if (c->budget && now > c->budget) {
    c->used.fetch_sub(bytes);
    throw Error(PDEHIP_ERR_OOM, "allocation of " + std::to_string(bytes) +
                                    " bytes exceeds the context's HBM budget");
  }
  size_t pk = c->peak.load();
  while (now > pk && !c->peak.compare_exchange_weak(pk, now)) {}
  void* p = nullptr;
  if (c->pool_cap) {
    std::lock_guard<std::mutex> g(c->pool_mu);
    auto it = c->pool.find(bytes);
    if (it != c->pool.end() && !it->second.empty()) {
      p = it->second.back();
      it->second.pop_back();
      c->pooled -= bytes;
    }
  }
  if (!p) {
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess && c->pooled) {   // give the pooled blocks back and try once more
      (void)hipGetLastError();
      c->trim_pool();
      e = hipMalloc(&p, bytes);
    }
    if (e != hipSuccess) {
      c->used.fetch_sub(bytes);
      hip_check(e, "hipMalloc");
    }
  }
  ctx_ = c;
  p_ = p;
  bytes_ = bytes;
}

void DBuf::release() {
  if (p_) {
    bool kept = false;
    if (ctx_->pool_cap) {
      std::lock_guard<std::mutex> g(ctx_->pool_mu);
      if (ctx_->pooled + bytes_ <= ctx_->pool_cap) {
        ctx_->pool[bytes_].push_back(p_);
        ctx_->pooled += bytes_;
        kept = true;
      }
    }
    if (!kept) (void)hipFree(p_);
    ctx_->used.fetch_sub(bytes_);
  }
  p_ = nullptr;
  bytes_ = 0;
}

// ------------------------------------------------------------------ geometry
Geo make_geo(const pdehip_problem& p) {
  Geo g;
  g.nx = p.nx;
  g.ny = p.ny;
  g.nz = std::max(1, p.nz);
  g.is3d = g.nz > 1;
  g.xmin = p.xmin; g.xmax = p.xmax; g.ymin = p.ymin; g.ymax = p.ymax;
  g.zmin = p.zmin; g.zmax = p.zmax;
  g.dx = (p.xmax - p.xmin) / static_cast<double>(g.nx - 1);
  g.dy = (p.ymax - p.ymin) / static_cast<double>(g.ny - 1);
  g.dz = g.is3d ? (p.zmax - p.zmin) / static_cast<double>(g.nz - 1) : 1.0;
  g.n = (long long)g.nx * g.ny * g.nz;
  return g;
}

Geo coarsen(const Geo& f) {
  Geo c = f;
  c.nx = (f.nx + 1) / 2;
  c.ny = (f.ny + 1) / 2;
  c.nz = f.is3d ? (f.nz + 1) / 2 : 1;
  c.dx = (c.xmax - c.xmin) / static_cast<double>(c.nx - 1);
  c.dy = (c.ymax - c.ymin) / static_cast<double>(c.ny - 1);
  c.dz = c.is3d ? (c.zmax - c.zmin) / static_cast<double>(c.nz - 1) : 1.0;
  c.n = (long long)c.nx * c.ny * c.nz;
  return c;
}

static bool can_coarsen(int n) { return n >= 5 && (n % 2 == 1); }  // multigrid.cpp:10-12

// Level rule of the reference V-cycle (multigrid.cpp:150-170), applied to every axis.
int level_count(const Geo& g0, int max_levels) {
  int levels = 1;
  Geo g = g0;
  while (levels < max_levels) {
    if (!can_coarsen(g.nx) || !can_coarsen(g.ny) || (g.is3d && !can_coarsen(g.nz))) break;
    const Geo c = coarsen(g);
    if (c.nx < 3 || c.ny < 3 || (g.is3d && c.nz < 3)) break;
    ++levels;
    g = c;
    if (c.nx <= 5 || c.ny <= 5 || (g.is3d && c.nz <= 5)) break;
  }
  return levels;
}

static size_t face_count(const Geo& g, int f) {
  if (f < 2) return (size_t)g.ny * g.nz;
  if (f < 4) return (size_t)g.nx * g.nz;
  return (size_t)g.nx * g.ny;
}

static double eval_expr(const pdehip_expr& e, size_t q, double x, double y, double z) {
  if (e.samples) return e.samples[q];
  return e.c + e.cx * x + e.cy * y + e.cz * z;
}

HostFaces sample_faces(const pdehip_problem& p, const Geo& g) {
  HostFaces h;
  h.convention = p.bc_convention;
  const int nf = g.is3d ? 6 : 4;
  for (int f = 0; f < 6; ++f) {
    h.kind[f] = f < nf ? p.bc[f].kind : PDEHIP_BC_DIRICHLET;
    h.h[f] = f < 2 ? g.dx : (f < 4 ? g.dy : g.dz);
    if (f >= nf) continue;
    const size_t m = face_count(g, f);
    h.value[f].resize(m); h.alpha[f].resize(m); h.beta[f].resize(m); h.gamma[f].resize(m);
    for (size_t q = 0; q < m; ++q) {
      double x, y, z;
      if (f < 2) {
        const int j = static_cast<int>(q % g.ny), k = static_cast<int>(q / g.ny);
        x = f == 0 ? g.xmin : g.xmax; y = g.ymin + j * g.dy; z = g.is3d ? g.zmin + k * g.dz : 0.0;
      } else if (f < 4) {
        const int i = static_cast<int>(q % g.nx), k = static_cast<int>(q / g.nx);
        x = g.xmin + i * g.dx; y = f == 2 ? g.ymin : g.ymax; z = g.is3d ? g.zmin + k * g.dz : 0.0;
      } else {
        const int i = static_cast<int>(q % g.nx), j = static_cast<int>(q / g.nx);
        x = g.xmin + i * g.dx; y = g.ymin + j * g.dy; z = f == 4 ? g.zmin : g.zmax;
      }
      const pdehip_face_bc& bc = p.bc[f];
      h.value[f][q] = eval_expr(bc.value, q, x, y, z);
      h.alpha[f][q] = eval_expr(bc.alpha, q, x, y, z);
      h.beta[f][q] = eval_expr(bc.beta, q, x, y, z);
      h.gamma[f][q] = eval_expr(bc.gamma, q, x, y, z);
    }
  }
  return h;
}

void fold_thetas(const pdehip_problem& p, const Geo& g, bool folded, double theta[6]) {
  for (int f = 0; f < 6; ++f) theta[f] = 0.0;
  if (!folded) return;
  const int nf = g.is3d ? 6 : 4;
  for (int f = 0; f < nf; ++f) {
    const pdehip_face_bc& bc = p.bc[f];
    const double h = f < 2 ? g.dx : (f < 4 ? g.dy : g.dz);
    if (bc.kind == PDEHIP_BC_NEUMANN) {
      theta[f] = 1.0;
    } else if (bc.kind == PDEHIP_BC_ROBIN) {
      const double al = bc.alpha.c, be = bc.beta.c;
      const bool high = (f & 1) != 0;
      if (p.bc_convention == PDEHIP_BCCONV_AXIS && high) theta[f] = -(be / h) / (al - be / h);
      else theta[f] = (be / h) / (al + be / h);
    }
  }
}

// Reference coefficient formulas (cpu_utils.cpp:293-297; solver.cpp:723-729); upwind
// first-order convection is an extension.
static void weights_from(double a, double b, double az, double c, double d, double dz, double e,
                         const Geo& g, int convection, double w[7]) {
  const double ax = a / (g.dx * g.dx);
  const double by = b / (g.dy * g.dy);
  const double cz = g.is3d ? az / (g.dz * g.dz) : 0.0;
  double C = g.is3d ? (-2.0 * ax - 2.0 * by - 2.0 * cz + e) : (-2.0 * ax - 2.0 * by + e);
  double W, E, S, N, B = 0.0, F = 0.0;
  if (convection == PDEHIP_CONV_UPWIND) {
    E = ax + std::min(c, 0.0) / g.dx;
    W = ax - std::max(c, 0.0) / g.dx;
    N = by + std::min(d, 0.0) / g.dy;
    S = by - std::max(d, 0.0) / g.dy;
    C += std::abs(c) / g.dx + std::abs(d) / g.dy;
    if (g.is3d) {
      F = cz + std::min(dz, 0.0) / g.dz;
      B = cz - std::max(dz, 0.0) / g.dz;
      C += std::abs(dz) / g.dz;
    }
  } else {
    const double cx = c / (2.0 * g.dx), dyc = d / (2.0 * g.dy);
    E = ax + cx; W = ax - cx; N = by + dyc; S = by - dyc;
    if (g.is3d) {
      const double dzc = dz / (2.0 * g.dz);
      F = cz + dzc; B = cz - dzc;
    }
  }
  w[0] = C; w[1] = W; w[2] = E; w[3] = S; w[4] = N; w[5] = B; w[6] = F;
}

void const_weights(const pdehip_problem& p, const Geo& g, double w[7]) {
  weights_from(p.a, p.b, p.az, p.c, p.d, p.dz, p.e, g, p.convection, w);
}

void DevFaces::upload(Ctx* c, const HostFaces& h, const Geo& g, hipStream_t s) {
  size_t total = 0;
  for (int f = 0; f < 6; ++f) total += 4 * h.value[f].size();
  std::vector<double> host(std::max<size_t>(total, 1));
  buf.alloc(c, host.size() * sizeof(double));
  double* base = buf.as<double>();
  size_t off = 0;
  for (int f = 0; f < 6; ++f) {
    fd.kind[f] = h.kind[f];
    fd.h[f] = h.h[f];
    const size_t m = h.value[f].size();
    const std::vector<double>* src[4] = {&h.value[f], &h.alpha[f], &h.beta[f], &h.gamma[f]};
    const double** dst[4] = {&fd.value[f], &fd.alpha[f], &fd.beta[f], &fd.gamma[f]};
    for (int t = 0; t < 4; ++t) {
      std::copy(src[t]->begin(), src[t]->end(), host.begin() + off);
      *dst[t] = m ? base + off : nullptr;
      off += m;
    }
  }
  fd.convention = h.convention;
  (void)g;
  HIPCHK(hipMemcpyAsync(base, host.data(), host.size() * sizeof(double), hipMemcpyHostToDevice, s));
  HIPCHK(hipStreamSynchronize(s));
}

// ------------------------------------------------------------------ operators
static DevOp make_devop(const Geo& g, int kind) {
  DevOp op{};
  op.nx = g.nx; op.ny = g.ny; op.nz = g.nz;
  op.kind = kind;
  op.sxy = g.sxy();
  op.n = g.n;
  for (double& v : op.phi) v = 1.0;
  return op;
}

// Fill a level's FP64 operator (and FP32 operator when `mixed`, scaled by sigma).
static void build_level_ops(Plan& P, Level& lv, const double* kappa64_dev, bool folded) {
  const pdehip_problem& p = P.prob;
  lv.op64 = make_devop(lv.g, p.op_kind);
  fold_thetas(p, lv.g, folded, lv.op64.theta);
  for (int f = 0; f < 6; ++f) lv.op64.phi[f] = 1.0 - lv.op64.theta[f];
  lv.op64.e = p.e;
  if (p.op_kind == PDEHIP_OP_CONST) {
    const_weights(p, lv.g, lv.op64.w);
  } else if (p.op_kind == PDEHIP_OP_FLUX) {
    lv.op64.kappa = kappa64_dev;
    lv.op64.fx = p.a / (lv.g.dx * lv.g.dx);
    lv.op64.fy = p.a / (lv.g.dy * lv.g.dy);
    lv.op64.fz = lv.g.is3d ? p.a / (lv.g.dz * lv.g.dz) : 0.0;
    lv.op64.e = p.e;
  }
  lv.op32 = lv.op64;
  if (P.mixed || P.k32) {
    const double s = lv.sigma;
    for (double& wv : lv.op32.w) wv *= s;
    lv.op32.fx *= s; lv.op32.fy *= s; lv.op32.fz *= s; lv.op32.e *= s;
    if (p.op_kind == PDEHIP_OP_FLUX) {
      lv.kap32.alloc(P.ctx, lv.g.n * sizeof(float));   // rho = 1/kappa (see face_coef)
      reciprocal_f32(P.stream(), kappa64_dev, lv.kap32.as<float>(), lv.g.n);
      lv.op32.kappa = lv.kap32.get();
    }
  }
}

// Host-side weights of every interior row of a (small) level, for the dense coarse solve.
static void host_level_weights(Plan& P, const Level& lv, std::vector<double>& W) {
  const Geo& g = lv.g;
  const pdehip_problem& p = P.prob;
  W.assign((size_t)g.n * 7, 0.0);
  std::vector<double> kap;
  if (p.op_kind == PDEHIP_OP_FLUX) {
    kap.resize(g.n);
    HIPCHK(hipMemcpy(kap.data(), lv.op64.kappa, g.n * sizeof(double), hipMemcpyDeviceToHost));
  }
  double wc[7];
  const_weights(p, g, wc);
  const int k0 = g.is3d ? 1 : 0, k1 = g.is3d ? g.nz - 1 : 1;
  for (int k = k0; k < k1; ++k)
    for (int j = 1; j < g.ny - 1; ++j)
      for (int i = 1; i < g.nx - 1; ++i) {
        const long long id = ((long long)k * g.ny + j) * g.nx + i;
        double w[7];
        if (p.op_kind == PDEHIP_OP_FLUX) {
          auto face = [&](long long nb, double fct) {
            const double kc = kap[id], kn = kap[nb], s = kc + kn;
            return s > 0.0 ? fct * (2.0 * kc * kn) / s : 0.0;
          };
          const double fx = lv.op64.fx, fy = lv.op64.fy, fz = lv.op64.fz;
          w[1] = face(id - 1, fx); w[2] = face(id + 1, fx);
          w[3] = face(id - g.nx, fy); w[4] = face(id + g.nx, fy);
          double sum = w[1] + w[2] + w[3] + w[4];
          w[5] = w[6] = 0.0;
          if (g.is3d) { w[5] = face(id - g.sxy(), fz); w[6] = face(id + g.sxy(), fz); sum += w[5] + w[6]; }
          w[0] = lv.op64.e - sum;
        } else {
          std::copy(wc, wc + 7, w);
        }
        const double* th = lv.op64.theta;
        if (i == 1) w[0] += w[1] * th[0];
        if (i == g.nx - 2) w[0] += w[2] * th[1];
        if (j == 1) w[0] += w[3] * th[2];
        if (j == g.ny - 2) w[0] += w[4] * th[3];
        if (g.is3d) {
          if (k == 1) w[0] += w[5] * th[4];
          if (k == g.nz - 2) w[0] += w[6] * th[5];
        }
        for (int t = 0; t < 7; ++t) W[(size_t)id * 7 + t] = w[t];
      }
}

// Dense inverse of the coarsest level's interior operator (scaled by sigma in mixed mode).
static bool build_dense(Plan& P, Level& lv) {
  const Geo& g = lv.g;
  std::vector<long> pos(g.n, -1);
  std::vector<int> map;
  const int k0 = g.is3d ? 1 : 0, k1 = g.is3d ? g.nz - 1 : 1;
  for (int k = k0; k < k1; ++k)
    for (int j = 1; j < g.ny - 1; ++j)
      for (int i = 1; i < g.nx - 1; ++i) {
        const long long id = ((long long)k * g.ny + j) * g.nx + i;
        pos[id] = static_cast<long>(map.size());
        map.push_back(static_cast<int>(id));
      }
  const size_t m = map.size();
  if (m == 0 || m > 1500) return false;
  std::vector<double> W;
  host_level_weights(P, lv, W);
  const double s = P.mixed ? lv.sigma : 1.0;
  std::vector<double> A(m * m, 0.0), inv(m * m, 0.0);
  for (size_t r = 0; r < m; ++r) {
    const long long id = map[r];
    const double* w = &W[(size_t)id * 7];
    A[r * m + r] = s * w[0];
    auto put = [&](long long nb, double wv) { if (pos[nb] >= 0) A[r * m + pos[nb]] += s * wv; };
    put(id - 1, w[1]); put(id + 1, w[2]); put(id - g.nx, w[3]); put(id + g.nx, w[4]);
    if (g.is3d) { put(id - g.sxy(), w[5]); put(id + g.sxy(), w[6]); }
    inv[r * m + r] = 1.0;
  }
  for (size_t c = 0; c < m; ++c) {  // Gauss-Jordan with partial pivoting
    size_t piv = c;
    for (size_t r = c + 1; r < m; ++r) if (std::abs(A[r * m + c]) > std::abs(A[piv * m + c])) piv = r;
    if (std::abs(A[piv * m + c]) < 1e-300) return false;
    if (piv != c)
      for (size_t q = 0; q < m; ++q) { std::swap(A[c * m + q], A[piv * m + q]); std::swap(inv[c * m + q], inv[piv * m + q]); }
    const double d = 1.0 / A[c * m + c];
    for (size_t q = 0; q < m; ++q) { A[c * m + q] *= d; inv[c * m + q] *= d; }
    for (size_t r = 0; r < m; ++r) {
      if (r == c) continue;
      const double fct = A[r * m + c];
      if (fct == 0.0) continue;
      for (size_t q = 0; q < m; ++q) { A[r * m + q] -= fct * A[c * m + q]; inv[r * m + q] -= fct * inv[c * m + q]; }
    }
  }
  lv.m = static_cast<int>(m);
  lv.ainv.alloc(P.ctx, m * m * sizeof(double));
  lv.map.alloc(P.ctx, m * sizeof(int));
  HIPCHK(hipMemcpy(lv.ainv.get(), inv.data(), m * m * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(lv.map.get(), map.data(), m * sizeof(int), hipMemcpyHostToDevice));
  return true;
}

// ------------------------------------------------------------------ plan creation
Plan::~Plan() {
  if (hsc) (void)hipHostFree(hsc);
  if (hstage) (void)hipHostFree(hstage);
  if (ev0) (void)hipEventDestroy(ev0);
  if (ev1) (void)hipEventDestroy(ev1);
  for (hipEvent_t e : evq)
    if (e) (void)hipEventDestroy(e);
}

static bool all_dirichlet(const pdehip_problem& p, const Geo& g) {
  for (int f = 0; f < (g.is3d ? 6 : 4); ++f)
    if (p.bc[f].kind != PDEHIP_BC_DIRICHLET) return false;
  return true;
}

static void validate(const pdehip_problem& p, const pdehip_options& o) {
  if (p.nx < 3 || p.ny < 3 || (p.nz > 1 && p.nz < 3))
    throw Error(PDEHIP_ERR_INVALID, "grid must be at least 3 nodes per axis");
  if (p.xmax <= p.xmin || p.ymax <= p.ymin || (p.nz > 1 && p.zmax <= p.zmin))
    throw Error(PDEHIP_ERR_INVALID, "domain bounds are invalid");
  if (p.op_kind == PDEHIP_OP_FLUX && !p.kappa)
    throw Error(PDEHIP_ERR_INVALID, "FLUX operator needs a kappa field");
  if (o.method < PDEHIP_JACOBI || o.method > PDEHIP_MGPCG)
    throw Error(PDEHIP_ERR_INVALID, "unknown method");
  if (o.method == PDEHIP_SOR && (o.sor_omega <= 0.0 || o.sor_omega >= 2.0))
    throw Error(PDEHIP_ERR_INVALID, "SOR omega must be in (0,2)");
}

static void upload(Ctx* c, DBuf& b, const double* host, size_t n, hipStream_t s) {
  b.alloc(c, n * sizeof(double));
  HIPCHK(hipMemcpyAsync(b.get(), host, n * sizeof(double), hipMemcpyHostToDevice, s));
}

std::unique_ptr<Plan> plan_create(Ctx* ctx, const pdehip_problem& p, const pdehip_options& o) {
  validate(p, o);
  auto P = std::make_unique<Plan>();
  P->ctx = ctx;
  P->prob = p;
  P->opt = o;
  P->g = make_geo(p);
  const Geo& g = P->g;
  const hipStream_t s = ctx->stream;
  const int m = o.method;
  const bool mg = (m == PDEHIP_MG && !o.mg_single_cycle) || m == PDEHIP_MGPCG;
  const bool krylov = (m == PDEHIP_CG || m == PDEHIP_BICGSTAB || m == PDEHIP_GMRES);
  P->folded = mg || (krylov && o.bc_rows == PDEHIP_ROWS_FOLDED);
  if (krylov && !P->folded && !all_dirichlet(p, g))
    throw Error(PDEHIP_ERR_UNSUPPORTED, "Krylov methods require Dirichlet boundaries on all sides");
  P->mixed = (m == PDEHIP_MGPCG) && o.precision == PDEHIP_MIXED;
  P->k32 = (m == PDEHIP_BICGSTAB || m == PDEHIP_GMRES) && o.precision == PDEHIP_MIXED;
  if (mg && p.op_kind == PDEHIP_OP_POINTWISE)
    throw Error(PDEHIP_ERR_UNSUPPORTED, "multigrid needs constant or FLUX coefficients");
  if (m == PDEHIP_MG && o.mg_single_cycle &&
      (g.is3d || !all_dirichlet(p, g) || p.op_kind != PDEHIP_OP_CONST || p.c != 0.0 || p.d != 0.0))
    throw Error(PDEHIP_ERR_UNSUPPORTED, "Multigrid requires Dirichlet boundaries on all sides");
  if (mg || (m == PDEHIP_MG)) {
    if (!can_coarsen(g.nx) || !can_coarsen(g.ny) || (g.is3d && !can_coarsen(g.nz)))
      throw Error(PDEHIP_ERR_UNSUPPORTED, "multigrid requires odd grid dimensions of at least 5");
  }
  for (int f = 0; f < 6 && P->folded; ++f) {
    const pdehip_face_bc& bc = p.bc[f];
    if (bc.kind == PDEHIP_BC_ROBIN && (bc.alpha.samples || bc.beta.samples || bc.alpha.cx != 0 ||
                                       bc.alpha.cy != 0 || bc.alpha.cz != 0 || bc.beta.cx != 0 ||
                                       bc.beta.cy != 0 || bc.beta.cz != 0))
      throw Error(PDEHIP_ERR_UNSUPPORTED, "folded Robin rows need constant alpha and beta per face");
  }
  P->nl.n = std::min(p.n_nonlinear, PDEHIP_MAX_NONLINEAR);
  for (int t = 0; t < P->nl.n; ++t) P->nl.t[t] = p.nonlinear[t];
  if (P->nl.n && (krylov || mg || m == PDEHIP_MG))
    throw Error(PDEHIP_ERR_UNSUPPORTED, "linear PDE required for this method");

  P->hfaces = sample_faces(p, g);
  P->faces.upload(ctx, P->hfaces, g, s);

  // ---- fine level operator
  const int nl = mg ? level_count(g, std::max(1, o.mg_max_levels))
                    : (m == PDEHIP_MG ? level_count(g, std::max(1, o.mg_max_levels)) : 1);
  P->L.resize(nl);
  P->L[0].g = g;
  if (p.op_kind == PDEHIP_OP_FLUX) upload(ctx, P->L[0].kap64, p.kappa, g.n, s);
  P->L[0].sigma = 1.0;
  build_level_ops(*P, P->L[0], P->L[0].kap64.as<double>(), P->folded);
  if (p.op_kind == PDEHIP_OP_POINTWISE) {
    // per-node weights (C, W, E, S, N, B, F) from the coefficient fields (relaxation.cpp:187-201)
    std::vector<double> W((size_t)g.n * 7, 0.0);
    for (int k = 0; k < g.nz; ++k)
      for (int j = 0; j < g.ny; ++j)
        for (int i = 0; i < g.nx; ++i) {
          const long long id = ((long long)k * g.ny + j) * g.nx + i;
          auto cf = [&](int c, double v) { return p.coef[c] ? p.coef[c][id] : v; };
          double w[7];
          weights_from(cf(PDEHIP_COEF_A, p.a), cf(PDEHIP_COEF_B, p.b), cf(PDEHIP_COEF_AZ, p.az),
                       cf(PDEHIP_COEF_C, p.c), cf(PDEHIP_COEF_D, p.d), cf(PDEHIP_COEF_DZ, p.dz),
                       cf(PDEHIP_COEF_E, p.e), g, p.convection, w);
          for (int t = 0; t < 7; ++t) W[(size_t)t * g.n + id] = w[t];
        }
    upload(ctx, P->L[0].pw, W.data(), W.size(), s);
    P->L[0].op64.pw = P->L[0].pw.as<double>();
    P->L[0].op32 = P->L[0].op64;
  }

  // ---- source, boundary data, right-hand side
  P->f_const = p.f;
  if (p.f_field) upload(ctx, P->f, p.f_field, g.n, s);
  const Launch L0{s, nullptr};
  P->x0.alloc(ctx, g.n * sizeof(double));
  fill(s, P64, P->x0.get(), g.n, 0.0);
  if (P->folded) closures_nr(L0, P->op(), P->faces.fd, P->x0.as<double>());
  closures_dirichlet(L0, P->op(), P->faces.fd, P->x0.as<double>());
  P->bvec.alloc(ctx, g.n * sizeof(double));
  build_rhs(s, P->op(), P->f.as<double>(), P->f_const, P->x0.as<double>(), P->bvec.as<double>());
  // Krylov on reference rows starts from x = 0 (krylov.cpp:25); folded solves start from the
  // boundary data so boundary rows of the residual vanish.
  if (!P->folded) fill(s, P64, P->x0.get(), g.n, 0.0);
  P->x.alloc(ctx, g.n * sizeof(double));

  // ---- hierarchy
  if (nl > 1 || mg) {
    for (int l = 1; l < nl; ++l) {
      Level& f = P->L[l - 1];
      Level& c = P->L[l];
      c.g = coarsen(f.g);
      c.sigma = P->mixed ? c.g.dx * c.g.dx : 1.0;
      if (p.op_kind == PDEHIP_OP_FLUX) {
        c.kap64.alloc(ctx, c.g.n * sizeof(double));
        const DevOp fo = make_devop(f.g, p.op_kind), co = make_devop(c.g, p.op_kind);
        restrict_kappa(s, fo, f.kap64.as<double>(), co, c.kap64.as<double>());
      }
      build_level_ops(*P, c, c.kap64.as<double>(), P->folded);
    }
    for (int l = 0; l < nl; ++l) {
      Level& lv = P->L[l];
      lv.prec = P->mixed ? (l == 0 || !mixed_fp16() ? P32 : P16) : P64;
      const size_t eb = prec_bytes(lv.prec);
      lv.r.alloc(ctx, lv.g.n * eb);
      fill(s, lv.prec, lv.r.get(), lv.g.n, 0.0);
      if (l > 0) {
        lv.u.alloc(ctx, lv.g.n * eb);
        lv.b.alloc(ctx, lv.g.n * eb);
        fill(s, lv.prec, lv.u.get(), lv.g.n, 0.0);
        fill(s, lv.prec, lv.b.get(), lv.g.n, 0.0);
      }
      if (l < nl - 1) {
        lv.alt.alloc(ctx, lv.g.n * eb);
        fill(s, lv.prec, lv.alt.get(), lv.g.n, 0.0);
      }
    }
    HIPCHK(hipStreamSynchronize(s));
    if (mg) build_dense(*P, P->L[nl - 1]);
  }

  // ---- work vectors
  const size_t n8 = g.n * sizeof(double), n4 = g.n * sizeof(float);
  auto z8 = [&](DBuf& b) { b.alloc(ctx, n8); fill(s, P64, b.get(), g.n, 0.0); };
  auto z4 = [&](DBuf& b) { b.alloc(ctx, n4); fill(s, P32, b.get(), g.n, 0.0); };
  if (m == PDEHIP_CG || m == PDEHIP_MGPCG) {
    z8(P->r);
    if (!P->mixed) { z8(P->p); z8(P->q); }
    if (m == PDEHIP_MGPCG && !P->mixed) z8(P->z);
    if (P->mixed) { z8(P->xb); z4(P->r32); z4(P->x32); z4(P->p32); z4(P->q32); z4(P->z32); }
    if (P->mixed && g.nz > 1) z4(P->p32b);
  } else if (m == PDEHIP_BICGSTAB && P->k32) {
    z8(P->r);
    const int nv = o.precond == PDEHIP_PC_JACOBI ? 8 : 6;
    for (int t = 0; t < nv; ++t) z4(P->v32[t]);
  } else if (m == PDEHIP_BICGSTAB) {
    z8(P->r); z8(P->p); z8(P->q); z8(P->z); z8(P->w1); z8(P->w2); z8(P->w3);
    if (o.precond == PDEHIP_PC_JACOBI) z8(P->w4);
  } else if (m == PDEHIP_GMRES) {
    const int mm = std::max(1, std::min(o.gmres_restart, std::max(1, o.max_iter)));
    if (P->k32) {
      z8(P->r);
      z4(P->v32[0]); z4(P->v32[1]);
      P->V32.alloc(ctx, (size_t)(mm + 1) * n4);
    } else {
      z8(P->r); z8(P->w1); z8(P->w2); z8(P->w3);
      P->V.alloc(ctx, (size_t)(mm + 1) * n8);
    }
    P->H.alloc(ctx, (size_t)(mm + 1) * mm * sizeof(double));
    P->cs.alloc(ctx, mm * sizeof(double));
    P->sn.alloc(ctx, mm * sizeof(double));
    P->gv.alloc(ctx, (mm + 1) * sizeof(double));
    P->yv.alloc(ctx, mm * sizeof(double));
    P->hcb.alloc(ctx, 2 * (size_t)(mm + 1) * sizeof(double));   // two staged CGS columns
    HIPCHK(hipMemsetAsync(P->H.get(), 0, P->H.bytes(), s));
  } else if (m == PDEHIP_MG) {
    z8(P->r); z8(P->q);
  } else if (m == PDEHIP_JACOBI) {
    z8(P->r);
  }

  // ---- reductions and scalars
  int np = std::max(partial_blocks_max(P->op()), vec_blocks(g.n));
  np = std::max(np, relax_blocks(P->op()));
  np = std::max(np, krylov_blocks(g.n));
  np = std::max(np, resnorm_blocks(P->op()));
  np = std::max(np, maxabs_blocks(g.n));
  P->npart = np;
  const int gm = (m == PDEHIP_GMRES) ? std::max(1, std::min(o.gmres_restart, std::max(1, o.max_iter))) + 1 : 1;
  P->part.alloc(ctx, (size_t)np * gm * sizeof(double));
  P->part2.alloc(ctx, (size_t)np * sizeof(double));
  P->part3.alloc(ctx, (size_t)2 * np * sizeof(double) + 64);
  P->scal.alloc(ctx, 3 * sizeof(KScalars));
  HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&P->hsc), 9 * sizeof(KScalars)));
  {
    DevOp rop = P->op();
    const size_t nst = (size_t)partial_blocks_max(P->op()) + 2 * (size_t)resnorm_blocks(rop) + 8;
    HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&P->hstage), nst * sizeof(double)));
  }
  HIPCHK(hipEventCreate(&P->ev0));
  HIPCHK(hipEventCreate(&P->ev1));
  for (hipEvent_t& e : P->evq) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  HIPCHK(hipStreamSynchronize(s));
  return P;
}

size_t plan_bytes(const Plan& P) {
  size_t t = P.faces.buf.bytes() + P.f.bytes() + P.bvec.bytes() + P.x.bytes() + P.x0.bytes() +
             P.bnd.bytes();
  for (const Level& lv : P.L)
    for (const DBuf* b : {&lv.kap64, &lv.kap32, &lv.pw, &lv.u, &lv.b, &lv.r, &lv.alt, &lv.ainv, &lv.map})
      t += b->bytes();
  for (const DBuf* b : {&P.r, &P.p, &P.q, &P.z, &P.w1, &P.w2, &P.w3, &P.w4, &P.xb, &P.r32, &P.x32, &P.p32, &P.p32b,
                        &P.q32, &P.z32, &P.V32, &P.V, &P.H, &P.cs, &P.sn, &P.gv, &P.yv, &P.hcb,
                        &P.part, &P.part2, &P.part3, &P.scal})
    t += b->bytes();
  for (const DBuf& b : P.v32) t += b.bytes();
  return t;
}

void plan_set_source(Plan& P, const double* f_host, double f_const) {
  const hipStream_t s = P.stream();
  P.f_const = f_const;
  if (f_host) {
    if (!P.f) P.f.alloc(P.ctx, P.g.n * sizeof(double));
    HIPCHK(hipMemcpyAsync(P.f.get(), f_host, P.g.n * sizeof(double), hipMemcpyHostToDevice, s));
  } else {
    P.f.release();
  }
  // Boundary rows of b carry the Dirichlet datum.  Folded plans keep it in x0 (their initial
  // guess); reference-row plans start from x = 0, so their datum is rebuilt from the faces.
  const double* bnd = P.x0.as<double>();
  if (!P.folded) {
    if (!P.bnd) P.bnd.alloc(P.ctx, P.g.n * sizeof(double));
    fill(s, P64, P.bnd.get(), P.g.n, 0.0);
    closures_dirichlet(Launch{s, nullptr}, P.op(), P.faces.fd, P.bnd.as<double>());
    bnd = P.bnd.as<double>();
  }
  build_rhs(s, P.op(), P.f.as<double>(), P.f_const, bnd, P.bvec.as<double>());
  HIPCHK(hipStreamSynchronize(s));
}

// ------------------------------------------------------------------ scalars helpers
static void init_scalars(Plan& P, int i, double eps) {
  KScalars k{};
  k.eps = eps;
  k.scale = 1.0;
  k.inv_scale = 1.0;
  P.hsc[i] = k;
  HIPCHK(hipMemcpyAsync(P.sc(i), &P.hsc[i], sizeof(KScalars), hipMemcpyHostToDevice, P.stream()));
}

// Scalar blocks 0 .. count-1 (eps[i] each) in one host-to-device copy.
static void init_scalars_n(Plan& P, int count, const double* eps) {
  for (int i = 0; i < count; ++i) {
    KScalars k{};
    k.eps = eps[i];
    k.scale = 1.0;
    k.inv_scale = 1.0;
    P.hsc[i] = k;
  }
  HIPCHK(hipMemcpyAsync(P.sc(0), P.hsc, count * sizeof(KScalars), hipMemcpyHostToDevice, P.stream()));
}

void plan_init_scalars(Plan& P, int i, double eps) { init_scalars(P, i, eps); }

static const KScalars& pull(Plan& P, int i) {
  HIPCHK(hipMemcpyAsync(&P.hsc[i], P.sc(i), sizeof(KScalars), hipMemcpyDeviceToHost, P.stream()));
  HIPCHK(hipStreamSynchronize(P.stream()));
  return P.hsc[i];
}

static bool cancelled(const Plan& P) { return P.opt.cancel && *P.opt.cancel; }

// Convergence checks that trail the device by one chunk.  After each chunk of iterations the
// plan's first `nsc` scalar blocks are copied to a pinned slot and an event is recorded; the host
// reads the previous chunk's copy while the current chunk runs, so the GPU does not idle through
// a host round trip per chunk.  When the residual trend says the chunk in flight may reach the
// target, the host waits for it before queuing the next one.  A chunk queued behind a converged
// one finds done set on the device and each of its kernels returns at once, so results never
// depend on how far ahead the host ran.
class TrailingCheck {
 public:
  // ratio(k): residual over its target after a chunk (from the copied scalar blocks)
  TrailingCheck(Plan& P, int nsc, int every) : P_(P), nsc_(nsc), every_(every) {}
  // Called after a chunk has been queued.  Returns true when the loop must end: a checked chunk
  // converged or broke down, or the solve was cancelled.
  template <typename Stop, typename Ratio>
  bool after_chunk(Stop stop, Ratio ratio) {
    const int cur = slot_;
    HIPCHK(hipMemcpyAsync(&P_.hsc[3 + 3 * cur], P_.sc(0), nsc_ * sizeof(KScalars),
                          hipMemcpyDeviceToHost, P_.stream()));
    HIPCHK(hipEventRecord(P_.evq[cur], P_.stream()));
    slot_ ^= 1;
    if (pending_ >= 0) {
      if (check(pending_, stop, ratio)) return true;
      pending_ = -1;
    }
    if (ahead_ok()) {
      pending_ = cur;
      return false;
    }
    return check(cur, stop, ratio);
  }

 private:
  template <typename Stop, typename Ratio>
  bool check(int sl, Stop stop, Ratio ratio) {
    HIPCHK(hipEventSynchronize(P_.evq[sl]));
    const KScalars* k = &P_.hsc[3 + 3 * sl];
    if (stop(k) || cancelled(P_)) return true;
    hist_.push_back(ratio(k));
    return false;
  }
  // the chunk in flight cannot converge even if the residual fell `kFast` per iteration (or
  // faster, as observed) from the last checked value: at least a factor 10 would remain
  bool ahead_ok() const {
    if (hist_.empty()) return false;
    double q = 0.05;
    for (size_t i = 1; i < hist_.size(); ++i)
      if (hist_[i - 1] > 0.0 && hist_[i] > 0.0)
        q = std::min(q, std::pow(hist_[i] / hist_[i - 1], 1.0 / every_));
    q = std::max(q, 1e-3);
    const double r = hist_.back();
    return std::isfinite(r) && r * std::pow(q, every_) > 10.0;
  }
  Plan& P_;
  int nsc_, every_;
  int slot_ = 0, pending_ = -1;
  std::vector<double> hist_;
};

static int check_every(const Plan& P) {
  if (P.opt.check_every > 0) return P.opt.check_every;
  if (P.g.n >= 2000000) return 1;
  if (P.g.n >= 200000) return 4;
  return 16;
}

// ------------------------------------------------------------------ multigrid cycle
// `sweeps` red-black sweeps on level lv (colour c1 first in each sweep), result in u.  Pairs of
// sweeps run as fused out-of-place sweeps u -> alt -> u; an odd last sweep runs in place.
// zero: u is known to be zero on entry (it need not hold zeros when sweeps >= 2).
// dot (optional): the last sweep, when it is a fused one, also leaves the partials of b . u
// there; *dot_np = their count (0: none written).
static void smooth(const Launch& L, Level& lv, const DevOp& op, const void* b, void* u,
                   int sweeps, int c1, bool zero, double* dot = nullptr, int dot_cap = 0,
                   int* dot_np = nullptr) {
  int done = 0;
  if (dot_np) *dot_np = 0;
  if (lv.alt && smooth_tuning().fused) {
    for (; done + 2 <= sweeps; done += 2) {
      rbgs_sweep(L, op, lv.prec, b, u, lv.alt.get(), c1, zero);
      if (dot && done + 2 == sweeps) {
        *dot_np = rbgs_sweep_dot(L, op, lv.prec, b, lv.alt.get(), u, c1, dot, dot_cap);
        if (*dot_np > 0) { zero = false; continue; }
      }
      rbgs_sweep(L, op, lv.prec, b, lv.alt.get(), u, c1, false);
      zero = false;
    }
  }
  if (zero && done < sweeps) fill(L.stream, lv.prec, u, lv.g.n, 0.0);
  for (; done < sweeps; ++done) {
    rbgs(L, op, lv.prec, b, u, c1);
    rbgs(L, op, lv.prec, b, u, 1 - c1);
  }
}

// Box smoothing (k_box.hip) serves this level: a coarse 3D FP32/FP16 level of a mixed V-cycle
// whose sides are all within the tuning bound.
static bool box_level(const Plan& P, const Level& lv, int l) {
  const SmoothTuning t = smooth_tuning();
  return P.mixed && l > 0 && t.box > 0 && t.fused && lv.alt && lv.g.nz > 1 && lv.prec != P64 &&
         lv.g.nx <= t.box && lv.g.ny <= t.box && lv.g.nz <= t.box &&
         (lv.op32.kind == PDEHIP_OP_CONST || lv.op32.kind == PDEHIP_OP_FLUX);
}

// `sweeps` box sweeps (colour c1 first) on level lv, two per launch on levels within the
// smooth.box2 bound (else one): from a zero guess (src
// null), or from src plus the prolongation `pin` of the next level's correction folded into the
// first launch.  Results alternate between u and alt and never land in src; returns the buffer
// that holds the last one.
struct ProlongIn {
  const DevOp* op;
  Prec prec;
  const void* u;
  double factor;
};
static void* box_smooth(const Launch& L, Level& lv, const DevOp& op, const void* b, void* src,
                        int sweeps, int c1, const ProlongIn* pin) {
  void* u = lv.u.get();
  void* alt = lv.alt.get();
  void* out = src == u ? alt : u;
  const int b2 = smooth_tuning().box2;
  const int per = lv.g.nx <= b2 && lv.g.ny <= b2 && lv.g.nz <= b2 ? 2 : 1;
  bool first = true;
  for (int left = sweeps; left > 0;) {
    const int ns = left >= per ? per : 1;
    const int mode = first ? (src ? (pin ? 2 : 0) : 1) : 0;
    const bool ok = box_sweeps(L, op, lv.prec, b, src, out, c1, ns, mode, pin ? pin->op : nullptr,
                               pin ? pin->prec : lv.prec, pin ? pin->u : nullptr,
                               pin ? pin->factor : 1.0, nullptr);
    if (!ok) throw Error(PDEHIP_ERR_DEVICE, "box smoothing refused a level it was chosen for");
    first = false;
    left -= ns;
    src = out;
    out = out == u ? alt : u;
  }
  return src;
}

// One V-cycle on b0 (level-0 right-hand side) into u0 (level-0 correction from a zero guess;
// u0's contents on entry are ignored).  dot (optional): partials of b0 . u0 left by the last
// level-0 sweep, *dot_np of them (0 when that sweep could not produce them).  Coarse levels
// that box_level() picks smooth on boxes: pre-smoothing from zero in one launch, and the
// prolongation of the coarser correction with the post-smoothing in one more; their correction
// then ends in u or alt, which the finer level's prolongation reads.
static void vcycle(Plan& P, const Launch& L, const void* b0, void* u0, double* dot = nullptr,
                   int* dot_np = nullptr) {
  if (dot_np) *dot_np = 0;
  const int nl = static_cast<int>(P.L.size());
  const int pre = std::max(0, P.opt.mg_pre), post = std::max(0, P.opt.mg_post);
  KScalars* sc = const_cast<KScalars*>(L.sc);
  std::vector<void*> corr(nl);   // level l's correction: after pre-smoothing, then after post-smoothing
  for (int l = 0; l < nl - 1; ++l) {
    Level& f = P.L[l];
    Level& c = P.L[l + 1];
    const DevOp& fo = P.mixed ? f.op32 : f.op64;
    const DevOp& co = P.mixed ? c.op32 : c.op64;
    const void* bl = l == 0 ? b0 : f.b.get();
    void* ul = l == 0 ? u0 : f.u.get();
    // every level's correction starts from zero; smooth() never reads it
    if (pre > 0 && box_level(P, f, l)) {
      ul = box_smooth(L, f, fo, bl, nullptr, pre, 0, nullptr);
    } else {
      smooth(L, f, fo, bl, ul, pre, 0, true);
      if (pre == 0) fill(L.stream, f.prec, ul, f.g.n, 0.0);
    }
    corr[l] = ul;
    const double* rscale = (P.mixed && l == 0) ? &sc->scale : nullptr;
    const int brr = smooth_tuning().box_rr;
    const bool fused_rr = P.mixed && f.g.nx <= brr && f.g.ny <= brr && f.g.nz <= brr &&
                          box_resrestrict(L, fo, f.prec, bl, ul, co, c.prec, c.b.get(), 1.0, rscale);
    if (!fused_rr) {
      residual(L, fo, f.prec, bl, ul, f.r.get(), nullptr);
      restrict_fw(L, fo, f.prec, f.r.get(), co, c.prec, c.b.get(), 1.0, rscale);
    }
  }
  Level& cl = P.L[nl - 1];
  const DevOp& cop = P.mixed ? cl.op32 : cl.op64;
  void* cu = nl == 1 ? u0 : cl.u.get();
  const void* cb = nl == 1 ? b0 : cl.b.get();
  // the dense coarsest solve writes every unknown of the level's own buffer, whose other nodes
  // nothing else writes (zero since plan creation); the iterative one starts from zero
  if (!(cl.m > 0 && nl > 1)) fill(L.stream, cl.prec, cu, cl.g.n, 0.0);
  if (cl.m > 0) {
    coarse_dense(L, cl.prec, cl.ainv.as<double>(), cl.map.as<int>(), cl.m, cb, cu);
  } else {
    const int half = std::max(1, P.opt.mg_coarse_iters / 2);
    for (int it = 0; it < half; ++it) { rbgs(L, cop, cl.prec, cb, cu, 0); rbgs(L, cop, cl.prec, cb, cu, 1); }
    for (int it = 0; it < half; ++it) { rbgs(L, cop, cl.prec, cb, cu, 1); rbgs(L, cop, cl.prec, cb, cu, 0); }
  }
  corr[nl - 1] = cu;
  for (int l = nl - 2; l >= 0; --l) {
    Level& f = P.L[l];
    Level& c = P.L[l + 1];
    const DevOp& fo = P.mixed ? f.op32 : f.op64;
    const DevOp& co = P.mixed ? c.op32 : c.op64;
    const void* bl = l == 0 ? b0 : f.b.get();
    void* ul = corr[l];
    if (post > 0 && box_level(P, f, l)) {
      const ProlongIn pin{&co, c.prec, corr[l + 1], c.sigma / f.sigma};
      corr[l] = box_smooth(L, f, fo, bl, ul, post, 1, &pin);
      continue;
    }
    prolong_add(L, co, c.prec, corr[l + 1], fo, f.prec, ul, c.sigma / f.sigma,
                (P.mixed && l == 0) ? &sc->inv_scale : nullptr);
    if (l == 0 && dot) smooth(L, f, fo, bl, ul, post, 1, false, dot, P.npart, dot_np);
    else smooth(L, f, fo, bl, ul, post, 1, false);
  }
}

// ------------------------------------------------------------------ solvers
static void report_from(const KScalars& k, pdehip_report* rep) {
  rep->iterations = k.iter;
  rep->final_metric = k.metric;
  rep->converged = k.done && !k.breakdown;
}

// PCG / CG on level-0 FP64 vectors.  use_mg: MG preconditioner; else plain CG
// (reference CgSolve, krylov.cpp:14-80).
static void solve_cg(Plan& P, bool use_mg, pdehip_report* rep) {
  const hipStream_t s = P.stream();
  const Geo& g = P.g;
  const DevOp& op = P.op();
  const long long n = g.n;
  init_scalars(P, 0, use_mg ? 1e-300 : 1e-30);
  KScalars* sc = P.sc(0);
  const Launch L{s, sc};
  double* part = P.part.as<double>();
  HIPCHK(hipMemcpyAsync(P.x.get(), P.x0.get(), n * sizeof(double), hipMemcpyDeviceToDevice, s));
  const Launch L0{s, nullptr};
  residual(L0, op, P64, P.bvec.get(), P.x.get(), P.r.get(), part);
  const int np0 = partial_blocks(op);
  finalize(L0, sc, part, np0, FIN_R0, P.opt.tol, P.opt.tol_mode, nullptr);
  void* z = use_mg ? P.z.get() : P.r.get();
  if (use_mg) {
    vcycle(P, L, P.r.get(), z);
    dot(L, P64, P.r.get(), z, n, part);
    finalize(L, sc, part, vec_blocks(n), FIN_CG_INIT, 0, 0, nullptr);
  } else {
    finalize(L, sc, part, np0, FIN_CG_INIT, 0, 0, nullptr);
  }
  HIPCHK(hipMemcpyAsync(P.p.get(), z, n * sizeof(double), hipMemcpyDeviceToDevice, s));
  const int every = check_every(P);
  const int max_iter = std::max(1, P.opt.max_iter);
  int launched = 0;
  TrailingCheck tc(P, 1, every);
  while (launched < max_iter) {
    const int chunk = std::min(every, max_iter - launched);
    for (int c = 0; c < chunk; ++c) {
      // The previous iteration's preconditioning and direction update open this one, so a
      // chunk ends at the stopping test.
      if (launched + c > 0) {
        if (use_mg) {
          vcycle(P, L, P.r.get(), z);
          dot(L, P64, P.r.get(), z, n, part);
        }
        finalize(L, sc, part, vec_blocks(n), FIN_CG_BETA, 0, 0, nullptr);
        cg_xpay(L, P64, z, P.p.get(), n);
      }
      apply(L, op, P64, P.p.get(), P.q.get(), part);
      finalize(L, sc, part, partial_blocks(op), FIN_CG_ALPHA, 0, 0, nullptr);
      cg_update(L, P64, P.x.get(), P.r.get(), P.p.get(), P.q.get(), n, part, nullptr);
      finalize(L, sc, part, vec_blocks(n), FIN_CG_RR, 0, 0, nullptr);
    }
    launched += chunk;
    if (tc.after_chunk([](const KScalars* k) { return k[0].done || k[0].breakdown; },
                       [](const KScalars* k) { return k[0].metric / k[0].target; }))
      break;
  }
  const KScalars& k = pull(P, 0);
  report_from(k, rep);
  if (cancelled(P) && !k.done) throw Error(PDEHIP_ERR_CANCELLED, "solve cancelled");
  if (k.breakdown) rep->status = PDEHIP_ERR_BREAKDOWN;
  else if (!k.done) rep->status = PDEHIP_ERR_NOT_CONVERGED;
}

// Mixed precision MG-PCG with residual replacement: one FP32 PCG recurrence on r / |r0|
// (V-cycles with FP16 or FP32 coarse levels under per-level sigma = h^2 scaling and a per-cycle
// input normalisation) whose correction is accumulated in FP32 and folded into the FP64
// solution whenever the FP32 residual is replaced by the true FP64 residual b - A x: once it has
// fallen by `delta` (options.inner_tol, default 1e-2) since the last replacement, or claims the
// target.  Unlike restarted refinement the Krylov space survives every replacement, so the
// iteration count stays that of the FP64 solve.  Only a true residual stops the solve.  The
// outer residual, the solution, the outer norms and the coarsest-level solve stay FP64.
static void solve_mixed(Plan& P, pdehip_report* rep) {
  const hipStream_t s = P.stream();
  const long long n = P.g.n;
  const DevOp& op64 = P.L[0].op64;
  const DevOp& op32 = P.L[0].op32;
  double* part = P.part.as<double>();
  double* part2 = P.part2.as<double>();
  const double eps3[3] = {1e-300, 1e-300, 0.0};
  init_scalars_n(P, 3, eps3);
  KScalars* so = P.sc(0);
  KScalars* si = P.sc(1);
  KScalars* gate = P.sc(2);   // done clear exactly when a replacement is due
  const Launch L0{s, nullptr};
  const Launch Li{s, si};
  const Launch Lg{s, gate};
  double* x64 = P.x.as<double>();
  double* r64 = P.r.as<double>();
  float* x32 = P.x32.as<float>();
  float* r32 = P.r32.as<float>();
  // fused replacement (3D plane marches): the solution alternates between P.x and P.xb
  double* const xs[2] = {x64, P.xb.as<double>()};
  const bool fused = P.xb && op64.nz > 1 && plane_stencils() && fused_replace();
  HIPCHK(hipMemcpyAsync(x64, P.x0.get(), n * sizeof(double), hipMemcpyDeviceToDevice, s));
  residual(L0, op64, P64, P.bvec.get(), x64, r64, part);
  finalize(L0, so, part, partial_blocks(op64), FIN_R0, P.opt.tol, P.opt.tol_mode, nullptr);
  const double delta = P.opt.inner_tol > 0 ? P.opt.inner_tol : 1e-2;
  finalize_ru(s, si, so, gate, part, 0, FIN_RU_START, delta);
  // r32 = r64 / snorm, x32 = 0, with the partials of r32.r32 and max |r32| in the same pass
  to_f32_scaled_stats(s, r64, r32, n, &so->snorm, x32, part, part2);
  finalize(Li, si, part, vec_blocks(n), FIN_R0, 0.0, PDEHIP_TOL_ABS, nullptr);
  fin_scale(s, part2, vec_blocks(n), si);
  // r . z comes from the V-cycle's last sweep where it can (else one dot launch)
  auto precondition = [&]() -> int {
    int np = 0;
    vcycle(P, Li, r32, P.z32.get(), part, &np);
    if (np == 0) {
      dot(Li, P32, r32, P.z32.get(), n, part);
      np = vec_blocks(n);
    }
    return np;
  };
  int npz = precondition();
  finalize(Li, si, part, npz, FIN_CG_INIT, 0, 0, nullptr);
  HIPCHK(hipMemcpyAsync(P.p32.get(), P.z32.get(), n * sizeof(float), hipMemcpyDeviceToDevice, s));
  // 3D: the direction update rides in the apply, p' into the other direction buffer
  void* pcur = P.p32.get();
  void* pnext = P.p32b.get();
  const bool fuse_xpay = pnext && fused_xpay();
  const int every = check_every(P);
  const int max_iter = std::max(1, P.opt.max_iter);
  int launched = 0;
  TrailingCheck tc(P, 2, every);
  while (launched < max_iter) {
    const int chunk = std::min(every, max_iter - launched);
    for (int c = 0; c < chunk; ++c) {
      // preconditioning and direction update of the previous iteration first: a chunk ends
      // at the replacement, the only place the solve can stop
      bool applied = false;
      if (launched + c > 0) {
        npz = precondition();
        finalize(Li, si, part, npz, FIN_CG_BETA, 0, 0, nullptr);
        if (fuse_xpay &&
            apply_xpay_plane(Li, op32, P32, P.z32.get(), pcur, pnext, P.q32.get(), part)) {
          std::swap(pcur, pnext);
          applied = true;
        } else {
          cg_xpay(Li, P32, P.z32.get(), pcur, n);
        }
      }
      if (!applied) apply(Li, op32, P32, pcur, P.q32.get(), part);
      finalize(Li, si, part, partial_blocks(op32), FIN_CG_ALPHA, 0, 0, nullptr);
      cg_update(Li, P32, x32, r32, pcur, P.q32.get(), n, part, part2,
                fused ? &so->fresh : nullptr);
      finalize_ru(s, si, so, gate, part, vec_blocks(n), FIN_RU_CHECK, delta, part2, vec_blocks(n));
      // replacement (only when due): x += snorm x32, r = b - A x in FP64, r32 = r / snorm
      if (fused) {
        replace_plane(Lg, op64, P.bvec.as<double>(), xs, x32, r64, r32, so, part);
        finalize_ru(s, si, so, gate, part, partial_blocks(op64), FIN_RU_APPLY_PP, delta);
      } else {
        flush_f32(s, x64, x32, n, &so->snorm, gate);
        residual(Lg, op64, P64, P.bvec.get(), x64, r64, part);
        to_f32_scaled_gated(s, r64, r32, n, &so->snorm, gate);
        finalize_ru(s, si, so, gate, part, partial_blocks(op64), FIN_RU_APPLY, delta);
      }
    }
    launched += chunk;
    // blocks 0 (outer) and 1 (inner): the inner residual is relative to the outer snorm
    if (tc.after_chunk([](const KScalars* k) { return k[1].done || k[1].breakdown; },
                       [](const KScalars* k) { return k[1].metric * k[0].snorm / k[0].target; }))
      break;
  }
  // Every decision after the loop is taken on the device, so the post-solve work queues behind
  // the last chunk without a host round trip; the report reads the scalars when it syncs.
  // An unconverged solve keeps the correction accumulated since the last replacement.
  if (!fused) {
    flush_f32(s, x64, x32, n, &so->snorm, so);
  } else {
    // the current solution buffer back into P.x, then (unconverged, correction not void) the
    // correction accumulated since the last replacement
    copy_if_xsel(s, x64, xs[1], n, so);
    flush_f32(s, x64, x32, n, &so->snorm, so, true);
  }
  // converged at a replacement: r64 and its partials are b - A x of the final x already
  fresh_gate(s, so, si, gate);
  P.resid_gate = gate;
  P.dev_report = true;
  (void)rep;
}

// Iterated V-cycles (extension) until ||b - A u||_2 meets the target.
static void solve_mg_iter(Plan& P, pdehip_report* rep) {
  const hipStream_t s = P.stream();
  const long long n = P.g.n;
  const DevOp& op = P.op();
  init_scalars(P, 0, 1e-300);
  KScalars* sc = P.sc(0);
  const Launch L0{s, nullptr};
  const Launch L{s, sc};
  double* part = P.part.as<double>();
  HIPCHK(hipMemcpyAsync(P.x.get(), P.x0.get(), n * sizeof(double), hipMemcpyDeviceToDevice, s));
  residual(L0, op, P64, P.bvec.get(), P.x.get(), P.r.get(), part);
  finalize(L0, sc, part, partial_blocks(op), FIN_R0, P.opt.tol, P.opt.tol_mode, nullptr);
  const int max_iter = std::max(1, P.opt.max_iter);
  const int every = check_every(P);
  int launched = 0;
  while (launched < max_iter) {
    const int chunk = std::min(every, max_iter - launched);
    for (int c = 0; c < chunk; ++c) {
      // correction form: e = V(r) from a zero guess, x += e
      vcycle(P, L, P.r.get(), P.q.get());
      axpy_dev(L, P64, P.x.get(), P.q.get(), n, nullptr, 1.0);
      residual(L, op, P64, P.bvec.get(), P.x.get(), P.r.get(), part);
      finalize(L, sc, part, partial_blocks(op), FIN_CG_RR, 0, 0, nullptr);
    }
    launched += chunk;
    const KScalars& k = pull(P, 0);
    if (k.done || cancelled(P)) break;
  }
  const KScalars& k = pull(P, 0);
  report_from(k, rep);
  if (!k.done) rep->status = PDEHIP_ERR_NOT_CONVERGED;
}

// The reference's single V-cycle (MultigridVcyclePoisson, multigrid.cpp:121-222): 2D, constant
// a/b/e, lexicographic Gauss-Seidel (wavefront order on the device), one cycle.
static void solve_mg_reference(Plan& P, pdehip_report* rep) {
  const hipStream_t s = P.stream();
  const pdehip_problem& p = P.prob;
  const Launch L0{s, nullptr};
  const int nl = static_cast<int>(P.L.size());
  // levels' constant coefficients at their spacing
  std::vector<Geo> G(nl);
  G[0] = P.g;
  for (int l = 1; l < nl; ++l) G[l] = coarsen(G[l - 1]);
  std::vector<DBuf> U(nl), B(nl), R(nl);
  for (int l = 0; l < nl; ++l) {
    const size_t bytes = G[l].n * sizeof(double);
    if (l > 0) { U[l].alloc(P.ctx, bytes); B[l].alloc(P.ctx, bytes); fill(s, P64, U[l].get(), G[l].n, 0.0); fill(s, P64, B[l].get(), G[l].n, 0.0); }
    R[l].alloc(P.ctx, bytes);
  }
  // xvec = Dirichlet values on a zero grid (solver.cpp:573-580)
  fill(s, P64, P.x.get(), P.g.n, 0.0);
  closures_dirichlet(L0, P.op(), P.faces.fd, P.x.as<double>());
  auto uptr = [&](int l) { return l == 0 ? P.x.as<double>() : U[l].as<double>(); };
  auto bptr = [&](int l) { return l == 0 ? P.bvec.as<double>() : B[l].as<double>(); };
  auto coef = [&](int l, double* ax, double* by, double* ct) {
    *ax = p.a / (G[l].dx * G[l].dx);
    *by = p.b / (G[l].dy * G[l].dy);
    *ct = -2.0 * (*ax) - 2.0 * (*by) + p.e;
  };
  for (int l = 0; l < nl; ++l) {
    double ax, by, ct;
    coef(l, &ax, &by, &ct);
    if (std::abs(ct) < 1e-12) throw Error(PDEHIP_ERR_DEGENERATE, "CPU multigrid solver failed");
  }
  for (int l = 0; l < nl - 1; ++l) {
    double ax, by, ct;
    coef(l, &ax, &by, &ct);
    lex_gs(s, G[l].nx, G[l].ny, ax, by, ct, bptr(l), uptr(l), std::max(0, P.opt.mg_pre));
    DevOp fo = make_devop(G[l], PDEHIP_OP_CONST);
    const_weights(p, G[l], fo.w);
    DevOp co = make_devop(G[l + 1], PDEHIP_OP_CONST);
    residual(L0, fo, P64, bptr(l), uptr(l), R[l].get(), nullptr);
    restrict_fw(L0, fo, P64, R[l].get(), co, P64, B[l + 1].get(), 1.0, nullptr);
    fill(s, P64, U[l + 1].get(), G[l + 1].n, 0.0);
  }
  {
    double ax, by, ct;
    coef(nl - 1, &ax, &by, &ct);
    lex_gs(s, G[nl - 1].nx, G[nl - 1].ny, ax, by, ct, bptr(nl - 1), uptr(nl - 1),
           std::max(1, P.opt.mg_coarse_iters));
  }
  for (int l = nl - 2; l >= 0; --l) {
    DevOp fo = make_devop(G[l], PDEHIP_OP_CONST);
    DevOp co = make_devop(G[l + 1], PDEHIP_OP_CONST);
    prolong_add(L0, co, P64, U[l + 1].get(), fo, P64, uptr(l), 1.0, nullptr);
    double ax, by, ct;
    coef(l, &ax, &by, &ct);
    lex_gs(s, G[l].nx, G[l].ny, ax, by, ct, bptr(l), uptr(l), std::max(0, P.opt.mg_post));
  }
  HIPCHK(hipStreamSynchronize(s));
  rep->iterations = 1;
  rep->converged = 1;
}

// Jacobi / Gauss-Seidel / SOR with the reference's closure order (relaxation.cpp:310-574,
// 52-308; solver.cpp:751-981).  Gauss-Seidel and SOR sweep red then black on the device.
static void solve_relax(Plan& P, pdehip_report* rep) {
  const hipStream_t s = P.stream();
  const Geo& g = P.g;
  const long long n = g.n;
  const DevOp& op = P.op();
  init_scalars(P, 0, 0.0);
  KScalars* sc = P.sc(0);
  const Launch L{s, sc};
  double* part = P.part.as<double>();
  double* maxslot = P.part3.as<double>() + P.npart;
  int* degenerate = reinterpret_cast<int*>(P.part3.as<double>() + P.npart + 1);
  HIPCHK(hipMemsetAsync(degenerate, 0, sizeof(int), s));
  const FaceData& fd = P.faces.fd;
  double* grid = P.x.as<double>();
  double* next = P.r.get() ? P.r.as<double>() : nullptr;
  fill(s, P64, grid, n, 0.0);
  const Launch L0{s, nullptr};
  closures_dirichlet(L0, op, fd, grid);
  if (g.is3d) { closures_nr(L0, op, fd, grid); closures_dirichlet(L0, op, fd, grid); }
  const bool jac = P.opt.method == PDEHIP_JACOBI;
  if (jac) HIPCHK(hipMemcpyAsync(next, grid, n * sizeof(double), hipMemcpyDeviceToDevice, s));
  const double omega = P.opt.method == PDEHIP_SOR ? P.opt.sor_omega : 1.0;
  const int max_iter = std::max(1, P.opt.max_iter);
  const int every = std::max(check_every(P), 32);
  const double tol = P.opt.tol;
  // relaxation stop rule: max |du| < tol, evaluated on the device each sweep
  int launched = 0;
  std::vector<double*> bufs = {grid, next};
  int cur = 0;
  while (launched < max_iter) {
    const int chunk = std::min(every, max_iter - launched);
    for (int c = 0; c < chunk; ++c) {
      double* u = bufs[cur];
      closures_nr(L, op, fd, u);
      if (jac) {
        double* v = bufs[cur ^ 1];
        jacobi_ref(L, op, P.f.as<double>(), P.f_const, P.nl, u, v, part, degenerate);
        closures_nr(L, op, fd, v);
        closures_dirichlet(L, op, fd, v);
      } else {
        sor_rb_ref(L, op, P.f.as<double>(), P.f_const, P.nl, u, omega, 0, part, degenerate);
        sor_rb_ref(L, op, P.f.as<double>(), P.f_const, P.nl, u, omega, 1, part, degenerate);
        if (g.is3d) { closures_nr(L, op, fd, u); closures_dirichlet(L, op, fd, u); }
      }
      finalize(L, sc, part, relax_blocks(op), FIN_RELAX, tol, 0, maxslot);
      if (jac) cur ^= 1;
    }
    launched += chunk;
    const KScalars& k = pull(P, 0);
    int deg = 0;
    HIPCHK(hipMemcpy(&deg, degenerate, sizeof(int), hipMemcpyDeviceToHost));
    if (deg) throw Error(PDEHIP_ERR_DEGENERATE, "degenerate PDE center coefficient");
    if (k.done || cancelled(P)) break;
  }
  const KScalars& k = pull(P, 0);
  if (cancelled(P) && !k.done) throw Error(PDEHIP_ERR_CANCELLED, "solve cancelled");
  // the sweep that set `done` is the last one that changed the grid
  const int sweeps = k.iter;
  const int final_buf = jac ? (sweeps % 2) : 0;
  if (final_buf != 0) HIPCHK(hipMemcpyAsync(grid, bufs[1], n * sizeof(double), hipMemcpyDeviceToDevice, s));
  rep->iterations = sweeps;
  rep->final_metric = k.metric;
  rep->converged = k.done;
}

// BiCGStab (reference krylov.cpp:82-185) with optional Jacobi right preconditioner.
// from_x: start from the current solution x (not x0) with the absolute target abs_target (the
// FP64 completion of a mixed solve).
static void solve_bicgstab(Plan& P, pdehip_report* rep, bool from_x = false, double abs_target = 0.0) {
  const hipStream_t s = P.stream();
  const long long n = P.g.n;
  const DevOp& op = P.op();
  init_scalars(P, 0, 1e-30);
  KScalars* sc = P.sc(0);
  const Launch L{s, sc};
  const Launch L0{s, nullptr};
  double* part = P.part.as<double>();
  double* part2 = P.part2.as<double>();
  double *x = P.x.as<double>(), *r = P.r.as<double>(), *pv = P.p.as<double>(), *v = P.q.as<double>();
  double *rhat = P.z.as<double>(), *sv = P.w1.as<double>(), *t = P.w2.as<double>(), *sh = P.w3.as<double>();
  const bool pc = P.opt.precond == PDEHIP_PC_JACOBI;
  double* ph = pc ? P.w4.as<double>() : pv;
  double* shv = pc ? sh : sv;
  if (!from_x) HIPCHK(hipMemcpyAsync(x, P.x0.get(), n * sizeof(double), hipMemcpyDeviceToDevice, s));
  residual(L0, op, P64, P.bvec.get(), x, r, part);
  if (from_x)
    finalize(L0, sc, part, partial_blocks(op), FIN_R0, abs_target, PDEHIP_TOL_ABS, nullptr);
  else
    finalize(L0, sc, part, partial_blocks(op), FIN_R0, P.opt.tol, P.opt.tol_mode, nullptr);
  HIPCHK(hipMemcpyAsync(rhat, r, n * sizeof(double), hipMemcpyDeviceToDevice, s));
  const int kb = krylov_blocks(n);
  const int max_iter = std::max(1, P.opt.max_iter);
  const int every = check_every(P);
  int launched = 0;
  while (launched < max_iter) {
    const int chunk = std::min(every, max_iter - launched);
    for (int c = 0; c < chunk; ++c) {
      dot(L, P64, rhat, r, n, part);
      fin_k(s, sc, part, nullptr, vec_blocks(n), KFIN_BI_RHO);
      bicg_p_pc(s, op, r, pv, v, ph, pc, sc);     // with the Jacobi scaling of p
      apply(L, op, P64, ph, v, nullptr);
      dot(L, P64, rhat, v, n, part);
      fin_k(s, sc, part, nullptr, vec_blocks(n), KFIN_BI_ALPHA);
      bicg_s_pc(s, op, r, v, sv, shv, pc, part, sc);   // with the Jacobi scaling of s
      fin_k(s, sc, part, nullptr, kb, KFIN_BI_S);
      bicg_xs(s, x, ph, n, sc);
      bi_sdone(s, sc);
      apply(L, op, P64, shv, t, nullptr);
      dot2(s, t, t, sv, n, part, part2, sc);
      fin_k(s, sc, part, part2, kb, KFIN_BI_OMEGA);
      bicg_x(s, x, ph, shv, n, sc);
      residual(L, op, P64, P.bvec.get(), x, r, part);
      fin_k(s, sc, part, nullptr, partial_blocks(op), KFIN_BI_R);
    }
    launched += chunk;
    const KScalars& k = pull(P, 0);
    if (k.done || k.breakdown || cancelled(P)) break;
  }
  const KScalars& k = pull(P, 0);
  report_from(k, rep);
  if (cancelled(P) && !k.done) throw Error(PDEHIP_ERR_CANCELLED, "solve cancelled");
  if (k.breakdown && !k.done) rep->status = PDEHIP_ERR_BREAKDOWN;
  else if (!k.done) rep->status = PDEHIP_ERR_NOT_CONVERGED;
}

// Restarted GMRES(m) (reference krylov.cpp:187-333) with CGS2 and optional Jacobi right
// preconditioner; the back-substitution uses the current cycle's step count.
static void solve_gmres(Plan& P, pdehip_report* rep) {
  const hipStream_t s = P.stream();
  const long long n = P.g.n;
  const DevOp& op = P.op();
  const int max_iter = std::max(1, P.opt.max_iter);
  const int mm = std::max(1, std::min(P.opt.gmres_restart, max_iter));
  init_scalars(P, 0, 1e-30);
  KScalars* sc = P.sc(0);
  const Launch L{s, sc};
  const Launch L0{s, nullptr};
  double* part = P.part.as<double>();
  double *x = P.x.as<double>(), *r = P.r.as<double>(), *w = P.w1.as<double>();
  double *zt = P.w2.as<double>(), *tmp = P.w3.as<double>();
  double* V = P.V.as<double>();
  double *H = P.H.as<double>(), *cs = P.cs.as<double>(), *sn = P.sn.as<double>();
  double *gv = P.gv.as<double>(), *yv = P.yv.as<double>();
  const bool pc = P.opt.precond == PDEHIP_PC_JACOBI;
  HIPCHK(hipMemcpyAsync(x, P.x0.get(), n * sizeof(double), hipMemcpyDeviceToDevice, s));
  residual(L0, op, P64, P.bvec.get(), x, r, part);
  finalize(L0, sc, part, partial_blocks(op), FIN_R0, P.opt.tol, P.opt.tol_mode, nullptr);
  const int kb = krylov_blocks(n);
  int iter_total = 0;
  while (iter_total < max_iter) {
    const KScalars& k0 = pull(P, 0);
    if (k0.done || k0.breakdown || cancelled(P)) break;
    const int m = std::min(mm, max_iter - iter_total);
    gm_start(s, V, r, n, gv, m, sc);
    for (int j = 0; j < m; ++j) {
      const double* vj = V + (long long)j * n;
      if (pc) { jacobi_pc(s, op, vj, zt, sc); apply(L, op, P64, zt, w, nullptr); }
      else apply(L, op, P64, vj, w, nullptr);
      double* hcol = P.hcb.as<double>();  // two passes of (j+1) coefficients, mm+1 apart
      for (int pass = 0; pass < 2; ++pass) {
        multidot(s, V, n, j + 1, w, n, part, kb, sc);
        fin_multi(s, part, kb, j + 1, hcol + (pass ? mm + 1 : 0), 0, sc);
        multiaxpy(s, V, n, j + 1, hcol + (pass ? mm + 1 : 0), w, n, -1.0, 1, sc);
      }
      gm_setcol(s, H, hcol, mm + 1, j, m, j + 1, sc);   // H[i][j] = h1[i] + h2[i]
      dot(L, P64, w, w, n, part);
      fin_norm(s, part, vec_blocks(n), H + (long long)(j + 1) * m + j, sc);
      gm_normalize(s, V + (long long)(j + 1) * n, w, n, H + (long long)(j + 1) * m + j, sc);
      givens(s, H, cs, sn, gv, j, m, sc);
    }
    backsub(s, H, gv, yv, m, sc);
    // x += M (V y)
    fill(s, P64, zt, n, 0.0);
    const KScalars& kc = pull(P, 0);
    const int steps = kc.steps;
    if (kc.breakdown) break;
    multiaxpy(s, V, n, steps, yv, zt, n, 1.0, 0, sc);
    if (pc) { jacobi_pc(s, op, zt, w, sc); axpy_dev(L, P64, x, w, n, nullptr, 1.0); }
    else axpy_dev(L, P64, x, zt, n, nullptr, 1.0);
    residual(L, op, P64, P.bvec.get(), x, r, part);
    fin_k(s, sc, part, nullptr, partial_blocks(op), KFIN_GM_R);
    iter_total += steps;
    HIPCHK(hipMemsetAsync(H, 0, P.H.bytes(), s));
  }
  const KScalars& k = pull(P, 0);
  rep->iterations = iter_total;
  rep->final_metric = k.metric;
  rep->converged = k.done;
  if (cancelled(P) && !k.done) throw Error(PDEHIP_ERR_CANCELLED, "solve cancelled");
  if (k.breakdown && !k.done) rep->status = PDEHIP_ERR_BREAKDOWN;
  else if (!k.done) rep->status = PDEHIP_ERR_NOT_CONVERGED;
}

void plan_residual_norms(Plan& P, const double* u_dev, double* l2, double* linf) {
  const hipStream_t s = P.stream();
  double* sums = P.part.as<double>();
  double* maxs = P.part2.as<double>();
  int* degenerate = reinterpret_cast<int*>(P.part3.as<double>() + P.npart + 1);
  HIPCHK(hipMemsetAsync(degenerate, 0, sizeof(int), s));
  DevOp op = P.op();
  for (double& t : op.theta) t = 0.0;
  residual_norms(s, op, P.faces.fd, u_dev, P.f.as<double>(), P.f_const, P.nl, sums, maxs, degenerate);
  const int nb = resnorm_blocks(op);
  std::vector<double> hs(nb), hm(nb);
  HIPCHK(hipMemcpyAsync(hs.data(), sums, nb * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(hm.data(), maxs, nb * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  double a = 0.0, m = 0.0;
  for (int i = 0; i < nb; ++i) { a += hs[i]; m = std::max(m, hm[i]); }
  *l2 = std::sqrt(std::max(0.0, a));
  *linf = m;
}

// Mixed-precision BiCGStab: FP32 BiCGStab (the recurrence r = s - omega t, CUDA semantics)
// whose x updates go straight into the FP64 solution, with FP64 residual replacement: at every
// check (check_every iterations: 1, 4 or 16 by size) r is replaced by (b - A x)/|r0| computed
// in FP64, and only that true residual stops the solve.  The Krylov state (p, r_hat, rho)
// carries on across replacements; a breakdown restarts it from the true residual.  FP32
// applies use the difference form.  On diffusion-dominated problems of 513^2 and more the FP32
// iteration can diverge or stall where FP64 BiCGStab converges: a true residual above 1e3 |r0|,
// or one that has not halved for 100 iterations, ends the FP32 pass, and the solve completes
// with the FP64 BiCGStab, from the current solution when the FP32 pass cut the true residual
// by 1e3 and from the initial guess otherwise (report: refinements = -1, inner_iterations = the
// FP32 iterations spent).  Restarted GMRES in mixed precision keeps the FP64 iteration count.
static void solve_bicgstab_mixed(Plan& P, pdehip_report* rep) {
  const hipStream_t s = P.stream();
  const long long n = P.g.n;
  const DevOp& op64 = P.op();
  const DevOp& op32 = P.L[0].op32;
  init_scalars(P, 0, 1e-30);
  init_scalars(P, 1, 1e-30);
  KScalars* so = P.sc(0);
  KScalars* si = P.sc(1);
  const Launch L0{s, nullptr};
  const Launch Li{s, si};
  double* part = P.part.as<double>();
  double* part2 = P.part2.as<double>();
  double* x = P.x.as<double>();
  double* r64 = P.r.as<double>();
  const bool pc = P.opt.precond == PDEHIP_PC_JACOBI;
  float* r = P.v32[0].as<float>();
  float* rhat = P.v32[1].as<float>();
  float* pv = P.v32[2].as<float>();
  float* v = P.v32[3].as<float>();
  float* sv = P.v32[4].as<float>();
  float* t = P.v32[5].as<float>();
  float* ph = pc ? P.v32[6].as<float>() : pv;
  float* sh = pc ? P.v32[7].as<float>() : sv;
  HIPCHK(hipMemcpyAsync(x, P.x0.get(), n * sizeof(double), hipMemcpyDeviceToDevice, s));
  residual(L0, op64, P64, P.bvec.get(), x, r64, part);
  finalize(L0, so, part, partial_blocks(op64), FIN_R0, P.opt.tol, P.opt.tol_mode, nullptr);
  const KScalars k0 = pull(P, 0);
  const int kb = krylov_blocks(n);
  const int max_iter = std::max(1, P.opt.max_iter);
  const int every = check_every(P);
  int replacements = 0, iter_base = 0;
  // stall window: a true residual that has not halved for this many iterations ends the FP32
  // pass (as a growth beyond 1e3 |r0| does); the FP64 completion continues from the current x
  constexpr int kStallWindow = 100;
  double best = k0.r0;
  int best_at = 0;
  if (!k0.done) {
    // inner vectors hold the residual divided by the fixed scale |r0|
    to_f32_scaled(s, r64, r, n, &so->r0);
    HIPCHK(hipMemcpyAsync(rhat, r, n * sizeof(float), hipMemcpyDeviceToDevice, s));
    dot(L0, P32, r, r, n, part);
    finalize(L0, si, part, vec_blocks(n), FIN_R0, k0.target / k0.r0, PDEHIP_TOL_ABS, nullptr);
    int launched = 0;
    while (launched < max_iter) {
      const int chunk = std::min(every, max_iter - launched);
      for (int c = 0; c < chunk; ++c) {
        dot(Li, P32, rhat, r, n, part);
        fin_k(s, si, part, nullptr, vec_blocks(n), KFIN_BI_RHO);
        bicg_p_pc<float>(s, op64, r, pv, v, ph, pc, si);
        apply(Li, op32, P32, ph, v, nullptr);
        dot(Li, P32, rhat, v, n, part);
        fin_k(s, si, part, nullptr, vec_blocks(n), KFIN_BI_ALPHA);
        bicg_s_pc<float>(s, op64, r, v, sv, sh, pc, part, si);
        fin_k(s, si, part, nullptr, kb, KFIN_BI_S);
        bicg_xs<float>(s, x, ph, n, si, &so->r0);
        bi_sdone(s, si);
        apply(Li, op32, P32, sh, t, nullptr);
        dot2<float>(s, t, t, sv, n, part, part2, si);
        fin_k(s, si, part, part2, kb, KFIN_BI_OMEGA);
        bicg_x<float>(s, x, ph, sh, n, si, &so->r0);
        bicg_r(s, sv, t, r, n, part, si);
        fin_k(s, si, part, nullptr, kb, KFIN_BI_R);
      }
      launched += chunk;
      if (cancelled(P)) break;
      KScalars ki = pull(P, 1);
      // replacement: the FP64 true residual decides
      residual(L0, op64, P64, P.bvec.get(), x, r64, part);
      finalize(L0, so, part, partial_blocks(op64), FIN_CG_RR, 0, 0, nullptr);
      ++replacements;
      const KScalars ko = pull(P, 0);
      if (ko.done || ko.breakdown || !(ko.metric < 1e3 * ko.r0)) break;   // converged / diverged
      if (ko.metric < 0.5 * best) {
        best = ko.metric;
        best_at = launched;
      } else if (launched - best_at >= kStallWindow) {
        break;   // stalled
      }
      to_f32_scaled(s, r64, r, n, &so->r0);
      if (ki.breakdown) {
        // rho or omega vanished: restart the recurrence from the true residual
        HIPCHK(hipMemcpyAsync(rhat, r, n * sizeof(float), hipMemcpyDeviceToDevice, s));
        iter_base += ki.iter;
        ki.iter = 0;
        ki.breakdown = 0;
      }
      ki.done = 0;
      ki.flag_s = 0;
      ki.metric = ko.metric / ko.r0;
      P.hsc[1] = ki;
      HIPCHK(hipMemcpyAsync(si, &P.hsc[1], sizeof(KScalars), hipMemcpyHostToDevice, s));
    }
  }
  const KScalars ko = pull(P, 0);
  const KScalars ki = pull(P, 1);
  const int mixed_iters = iter_base + ki.iter;
  if (!ko.done && !cancelled(P)) {
    // the FP32 iteration diverged, stalled or ran out of iterations: the FP64 BiCGStab
    // (reference semantics) continues from the current FP64 solution when the FP32 pass cut the
    // true residual by 1e3 (else from the initial guess), to the first pass's target; its
    // vectors are allocated on the first such case
    const size_t n8 = n * sizeof(double);
    for (DBuf* b : {&P.p, &P.q, &P.z, &P.w1, &P.w2, &P.w3, &P.w4}) {
      if (b == &P.w4 && !pc) continue;
      if (!*b) { b->alloc(P.ctx, n8); fill(s, P64, b->get(), n, 0.0); }
    }
    pdehip_report r64rep{};
    const bool keep_x = std::isfinite(ko.metric) && ko.metric < 1e-3 * ko.r0;
    solve_bicgstab(P, &r64rep, keep_x, ko.target);
    if (keep_x) {
      // the report's relative residual is against the first pass's initial residual
      KScalars kf = pull(P, 0);
      kf.r0 = ko.r0;
      P.hsc[0] = kf;
      HIPCHK(hipMemcpyAsync(P.sc(0), &P.hsc[0], sizeof(KScalars), hipMemcpyHostToDevice, s));
    }
    *rep = r64rep;
    rep->iterations = mixed_iters + r64rep.iterations;
    rep->inner_iterations = mixed_iters;
    rep->refinements = -1;   // marks the FP64 completion
    return;
  }
  rep->iterations = mixed_iters;
  rep->refinements = replacements;
  rep->inner_iterations = mixed_iters;
  rep->final_metric = ko.metric;
  rep->converged = ko.done;
  if (cancelled(P) && !ko.done) throw Error(PDEHIP_ERR_CANCELLED, "solve cancelled");
  if (!ko.done) rep->status = PDEHIP_ERR_NOT_CONVERGED;
}

// Mixed-precision restarted GMRES: the Arnoldi basis, the operator applies and the
// preconditioner run in FP32 (dot products and the Hessenberg/Givens algebra in FP64); each
// restart cycle adds its correction to the FP64 solution and recomputes the FP64 residual, so
// the restarts are the refinement steps.
static void solve_gmres_mixed(Plan& P, pdehip_report* rep) {
  const hipStream_t s = P.stream();
  const long long n = P.g.n;
  const DevOp& op64 = P.op();
  const DevOp& op32 = P.L[0].op32;
  const int max_iter = std::max(1, P.opt.max_iter);
  const int mm = std::max(1, std::min(P.opt.gmres_restart, max_iter));
  init_scalars(P, 0, 1e-30);
  KScalars* sc = P.sc(0);
  const Launch L{s, sc};
  const Launch L0{s, nullptr};
  double* part = P.part.as<double>();
  double* x = P.x.as<double>();
  double* r = P.r.as<double>();
  float* w = P.v32[0].as<float>();
  float* zt = P.v32[1].as<float>();
  float* V = P.V32.as<float>();
  double *H = P.H.as<double>(), *cs = P.cs.as<double>(), *sn = P.sn.as<double>();
  double *gv = P.gv.as<double>(), *yv = P.yv.as<double>();
  const bool pc = P.opt.precond == PDEHIP_PC_JACOBI;
  HIPCHK(hipMemcpyAsync(x, P.x0.get(), n * sizeof(double), hipMemcpyDeviceToDevice, s));
  residual(L0, op64, P64, P.bvec.get(), x, r, part);
  finalize(L0, sc, part, partial_blocks(op64), FIN_R0, P.opt.tol, P.opt.tol_mode, nullptr);
  const int kb = krylov_blocks(n);
  int iter_total = 0, cycles = 0;
  while (iter_total < max_iter) {
    const KScalars& k0 = pull(P, 0);
    if (k0.done || k0.breakdown || cancelled(P)) break;
    const int m = std::min(mm, max_iter - iter_total);
    ++cycles;
    gm_start<float>(s, V, r, n, gv, m, sc);
    for (int j = 0; j < m; ++j) {
      const float* vj = V + (long long)j * n;
      if (pc) { jacobi_pc<float>(s, op64, vj, zt, sc); apply(L, op32, P32, zt, w, nullptr); }
      else apply(L, op32, P32, vj, w, nullptr);
      double* hcol = P.hcb.as<double>();
      for (int pass = 0; pass < 2; ++pass) {
        multidot<float>(s, V, n, j + 1, w, n, part, kb, sc);
        fin_multi(s, part, kb, j + 1, hcol + (pass ? mm + 1 : 0), 0, sc);
        multiaxpy<float>(s, V, n, j + 1, hcol + (pass ? mm + 1 : 0), w, n, -1.0, 1, sc);
      }
      gm_setcol(s, H, hcol, mm + 1, j, m, j + 1, sc);
      dot(L, P32, w, w, n, part);
      fin_norm(s, part, vec_blocks(n), H + (long long)(j + 1) * m + j, sc);
      gm_normalize<float>(s, V + (long long)(j + 1) * n, w, n, H + (long long)(j + 1) * m + j, sc);
      givens(s, H, cs, sn, gv, j, m, sc);
    }
    backsub(s, H, gv, yv, m, sc);
    fill(s, P32, zt, n, 0.0);
    const KScalars& kc = pull(P, 0);
    const int steps = kc.steps;
    if (kc.breakdown) break;
    multiaxpy<float>(s, V, n, steps, yv, zt, n, 1.0, 0, sc);
    if (pc) { jacobi_pc<float>(s, op64, zt, w, sc); add_f32(s, x, w, n, nullptr, 1.0); }
    else add_f32(s, x, zt, n, nullptr, 1.0);
    residual(L, op64, P64, P.bvec.get(), x, r, part);
    fin_k(s, sc, part, nullptr, partial_blocks(op64), KFIN_GM_R);
    iter_total += steps;
    HIPCHK(hipMemsetAsync(H, 0, P.H.bytes(), s));
  }
  const KScalars& k = pull(P, 0);
  rep->iterations = iter_total;
  rep->refinements = cycles;
  rep->inner_iterations = iter_total;
  rep->final_metric = k.metric;
  rep->converged = k.done;
  if (cancelled(P) && !k.done) throw Error(PDEHIP_ERR_CANCELLED, "solve cancelled");
  if (k.breakdown && !k.done) rep->status = PDEHIP_ERR_BREAKDOWN;
  else if (!k.done) rep->status = PDEHIP_ERR_NOT_CONVERGED;
}

void plan_solve(Plan& P, pdehip_report* rep) {
  std::memset(rep, 0, sizeof(*rep));
  const hipStream_t s = P.stream();
  const double eps2[2] = {0.0, 0.0};
  init_scalars_n(P, 2, eps2);
  HIPCHK(hipEventRecord(P.ev0, s));
  P.resid_fresh = false;
  P.resid_gate = nullptr;
  P.dev_report = false;
  switch (P.opt.method) {
    case PDEHIP_JACOBI:
    case PDEHIP_GAUSS_SEIDEL:
    case PDEHIP_SOR: solve_relax(P, rep); break;
    case PDEHIP_CG: solve_cg(P, false, rep); break;
    case PDEHIP_BICGSTAB:
      if (P.k32) solve_bicgstab_mixed(P, rep);
      else solve_bicgstab(P, rep);
      break;
    case PDEHIP_GMRES:
      if (P.k32) solve_gmres_mixed(P, rep);
      else solve_gmres(P, rep);
      break;
    case PDEHIP_MG:
      if (P.opt.mg_single_cycle) solve_mg_reference(P, rep);
      else solve_mg_iter(P, rep);
      break;
    case PDEHIP_MGPCG:
      if (P.mixed) solve_mixed(P, rep);
      else solve_cg(P, true, rep);
      break;
  }
  HIPCHK(hipEventRecord(P.ev1, s));
  plan_finish_enqueue(P);
  HIPCHK(hipStreamSynchronize(s));
  plan_finish_collect(P, rep);
}

// Post-solve work on the plan's stream, no host wait: the solver system's true residual,
// boundary reconstruction of folded solves (closures then Dirichlet, reference order), the
// reference residual norms, and the copies the report needs into the plan's pinned staging.
void plan_finish_enqueue(Plan& P) {
  const hipStream_t s = P.stream();
  const Launch L0{s, nullptr};
  double* part = P.part.as<double>();
  const int mth = P.opt.method;
  P.fin_resid = P.r && mth != PDEHIP_JACOBI && mth != PDEHIP_GAUSS_SEIDEL && mth != PDEHIP_SOR;
  if (P.fin_resid && P.resid_gate)
    residual(Launch{s, P.resid_gate}, P.op(), P64, P.bvec.get(), P.x.get(), P.r.get(), part);
  else if (P.fin_resid && !P.resid_fresh)
    residual(L0, P.op(), P64, P.bvec.get(), P.x.get(), P.r.get(), part);
  if (P.folded) {
    closures_nr(L0, P.op(), P.faces.fd, P.x.as<double>());
    closures_dirichlet(L0, P.op(), P.faces.fd, P.x.as<double>());
  }
  double* sums = P.part3.as<double>();
  double* maxs = sums + P.npart;
  int* degenerate = reinterpret_cast<int*>(sums + 2 * P.npart + 1);
  HIPCHK(hipMemsetAsync(degenerate, 0, sizeof(int), s));
  DevOp rop = P.op();
  for (double& t : rop.theta) t = 0.0;
  residual_norms(s, rop, P.faces.fd, P.x.as<double>(), P.f.as<double>(), P.f_const, P.nl, sums,
                 maxs, degenerate);
  const int np = partial_blocks(P.op()), nb = resnorm_blocks(rop);
  double* h = P.hstage;
  if (P.fin_resid) HIPCHK(hipMemcpyAsync(h, part, np * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(h + np, sums, nb * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(h + np + nb, maxs, nb * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(&P.hsc[0], P.sc(0), (P.dev_report ? 2 : 1) * sizeof(KScalars),
                        hipMemcpyDeviceToHost, s));
}

// Report fields from the staged copies (after the stream has been synchronised).
void plan_finish_collect(Plan& P, pdehip_report* rep) {
  DevOp rop = P.op();
  const int np = partial_blocks(P.op()), nb = resnorm_blocks(rop);
  const double* h = P.hstage;
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, P.ev0, P.ev1) == hipSuccess) rep->solve_ms = ms;
  if (P.fin_resid) {
    double rr = 0.0;
    for (int i = 0; i < np; ++i) rr += h[i];
    // relative to the solve's initial residual (||b|| for reference rows, which start at x = 0)
    const double r0 = P.hsc[0].r0;
    rep->rel_residual = r0 > 0 ? std::sqrt(rr) / r0 : std::sqrt(rr);
  }
  double a = 0.0, m = 0.0;
  for (int i = 0; i < nb; ++i) {
    a += h[np + i];
    const double v = h[np + nb + i];
    if (v != v || v > m) m = v;   // NaN propagates
  }
  rep->residual_l2 = std::sqrt(a);
  rep->residual_linf = m;
  if (P.dev_report) {
    const KScalars& ko = P.hsc[0];
    const KScalars& ki = P.hsc[1];
    rep->iterations = ki.iter;
    rep->refinements = ko.steps;
    rep->inner_iterations = ki.iter;
    rep->final_metric = ko.metric;
    rep->converged = ko.done && !ko.breakdown;
    if (cancelled(P) && !ko.done) throw Error(PDEHIP_ERR_CANCELLED, "solve cancelled");
    if (ko.breakdown || ki.breakdown) rep->status = PDEHIP_ERR_BREAKDOWN;
    else if (!ko.done) rep->status = PDEHIP_ERR_NOT_CONVERGED;
  }
}

void plan_download(Plan& P, double* u_out) {
  HIPCHK(hipMemcpyAsync(u_out, P.x.get(), P.g.n * sizeof(double), hipMemcpyDeviceToHost, P.stream()));
  HIPCHK(hipStreamSynchronize(P.stream()));
}

// ------------------------------------------------------------------ apply (reference rows)
void op_apply(Ctx* ctx, const pdehip_problem& p, const double* x, double* y) {
  pdehip_options o{};
  o.method = PDEHIP_JACOBI;
  o.max_iter = 1;
  o.sor_omega = 1.5;
  auto P = plan_create(ctx, p, o);
  const hipStream_t s = ctx->stream;
  DBuf dx(ctx, P->g.n * sizeof(double)), dy(ctx, P->g.n * sizeof(double));
  HIPCHK(hipMemcpyAsync(dx.get(), x, P->g.n * sizeof(double), hipMemcpyHostToDevice, s));
  apply(Launch{s, nullptr}, P->op(), P64, dx.get(), dy.get(), nullptr);
  HIPCHK(hipMemcpyAsync(y, dy.get(), P->g.n * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
}

// ------------------------------------------------------------------ time series
// SolvePDETimeSeries (solver.cpp:989-1329): 2D explicit steps, closures after each step.
namespace {

void stage(hipStream_t s, const DevOp& op, const double* v, const double* un, double* acc,
           double* out, const double* f, double fc, const NonlinearTerms& nl, double ut, double dt,
           double c1, int mode) {
  time_stage(s, op, v, un, acc, out, f, fc, nl, ut, dt, c1, mode);
}
void stage(hipStream_t s, const DevOp& op, const float* v, const float* un, float* acc,
           float* out, const double* f, double fc, const NonlinearTerms& nl, double ut, double dt,
           double c1, int mode) {
  time_stage32(s, op, v, un, acc, out, f, fc, nl, ut, dt, c1, mode);
}

// The integration loop with state type S (double, or float for the FP32 state); stage
// arithmetic is FP64 in both.  Frames and the result are widened to FP64 on the device.
template <typename S>
void run_series(Ctx* ctx, Plan& PL, const pdehip_problem& p, const pdehip_time_config& tc,
                pdehip_frame_fn cb, void* user, double* u_out, pdehip_report* rep) {
  const bool has_ut = std::abs(p.ut) > 1e-12, has_utt = std::abs(p.utt) > 1e-12;
  constexpr Prec PS = sizeof(S) == 8 ? P64 : P32;
  const Geo& g = PL.g;
  const hipStream_t s = ctx->stream;
  const long long n = g.n;
  const DevOp& op = PL.op();
  const FaceData& fd = PL.faces.fd;
  const Launch L0{s, nullptr};
  const size_t es = sizeof(S);
  DBuf u(ctx, n * es), un(ctx, n * es), acc(ctx, n * es), ta(ctx, n * es), tb(ctx, n * es), vel;
  DBuf stage64;
  if (p.initial) {
    if constexpr (PS == P64) {
      HIPCHK(hipMemcpyAsync(u.get(), p.initial, n * 8, hipMemcpyHostToDevice, s));
    } else {
      stage64.alloc(ctx, n * 8);
      HIPCHK(hipMemcpyAsync(stage64.get(), p.initial, n * 8, hipMemcpyHostToDevice, s));
      convert(s, P64, stage64.get(), P32, u.get(), n, 1.0);
    }
  } else {
    fill(s, PS, u.get(), n, 0.0);
  }
  if (has_utt) {
    vel.alloc(ctx, n * 8);
    if (p.initial_velocity) HIPCHK(hipMemcpyAsync(vel.get(), p.initial_velocity, n * 8, hipMemcpyHostToDevice, s));
    else fill(s, P64, vel.get(), n, 0.0);
  }
  closures_dirichlet(L0, op, fd, u.as<S>());
  closures_nr(L0, op, fd, u.as<S>());
  closures_dirichlet(L0, op, fd, u.as<S>());
  const int frames = std::max(1, tc.frames);
  const bool integ = !has_utt && tc.integrator != PDEHIP_TI_FORWARD_EULER;
  const double dt = tc.dt;
  const double* f = PL.f.as<double>();
  const double fc = PL.f_const, ut = p.ut;
  const NonlinearTerms& nl = PL.nl;
  // frame buffering: up to buffer_frames frames stay on the device before one batched copy
  int buffer = tc.buffer_frames > 0 ? tc.buffer_frames : 1;
  buffer = std::max(1, std::min(buffer, frames));
  DBuf fbuf(ctx, (size_t)(cb ? buffer : 1) * n * 8);
  std::vector<double> host(cb ? (size_t)buffer * n : 0), vhost;
  int in_buf = 0, buf_first = 0;
  bool stop = false;
  auto widen = [&](const S* src, double* dst) {
    if constexpr (PS == P64) HIPCHK(hipMemcpyAsync(dst, src, n * 8, hipMemcpyDeviceToDevice, s));
    else convert(s, P32, src, P64, dst, n, 1.0);
  };
  auto flush = [&]() {
    if (in_buf == 0) return;
    HIPCHK(hipMemcpyAsync(host.data(), fbuf.get(), (size_t)in_buf * n * 8, hipMemcpyDeviceToHost, s));
    if (has_utt) { vhost.resize(n); HIPCHK(hipMemcpyAsync(vhost.data(), vel.get(), n * 8, hipMemcpyDeviceToHost, s)); }
    HIPCHK(hipStreamSynchronize(s));
    for (int q = 0; q < in_buf && !stop; ++q) {
      const int fr = buf_first + q;
      if (cb && !cb(user, fr, tc.t_start + fr * dt, host.data() + (size_t)q * n,
                    has_utt ? vhost.data() : nullptr))
        stop = true;
    }
    in_buf = 0;
  };
  auto t0 = std::chrono::steady_clock::now();
  int steps = 0;
  for (int frame = 0; frame < frames && !stop; ++frame) {
    if (cb) {
      if (in_buf == 0) buf_first = frame;
      widen(u.as<S>(), fbuf.as<double>() + (size_t)in_buf * n);
      ++in_buf;
      // velocity is reported at the flush only; flush every frame for u_tt to keep it exact
      if (in_buf == buffer || has_utt) flush();
    }
    if (frame + 1 >= frames) break;
    S* U = u.as<S>();
    S* A = acc.as<S>();
    S* Ta = ta.as<S>();
    S* Tb = tb.as<S>();
    bool closed = false;   // the step's kernel applied the closures itself
    if (integ) {
      switch (tc.integrator) {
        case PDEHIP_TI_RK2:
          stage(s, op, U, U, A, Ta, f, fc, nl, ut, dt, dt, 0);
          stage(s, op, Ta, U, A, U, f, fc, nl, ut, dt, 0.5 * dt, 3);
          break;
        case PDEHIP_TI_RK4:
          if (fused_steps()) {
            // the whole step and its closures in one launch (same bits)
            time_rk4_step(s, op, fd, U, Ta, f, fc, nl, ut, dt);
            std::swap(u, ta);
            closed = true;
            break;
          }
          stage(s, op, U, U, A, Ta, f, fc, nl, ut, dt, 0.5 * dt, 0);
          stage(s, op, Ta, U, A, Tb, f, fc, nl, ut, dt, 0.5 * dt, 1);
          stage(s, op, Tb, U, A, Ta, f, fc, nl, ut, dt, dt, 1);
          stage(s, op, Ta, U, A, U, f, fc, nl, ut, dt, dt / 6.0, 2);
          break;
        case PDEHIP_TI_SSPRK2:
          HIPCHK(hipMemcpyAsync(un.get(), U, n * es, hipMemcpyDeviceToDevice, s));
          stage(s, op, U, U, static_cast<S*>(nullptr), Ta, f, fc, nl, ut, dt, dt, 0);
          stage(s, op, Ta, un.as<S>(), static_cast<S*>(nullptr), U, f, fc, nl, ut, dt, 0.0, 4);
          break;
        case PDEHIP_TI_SSPRK3:
          HIPCHK(hipMemcpyAsync(un.get(), U, n * es, hipMemcpyDeviceToDevice, s));
          stage(s, op, U, U, static_cast<S*>(nullptr), Ta, f, fc, nl, ut, dt, dt, 0);
          stage(s, op, Ta, un.as<S>(), static_cast<S*>(nullptr), Tb, f, fc, nl, ut, dt, 0.0, 5);
          stage(s, op, Tb, un.as<S>(), static_cast<S*>(nullptr), U, f, fc, nl, ut, dt, 0.0, 6);
          break;
        default:
          stage(s, op, U, U, static_cast<S*>(nullptr), Ta, f, fc, nl, ut, dt, dt, 0);
          HIPCHK(hipMemcpyAsync(U, Ta, n * es, hipMemcpyDeviceToDevice, s));
      }
    } else if (has_utt) {
      if constexpr (PS == P64) {
        time_utt(s, op, U, vel.as<double>(), Ta, f, fc, nl, ut, p.utt, dt);
        std::swap(u, ta);
      }
    } else {
      stage(s, op, U, U, static_cast<S*>(nullptr), Ta, f, fc, nl, ut, dt, dt, 0);
      std::swap(u, ta);
    }
    if (!closed) {
      closures_nr(L0, op, fd, u.as<S>());
      closures_dirichlet(L0, op, fd, u.as<S>());
    }
    ++steps;
  }
  flush();
  HIPCHK(hipStreamSynchronize(s));
  auto t1 = std::chrono::steady_clock::now();
  if (u_out) {
    widen(u.as<S>(), fbuf.as<double>());
    HIPCHK(hipMemcpyAsync(u_out, fbuf.get(), n * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
  }
  if (rep) {
    std::memset(rep, 0, sizeof(*rep));
    rep->iterations = steps;
    rep->converged = 1;
    rep->solve_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  }
}

}  // namespace

// Time series (SolvePDETimeSeries, solver.cpp:989-1329).  precision PDEHIP_MIXED keeps the
// state in FP32 (stage arithmetic FP64); u_tt problems always run in FP64.
void time_series(Ctx* ctx, const pdehip_problem& p, const pdehip_time_config& tc,
                 pdehip_frame_fn cb, void* user, double* u_out, pdehip_report* rep) {
  if (p.nz > 1) throw Error(PDEHIP_ERR_UNSUPPORTED, "3D Time Series Solver not yet refactored.");
  if (tc.dt <= 0.0) throw Error(PDEHIP_ERR_INVALID, "time step dt must be positive");
  const bool has_ut = std::abs(p.ut) > 1e-12, has_utt = std::abs(p.utt) > 1e-12;
  if (!has_ut && !has_utt) throw Error(PDEHIP_ERR_INVALID, "time-dependent solve requires u_t or u_tt term");
  if (has_utt && p.initial && !p.initial_velocity)
    throw Error(PDEHIP_ERR_INVALID, "checkpoint restart for u_tt requires velocity data");
  pdehip_options o{};
  o.method = PDEHIP_JACOBI;
  o.max_iter = 1;
  o.sor_omega = 1.5;
  auto P = plan_create(ctx, p, o);
  if (tc.precision == PDEHIP_MIXED && !has_utt)
    run_series<float>(ctx, *P, p, tc, cb, user, u_out, rep);
  else
    run_series<double>(ctx, *P, p, tc, cb, user, u_out, rep);
}

// ------------------------------------------------------------------ kernel timing
double time_kernel(Plan& P, int kernel, int precision, int reps, double* bytes) {
  const hipStream_t s = P.stream();
  const long long n = P.g.n;
  const Launch L{s, nullptr};
  const bool p32 = precision == PDEHIP_MIXED;
  const Prec pr = p32 ? P32 : P64;
  const size_t es = prec_bytes(pr);
  const DevOp& op = p32 ? P.L[0].op32 : P.L[0].op64;
  if (p32 && P.prob.op_kind == PDEHIP_OP_FLUX && !P.L[0].kap32)
    throw Error(PDEHIP_ERR_INVALID, "plan holds no FP32 operator");
  const size_t kb = P.prob.op_kind == PDEHIP_OP_FLUX ? kappa_bytes(pr) : 0;
  // The plan's own level-0 work vectors when it has them (MG-PCG plans): the timed launches
  // then touch the same buffers, and the same cache footprint, as the solver's launches.
  // Their contents are scratch between solves.  Otherwise three temporary vectors.
  DBuf ta, tb, tc;
  void *a = nullptr, *b = nullptr, *c = nullptr, *alt = nullptr;
  if (P.opt.method == PDEHIP_MGPCG && (p32 == P.mixed)) {
    a = p32 ? P.z32.get() : P.z.get();      // V-cycle correction (smoother x)
    c = p32 ? P.r32.get() : P.r.get();      // V-cycle right-hand side
    b = P.L[0].r.get();                     // level-0 residual
    alt = P.L[0].alt.get();
  }
  if (!a || !b || !c) {
    ta.alloc(P.ctx, n * es); tb.alloc(P.ctx, n * es); tc.alloc(P.ctx, n * es);
    a = ta.get(); b = tb.get(); c = tc.get();
    fill(s, pr, c, n, 0.0);
  }
  if (!alt) alt = b;
  fill(s, pr, a, n, 1.0);
  auto once = [&]() {
    switch (kernel) {
      case 0: apply(L, op, pr, a, b, nullptr); break;
      case 1: rbgs(L, op, pr, c, a, 0); rbgs(L, op, pr, c, a, 1); break;
      case 2: residual(L, op, pr, c, a, b, nullptr); break;
      case 4: rbgs_sweep(L, op, pr, c, a, alt, 0, false); break;
      default: convert(s, pr, a, pr, b, n, 1.0); break;
    }
  };
  switch (kernel) {
    case 0: *bytes = (double)n * (2 * es + kb); break;          // read x (+kappa), write y
    case 1:
    case 4: *bytes = (double)n * (3 * es + kb); break;          // read u, b (+kappa), write u
    case 2: *bytes = (double)n * (3 * es + kb); break;          // read b, x (+kappa), write r
    default: *bytes = (double)n * 2 * es; break;                // read + write
  }
  once();
  HIPCHK(hipStreamSynchronize(s));
  HIPCHK(hipEventRecord(P.ev0, s));
  for (int i = 0; i < reps; ++i) once();
  HIPCHK(hipEventRecord(P.ev1, s));
  HIPCHK(hipEventSynchronize(P.ev1));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, P.ev0, P.ev1));
  return ms / std::max(1, reps);
}

}  // namespace pdehip
