// noalloc.hip — pdehip_residual_norms_noalloc: ComputeResidualNorms (residual.cpp:153-583) on
// device-resident inputs with no device allocation.  Every buffer the call needs beyond the
// caller's inputs and outputs lives in one file-scope static device array, `scratch`:
//
//   [0, kParts)            block partial sums of r^2
//   [kParts, 2 kParts)     block partial maxima of |r|
//   [2 kParts]             degenerate-row flag (int)
//   [2 kParts + 8, +2)     the reduced |r|_2^2 and |r|_inf
//   [kFaceOff, ...)        2D boundary data sampled per face node: 4 faces x {value, alpha,
//                          beta, gamma} x kMaxFace doubles
//
// The array's device address is looked up once per device (hipGetSymbolAddress, at context
// creation with that device current) and only that address is handed to kernels and copies.
// Calls that overlap on one device share the array and race; the array keeps the last call's
// contents, and each call clears what it accumulates into.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>

#include "runtime.hpp"

namespace {

constexpr int kParts = 4096;                 // resnorm_blocks() never exceeds this (kMaxPartials)
constexpr int kMaxFace = 16385;              // 2D: longest grid side with face data in scratch
constexpr size_t kFaceOff = 2 * kParts + 16;
constexpr size_t kScratchDoubles = kFaceOff + 16 * (size_t)kMaxFace;
constexpr int kMaxDevices = 64;

}  // namespace

// 2.16 MB per device, held until the process ends.
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace pdehip {
namespace {

std::mutex g_mu;
double* g_ptr[kMaxDevices] = {};

// c + cx x + cy y + cz z at the nodes of one 2D face, with the host's rounding (no FMA
// contraction): the existing entry samples the same expressions on the host.
__global__ void k_sample_face2(double* out, int m, int face, double c, double cx, double cy,
                               double cz, double xmin, double xmax, double ymin, double ymax,
                               double dx, double dy) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= m) return;
  double x, y;
  if (face < 2) { x = face == 0 ? xmin : xmax; y = __dadd_rn(ymin, __dmul_rn(q, dy)); }
  else { x = __dadd_rn(xmin, __dmul_rn(q, dx)); y = face == 2 ? ymin : ymax; }
  const double z = 0.0;
  out[q] = __dadd_rn(__dadd_rn(__dadd_rn(c, __dmul_rn(cx, x)), __dmul_rn(cy, y)), __dmul_rn(cz, z));
}

__global__ void k_clear_flag(int* flag) {
  if (threadIdx.x == 0) *flag = 0;
}

// One thread sums the block partials in index order (the existing entry's host loop order),
// so the two entries agree bit for bit.
__global__ void k_reduce_parts(const double* sums, const double* maxs, int nb, double* out) {
  if (threadIdx.x != 0) return;
  double a = 0.0, m = 0.0;
  for (int i = 0; i < nb; ++i) {
    a += sums[i];
    m = fmax(m, maxs[i]);
  }
  out[0] = a;
  out[1] = m;
}

double* device_scratch(int device) {
  std::lock_guard<std::mutex> g(g_mu);
  if (device < 0 || device >= kMaxDevices || !g_ptr[device])
    throw Error(PDEHIP_ERR_DEVICE, "scratch array not initialised for this device");
  return g_ptr[device];
}

}  // namespace

void noalloc_init(int device) {
  std::lock_guard<std::mutex> g(g_mu);
  if (device < 0 || device >= kMaxDevices)
    throw Error(PDEHIP_ERR_DEVICE, "device index out of range for the scratch table");
  if (g_ptr[device]) return;
  void* p = nullptr;
  HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  g_ptr[device] = static_cast<double*>(p);
}

void residual_norms_noalloc(Ctx* ctx, const pdehip_problem& p, const double* u, double* l2,
                            double* linf) {
  if (p.nx < 3 || p.ny < 3 || (p.nz > 1 && p.nz < 3))
    throw Error(PDEHIP_ERR_INVALID, "grid must be at least 3 nodes per axis");
  if (p.xmax <= p.xmin || p.ymax <= p.ymin || (p.nz > 1 && p.zmax <= p.zmin))
    throw Error(PDEHIP_ERR_INVALID, "domain bounds are invalid");
  if (p.op_kind == PDEHIP_OP_POINTWISE)
    throw Error(PDEHIP_ERR_UNSUPPORTED,
                "the allocation-free residual norms take CONST or FLUX operators (pointwise "
                "weights are assembled on the host)");
  if (p.op_kind == PDEHIP_OP_FLUX && !p.kappa)
    throw Error(PDEHIP_ERR_INVALID, "FLUX operator needs a kappa field");
  const Geo g = make_geo(p);
  if (!g.is3d && (g.nx > kMaxFace || g.ny > kMaxFace))
    throw Error(PDEHIP_ERR_INVALID, "allocation-free residual norms: a 2D grid side exceeds " +
                                        std::to_string(kMaxFace) + " nodes (scratch limit)");
  double* base = device_scratch(ctx->device);
  const hipStream_t s = ctx->stream;

  // operator: the reference rows (no folded faces), as the allocating entry builds it
  DevOp op{};
  op.nx = g.nx; op.ny = g.ny; op.nz = g.nz;
  op.kind = p.op_kind;
  op.sxy = g.sxy();
  op.n = g.n;
  for (double& v : op.phi) v = 1.0;
  op.e = p.e;
  if (p.op_kind == PDEHIP_OP_CONST) {
    const_weights(p, g, op.w);
  } else {
    op.kappa = p.kappa;
    op.fx = p.a / (g.dx * g.dx);
    op.fy = p.a / (g.dy * g.dy);
    op.fz = g.is3d ? p.a / (g.dz * g.dz) : 0.0;
  }
  // boundary data (2D boundary rows; 3D norms cover interior rows only)
  FaceData fd{};
  fd.convention = p.bc_convention;
  for (int f = 0; f < 6; ++f) {
    fd.kind[f] = (f < 4 || g.is3d) ? p.bc[f].kind : PDEHIP_BC_DIRICHLET;
    fd.h[f] = f < 2 ? g.dx : (f < 4 ? g.dy : g.dz);
  }
  if (!g.is3d) {
    for (int f = 0; f < 4; ++f) {
      const int m = f < 2 ? g.ny : g.nx;
      const pdehip_expr* ex[4] = {&p.bc[f].value, &p.bc[f].alpha, &p.bc[f].beta, &p.bc[f].gamma};
      const double** dst[4] = {&fd.value[f], &fd.alpha[f], &fd.beta[f], &fd.gamma[f]};
      for (int t = 0; t < 4; ++t) {
        if (ex[t]->samples) { *dst[t] = ex[t]->samples; continue; }   // caller's device array
        double* out = base + kFaceOff + (size_t)(4 * f + t) * kMaxFace;
        hipLaunchKernelGGL(k_sample_face2, dim3((m + 255) / 256), dim3(256), 0, s, out, m, f,
                           ex[t]->c, ex[t]->cx, ex[t]->cy, ex[t]->cz, g.xmin, g.xmax, g.ymin,
                           g.ymax, g.dx, g.dy);
        *dst[t] = out;
      }
    }
  }
  NonlinearTerms nl{};
  nl.n = std::max(0, std::min(p.n_nonlinear, PDEHIP_MAX_NONLINEAR));
  for (int t = 0; t < nl.n; ++t) nl.t[t] = p.nonlinear[t];

  double* sums = base;
  double* maxs = base + kParts;
  int* degenerate = reinterpret_cast<int*>(base + 2 * kParts);
  double* red = base + 2 * kParts + 8;
  const int nb = resnorm_blocks(op);
  if (nb > kParts) throw Error(PDEHIP_ERR_DEVICE, "residual partial count exceeds the scratch");
  hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(64), 0, s, degenerate);
  residual_norms(s, op, fd, u, p.f_field, p.f, nl, sums, maxs, degenerate);
  hipLaunchKernelGGL(k_reduce_parts, dim3(1), dim3(64), 0, s, sums, maxs, nb, red);
  HIPCHK(hipGetLastError());
  double host[2] = {0.0, 0.0};
  HIPCHK(hipMemcpyAsync(host, red, sizeof(host), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  *l2 = std::sqrt(std::max(0.0, host[0]));
  *linf = host[1];
}

}  // namespace pdehip
