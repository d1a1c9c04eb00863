// kernels_krylov.hpp — launchers of k_krylov.hip (BiCGStab, GMRES, Jacobi preconditioner).
#pragma once

#include "kernels.hpp"

namespace pdehip {

enum KFinOp { KFIN_BI_RHO = 0, KFIN_BI_ALPHA = 1, KFIN_BI_S = 2, KFIN_BI_OMEGA = 3, KFIN_BI_R = 4,
              KFIN_GM_R = 5 };

int krylov_blocks(long long n);
// Vector element type S: double (FP64 path) or float (mixed path).  x stays FP64; scale (device
// pointer or null) multiplies the x updates of the mixed path.
template <typename S> void bicg_xs(hipStream_t s, double* x, const S* ph, long long n, const KScalars* sc, const double* scale = nullptr);
template <typename S> void dot2(hipStream_t s, const S* a, const S* b, const S* c, long long n, double* p1, double* p2, const KScalars* sc);
template <typename S> void bicg_x(hipStream_t s, double* x, const S* ph, const S* sh, long long n, const KScalars* sc, const double* scale = nullptr);
// Mixed path: r = s - omega t with partials of r.r (the recurrence; the FP64 residual replaces it
// at every refinement).
void bicg_r(hipStream_t s, const float* sv, const float* t, float* r, long long n, double* partials, const KScalars* sc);
void bi_sdone(hipStream_t s, KScalars* sc);
void fin_k(hipStream_t s, KScalars* sc, const double* p1, const double* p2, int np, int op);
// bicg_p_pc: p = r + beta (p - omega v) (p = r on the first iteration), plus ph = D^-1 p when pc;
// bicg_s_pc: s = r - alpha v with partials of s.s, plus sh = D^-1 s when pc.
template <typename S> void bicg_p_pc(hipStream_t s, const DevOp& op, const S* r, S* p, const S* v, S* ph, bool pc, const KScalars* sc);
template <typename S> void bicg_s_pc(hipStream_t s, const DevOp& op, const S* r, const S* v, S* sv, S* sh, bool pc, double* partials, const KScalars* sc);
template <typename S> void jacobi_pc(hipStream_t s, const DevOp& op, const S* in, S* out, const KScalars* sc);
template <typename S> void multidot(hipStream_t s, const S* V, long long ld, int cnt, const S* w, long long n, double* partials, int nb, const KScalars* sc);
template <typename S> void multiaxpy(hipStream_t s, const S* V, long long ld, int cnt, const double* h, S* w, long long n, double sign, int check_cycle, const KScalars* sc);
void fin_multi(hipStream_t s, const double* partials, int nb, int cnt, double* hcol, int add, const KScalars* sc);
template <typename S> void gm_normalize(hipStream_t s, S* vnext, const S* w, long long n, const double* hval, const KScalars* sc);
void fin_norm(hipStream_t s, const double* partials, int nb, double* out, const KScalars* sc);
void givens(hipStream_t s, double* H, double* cs, double* sn, double* g, int j, int m, KScalars* sc);
void backsub(hipStream_t s, const double* H, const double* g, double* y, int m, KScalars* sc);
// hc holds two staged columns of cnt coefficients, `stride` doubles apart.
void gm_setcol(hipStream_t s, double* H, const double* hc, int stride, int j, int m, int cnt,
               const KScalars* sc);
template <typename S> void gm_start(hipStream_t s, S* v0, const double* r, long long n, double* g, int m, KScalars* sc);

}  // namespace pdehip
