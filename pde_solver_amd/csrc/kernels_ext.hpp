// kernels_ext.hpp — launchers of k_time.hip (time stages, residual norms, lexicographic GS).
#pragma once

#include "kernels.hpp"

namespace pdehip {

void time_stage(hipStream_t s, const DevOp& op, const double* v, const double* un, double* acc,
                double* out, const double* f, double f_const, const NonlinearTerms& nl, double ut,
                double dt, double c1, int mode);
// One whole RK4 step (four stages and the closures that follow) in one launch, u -> out
// (2D; the same bits as the per-stage launches).  fused_steps: whether time series use it.
void time_rk4_step(hipStream_t s, const DevOp& op, const FaceData& fd, const double* u, double* out,
                   const double* f, double fc, const NonlinearTerms& nl, double ut, double dt);
void time_rk4_step(hipStream_t s, const DevOp& op, const FaceData& fd, const float* u, float* out,
                   const double* f, double fc, const NonlinearTerms& nl, double ut, double dt);
void set_fused_steps(bool on);
bool fused_steps();
void time_stage32(hipStream_t s, const DevOp& op, const float* v, const float* un, float* acc,
                  float* out, const double* f, double f_const, const NonlinearTerms& nl, double ut,
                  double dt, double c1, int mode);
void time_utt(hipStream_t s, const DevOp& op, const double* u, double* vel, double* next,
              const double* f, double f_const, const NonlinearTerms& nl, double ut, double utt,
              double dt);
int resnorm_blocks(const DevOp& op);
void residual_norms(hipStream_t s, const DevOp& op, const FaceData& fd, const double* u,
                    const double* f, double f_const, const NonlinearTerms& nl, double* sums,
                    double* maxs, int* degenerate);
void lex_gs(hipStream_t s, int nx, int ny, double ax, double by, double center, const double* b,
            double* u, int iters);

void build_rhs(hipStream_t s, const DevOp& op, const double* f, double f_const, const double* bnd,
               double* b);
void add_f32(hipStream_t s, double* x, const float* e, long long n, const double* scale, double hs);
void to_f32_scaled(hipStream_t s, const double* in, float* out, long long n, const double* norm);
// to_f32_scaled that also clears zero_out (n floats) and leaves vec_blocks(n) partials of
// out.out (the dot kernel's per-element products) in sums and of max |out| in maxs: the start of
// a mixed solve in one pass.
void to_f32_scaled_stats(hipStream_t s, const double* in, float* out, long long n,
                         const double* norm, float* zero_out, double* sums, double* maxs);
// Mixed CG with residual replacement (single problem).  `gate` is a scalars block whose done
// flag is clear exactly when a replacement is due: kernels launched with it (residual, the two
// below) run only then.
void to_f32_scaled_gated(hipStream_t s, const double* in, float* out, long long n,
                         const double* norm, const KScalars* gate);
// x64 += (*scale) x32, x32 = 0 (folds the FP32 correction accumulated since the last
// replacement into the FP64 solution).  skip_fresh: also skipped when gate->fresh is set (the
// correction was voided by a fused replacement).
void flush_f32(hipStream_t s, double* x64, float* x32, long long n, const double* scale,
               const KScalars* gate, bool skip_fresh = false);
// dst = src (n doubles) when so->xsel is set, else nothing.
void copy_if_xsel(hipStream_t s, double* dst, const double* src, long long n, const KScalars* so);
// gate->done = (so converged at a replacement: so->done, no breakdown in so or si),
// gate->breakdown = 0: kernels launched with the gate skip exactly then.
void fresh_gate(hipStream_t s, const KScalars* so, const KScalars* si, KScalars* gate);
// FIN_RU_START / FIN_RU_CHECK (tol = the replacement threshold delta) / FIN_RU_APPLY on the
// inner (si) and outer (so) scalars; gate->done = !(replacement due).  With maxp, the same
// launch then does fin_scale(maxp, nmax, si).
void finalize_ru(hipStream_t s, KScalars* si, KScalars* so, KScalars* gate,
                 const double* partials, int np, int op, double delta,
                 const double* maxp = nullptr, int nmax = 0);
int maxabs_blocks(long long n);
void maxabs32(hipStream_t s, const float* a, long long n, double* partials);
void fin_scale(hipStream_t s, const double* partials, int np, KScalars* sc);

}  // namespace pdehip
