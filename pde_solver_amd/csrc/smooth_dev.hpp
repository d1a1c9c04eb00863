// smooth_dev.hpp — per-node pieces of the red-black Gauss-Seidel smoothers shared by the
// plane-march sweeps (k_smooth.hip) and the box sweeps (k_box.hip): element types per level
// precision, loads and rounding to the storage type, and the smoother weights of an interior
// row.  One definition, so every sweep kernel forms each update with the same operations and
// the kernels agree bit for bit.
#pragma once

#include "kernels.hpp"

namespace pdehip {
namespace {

template <Prec P> struct PT;
template <> struct PT<P64> { using T = double; using S = double; };
template <> struct PT<P32> { using T = float; using S = float; };
template <> struct PT<P16> { using T = float; using S = __half; };

template <typename S> __device__ __forceinline__ auto ldS(const S* p, long long i);
template <> __device__ __forceinline__ auto ldS<double>(const double* p, long long i) { return p[i]; }
template <> __device__ __forceinline__ auto ldS<float>(const float* p, long long i) { return p[i]; }
template <> __device__ __forceinline__ auto ldS<__half>(const __half* p, long long i) { return __half2float(p[i]); }

// Value as the in-place sweep would read it back: rounded to the storage type.
template <typename S, typename T> __device__ __forceinline__ T rounded(T v);
template <> __device__ __forceinline__ double rounded<double, double>(double v) { return v; }
template <> __device__ __forceinline__ float rounded<float, float>(float v) { return v; }
template <> __device__ __forceinline__ float rounded<__half, float>(float v) { return __half2float(__float2half(v)); }

template <typename S, typename T> __device__ __forceinline__ S to_store(T v);
template <> __device__ __forceinline__ double to_store<double, double>(double v) { return v; }
template <> __device__ __forceinline__ float to_store<float, float>(float v) { return v; }
template <> __device__ __forceinline__ __half to_store<__half, float>(float v) { return __float2half(v); }

__device__ __forceinline__ bool skip(const KScalars* sc) { return sc && (sc->done | sc->breakdown); }

// Smoother weights of an interior row from face coefficients already at hand (FLUX: W, E, S,
// N, B, F faces computed by face_smooth), or from the operator (CONST, POINTWISE).  The sum and
// the centre follow row_weights<..., SMOOTH=true> operation for operation.
template <int DIM, int KIND, typename T>
__device__ __forceinline__ W7<T> weights_at(const DevOp& op, long long id, T fw, T fe, T fs,
                                            T fn, T fb, T ff) {
  W7<T> w;
  if constexpr (KIND == PDEHIP_OP_CONST) {
    w.C = T(op.w[0]); w.W = T(op.w[1]); w.E = T(op.w[2]); w.S = T(op.w[3]); w.N = T(op.w[4]);
    w.B = T(op.w[5]); w.F = T(op.w[6]);
  } else if constexpr (KIND == PDEHIP_OP_POINTWISE) {
    const double* p = op.pw;
    const long long n = op.n;
    w.C = T(p[id]); w.W = T(p[n + id]); w.E = T(p[2 * n + id]); w.S = T(p[3 * n + id]);
    w.N = T(p[4 * n + id]);
    if constexpr (DIM == 3) { w.B = T(p[5 * n + id]); w.F = T(p[6 * n + id]); }
    else { w.B = T(0); w.F = T(0); }
  } else {
    w.W = fw; w.E = fe; w.S = fs; w.N = fn;
    T sum = w.W + w.E + w.S + w.N;
    if constexpr (DIM == 3) {
      w.B = fb; w.F = ff;
      sum += w.B + w.F;
    } else {
      w.B = T(0); w.F = T(0);
    }
    w.C = T(op.e) - sum;
  }
  return w;
}

// Tile index with XCD-contiguous numbering: dispatch slot b goes to XCD b % 8; tile
// (b % 8) * per + b / 8 gives XCD x the tiles [x*per, (x+1)*per).  gridDim.x = 8 * per.
__device__ __forceinline__ int xcd_tile() {
  const int per = static_cast<int>(gridDim.x >> 3);
  return static_cast<int>(blockIdx.x & 7) * per + static_cast<int>(blockIdx.x >> 3);
}

}  // namespace
}  // namespace pdehip
