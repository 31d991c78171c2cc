// Streaming vector kernels for the Krylov recurrences (any basis).
//
// The Lanczos recurrence keeps unnormalised vectors u_j with norms
// nrm[j] = |u_j| in device memory, so no step waits on the host:
//   apply step j : w       = H (u_j / nrm[j]) - (nrm[j] / nrm[j-1]) u_{j-1}
//                  alpha_j = <q_j, w>                       (fused reduction)
//   orth step j  : u_{j+1} = w - (alpha_j / nrm[j]) u_j,  nrm[j+1] = |u_{j+1}|
// beta_j = nrm[j+1].  The replay pass re-runs the same kernels with the
// stored alpha/nrm and accumulates Ritz vectors x_k += (y_k[j]/nrm[j]) u_j.
#pragma once
#include "common.hpp"

namespace ryd {

constexpr int STREAM_THREADS = 256;
#ifndef RYD_LZ_UNROLL
#define RYD_LZ_UNROLL 4
#endif
constexpr int LZ_UNROLL = RYD_LZ_UNROLL;

template <class V>
__global__ __launch_bounds__(256) void k_lz_orth(const V* __restrict__ uj, V* __restrict__ w,
                                                 int64_t D, const double* alpha, const double* nrm,
                                                 int j, V* ritz0, V* ritz1, const double* yc0,
                                                 const double* yc1, const int* stop, int reduce,
                                                 Reduce rd) {
  using RT = typename VT<V>::R;
  if (stop && load_coherent(stop)) return;
  const double inv = 1.0 / nrm[j];
  const RT coef = (RT)(-alpha[j] * inv);
  const RT r0c = ritz0 ? (RT)(yc0[j] * inv) : (RT)0;
  const RT r1c = ritz1 ? (RT)(yc1[j] * inv) : (RT)0;
  double acc = 0.0;
  // LZ_UNROLL elements per thread and pass, every load issued before the
  // first store (more reads in flight per CU than one element at a time)
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i0 < D;
       i0 += stride * LZ_UNROLL) {
    V u[LZ_UNROLL], v[LZ_UNROLL];
#pragma unroll
    for (int k = 0; k < LZ_UNROLL; ++k) {
      const int64_t i = i0 + k * stride;
      u[k] = i < D ? uj[i] : vzero(u[0]);
      v[k] = i < D ? w[i] : vzero(u[0]);
    }
#pragma unroll
    for (int k = 0; k < LZ_UNROLL; ++k) {
      const int64_t i = i0 + k * stride;
      if (i >= D) break;
      vfma(v[k], coef, u[k]);
      w[i] = v[k];
      acc += nrm2(v[k]);
      if (ritz0) { V a = ritz0[i]; vfma(a, r0c, u[k]); ritz0[i] = a; }
      if (ritz1) { V a = ritz1[i]; vfma(a, r1c, u[k]); ritz1[i] = a; }
    }
  }
  if (reduce) block_reduce_store(0.0, acc, 0.0, 0.0, rd.partials);
}

// Ritz vectors from a stored Krylov basis: out_k (+)= Σ_j (y_k[j] / nrm[j0+j]) u_{j0+j}
// over `count` rows of `basis` (row stride ld); coefficients precomputed per
// block in LDS (count <= RITZ_MAX).
constexpr int RITZ_MAX = 2048;
template <class V>
__global__ __launch_bounds__(256) void k_ritz_stored(const V* __restrict__ basis, int64_t ld,
                                                     int64_t dim, int count, const double* y,
                                                     int two, const double* nrm, int64_t j0,
                                                     int accumulate, V* __restrict__ out0,
                                                     V* __restrict__ out1) {
  using RT = typename VT<V>::R;
  __shared__ RT c0[RITZ_MAX], c1[RITZ_MAX];
  for (int j = threadIdx.x; j < count; j += blockDim.x) {
    const double inv = 1.0 / nrm[j0 + j];
    c0[j] = (RT)(y[j] * inv);
    c1[j] = two ? (RT)(y[RITZ_MAX + j] * inv) : (RT)0;
  }
  __syncthreads();
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < dim; i += stride) {
    V a0 = accumulate ? out0[i] : vzero(basis[i]);
    V a1 = vzero(basis[i]);
    if (two) a1 = accumulate ? out1[i] : a1;
#pragma unroll 4
    for (int j = 0; j < count; ++j) {
      const V u = basis[(int64_t)j * ld + i];
      vfma(a0, c0[j], u);
      if (two) vfma(a1, c1[j], u);
    }
    out0[i] = a0;
    if (two) out1[i] = a1;
  }
}

// The same combination for small dims (a few thousand elements, where one
// thread per element walking every row is bound by load latency): a block is
// RITZ_E consecutive elements × RITZ_S row slices (slice s takes rows
// s, s + RITZ_S, ...), the slices' sums added in LDS in slice order.
constexpr int RITZ_E = 16, RITZ_S = 16;
constexpr int64_t RITZ_SLICED_MAX_DIM = 1 << 20;
template <class V>
__global__ __launch_bounds__(256) void k_ritz_stored_sliced(const V* __restrict__ basis,
                                                            int64_t ld, int64_t dim, int count,
                                                            const double* y, int two,
                                                            const double* nrm, int64_t j0,
                                                            int accumulate, V* __restrict__ out0,
                                                            V* __restrict__ out1) {
  using RT = typename VT<V>::R;
  static_assert(RITZ_E * RITZ_S == 256, "one thread per element and slice");
  __shared__ RT c0[RITZ_MAX], c1[RITZ_MAX];
  __shared__ V p0[RITZ_S][RITZ_E], p1[RITZ_S][RITZ_E];
  for (int j = threadIdx.x; j < count; j += blockDim.x) {
    const double inv = 1.0 / nrm[j0 + j];
    c0[j] = (RT)(y[j] * inv);
    c1[j] = two ? (RT)(y[RITZ_MAX + j] * inv) : (RT)0;
  }
  __syncthreads();
  const int e = threadIdx.x % RITZ_E, sl = threadIdx.x / RITZ_E;
  const int64_t i = (int64_t)blockIdx.x * RITZ_E + e;
  V a0 = vzero(V{}), a1 = vzero(V{});
  if (i < dim) {
#pragma unroll 4
    for (int j = sl; j < count; j += RITZ_S) {
      const V u = basis[(int64_t)j * ld + i];
      vfma(a0, c0[j], u);
      if (two) vfma(a1, c1[j], u);
    }
  }
  p0[sl][e] = a0;
  p1[sl][e] = a1;
  __syncthreads();
  if (sl == 0 && i < dim) {
    V r0 = accumulate ? out0[i] : vzero(V{});
    V r1 = (two && accumulate) ? out1[i] : vzero(V{});
    for (int k = 0; k < RITZ_S; ++k) {
      vfma(r0, (RT)1, p0[k][e]);
      if (two) vfma(r1, (RT)1, p1[k][e]);
    }
    out0[i] = r0;
    if (two) out1[i] = r1;
  }
}

// sums: <a, b>, |b|^2, |a|^2 ; used for norms and Rayleigh quotients
template <class V>
__global__ __launch_bounds__(256) void k_dots(const V* __restrict__ a, const V* __restrict__ b,
                                              int64_t D, Reduce rd) {
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < D; i += stride) {
    const V bv = b[i];
    s1 += nrm2(bv);
    if (a) { const V av = a[i]; s0 += rdot(av, bv); s2 += nrm2(av); }
  }
  block_reduce_store(s0, s1, s2, 0.0, rd.partials);
}

// y = a * x + b * y  (real a, b); x may be null (scale only)
template <class V>
__global__ __launch_bounds__(256) void k_axpby(const V* __restrict__ x, V* __restrict__ y,
                                               int64_t D, double a, double b) {
  using RT = typename VT<V>::R;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < D; i += stride) {
    V v = vscale(y[i], (RT)b);
    if (x) vfma(v, (RT)a, x[i]);
    y[i] = v;
  }
}

// precision casts between state-vector types of the same kind
template <class A, class B>
__global__ __launch_bounds__(256) void k_cast(const A* __restrict__ x, B* __restrict__ y, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    y[i] = (B)x[i];
}

// probe bits -> ±1 vector (real part for complex element types)
template <class V>
__global__ __launch_bounds__(256) void k_probe_expand(const uint32_t* __restrict__ bits,
                                                      V* __restrict__ y, int64_t D) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < D; i += stride)
    y[i] = vfrom_real<V>(((bits[i >> 5] >> (i & 31)) & 1u) ? 1.0 : -1.0);
}

}  // namespace ryd
