// Lanczos with full reorthogonalisation over a stored basis of unit rows
// q_0 .. q_j (row stride ld), every coefficient in device memory:
//   w        = H q_j                       (H·v launch, alpha_j = Re<q_j, w>)
//   w       -= alpha_j q_j + beta_{j-1} q_{j-1}
//   w       -= Q (Q^H w)   twice           (classical Gram-Schmidt, two passes)
//   beta_j   = |w|,  q_{j+1} = w / beta_j
// (the recurrence of tomisnot/energy-level src/core/manybody/spectrum.py:74-95,
// whose two sweeps are modified Gram-Schmidt; both reach orthogonality to
// rounding, so alpha/beta agree to rounding).  Q^H w is one launch over
// element chunks (each block takes a chunk of w and every row, partial sums
// in fixed order) and a finishing launch per row; w -= Q c keeps c in LDS.
#pragma once
#include "common.hpp"
#include "measure.hpp"   // cdot_acc: conj(a)·b accumulated in FP64

namespace ryd {

constexpr int GS_CHUNK = 4096;        // elements of w per block of k_gs_dots
constexpr int GS_MAX_ROWS = 2048;     // rows per k_gs_axpy launch (coefficients in LDS)
constexpr int GS_GROUP = 8;           // rows reduced together per barrier

// w -= (cr + i ci) q
__device__ __forceinline__ void csub(float& w, double cr, double, float q) { w -= (float)cr * q; }
__device__ __forceinline__ void csub(double& w, double cr, double, double q) { w -= cr * q; }
__device__ __forceinline__ void csub(c32& w, double cr, double ci, c32 q) {
  w.re -= (float)cr * q.re - (float)ci * q.im;
  w.im -= (float)cr * q.im + (float)ci * q.re;
}
__device__ __forceinline__ void csub(c64& w, double cr, double ci, c64 q) {
  w.re -= cr * q.re - ci * q.im;
  w.im -= cr * q.im + ci * q.re;
}

// partials[(row * nchunk + chunk) * 2 + {0, 1}] = Σ_{e in chunk} conj(q_row[e]) w[e]
template <class V>
__global__ __launch_bounds__(256) void k_gs_dots(const V* __restrict__ basis, int64_t ld,
                                                 int64_t dim, int count, const V* __restrict__ w,
                                                 double* __restrict__ partials, const int* stop) {
  __shared__ double red[4][2 * GS_GROUP];
  if (stop && load_coherent(stop)) return;
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  const int64_t nchunk = gridDim.x;
  const int64_t e0 = (int64_t)blockIdx.x * GS_CHUNK;
  constexpr int PER = GS_CHUNK / 256;
  V wv[PER];
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    const int64_t e = e0 + p * 256 + t;
    wv[p] = e < dim ? w[e] : V();
  }
  for (int r0 = 0; r0 < count; r0 += GS_GROUP) {
    double re[GS_GROUP], im[GS_GROUP];
#pragma unroll
    for (int g = 0; g < GS_GROUP; ++g) { re[g] = 0.0; im[g] = 0.0; }
#pragma unroll
    for (int g = 0; g < GS_GROUP; ++g) {
      const int r = r0 + g;
      if (r >= count) break;
      const V* q = basis + (int64_t)r * ld;
#pragma unroll
      for (int p = 0; p < PER; ++p) {
        const int64_t e = e0 + p * 256 + t;
        if (e < dim) cdot_acc(q[e], wv[p], re[g], im[g]);
      }
    }
#pragma unroll
    for (int g = 0; g < GS_GROUP; ++g) {
      re[g] = wave_sum(re[g]);
      im[g] = wave_sum(im[g]);
    }
    if (lane == 0) {
#pragma unroll
      for (int g = 0; g < GS_GROUP; ++g) { red[wid][2 * g] = re[g]; red[wid][2 * g + 1] = im[g]; }
    }
    __syncthreads();
    if (t < 2 * GS_GROUP) {
      const int g = t >> 1, r = r0 + g;
      if (r < count)
        partials[((int64_t)r * nchunk + blockIdx.x) * 2 + (t & 1)] =
            red[0][t] + red[1][t] + red[2][t] + red[3][t];
    }
    __syncthreads();
  }
}

// coef[row * 2 + {0, 1}] = Σ_chunk partials (fixed order), one block per row
__global__ __launch_bounds__(256) void k_gs_finish(const double* __restrict__ partials,
                                                   int64_t nchunk, double* __restrict__ coef,
                                                   const int* stop) {
  __shared__ double red[4][2];
  if (stop && load_coherent(stop)) return;
  const int r = blockIdx.x, t = threadIdx.x, lane = t & 63, wid = t >> 6;
  double re = 0.0, im = 0.0;
  for (int64_t c = t; c < nchunk; c += blockDim.x) {
    re += partials[((int64_t)r * nchunk + c) * 2];
    im += partials[((int64_t)r * nchunk + c) * 2 + 1];
  }
  re = wave_sum(re);
  im = wave_sum(im);
  if (lane == 0) { red[wid][0] = re; red[wid][1] = im; }
  __syncthreads();
  if (t == 0) {
    coef[2 * r] = red[0][0] + red[1][0] + red[2][0] + red[3][0];
    coef[2 * r + 1] = red[0][1] + red[1][1] + red[2][1] + red[3][1];
  }
}

// w -= Σ_{r < count} coef_r q_r   (count <= GS_MAX_ROWS)
template <class V>
__global__ __launch_bounds__(256) void k_gs_axpy(const V* __restrict__ basis, int64_t ld,
                                                 int64_t dim, int count,
                                                 const double* __restrict__ coef,
                                                 V* __restrict__ w, const int* stop) {
  __shared__ double cs[2 * GS_MAX_ROWS];
  if (stop && load_coherent(stop)) return;
  for (int i = threadIdx.x; i < 2 * count; i += blockDim.x) cs[i] = coef[i];
  __syncthreads();
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < dim; e += stride) {
    V v = w[e];
    for (int r = 0; r < count; ++r) csub(v, cs[2 * r], cs[2 * r + 1], basis[(int64_t)r * ld + e]);
    w[e] = v;
  }
}

// w -= alpha_j q_j + beta_{j-1} q_{j-1}   (nrm[j] = beta_{j-1})
template <class V>
__global__ __launch_bounds__(256) void k_rl_three_term(V* __restrict__ w, const V* __restrict__ qj,
                                                       const V* __restrict__ qjm, int64_t dim,
                                                       const double* alpha, const double* nrm,
                                                       int j, const int* stop) {
  using RT = typename VT<V>::R;
  if (stop && load_coherent(stop)) return;
  const RT a = (RT)(-alpha[j]);
  const RT b = j > 0 ? (RT)(-nrm[j]) : (RT)0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < dim; e += stride) {
    V v = w[e];
    vfma(v, a, qj[e]);
    if (j > 0) vfma(v, b, qjm[e]);
    w[e] = v;
  }
}

// out = in / nrm[j + 1]
template <class V>
__global__ __launch_bounds__(256) void k_rl_scale(V* __restrict__ out, const V* __restrict__ in,
                                                  int64_t dim, const double* nrm, int j,
                                                  const int* stop) {
  using RT = typename VT<V>::R;
  if (stop && load_coherent(stop)) return;
  const RT s = (RT)(1.0 / nrm[j + 1]);
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < dim; e += stride)
    out[e] = vscale(in[e], s);
}

// x *= 1/sqrt(s2[1]) (s2[1] = |x|^2 from a k_dots reduction); left as is
// when the norm is zero
template <class V>
__global__ __launch_bounds__(256) void k_scale_rsqrt(V* __restrict__ x, int64_t D,
                                                     const double* __restrict__ s2) {
  using RT = typename VT<V>::R;
  const double q = s2[1];
  if (!(q > 0.0)) return;
  const RT f = (RT)(1.0 / sqrt(q));
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < D; i += stride)
    x[i] = vscale(x[i], f);
}

// Locked orth step: u_pre = w − (alpha_j / nrm_j) u_j in place of w, with the
// sums {Re<g, u_pre>, |u_pre|^2, Im<g, u_pre>} (POST_LZ_LOCK then forms
// γ = <g, u_pre> and the norm of u_pre − γ g); the replay skips the sums and
// accumulates the Ritz vectors from u_j as k_lz_orth does.
template <class V>
__global__ __launch_bounds__(256) void k_lz_orth_lock(const V* __restrict__ uj, V* __restrict__ w,
                                                      const V* __restrict__ g, int64_t D,
                                                      const double* alpha, const double* nrm,
                                                      int j, V* ritz0, V* ritz1,
                                                      const double* yc0, const double* yc1,
                                                      const int* stop, int reduce, Reduce rd) {
  using RT = typename VT<V>::R;
  if (stop && load_coherent(stop)) return;
  const double inv = 1.0 / nrm[j];
  const RT coef = (RT)(-alpha[j] * inv);
  const RT r0c = ritz0 ? (RT)(yc0[j] * inv) : (RT)0;
  const RT r1c = ritz1 ? (RT)(yc1[j] * inv) : (RT)0;
  double nn = 0.0, gre = 0.0, gim = 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < D; i += stride) {
    const V u = uj[i];
    V v = w[i];
    vfma(v, coef, u);
    w[i] = v;
    if (reduce) {
      nn += nrm2(v);
      cdot_acc(g[i], v, gre, gim);
    }
    if (ritz0) { V a = ritz0[i]; vfma(a, r0c, u); ritz0[i] = a; }
    if (ritz1) { V a = ritz1[i]; vfma(a, r1c, u); ritz1[i] = a; }
  }
  if (reduce) block_reduce_store(gre, nn, gim, 0.0, rd.partials);
}

// u_{j+1} = u_pre − γ g, γ = gam[0] + i·gam[1] (real types use gam[0])
template <class V>
__global__ __launch_bounds__(256) void k_lz_unlock(V* __restrict__ w, const V* __restrict__ g,
                                                   int64_t D, const double* gam, const int* stop) {
  if (stop && load_coherent(stop)) return;
  const double gr = gam[0], gi = gam[1];
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < D; i += stride) {
    V v = w[i];
    csub(v, gr, gi, g[i]);
    w[i] = v;
  }
}

}  // namespace ryd
