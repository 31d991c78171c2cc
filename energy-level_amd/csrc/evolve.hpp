// Piecewise-constant evolution on the full basis: the fused Chebyshev step of
// the forward propagator and of its adjoint (Clenshaw) recurrence.
//
//   H(Ω, Δ) = om · F + D + dadd · N
// with F the plan's flip operator (amplitudes Ω_site/2), D the plan's diagonal
// (delta_site n_i + V n_i n_j), N = Σ_i n_i and per-slice scalars om, dadd.
// One launch computes, for x~ = s·x,
//   y    = ch·(om·F x~ + D x~ + dadd·N x~) + cx·x~ + cz·z + cw·w
//   acc += ca·y                                   (complex cw, ca; optional)
//   sums = {Re<q, F x~>, Re<q, N x~>, |y|^2, Re<q, y>}   (optional q)
// Forward slice: t_{j+1} = 2H~t_j − t_{j−1}, ψ_out = Σ a_j t_j (acc).
// Adjoint slice: b_j = conj(a_j) g + 2H~ b_{j+1} − b_{j+2} with q = t_j gives
// the gradient dots <F b_{j+1}, t_j> and <N b_{j+1}, t_j>.
// Same tiling as k_full_tiled (single pass, n <= 22), or one element per
// thread below the tile size.
#pragma once
#include "common.hpp"
#include "hv_full.hpp"

namespace ryd {

template <class V>
struct EvoIO {
  const V* x;
  const V* z;
  const V* w;
  V* y;
  V* acc;
  const V* q;
  double om, dadd, s, ch, cx, cz;
  double cw_re, cw_im, ca_re, ca_im;
  int reduce;
};

template <class V>
__device__ __forceinline__ void evo_finish(const EvoIO<V>& io, uint64_t g, V fl, V xe, double d,
                                           double pop, double& s0, double& s1, double& s2,
                                           double& s3) {
  using RT = typename VT<V>::R;
  V hx = vscale(fl, (RT)io.om);
  vfma(hx, (RT)(d + io.dadd * pop), xe);
  V v = vscale(hx, (RT)io.ch);
  vfma(v, (RT)io.cx, xe);
  if (io.z) vfma(v, (RT)io.cz, io.z[g]);
  if (io.w) vcfma(v, (RT)io.cw_re, (RT)io.cw_im, io.w[g]);
  io.y[g] = v;
  if (io.acc) {
    V a = io.acc[g];
    vcfma(a, (RT)io.ca_re, (RT)io.ca_im, v);
    io.acc[g] = a;
  }
  if (io.reduce) {
    s2 += nrm2(v);
    if (io.q) {
      const V qv = io.q[g];
      s0 += rdot(qv, fl);
      s1 += pop * rdot(qv, xe);
      s3 += rdot(qv, v);
    }
  }
}

#ifndef RYD_EVO_FAR_BATCH
#define RYD_EVO_FAR_BATCH 8
#endif

// RR r-groups per block: RR = tile_r (one tile per block, bits 10.. of the
// tile in registers) or RR = 1 (a tile spread over tile_r blocks, so a small
// vector still fills the chip; the tile's bits 10.. then come from the other
// groups through L2 like the far bits).
template <class V, int RR>
__global__ __launch_bounds__(256) void k_evo_tiled(FullModel M, EvoIO<V> io, Reduce rd) {
  using RT = typename VT<V>::R;
  constexpr int R = tile_r<V>();
  constexpr int TB = tile_bits<V>();
  constexpr int SUB = R / RR;          // blocks per tile
  static_assert(RR == R || RR == 1, "a block holds one group or the whole tile");
  __shared__ V lds[RR * 1024];
  __shared__ double Fx[16];
  __shared__ double T0[256];
  __shared__ double T1[16];
  const RT sr = (RT)io.s;
  const int t = threadIdx.x;
  const int nhi = M.n - TB;
  const uint64_t nblk = (1ull << nhi) * SUB;
  const uint64_t b = blockIdx.x;
  const uint64_t blk = nblk >= 8 ? ((b & 7ull) * (nblk >> 3) + (b >> 3)) : b;
  const uint64_t tile = blk / SUB;
  const int r0 = (int)(blk % SUB) * RR;
  const uint64_t base = tile << TB;
  V xs[RR][4];
#pragma unroll
  for (int rr = 0; rr < RR; ++rr) {
    V4<V> qv = *reinterpret_cast<const V4<V>*>(io.x + base + (r0 + rr) * 1024 + t * 4);
#pragma unroll
    for (int c = 0; c < 4; ++c) xs[rr][c] = vscale(qv.v[c], sr);
    V4<V> wv;
#pragma unroll
    for (int c = 0; c < 4; ++c) wv.v[c] = xs[rr][c];
    *reinterpret_cast<V4<V>*>(lds + rr * 1024 + t * 4) = wv;
  }
  if (t < TB) {
    double f = 0.0;
    const double* row = M.vb + (size_t)t * M.n;
    for (int q = TB; q < M.n; ++q)
      if ((tile >> (q - TB)) & 1ull) f += row[q];
    Fx[t] = f;
  }
  __syncthreads();
  {
    double v = 0.0;
#pragma unroll
    for (int p = 0; p < 8; ++p)
      if ((t >> p) & 1) v += Fx[p];
    T0[t] = v;
    if (t < (1 << (TB - 8))) {
      double w = 0.0;
#pragma unroll
      for (int p = 8; p < TB; ++p)
        if ((t >> (p - 8)) & 1) w += Fx[p];
      T1[t] = w;
    }
  }
  const double eh = M.ehi[tile];
  const int pop_hi = __popcll(tile);
  __syncthreads();
  // flip sum (plan amplitudes, no diagonal)
  V fl[RR][4];
#pragma unroll
  for (int rr = 0; rr < RR; ++rr)
#pragma unroll
    for (int c = 0; c < 4; ++c) fl[rr][c] = vzero(xs[rr][c]);
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    const RT ar = (RT)M.ar[p], ai = (RT)M.ai[p];
#pragma unroll
    for (int rr = 0; rr < RR; ++rr)
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int bit = (c >> p) & 1;
        vcfma(fl[rr][c], ar, bit ? ai : -ai, xs[rr][c ^ (1 << p)]);
      }
  }
  // bits 10..TB-1: the tile's other groups (registers, or global for RR = 1)
#pragma unroll
  for (int k = 0; (1 << k) < R; ++k) {
    const int p = 10 + k;
    if constexpr (RR == R) {
      const RT ar = (RT)M.ar[p], ai = (RT)M.ai[p];
#pragma unroll
      for (int rr = 0; rr < RR; ++rr) {
        const int bit = (rr >> k) & 1;
#pragma unroll
        for (int c = 0; c < 4; ++c) vcfma(fl[rr][c], ar, bit ? ai : -ai, xs[rr ^ (1 << k)][c]);
      }
    } else {
      const int bit = (r0 >> k) & 1;
      const RT ar = (RT)(M.ar[p] * io.s);
      const RT aim = (RT)((bit ? M.ai[p] : -M.ai[p]) * io.s);
      V4<V> qv = *reinterpret_cast<const V4<V>*>(io.x + base + (r0 ^ (1 << k)) * 1024 + t * 4);
#pragma unroll
      for (int c = 0; c < 4; ++c) vcfma(fl[0][c], ar, aim, qv.v[c]);
    }
  }
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int p = 2 + k;
    const RT ar = (RT)M.ar[p], ai = (RT)M.ai[p];
    const int bit = (t >> k) & 1;
    const RT aim = bit ? ai : -ai;
    const int tp = t ^ (1 << k);
#pragma unroll
    for (int rr = 0; rr < RR; ++rr) {
      V4<V> qv = *reinterpret_cast<const V4<V>*>(lds + rr * 1024 + tp * 4);
#pragma unroll
      for (int c = 0; c < 4; ++c) vcfma(fl[rr][c], ar, aim, qv.v[c]);
    }
  }
  if constexpr (RR == 1 && RYD_EVO_FAR_BATCH > 1) {
    // small vectors (a tile spread over several blocks): the far partners
    // are L2 hits, so the loads of RYD_EVO_FAR_BATCH far bits go out
    // together instead of one round trip per bit
    constexpr int FB = RYD_EVO_FAR_BATCH;
    for (int k0 = 0; k0 < nhi; k0 += FB) {
      V4<V> q[FB];
#pragma unroll
      for (int u = 0; u < FB; ++u) {
        const int k = k0 + u;
        if (k < nhi)
          q[u] = *reinterpret_cast<const V4<V>*>(io.x + (((tile ^ (1ull << k)) << TB) + r0 * 1024 + t * 4));
      }
#pragma unroll
      for (int u = 0; u < FB; ++u) {
        const int k = k0 + u;
        if (k >= nhi) break;
        const int p = TB + k;
        const int bit = (int)((tile >> k) & 1ull);
        const RT ar = (RT)(M.ar[p] * io.s);
        const RT aim = (RT)((bit ? M.ai[p] : -M.ai[p]) * io.s);
#pragma unroll
        for (int c = 0; c < 4; ++c) vcfma(fl[0][c], ar, aim, q[u].v[c]);
      }
    }
  } else {
  for (int k = 0; k < nhi; ++k) {
    const int p = TB + k;
    const uint64_t pbase = (tile ^ (1ull << k)) << TB;
    const int bit = (int)((tile >> k) & 1ull);
    const RT ar = (RT)(M.ar[p] * io.s);
    const RT aim = (RT)((bit ? M.ai[p] : -M.ai[p]) * io.s);
#pragma unroll
    for (int rr = 0; rr < RR; ++rr) {
      V4<V> qv = *reinterpret_cast<const V4<V>*>(io.x + pbase + (r0 + rr) * 1024 + t * 4);
#pragma unroll
      for (int c = 0; c < 4; ++c) vcfma(fl[rr][c], ar, aim, qv.v[c]);
    }
  }
  }
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
#pragma unroll
  for (int rr = 0; rr < RR; ++rr) {
    const int e0 = (r0 + rr) * 1024 + t * 4;
    V4<double> el = *reinterpret_cast<const V4<double>*>(M.elo + e0);
    const double t1 = T1[e0 >> 8];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int e = e0 + c;
      const double d = eh + el.v[c] + T0[e & 255] + t1;
      const double pop = (double)(pop_hi + __popc((unsigned)e));
      evo_finish(io, base + e, fl[rr][c], xs[rr][c], d, pop, s0, s1, s2, s3);
    }
  }
  if (io.reduce) block_reduce_store(s0, s1, s2, s3, rd.partials);
}

template <class V>
__global__ __launch_bounds__(256) void k_evo_small(FullModel M, EvoIO<V> io, Reduce rd) {
  using RT = typename VT<V>::R;
  const uint64_t D = 1ull << M.n;
  const uint64_t g = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  if (g < D) {
    const V xe = vscale(io.x[g], (RT)io.s);
    V fl = vzero(xe);
    for (int p = 0; p < M.n; ++p) {
      const int bit = (int)((g >> p) & 1ull);
      vcfma(fl, (RT)(M.ar[p] * io.s), (RT)((bit ? M.ai[p] : -M.ai[p]) * io.s),
            io.x[g ^ (1ull << p)]);
    }
    const double d = M.elo[g & ((1ull << M.tb) - 1)] + M.ehi[g >> M.tb];
    evo_finish(io, g, fl, xe, d, (double)__popcll(g), s0, s1, s2, s3);
  }
  if (io.reduce) block_reduce_store(s0, s1, s2, s3, rd.partials);
}

// y = a·x (complex a; y may alias nothing)
template <class V>
__global__ __launch_bounds__(256) void k_cscale(const V* __restrict__ x, V* __restrict__ y,
                                                int64_t D, double are, double aim) {
  using RT = typename VT<V>::R;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < D; i += stride) {
    V v = vzero(x[i]);
    vcfma(v, (RT)are, (RT)aim, x[i]);
    y[i] = v;
  }
}

// y += a·x (complex a)
template <class V>
__global__ __launch_bounds__(256) void k_caxpy(const V* __restrict__ x, V* __restrict__ y,
                                               int64_t D, double are, double aim) {
  using RT = typename VT<V>::R;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < D; i += stride) {
    V v = y[i];
    vcfma(v, (RT)are, (RT)aim, x[i]);
    y[i] = v;
  }
}

// One Taylor term of psi <- exp(-tau·(i·H + G/2)) psi, the propagator of the
// non-Hermitian H - (i/2)·G with the diagonal G = Σ_i gamma_i n_i:
//   v_k = (tau/k)·(-i·(H v_{k-1}) - (G/2)·v_{k-1}),   out += v_k.
// w holds H v_{k-1}; v is overwritten with v_k; gb[p] = gamma/2 of the site
// at bit position p, summed over the set bits of each basis code.
struct TaylorDiag {
  double gb[MAXN];
};

template <class V>
__global__ __launch_bounds__(256) void k_taylor_term(ObsBasis B, TaylorDiag T,
                                                     const V* __restrict__ w, V* __restrict__ v,
                                                     V* __restrict__ out, double f, int64_t D) {
  using RT = typename VT<V>::R;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < D; i += stride) {
    double g = 0.0;
    for (uint64_t b = obs_code(B, i); b; b &= b - 1) g += T.gb[__builtin_ctzll(b)];
    const V hw = w[i], vv = v[i];
    V r;
    r.re = (RT)(f * ((double)hw.im - g * (double)vv.re));
    r.im = (RT)(f * (-(double)hw.re - g * (double)vv.im));
    v[i] = r;
    V o = out[i];
    o.re += r.re;
    o.im += r.im;
    out[i] = o;
  }
}

// y += -(i/2)·G·x with G = Σ_i gamma_i n_i from each rank's code (gb[p] =
// gamma/2 of the site at bit position p): the loss diagonal of a
// non-Hermitian H added to a plan's H·v.
template <class V>
__global__ __launch_bounds__(256) void k_loss_axpy(ObsBasis B, TaylorDiag T,
                                                   const V* __restrict__ x, V* __restrict__ y,
                                                   int64_t D) {
  using RT = typename VT<V>::R;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < D; i += stride) {
    double g = 0.0;
    for (uint64_t b = obs_code(B, i); b; b &= b - 1) g += T.gb[__builtin_ctzll(b)];
    const V xv = x[i];
    V o = y[i];
    o.re += (RT)(g * (double)xv.im);      // -(i)·g·(a + ib) = g·b - i·g·a
    o.im -= (RT)(g * (double)xv.re);
    y[i] = o;
  }
}

}  // namespace ryd
