// Piecewise-constant evolution on the full basis: the fused Chebyshev step of
// the forward propagator and of its adjoint (Clenshaw) recurrence.
//
//   H(Ω, Δ) = om · F + D + dadd · N
// with F the plan's flip operator (amplitudes Ω_site/2), D the plan's diagonal
// (delta_site n_i + V n_i n_j), N = Σ_i n_i and per-slice scalars om, dadd.
// One launch computes, for x~ = s·x,
//   y    = ch·(om·F x~ + D x~ + dadd·N x~) + cx·x~ + cz·z + cw·w
//   acc += ca·y                                   (complex cw, ca; optional)
//   acc  = c0·x~ + ca·y                           (acc_init: the first term)
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
  double c0_re, c0_im;      // acc_init: acc starts from c0·x~ instead of its old value
  int acc_init;
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
    V a;
    if (io.acc_init) {
      a = vzero(xe);
      vcfma(a, (RT)io.c0_re, (RT)io.c0_im, xe);
    } else {
      a = io.acc[g];
    }
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

// Small complex64 vectors (n in [11, 21]): one 2^11-code tile per
// 1024-thread block.  A thread owns 2 consecutive codes (bit 0 in
// registers), bits 1–10 go through LDS, and the far partners (bits 11..n−1,
// at most EVO_WIDE_FAR) are requested together with the thread's own codes,
// before the block's barrier, and folded into the flip sum at once; the z,
// w, q and acc operands are requested right after the first barrier, so
// their latency hides behind the LDS phase.  A launch then waits on about
// two memory round trips instead of one per far bit (the 256-thread tiles
// spread over several blocks also fetch bits 10–11 from other blocks).
// Tile map XCD-aware as in k_evo_tiled.
constexpr int EVO_WIDE_TB = 11;
constexpr int EVO_WIDE_FAR = 10;
constexpr int EVO_WIDE_MAXN = EVO_WIDE_TB + EVO_WIDE_FAR;

template <class V> struct alignas(2 * sizeof(V)) V2 { V v[2]; };

template <class V>
__global__ __launch_bounds__(1024, 8) void k_evo_wide(FullModel M, EvoIO<V> io, Reduce rd) {
  using RT = typename VT<V>::R;
  constexpr int TB = EVO_WIDE_TB;
  __shared__ V2<V> lds[1024];
  __shared__ double Fx[TB];
  __shared__ double T0[256];
  __shared__ double T1[1 << (TB - 8)];
  const int t = threadIdx.x;
  const int nhi = M.n - TB;
  const uint64_t nblk = 1ull << nhi;
  const uint64_t b = blockIdx.x;
  const uint64_t tile = nblk >= 8 ? ((b & 7ull) * (nblk >> 3) + (b >> 3)) : b;
  const uint64_t base = tile << TB;
  const uint64_t g0 = base + (uint64_t)t * 2;
  const RT sr = (RT)io.s;
  const V2<V> own = *reinterpret_cast<const V2<V>*>(io.x + g0);
  // far partners: bit TB+k of the tile flipped; beyond nhi the load rereads
  // the thread's own codes (in bounds) and its coefficient is zero
  V2<V> far[EVO_WIDE_FAR];
#pragma unroll
  for (int k = 0; k < EVO_WIDE_FAR; ++k) {
    const uint64_t src = k < nhi ? (((tile ^ (1ull << k)) << TB) + (uint64_t)t * 2) : g0;
    far[k] = *reinterpret_cast<const V2<V>*>(io.x + src);
  }
  V xs[2];
#pragma unroll
  for (int c = 0; c < 2; ++c) xs[c] = vscale(own.v[c], sr);
  {
    V2<V> wv;
    wv.v[0] = xs[0];
    wv.v[1] = xs[1];
    lds[t] = wv;
  }
  if (t < TB) {
    double f = 0.0;
    const double* row = M.vb + (size_t)t * M.n;
    for (int q = TB; q < M.n; ++q)
      if ((tile >> (q - TB)) & 1ull) f += row[q];
    Fx[t] = f;
  }
  V fl[2];
#pragma unroll
  for (int c = 0; c < 2; ++c) fl[c] = vzero(xs[c]);
#pragma unroll
  for (int k = 0; k < EVO_WIDE_FAR; ++k) {
    const int p = k < nhi ? TB + k : 0;
    const int bit = (int)((tile >> k) & 1ull);
    const double on = k < nhi ? io.s : 0.0;
    const RT ar = (RT)(M.ar[p] * on);
    const RT aim = (RT)((bit ? M.ai[p] : -M.ai[p]) * on);
#pragma unroll
    for (int c = 0; c < 2; ++c) vcfma(fl[c], ar, aim, far[k].v[c]);
  }
  // bit 0 in registers
  {
    const RT ar = (RT)M.ar[0], ai = (RT)M.ai[0];
    vcfma(fl[0], ar, -ai, xs[1]);
    vcfma(fl[1], ar, ai, xs[0]);
  }
  __syncthreads();
  V2<V> zv, wv, qv, av;
  if (io.z) zv = *reinterpret_cast<const V2<V>*>(io.z + g0);
  if (io.w) wv = *reinterpret_cast<const V2<V>*>(io.w + g0);
  if (io.q) qv = *reinterpret_cast<const V2<V>*>(io.q + g0);
  if (io.acc && !io.acc_init) av = *reinterpret_cast<const V2<V>*>(io.acc + g0);
  if (t < 256) {
    double v = 0.0;
#pragma unroll
    for (int p = 0; p < 8; ++p)
      if ((t >> p) & 1) v += Fx[p];
    T0[t] = v;
  }
  if (t < (1 << (TB - 8))) {
    double w = 0.0;
#pragma unroll
    for (int p = 8; p < TB; ++p)
      if ((t >> (p - 8)) & 1) w += Fx[p];
    T1[t] = w;
  }
  // bits 1–10 through LDS
#pragma unroll
  for (int k = 0; k < TB - 1; ++k) {
    const int p = 1 + k;
    const RT ar = (RT)M.ar[p], ai = (RT)M.ai[p];
    const RT aim = ((t >> k) & 1) ? ai : -ai;
    const V2<V> pv = lds[t ^ (1 << k)];
#pragma unroll
    for (int c = 0; c < 2; ++c) vcfma(fl[c], ar, aim, pv.v[c]);
  }
  __syncthreads();
  const double eh = M.ehi[tile];
  const int pop_hi = __popcll(tile);
  const int e0 = t * 2;
  const double el0 = M.elo[e0], el1 = M.elo[e0 + 1];
  const double t1 = T1[e0 >> 8];
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  V2<V> yv, ov;
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    const int e = e0 + c;
    const double d = eh + (c ? el1 : el0) + T0[e & 255] + t1;
    const double pop = (double)(pop_hi + __popc((unsigned)e));
    V hx = vscale(fl[c], (RT)io.om);
    vfma(hx, (RT)(d + io.dadd * pop), xs[c]);
    V v = vscale(hx, (RT)io.ch);
    vfma(v, (RT)io.cx, xs[c]);
    if (io.z) vfma(v, (RT)io.cz, zv.v[c]);
    if (io.w) vcfma(v, (RT)io.cw_re, (RT)io.cw_im, wv.v[c]);
    yv.v[c] = v;
    if (io.acc) {
      V a;
      if (io.acc_init) {
        a = vzero(v);
        vcfma(a, (RT)io.c0_re, (RT)io.c0_im, xs[c]);
      } else {
        a = av.v[c];
      }
      vcfma(a, (RT)io.ca_re, (RT)io.ca_im, v);
      ov.v[c] = a;
    }
    if (io.reduce) {
      s2 += nrm2(v);
      if (io.q) {
        s0 += rdot(qv.v[c], fl[c]);
        s1 += pop * rdot(qv.v[c], xs[c]);
        s3 += rdot(qv.v[c], v);
      }
    }
  }
  *reinterpret_cast<V2<V>*>(io.y + g0) = yv;
  if (io.acc) *reinterpret_cast<V2<V>*>(io.acc + g0) = ov;
  if (io.reduce) block_reduce_store(s0, s1, s2, s3, rd.partials);
}

// Fixed-order sums of per-term block partials: block j sums the nblocks
// partials of term j (partials + j·nblocks·NRED) into out[j·NRED ..].
__global__ __launch_bounds__(256) void k_reduce_terms(const double* __restrict__ partials,
                                                      unsigned nblocks, double* __restrict__ out) {
  __shared__ double red_lds[4 * NRED];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const double* p = partials + (size_t)blockIdx.x * nblocks * NRED;
  double s[NRED] = {0.0, 0.0, 0.0, 0.0};
  for (unsigned b = threadIdx.x; b < nblocks; b += blockDim.x) {
#pragma unroll
    for (int k = 0; k < NRED; ++k) s[k] += p[(size_t)b * NRED + k];
  }
#pragma unroll
  for (int k = 0; k < NRED; ++k) s[k] = wave_sum(s[k]);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NRED; ++k) red_lds[wid * NRED + k] = s[k];
  }
  __syncthreads();
  if (threadIdx.x < NRED) {
    const int k = threadIdx.x;
    out[(size_t)blockIdx.x * NRED + k] =
        red_lds[k] + red_lds[NRED + k] + red_lds[2 * NRED + k] + red_lds[3 * NRED + k];
  }
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
