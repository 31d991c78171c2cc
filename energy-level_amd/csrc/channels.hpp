// Channel H·v with pair channels (J exchange, A pair creation): any basis
// given as sorted codes (k_chan_apply, a gather per channel and row), and the
// full basis (k_chan_full_tiled, LDS tiles; k_chan_apply below the tile width).
//
// Channel semantics of tomisnot/energy-level src/core/manybody/channels.py:91-132,
// in gather form for the row of code c (source s = c ^ mask):
//   Ω on bit p   : s = c ^ 2^p, amplitude Ω/2 if c has bit p set (g→r),
//                  conj(Ω)/2 otherwise;
//   J on (i, j)  : s = c ^ (2^pi | 2^pj), source bits (s_i, s_j) = (1, 0) give
//                  conj(J), (0, 1) give J;
//   A on (i, j)  : source (0, 0) gives A, (1, 1) gives conj(A);
//   diagonal     : Σ delta_p b_p + Σ_{p<q} V_pq b_p b_q, rebuilt per code from
//                  V held in LDS.
// A source outside the basis is dropped (the projected operator, as the
// reference's restricted-basis matvec).  Ranks: identity (full basis), the
// combinatorial number system (a complete fixed-popcount sector, ascending),
// a binary search over the sorted codes, or the split tables of an
// independent-set basis (graph.hpp).
#pragma once
#include "common.hpp"
#include "graph.hpp"
#include "hv_full.hpp"
#include "observables.hpp"

namespace ryd {

enum RankMode : int { RANK_FULL = 0, RANK_NUMBER = 1, RANK_SEARCH = 2, RANK_GRAPH = 3 };

struct PairChan {
  int pi, pj;             // bit positions of sites i, j
  double jr, ji;          // J
  double ar, ai;          // A
};

constexpr int CHAN_MAX_BLOCKS = 16384;
constexpr int BINOM_LD = 65;          // binom[p * BINOM_LD + i] = C(p, i)
constexpr int RS_MAXN = 32;           // number-sector split rank tables up to this n

struct ChanModel {
  int n;
  int mode;                    // RankMode
  int m;                       // popcount of a number sector
  int npairs;
  int64_t dim;
  const int64_t* codes;        // [dim] ascending (modes NUMBER, SEARCH)
  const int64_t* binom;        // [64][BINOM_LD]
  int rsL;                     // RANK_NUMBER split tables (n <= RS_MAXN): low bits
  const int64_t* rsplit;       // [2^rsL] low part, then [2^(n-rsL)][rsL+1] high part
  const PairChan* pairs;       // [npairs]
  const double* vb;            // [n*n] V by bit position
  // RANK_GRAPH tables (graph.hpp)
  int gl;                      // low bits L
  const uint64_t* hpk;         // [2^H + 1] first rank << 24 | class
  const uint32_t* lo_idx;      // [2^L] index of a low set or GR_NONE
  const uint32_t* fc;          // [classes] low bits blocked by the class
  const uint32_t* pre;         // [classes][tl + 1] compatible low sets before t
  int64_t tl;                  // number of low sets
  double dl[MAXN];             // delta by bit position
  double ar[MAXN], ai[MAXN];   // Ω/2 by bit position (0: no Ω channel)
};

// code and rank helpers take any struct with the rank-table fields of ChanModel
template <class M_>
__device__ __forceinline__ uint64_t chan_code(const M_& M, int64_t k) {
  return M.mode == RANK_FULL ? (uint64_t)k : (uint64_t)M.codes[k];
}

// rank of code s in the basis, -1 when s is not a basis code
template <class M_>
__device__ inline int64_t chan_rank(const M_& M, uint64_t s) {
  if (M.n < 64 && (s >> M.n)) return -1;
  if (M.mode == RANK_FULL) return (int64_t)s;
  if (M.mode == RANK_GRAPH) {
    const uint64_t e = M.hpk[s >> M.gl];
    const uint64_t c = e & GR_CLS;
    if (c == GR_CLS) return -1;
    const uint32_t l = (uint32_t)(s & ((1ull << M.gl) - 1ull));
    const uint32_t t = M.lo_idx[l];
    if (t == GR_NONE || (l & M.fc[c])) return -1;
    return (int64_t)(e >> 24) + (int64_t)M.pre[(size_t)c * (size_t)(M.tl + 1) + t];
  }
  if (M.mode == RANK_NUMBER) {
    if (__popcll(s) != M.m) return -1;
    if (M.rsplit) {       // r = Σ_i C(p_i, i) split at bit rsL: two table loads
      const uint64_t l = s & ((1ull << M.rsL) - 1ull);
      return M.rsplit[l] + M.rsplit[(1ll << M.rsL) + (int64_t)(s >> M.rsL) * (M.rsL + 1) + __popcll(l)];
    }
    int64_t r = 0;
    int i = 1;
    for (uint64_t c = s; c; c &= c - 1, ++i) r += M.binom[__builtin_ctzll(c) * BINOM_LD + i];
    return r;
  }
  int64_t lo = 0, hi = M.dim - 1;
  while (lo <= hi) {
    const int64_t mid = (lo + hi) >> 1;
    const uint64_t v = (uint64_t)M.codes[mid];
    if (v == s) return mid;
    if (v < s) lo = mid + 1;
    else hi = mid - 1;
  }
  return -1;
}

__device__ inline double chan_diag(const ChanModel& M, const double* vs, uint64_t c) {
  double d = 0.0;
  for (uint64_t a = c; a; a &= a - 1) {
    const int p = __builtin_ctzll(a);
    d += M.dl[p];
    const double* row = vs + (size_t)p * M.n;
    for (uint64_t b = a & (a - 1); b; b &= b - 1) d += row[__builtin_ctzll(b)];
  }
  return d;
}

// V (n*n doubles) staged in LDS by the whole block
__device__ inline void chan_stage_v(const ChanModel& M, double* vs) {
  for (int k = threadIdx.x; k < M.n * M.n; k += blockDim.x) vs[k] = M.vb[k];
  __syncthreads();
}

// Pair channels per gather batch.  Kagome N=24 XY (276 J pairs, full basis)
// FP64 H·v: 1 at a time 11.5 ms, 4: 10.3, 8: 9.9, 16: 9.6 ms.
#ifndef RYD_CHAN_PB
#define RYD_CHAN_PB 16
#endif
constexpr int CHAN_PB = RYD_CHAN_PB;

template <class V>
__global__ __launch_bounds__(256) void k_chan_apply(ChanModel M, ApplyIO<V> io, Reduce rd) {
  using RT = typename VT<V>::R;
  __shared__ double vs[MAXN * MAXN];
  if (io.stop && load_coherent(io.stop)) return;
  double s, ch, cx, cz;
  apply_scalars(io, s, ch, cx, cz);
  const RT sr = (RT)s;
  chan_stage_v(M, vs);
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  const RT chr = (RT)ch, cxr = (RT)cx, czr = (RT)cz;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < M.dim; k += stride) {
    const uint64_t c = chan_code(M, k);
    const V xe = vscale(io.x[k], sr);
    V acc = vscale(xe, (RT)chan_diag(M, vs, c));
    // number sector with split rank tables: the row's own two terms, reused
    // by the pairs whose bits lie on one side of the split
    const bool nsplit = M.mode == RANK_NUMBER && M.rsplit != nullptr;
    const uint64_t lmask = (1ull << M.rsL) - 1ull;
    int64_t rlo = 0, rhi = 0;
    int pcl = 0;
    if (nsplit) {
      pcl = __popcll(c & lmask);
      rlo = M.rsplit[c & lmask];
      rhi = M.rsplit[(1ll << M.rsL) + (int64_t)(c >> M.rsL) * (M.rsL + 1) + pcl];
    }
    for (int p0 = 0; p0 < M.n; p0 += CHAN_PB) {    // Ω flips, batched like the pairs
      int64_t r[CHAN_PB];
#pragma unroll
      for (int u = 0; u < CHAN_PB; ++u) {
        const int p = p0 + u < M.n ? p0 + u : 0;
        r[u] = (p0 + u < M.n && (M.ar[p] != 0.0 || M.ai[p] != 0.0)) ? chan_rank(M, c ^ (1ull << p)) : -1;
      }
      V xv[CHAN_PB];
#pragma unroll
      for (int u = 0; u < CHAN_PB; ++u) xv[u] = r[u] >= 0 ? io.x[r[u]] : vzero(xe);
#pragma unroll
      for (int u = 0; u < CHAN_PB; ++u) {
        if (r[u] < 0) continue;
        const int p = p0 + u;
        const int bit = (int)((c >> p) & 1ull);
        vcfma(acc, (RT)(M.ar[p] * s), (RT)((bit ? M.ai[p] : -M.ai[p]) * s), xv[u]);
      }
    }
    // pair channels in groups of CHAN_PB: the ranks of a group first, then
    // all its gathers in flight together, then the FMAs
    for (int q0 = 0; q0 < M.npairs; q0 += CHAN_PB) {
      int64_t r[CHAN_PB];
      double are[CHAN_PB], aim[CHAN_PB];
#pragma unroll
      for (int u = 0; u < CHAN_PB; ++u) {
        r[u] = -1;
        are[u] = aim[u] = 0.0;
        const int q = q0 + u;
        if (q >= M.npairs) continue;
        const PairChan P = M.pairs[q];
        const uint64_t src = c ^ ((1ull << P.pi) | (1ull << P.pj));
        const int bi = (int)((src >> P.pi) & 1ull), bj = (int)((src >> P.pj) & 1ull);
        if (bi != bj) {                     // exchange: (1,0) conj(J), (0,1) J
          are[u] = P.jr;
          aim[u] = bi ? -P.ji : P.ji;
        } else {                            // pair creation: (0,0) A, (1,1) conj(A)
          are[u] = P.ar;
          aim[u] = bi ? -P.ai : P.ai;
        }
        if (are[u] == 0.0 && aim[u] == 0.0) continue;
        if (nsplit) {
          // equal source bits change the popcount by 2: outside the sector
          if (bi == bj) continue;
          const uint64_t msk = (1ull << P.pi) | (1ull << P.pj);
          const int64_t hb = (1ll << M.rsL) + (int64_t)(src >> M.rsL) * (M.rsL + 1);
          if (!(msk & ~lmask)) r[u] = M.rsplit[src & lmask] + rhi;      // both bits low
          else if (!(msk & lmask)) r[u] = rlo + M.rsplit[hb + pcl];     // both bits high
          else r[u] = M.rsplit[src & lmask] + M.rsplit[hb + __popcll(src & lmask)];
        } else {
          r[u] = chan_rank(M, src);
        }
      }
      V xv[CHAN_PB];
#pragma unroll
      for (int u = 0; u < CHAN_PB; ++u) xv[u] = r[u] >= 0 ? io.x[r[u]] : vzero(xe);
#pragma unroll
      for (int u = 0; u < CHAN_PB; ++u)
        if (r[u] >= 0) vcfma(acc, (RT)(are[u] * s), (RT)(aim[u] * s), xv[u]);
    }
    V v = vscale(acc, chr);
    vfma(v, cxr, xe);
    if (io.z) vfma(v, czr, io.z[k]);
    io.y[k] = v;
    if (io.reduce) {
      s0 += rdot(xe, v);
      s1 += nrm2(v);
      s2 += nrm2(xe);
      if (io.probe) s3 += zdot(((io.probe[k >> 5] >> (k & 31)) & 1u) ? 1.0 : -1.0, v);
    }
  }
  if (io.reduce) block_reduce_store(s0, s1, s2, s3, rd.partials);
}

// Full basis with pair channels, tiled: a block holds 2^TB consecutive codes
// (the single-pass tile width of the element type) scaled by s in LDS.
// Diagonal from the plan's E_lo/E_hi tables of that width plus the cross
// field of the tile's high bits (as k_full_tiled).  A flip or pair whose bits
// all lie below TB reads LDS; one that touches high bits reads the partner
// tile (tile ^ high part of the mask) at position e ^ (low part) through the
// cache hierarchy.  A pair with both bits high has one coefficient for the
// whole tile (its source bits are the tile's); a pair with one low and one
// high bit loads only for the codes whose coefficient is non-zero.
struct ChanFullModel {
  int n;
  int npairs;
  const double* elo;           // [2^TB] diagonal of the bits below TB
  const double* ehi;           // [2^(n-TB)] diagonal of the bits from TB up
  const double* vb;            // [n*n] V by bit position
  const PairChan* pairs;       // [npairs]
  double ar[MAXN], ai[MAXN];   // Ω/2 by bit position (0: no Ω channel)
};

// gather-form coefficient of a pair for source bits (bi, bj) at (pi, pj):
// exchange (1,0) conj(J), (0,1) J; pair creation (0,0) A, (1,1) conj(A)
__device__ __forceinline__ void pair_coef(const PairChan& P, int bi, int bj, double& cr,
                                          double& ci) {
  const bool ex = bi != bj;
  cr = ex ? P.jr : P.ar;
  const double im = ex ? P.ji : P.ai;
  ci = bi ? -im : im;
}

template <class V>
__global__ __launch_bounds__(256) void k_chan_full_tiled(ChanFullModel M, ApplyIO<V> io,
                                                         Reduce rd) {
  using RT = typename VT<V>::R;
  constexpr int TB = tile_bits<V>();
  constexpr int TILE = 1 << TB;
  constexpr int R = TILE / 256;
  __shared__ V lx[TILE];
  __shared__ double Fx[16];
  __shared__ double T0[256];
  __shared__ double T1[1 << (TB - 8)];

  if (io.stop && load_coherent(io.stop)) return;
  double s, ch, cx, cz;
  apply_scalars(io, s, ch, cx, cz);
  const RT sr = (RT)s;
  const int t = threadIdx.x;
  const uint64_t ntiles = 1ull << (M.n - TB);
  const uint64_t b = blockIdx.x;
  const uint64_t tile = ntiles >= 8 ? ((b & 7ull) * (ntiles >> 3) + (b >> 3)) : b;
  const uint64_t base = tile << TB;

#pragma unroll
  for (int r = 0; r < R; ++r) lx[r * 256 + t] = vscale(io.x[base + r * 256 + t], sr);
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
  __syncthreads();

  V acc[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int e = r * 256 + t;
    const double d = eh + M.elo[e] + T0[e & 255] + T1[e >> 8];
    acc[r] = vscale(lx[e], (RT)d);
  }

  // Element e = (r << 8) | t: a bit below 8 is the thread's, constant over r;
  // a bit from 8 up is r's, uniform per unrolled r.  Each case below keeps
  // the per-element work to an LDS or buffer load and an FMA.
  constexpr uint32_t VB = sizeof(V);
  // Ω flips
  for (int p = 0; p < M.n; ++p) {
    const double arp = M.ar[p], aip = M.ai[p];
    if (arp == 0.0 && aip == 0.0) continue;
    if (p < 8) {
      const int bit = (t >> p) & 1;
      const RT a = (RT)arp, ai = (RT)(bit ? aip : -aip);
      const int tp = t ^ (1 << p);
#pragma unroll
      for (int r = 0; r < R; ++r) vcfma(acc[r], a, ai, lx[(r << 8) + tp]);
    } else if (p < TB) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int bit = (r >> (p - 8)) & 1;
        vcfma(acc[r], (RT)arp, (RT)(bit ? aip : -aip), lx[((r ^ (1 << (p - 8))) << 8) + t]);
      }
    } else {
      const int bit = (int)((tile >> (p - TB)) & 1ull);
      const auto rs = buf_rsrc(io.x + ((tile ^ (1ull << (p - TB))) << TB), TILE * VB);
      const RT a = (RT)(arp * s), ai = (RT)((bit ? aip : -aip) * s);
      V xv[R];
#pragma unroll
      for (int r = 0; r < R; ++r) xv[r] = buf_load_so<V>(rs, t * VB, (r << 8) * VB);
#pragma unroll
      for (int r = 0; r < R; ++r) vcfma(acc[r], a, ai, xv[r]);
    }
  }

  // pair channels
  for (int q = 0; q < M.npairs; ++q) {
    const PairChan P = M.pairs[q];
    const int lo = P.pi < P.pj ? P.pi : P.pj;
    const int hi = P.pi < P.pj ? P.pj : P.pi;
    if (hi < TB) {
      const int m = (1 << P.pi) | (1 << P.pj);
      const int tl = t ^ (m & 255);              // thread part of the source index
      const int mh = m >> 8;
      if (hi < 8) {                              // one coefficient per thread
        double cr, ci;
        pair_coef(P, (tl >> P.pi) & 1, (tl >> P.pj) & 1, cr, ci);
#pragma unroll
        for (int r = 0; r < R; ++r) vcfma(acc[r], (RT)cr, (RT)ci, lx[(r << 8) + tl]);
      } else if (lo >= 8) {                      // one coefficient per r
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int rs = r ^ mh;
          double cr, ci;
          pair_coef(P, (rs >> (P.pi - 8)) & 1, (rs >> (P.pj - 8)) & 1, cr, ci);
          if (cr != 0.0 || ci != 0.0) vcfma(acc[r], (RT)cr, (RT)ci, lx[(rs << 8) + tl]);
        }
      } else {                                   // low site in t, high site in r
        const int bt = (tl >> lo) & 1;
        double c0r, c0i, c1r, c1i;
        pair_coef(P, P.pi == lo ? bt : 0, P.pj == lo ? bt : 0, c0r, c0i);
        pair_coef(P, P.pi == lo ? bt : 1, P.pj == lo ? bt : 1, c1r, c1i);
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int rs = r ^ mh;
          const int bh = (rs >> (hi - 8)) & 1;
          vcfma(acc[r], (RT)(bh ? c1r : c0r), (RT)(bh ? c1i : c0i), lx[(rs << 8) + tl]);
        }
      }
    } else if (lo >= TB) {
      const uint64_t st = tile ^ ((1ull << (P.pi - TB)) | (1ull << (P.pj - TB)));
      double cr, ci;
      pair_coef(P, (int)((st >> (P.pi - TB)) & 1ull), (int)((st >> (P.pj - TB)) & 1ull), cr, ci);
      if (cr == 0.0 && ci == 0.0) continue;
      const auto rs = buf_rsrc(io.x + (st << TB), TILE * VB);
      const RT a = (RT)(cr * s), ai = (RT)(ci * s);
      V xv[R];
#pragma unroll
      for (int r = 0; r < R; ++r) xv[r] = buf_load_so<V>(rs, t * VB, (r << 8) * VB);
#pragma unroll
      for (int r = 0; r < R; ++r) vcfma(acc[r], a, ai, xv[r]);
    } else {
      const uint64_t st = tile ^ (1ull << (hi - TB));
      const int bh = (int)((st >> (hi - TB)) & 1ull);      // source bit of the high site
      const auto rs = buf_rsrc(io.x + (st << TB), TILE * VB);
      // coefficients for source bit 0 and 1 of the low site
      double c0r, c0i, c1r, c1i;
      pair_coef(P, P.pi == lo ? 0 : bh, P.pj == lo ? 0 : bh, c0r, c0i);
      pair_coef(P, P.pi == lo ? 1 : bh, P.pj == lo ? 1 : bh, c1r, c1i);
      const bool use0 = c0r != 0.0 || c0i != 0.0, use1 = c1r != 0.0 || c1i != 0.0;
      if (!use0 && !use1) continue;
      V xv[R];
      if (lo < 8) {                              // source bit of the low site per thread
        const int bl = ((t >> lo) & 1) ^ 1;
        const bool use = bl ? use1 : use0;
        const RT a = (RT)((bl ? c1r : c0r) * s), ai = (RT)((bl ? c1i : c0i) * s);
        const uint32_t vo = (uint32_t)(t ^ (1 << lo)) * VB;
        if (use) {
#pragma unroll
          for (int r = 0; r < R; ++r) xv[r] = buf_load_so<V>(rs, vo, (r << 8) * VB);
#pragma unroll
          for (int r = 0; r < R; ++r) vcfma(acc[r], a, ai, xv[r]);
        }
      } else {                                   // per r
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int bl = ((r >> (lo - 8)) & 1) ^ 1;
          xv[r] = (bl ? use1 : use0)
                      ? buf_load_so<V>(rs, t * VB, ((r ^ (1 << (lo - 8))) << 8) * VB)
                      : vzero(acc[r]);
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int bl = ((r >> (lo - 8)) & 1) ^ 1;
          vcfma(acc[r], (RT)((bl ? c1r : c0r) * s), (RT)((bl ? c1i : c0i) * s), xv[r]);
        }
      }
    }
  }

  // combine, store, reduce
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  const RT chr = (RT)ch, cxr = (RT)cx, czr = (RT)cz;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint64_t g = base + r * 256 + t;
    const V xe = lx[r * 256 + t];
    V v = vscale(acc[r], chr);
    vfma(v, cxr, xe);
    if (io.z) vfma(v, czr, io.z[g]);
    io.y[g] = v;
    if (io.reduce) {
      s0 += rdot(xe, v);
      s1 += nrm2(v);
      s2 += nrm2(xe);
      if (io.probe) s3 += zdot(((io.probe[g >> 5] >> (g & 31)) & 1u) ? 1.0 : -1.0, v);
    }
  }
  if (io.reduce) block_reduce_store(s0, s1, s2, s3, rd.partials);
}

// H·v with a caller's sparse matrix in CSR form (rows and columns in the
// order of the vectors): one row per thread, grid-stride, the same fused
// scalars and FP64 reductions as the other apply kernels.
struct CsrModel {
  int64_t dim;
  const int64_t* ptr;          // [dim + 1]
  const int32_t* col;          // [nnz]
  const double* vre;           // [nnz]
  const double* vim;           // [nnz], null for a real matrix
};

template <class V>
__global__ __launch_bounds__(256) void k_csr_apply(CsrModel M, ApplyIO<V> io, Reduce rd) {
  using RT = typename VT<V>::R;
  if (io.stop && load_coherent(io.stop)) return;
  double s, ch, cx, cz;
  apply_scalars(io, s, ch, cx, cz);
  const RT sr = (RT)s;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  const RT chr = (RT)ch, cxr = (RT)cx, czr = (RT)cz;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < M.dim; k += stride) {
    const V xe = vscale(io.x[k], sr);
    V acc = vzero(xe);
    const int64_t a = M.ptr[k], b = M.ptr[k + 1];
    for (int64_t q = a; q < b; ++q)
      vcfma(acc, (RT)(M.vre[q] * s), (RT)(M.vim ? M.vim[q] * s : 0.0), io.x[M.col[q]]);
    V v = vscale(acc, chr);
    vfma(v, cxr, xe);
    if (io.z) vfma(v, czr, io.z[k]);
    io.y[k] = v;
    if (io.reduce) {
      s0 += rdot(xe, v);
      s1 += nrm2(v);
      s2 += nrm2(xe);
      if (io.probe) s3 += zdot(((io.probe[k >> 5] >> (k & 31)) & 1u) ? 1.0 : -1.0, v);
    }
  }
  if (io.reduce) block_reduce_store(s0, s1, s2, s3, rd.partials);
}

// site occupations over a code basis: partials[block][p] = Σ |x_k|^2 b_p(code_k),
// partials[block][MAXN_OCC] = Σ |x_k|^2 (deterministic: fixed chunk per block)
template <class V>
__global__ __launch_bounds__(256) void k_chan_occ(ChanModel M, const V* __restrict__ x,
                                                  double* partials) {
  __shared__ double red[4][MAXN_OCC + 1];
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  const int64_t per = (M.dim + gridDim.x - 1) / gridDim.x;
  const int64_t k0 = (int64_t)blockIdx.x * per;
  const int64_t k1 = k0 + per < M.dim ? k0 + per : M.dim;
  double* out = partials + (size_t)blockIdx.x * (MAXN_OCC + 1);
  for (int p = -1; p < M.n; ++p) {
    double acc = 0.0;
    for (int64_t k = k0 + t; k < k1; k += 256) {
      const double w = nrm2(x[k]);
      if (p < 0) acc += w;
      else if ((chan_code(M, k) >> p) & 1ull) acc += w;
    }
    acc = wave_sum(acc);
    if (lane == 0) red[wid][p < 0 ? MAXN_OCC : p] = acc;
  }
  __syncthreads();
  if (t <= MAXN_OCC) {
    const bool used = t == MAXN_OCC || t < M.n;
    out[t] = used ? red[0][t] + red[1][t] + red[2][t] + red[3][t] : 0.0;
  }
}

// codes[k] = basis code of rank ranks[k] (or of r0 + k), -1 outside the basis
__global__ __launch_bounds__(256) void k_chan_codes(ChanModel M, int64_t r0, int64_t count,
                                                    const int64_t* __restrict__ ranks,
                                                    int64_t* __restrict__ codes) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < count; k += stride) {
    const int64_t r = ranks ? ranks[k] : r0 + k;
    codes[k] = (r < 0 || r >= M.dim) ? -1 : (int64_t)chan_code(M, r);
  }
}

__global__ __launch_bounds__(256) void k_chan_rank(ChanModel M, const int64_t* __restrict__ codes,
                                                   int64_t count, int64_t* __restrict__ ranks) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < count; k += stride)
    ranks[k] = codes[k] < 0 ? -1 : chan_rank(M, (uint64_t)codes[k]);
}

}  // namespace ryd
