// Chain nearest-neighbour blockade sector ("Fibonacci" basis) on the device.
//
// Basis: all n-site codes with no two adjacent excitations, in ascending code
// order — the order of the reference's independent_states for the path graph
// (tomisnot/energy-level src/core/manybody/basis.py:24-47, dimension
// F(n+2)).  With bit position p (site n-1-p) weighted W_p = F(p+2), the rank of
// a code is Σ_{b_p = 1} W_p (Zeckendorf), so
//   * clearing bit p moves the rank by -W_p (always inside the sector),
//   * setting bit p moves it by +W_p (inside iff bits p-1 and p+1 are clear),
// and H·v needs no code table and no search.
//
// Codes are rebuilt per tile: the bits >= L of a rank come from a greedy
// Zeckendorf split (once per tile, at most two high parts per tile because a
// tile is shorter than F(L+1)), the bits < L from a table of the L-bit system.
// The diagonal is E_hi(h) + E_lo[lo] + cross(h, lo) with the cross field of
// the high part on the low bits held in per-tile byte tables (LDS).
#pragma once
#include "common.hpp"
#include "hv_full.hpp"

namespace ryd {

constexpr int FIB_L = 18;          // low bits decoded through the table
constexpr int FIB_T = 2048;        // ranks per tile (< F(FIB_L + 1) = 4181)
constexpr int FIB_E = FIB_T / 256; // ranks per thread

struct FibModel {
  int n;
  int L;                      // min(n, FIB_L)
  int64_t dim;                // F(n+2)
  const uint32_t* lo_code;    // [F(L+2)] L-bit sector codes, ascending
  const double* lo_diag;      // [F(L+2)] diagonal from bits < L
  const double* vb;           // [n*n] V by bit position
  const double* dl;           // [n] delta by bit position
  int64_t W[MAXN];            // W_p = F(p+2)
  int64_t cnt_lo[2];          // F(L+2), F(L+1): low-part counts after hi bit L clear/set
  double ar[MAXN], ai[MAXN];  // Ω/2 by bit position
};

// rank -> (high part code, low-part rank); bits >= L only
__device__ inline void fib_split(const FibModel& M, int64_t r, uint64_t& h, int64_t& rho) {
  uint64_t code = 0;
  int prev = 0;
  for (int p = M.n - 1; p >= M.L; --p) {
    if (!prev && r >= M.W[p]) {
      code |= 1ull << p;
      r -= M.W[p];
      prev = 1;
    } else {
      prev = 0;
    }
  }
  h = code;
  rho = r;
}

// per-tile tables for one high part h (thread-cooperative, 256 threads)
__device__ inline void fib_hi_tables(const FibModel& M, uint64_t h, double* Fx, double* Ec,
                                     double* T0, double* T1, double* T2, double* ehi) {
  const int t = threadIdx.x;
  if (t < M.L) {                       // cross field of the high part on low bit t
    double f = 0.0;
    const double* row = M.vb + (size_t)t * M.n;
    for (int q = M.L; q < M.n; ++q)
      if ((h >> q) & 1ull) f += row[q];
    Fx[t] = f;
  } else if (t < M.n) {                // energy of high bit t with the high bits above it
    double e = 0.0;
    if ((h >> t) & 1ull) {
      e = M.dl[t];
      const double* row = M.vb + (size_t)t * M.n;
      for (int q = t + 1; q < M.n; ++q)
        if ((h >> q) & 1ull) e += row[q];
    }
    Ec[t] = e;
  }
  __syncthreads();
  if (t == 0) {
    double e = 0.0;
    for (int p = M.L; p < M.n; ++p) e += Ec[p];
    *ehi = e;
  }
  double v0 = 0.0, v1 = 0.0;
  for (int p = 0; p < 8; ++p) {
    if ((t >> p) & 1) {
      if (p < M.L) v0 += Fx[p];
      if (p + 8 < M.L) v1 += Fx[p + 8];
    }
  }
  T0[t] = v0;
  T1[t] = v1;
  if (t < 4) {
    double v2 = 0.0;
    for (int p = 0; p < 2; ++p)
      if (((t >> p) & 1) && p + 16 < M.L) v2 += Fx[p + 16];
    T2[t] = v2;
  }
  __syncthreads();
}

// Tile header: high parts h0, h1 and the low-part rank/count of the tile start.
__device__ inline void fib_tile_head(const FibModel& M, int64_t r0, uint64_t* hs, int64_t* meta) {
  uint64_t h0, h1;
  int64_t rho0, rho1;
  fib_split(M, r0, h0, rho0);
  const int64_t cnt0 = M.cnt_lo[(h0 >> M.L) & 1ull];
  h1 = h0;
  const int64_t r1 = r0 + (cnt0 - rho0);
  if (r1 < M.dim) fib_split(M, r1, h1, rho1);
  hs[0] = h0;
  hs[1] = h1;
  meta[0] = rho0;
  meta[1] = cnt0;
}

// Flips of bits below FIB_NEAR stay within FIB_HALO = F(FIB_NEAR + 1) ranks,
// so they read an LDS copy of the tile with that halo on both sides.
constexpr int FIB_NEAR = 12;
constexpr int FIB_HALO = 233;      // F(13)

template <class V>
__global__ __launch_bounds__(256) void k_fib_apply(FibModel M, ApplyIO<V> io, Reduce rd) {
  using RT = typename VT<V>::R;
  __shared__ double Fx[2][FIB_L];
  __shared__ double Ec[2][MAXN];
  __shared__ double T0[2][256], T1[2][256], T2[2][4];
  __shared__ double ehi[2];
  __shared__ uint64_t hs[2];
  __shared__ int64_t meta[2];   // rho0, cnt0
  __shared__ V xt[FIB_T + 2 * FIB_HALO];
  if (io.stop && load_coherent(io.stop)) return;
  double s, ch, cx, cz;
  apply_scalars(io, s, ch, cx, cz);
  const RT sr = (RT)s;
  const int t = threadIdx.x;
  const int64_t ntiles = (M.dim + FIB_T - 1) / FIB_T;
  const int64_t b = blockIdx.x;
  // XCD-aware: XCD b%8 sweeps a contiguous eighth of the tiles
  const int64_t per = (ntiles + 7) / 8;
  const int64_t tile = ntiles >= 64 ? ((b & 7) * per + (b >> 3)) : b;
  if (tile >= ntiles) {   // tail of the XCD remap: idle block still joins the reduction
    if (io.reduce) block_reduce_store(0.0, 0.0, 0.0, 0.0, rd.partials);
    return;
  }
  const int64_t r0 = tile * FIB_T;
  // stage scaled x over [r0 - HALO, r0 + T + HALO) clipped to the sector
  for (int k = t; k < FIB_T + 2 * FIB_HALO; k += 256) {
    const int64_t r = r0 - FIB_HALO + k;
    xt[k] = (r >= 0 && r < M.dim) ? vscale(io.x[r], sr) : vzero(xt[0]);
  }
  if (t == 0) fib_tile_head(M, r0, hs, meta);
  __syncthreads();
  fib_hi_tables(M, hs[0], Fx[0], Ec[0], T0[0], T1[0], T2[0], &ehi[0]);
  fib_hi_tables(M, hs[1], Fx[1], Ec[1], T0[1], T1[1], T2[1], &ehi[1]);
  const int64_t rho0 = meta[0], cnt0 = meta[1];
  uint64_t code[FIB_E];
  V acc[FIB_E];
  bool ok[FIB_E];
#pragma unroll
  for (int e = 0; e < FIB_E; ++e) {
    const int k = e * 256 + t;
    const int64_t r = r0 + k;
    ok[e] = r < M.dim;
    code[e] = 0;
    acc[e] = vzero(xt[0]);
    if (!ok[e]) continue;
    const int64_t rl = rho0 + k;
    const int w = rl < cnt0 ? 0 : 1;
    const int64_t lr = w ? rl - cnt0 : rl;
    const uint32_t lo = M.lo_code[lr];
    code[e] = hs[w] | (uint64_t)lo;
    const double d = ehi[w] + M.lo_diag[lr] + T0[w][lo & 255u] + T1[w][(lo >> 8) & 255u] +
                     T2[w][(lo >> 16) & 3u];
    acc[e] = vscale(xt[FIB_HALO + k], (RT)d);
  }
  // near flips from LDS
  const int pn = M.n < FIB_NEAR ? M.n : FIB_NEAR;
  for (int p = 0; p < pn; ++p) {
    const int W = (int)M.W[p];
    const RT ar = (RT)M.ar[p], ai = (RT)M.ai[p];
#pragma unroll
    for (int e = 0; e < FIB_E; ++e) {
      if (!ok[e]) continue;
      const uint64_t c = code[e];
      const int bit = (int)((c >> p) & 1ull);
      const bool free_ = !(((c >> (p + 1)) | (p > 0 ? (c >> (p - 1)) : 0ull)) & 1ull);
      if (!bit && !free_) continue;
      const int k = e * 256 + t + FIB_HALO + (bit ? -W : W);
      vcfma(acc[e], ar, bit ? ai : -ai, xt[k]);
    }
  }
  // far flips from global memory: a thread's FIB_E gathers are issued together
  for (int p = pn; p < M.n; ++p) {
    const int64_t W = M.W[p];
    const RT ar = (RT)(M.ar[p] * s), ai = (RT)(M.ai[p] * s);
    V q[FIB_E];
    bool use[FIB_E];
#pragma unroll
    for (int e = 0; e < FIB_E; ++e) {
      const uint64_t c = code[e];
      const int bit = (int)((c >> p) & 1ull);
      const bool free_ = !(((c >> (p + 1)) | (c >> (p - 1))) & 1ull);
      use[e] = ok[e] && (bit || free_);
      const int64_t src = r0 + e * 256 + t + (bit ? -W : W);
      q[e] = use[e] ? io.x[src] : vzero(xt[0]);
    }
#pragma unroll
    for (int e = 0; e < FIB_E; ++e) {
      if (!use[e]) continue;
      const int bit = (int)((code[e] >> p) & 1ull);
      vcfma(acc[e], ar, bit ? ai : -ai, q[e]);
    }
  }
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  const RT chr = (RT)ch, cxr = (RT)cx, czr = (RT)cz;
#pragma unroll
  for (int e = 0; e < FIB_E; ++e) {
    if (!ok[e]) continue;
    const int k = e * 256 + t;
    const int64_t r = r0 + k;
    const V xe = xt[FIB_HALO + k];
    V v = vscale(acc[e], chr);
    vfma(v, cxr, xe);
    if (io.z) vfma(v, czr, io.z[r]);
    io.y[r] = v;
    if (io.reduce) {
      s0 += rdot(xe, v);
      s1 += nrm2(v);
      s2 += nrm2(xe);
      if (io.probe) s3 += zdot(((io.probe[r >> 5] >> (r & 31)) & 1u) ? 1.0 : -1.0, v);
    }
  }
  if (io.reduce) block_reduce_store(s0, s1, s2, s3, rd.partials);
}

// Site occupations over the sector.  Low bits (< L) accumulate per thread in
// registers; the high part is constant per (tile, h) so its weight is reduced
// per tile and added for the set high bits.  partials: [grid][65], column 64
// holds the total weight.
template <class V>
__global__ __launch_bounds__(256) void k_fib_occ(FibModel M, const V* __restrict__ x,
                                                 double* partials) {
  __shared__ double acc[64 + 1];
  __shared__ double red[4][2];
  __shared__ uint64_t hs[2];
  __shared__ int64_t meta[2];
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  if (t <= 64) acc[t] = 0.0;
  double lo_acc[FIB_L];
#pragma unroll
  for (int p = 0; p < FIB_L; ++p) lo_acc[p] = 0.0;
  const int64_t ntiles = (M.dim + FIB_T - 1) / FIB_T;
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int64_t r0 = tile * FIB_T;
    __syncthreads();
    if (t == 0) fib_tile_head(M, r0, hs, meta);
    __syncthreads();
    double hw0 = 0.0, hw1 = 0.0;
    for (int e = 0; e < FIB_E; ++e) {
      const int64_t k = (int64_t)e * 256 + t;
      const int64_t r = r0 + k;
      if (r >= M.dim) break;
      const int64_t rl = meta[0] + k;
      const int w = rl < meta[1] ? 0 : 1;
      const uint32_t lo = M.lo_code[w ? rl - meta[1] : rl];
      const double pw = nrm2(x[r]);
      if (w) hw1 += pw; else hw0 += pw;
#pragma unroll
      for (int p = 0; p < FIB_L; ++p)
        if ((lo >> p) & 1u) lo_acc[p] += pw;
    }
    hw0 = wave_sum(hw0);
    hw1 = wave_sum(hw1);
    if (lane == 0) { red[wid][0] = hw0; red[wid][1] = hw1; }
    __syncthreads();
    if (t == 0) {
      for (int w = 0; w < 2; ++w) {
        const double T = red[0][w] + red[1][w] + red[2][w] + red[3][w];
        acc[64] += T;
        for (uint64_t c = hs[w]; c; c &= c - 1) acc[__builtin_ctzll(c)] += T;
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int p = 0; p < FIB_L; ++p) {
    const double v = wave_sum(lo_acc[p]);
    if (lane == 0) atomicAdd(&acc[p], v);
  }
  __syncthreads();
  if (t <= 64) partials[(size_t)blockIdx.x * (64 + 1) + t] = acc[t];
}

// codes[k] = code of rank r0 + k (ranks >= dim give -1)
__global__ __launch_bounds__(256) void k_fib_unrank(FibModel M, int64_t r0, int64_t count,
                                                    const int64_t* __restrict__ ranks,
                                                    int64_t* __restrict__ codes) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < count; k += stride) {
    int64_t r = ranks ? ranks[k] : r0 + k;
    if (r < 0 || r >= M.dim) { codes[k] = -1; continue; }
    uint64_t code = 0;
    int prev = 0;
    for (int p = M.n - 1; p >= 0; --p) {
      if (!prev && r >= M.W[p]) {
        code |= 1ull << p;
        r -= M.W[p];
        prev = 1;
      } else {
        prev = 0;
      }
    }
    codes[k] = (int64_t)code;
  }
}

// ranks[k] = Σ_{b_p=1} W_p for a sector code, -1 for a code outside the sector
__global__ __launch_bounds__(256) void k_fib_rank(FibModel M, const int64_t* __restrict__ codes,
                                                  int64_t count, int64_t* __restrict__ ranks) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < count; k += stride) {
    const uint64_t c = (uint64_t)codes[k];
    if (codes[k] < 0 || (M.n < 64 && (c >> M.n)) || (c & (c >> 1))) { ranks[k] = -1; continue; }
    int64_t r = 0;
    for (uint64_t b = c; b; b &= b - 1) r += M.W[__builtin_ctzll(b)];
    ranks[k] = r;
  }
}

}  // namespace ryd
