// Full 2^n basis: fused diagonal + bit-flip H·v with an optional
// three-term update and FP64 reductions, for gfx950.
//
//   y = ch · H(s·x) + cx · (s·x) + cz · z
//   sums: Re<s·x, y>, |y|^2, |s·x|^2, Σ_e z_e·Re(y_e)  (z_e = ±1 probe bits)
//
// Channel semantics follow tomisnot/energy-level
// src/core/manybody/channels.py:58-77 (diagonal +delta_i n_i + V n_i n_j) and
// :114-118 (Ω flips: target bit 1 takes Ω/2, target bit 0 takes conj(Ω)/2).
// Bit position p (LSB = 0) is site n-1-p (site 0 is the MSB).
//
// Tiling (per block of 2^TT threads, one tile of 2^TB consecutive codes):
//   element e = r·4·2^TT + t·4 + c  (t = thread, c = bits 0..1, r = bits
//   2+TT..TB-1); bits 0..1 and the r bits flip inside a thread's registers,
//   bits 2..TT+1 flip across threads through LDS, bits >= TB read the partner
//   tile through L2/MALL.  The diagonal is rebuilt on the fly:
//     d(e) = E_hi[tile] + E_lo[e] + T0[e & 255] + T1[e >> 8]
//   where T0/T1 hold the cross field Σ_{q>=TB} V_pq b_q of this tile on the
//   low bits, built per tile in LDS.  Tiles are mapped XCD-aware: block b runs
//   on XCD b % 8, which sweeps its own contiguous eighth of the vector.
#pragma once
#include "common.hpp"

namespace ryd {

constexpr int MAXN = 64;

struct FullModel {
  int n;              // number of sites (bits)
  int tb;             // tile bits of the tables below
  const double* elo;  // [2^tb] diagonal from bits < tb (delta and V inside)
  const double* ehi;  // [2^(n-tb)] diagonal from bits >= tb
  const double* vb;   // [n*n] V between bit positions (symmetric)
  double ar[MAXN];    // Re(Ω)/2 per bit position
  double ai[MAXN];    // Im(Ω)/2 per bit position
};

enum ScalarMode : int { SM_CONST = 0, SM_LANCZOS = 1 };

template <class V>
struct ApplyIO {
  const V* x;
  const V* z;               // optional third term
  V* y;
  const uint32_t* probe;    // optional probe bits (bit set = +1)
  int smode;
  double s, ch, cx, cz;     // SM_CONST
  const double* nrm;        // SM_LANCZOS: s = 1/nrm[j], cz = -nrm[j]/nrm[j-1]
  int j;
  const int* stop;          // optional early exit flag (device)
  int reduce;
  int nfar;                 // far bits the tiled kernel handles (bits TB .. TB+nfar-1)
  int partial;              // 1: write the raw partial sum (pass A of a split H·v)
};

template <class V>
__device__ __forceinline__ void apply_scalars(const ApplyIO<V>& io, double& s, double& ch,
                                              double& cx, double& cz) {
  if (io.smode == SM_LANCZOS) {
    const double nj = io.nrm[io.j];
    s = 1.0 / nj;
    ch = 1.0;
    cx = 0.0;
    cz = io.j > 0 ? -nj / io.nrm[io.j - 1] : 0.0;
  } else {
    s = io.s; ch = io.ch; cx = io.cx; cz = io.cz;
  }
}

// Groups of 4 codes per thread of a tile (1024 codes per group): FP32 8
// (13-bit tiles), FP64 and complex64 4 (12-bit), complex128 2 (11-bit).
// Measured at n = 28: FP32 13-bit tiles 1.47 ms per H·v against 1.52 ms at
// 12 bits (and C2's KPM steps 3 % faster); FP64 11-bit tiles or an unrolled
// far-bit loop were slower (with the 1024-thread pass A, unrolled by 2 or 4:
// 2.58 against 2.52 ms per FP64 H·v).  Halving them (RYD_TILE_R8 = 2,
// RYD_TILE_R4 = 4: pass-A tiles of 2^13 FP64 / 2^14 FP32 codes, two blocks
// per CU, one more far bit) loses: 2.25 -> 2.35 ms FP64, 1.19 -> 1.26 ms FP32.
#ifndef RYD_TILE_R4
#define RYD_TILE_R4 8
#endif
#ifndef RYD_TILE_R8
#define RYD_TILE_R8 4
#endif
#ifndef RYD_FAR_UNROLL
#define RYD_FAR_UNROLL 1
#endif
// Far bits from the highest down (the partner tile furthest away first):
// n = 28, 2.36 -> 2.34 ms FP64, 1.21 -> 1.19 ms FP32 per H·v; 512-thread
// pass-A tiles (2^13 codes, two blocks per CU) lose (2.44 ms FP64).
#ifndef RYD_FAR_REV
#define RYD_FAR_REV 1
#endif
// Streaming accesses of the split H·v (n = 28, ms per H·v FP64 / FP32):
// pass A's partial as 16-byte non-temporal stores, pass B's single reads of
// x, the partial and z and its write of y non-temporal: 2.50 / 1.27 ->
// 2.36 / 1.21.  Batching pass A's far-partner loads (2 or 3 bits per round
// trip) lost (2.55), and so did a persistent pass A that prefetched the next
// tile into registers (2.75: the prefetch sits ahead of the far-partner
// loads in the in-order load queue, and the blocks of an XCD drift apart).
#ifndef RYD_TOP_NT
#define RYD_TOP_NT 1
#endif
// Pass A of a split H·v (n >= 23) runs blocks of 2^TT threads, so its tiles
// are 2^(TT-8) times the single-pass tile and fewer bits are far: 1024
// threads give 15-bit FP32, 14-bit FP64 and 13-bit complex128 tiles (128 KB
// of LDS, one block per CU).  Each far bit re-reads x through L2 (about
// 0.085 ms per bit for FP64 at n = 28), which the wider tile saves twice
// over: 1024 threads beat 256 by 0.07 ms (FP64) and 0.10 ms (FP32) per H·v.
#ifndef RYD_PASSA_TT4
#define RYD_PASSA_TT4 10
#endif
#ifndef RYD_PASSA_TT8
#define RYD_PASSA_TT8 10
#endif
#ifndef RYD_PASSA_TT16
#define RYD_PASSA_TT16 10
#endif
template <class V> constexpr int passa_tt() {
  return sizeof(V) == 4 ? RYD_PASSA_TT4 : (sizeof(V) == 16 ? RYD_PASSA_TT16 : RYD_PASSA_TT8);
}
constexpr int NTBW = 6;          // diagonal table widths 10..15

template <class V> constexpr int tile_r() {
  return sizeof(V) >= 16 ? 2 : (sizeof(V) == 8 ? RYD_TILE_R8 : RYD_TILE_R4);
}
// Tile bits of a block of 2^TT threads (TT = 8: the single-pass tile).
template <class V, int TT = 8> constexpr int tile_bits() {
  return 2 + TT + (tile_r<V>() >= 8 ? 3 : (tile_r<V>() >= 4 ? 2 : (tile_r<V>() >= 2 ? 1 : 0)));
}

// Register groups of 16 contiguous bytes: EPG = 16 / sizeof(V) codes per
// thread and group (FP32 4, FP64 and complex64 2, complex128 1), so every
// global load and store of a wavefront is one contiguous 1 KB run and every
// LDS access a conflict-free 16-byte read.  (Four codes per thread made the
// 8- and 16-byte types issue 16-byte accesses at a 32- or 64-byte stride:
// a load-and-store-only pass A moved 4.4 TB/s in FP64 against 5.7 in FP32.)
template <class V> struct alignas(16) G16 { V v[16 / sizeof(V)]; };
template <class V> constexpr int epg() { return (int)(16 / sizeof(V)); }
template <class V> constexpr int epg_bits() { return sizeof(V) == 4 ? 2 : (sizeof(V) == 8 ? 1 : 0); }

template <class V, int TT = 8>
__global__ __launch_bounds__(1 << TT) void k_full_tiled(FullModel M, ApplyIO<V> io, Reduce rd) {
  using RT = typename VT<V>::R;
  constexpr int TB = tile_bits<V, TT>();
  constexpr int EPG = epg<V>();                  // codes per thread and group
  constexpr int EB = epg_bits<V>();
  constexpr int GS = EPG << TT;                  // codes per register group
  constexpr int R = (1 << TB) / GS;              // groups per thread
  constexpr int RB = TB - EB - TT;               // bits carried by the group index
  static_assert(TT >= 8 && TT <= 10 && TB <= 15, "tiles: 256..1024 threads, at most 2^15 codes");
  static_assert(R * GS == (1 << TB) && (1 << RB) == R, "tile: whole groups");
  __shared__ G16<V> lds[R << TT];
  __shared__ double Fx[16];
  __shared__ double T0[256];
  __shared__ double T1[1 << (TB - 8)];

  if (io.stop && load_coherent(io.stop)) return;
  double s, ch, cx, cz;
  apply_scalars(io, s, ch, cx, cz);
  const RT sr = (RT)s;

  const int t = threadIdx.x;
  const int nhi = M.n - TB;
  const uint64_t ntiles = 1ull << nhi;
  const uint64_t b = blockIdx.x;
  const uint64_t tile = ntiles >= 8 ? ((b & 7ull) * (ntiles >> 3) + (b >> 3)) : b;
  const uint64_t base = tile << TB;

  // stage the tile (scaled) in registers and LDS
  V xs[R][EPG];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    G16<V> q = *reinterpret_cast<const G16<V>*>(io.x + base + r * GS + t * EPG);
#pragma unroll
    for (int c = 0; c < EPG; ++c) xs[r][c] = vscale(q.v[c], sr);
    G16<V> w;
#pragma unroll
    for (int c = 0; c < EPG; ++c) w.v[c] = xs[r][c];
    lds[(r << TT) + t] = w;
  }

  // per-tile cross field of the high bits on the low bits
  if (t < TB) {
    double f = 0.0;
    const double* row = M.vb + (size_t)t * M.n;
    for (int q = TB; q < M.n; ++q)
      if ((tile >> (q - TB)) & 1ull) f += row[q];
    Fx[t] = f;
  }
  __syncthreads();
  if (t < 256) {
    double v = 0.0;
#pragma unroll
    for (int p = 0; p < 8; ++p)
      if ((t >> p) & 1) v += Fx[p];
    T0[t] = v;
  }
  {
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

  // diagonal
  V acc[R][EPG];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int e0 = r * GS + t * EPG;
    const double t1 = T1[e0 >> 8];
#pragma unroll
    for (int c = 0; c < EPG; ++c) {
      const double d = eh + M.elo[e0 + c] + T0[(e0 + c) & 255] + t1;
      acc[r][c] = vscale(xs[r][c], (RT)d);
    }
  }

  // bits 0..EB-1: inside the thread's group
#pragma unroll
  for (int p = 0; p < EB; ++p) {
    const RT ar = (RT)M.ar[p], ai = (RT)M.ai[p];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int c = 0; c < EPG; ++c) {
        const int bit = (c >> p) & 1;
        vcfma(acc[r][c], ar, bit ? ai : -ai, xs[r][c ^ (1 << p)]);
      }
  }
  // bits EB+TT..TB-1: across the thread's register groups
#pragma unroll
  for (int k = 0; k < RB; ++k) {
    const int p = EB + TT + k;
    const RT ar = (RT)M.ar[p], ai = (RT)M.ai[p];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int bit = (r >> k) & 1;
#pragma unroll
      for (int c = 0; c < EPG; ++c) vcfma(acc[r][c], ar, bit ? ai : -ai, xs[r ^ (1 << k)][c]);
    }
  }
  // bits EB..EB+TT-1: partner thread through LDS
#pragma unroll
  for (int k = 0; k < TT; ++k) {
    const int p = EB + k;
    const RT ar = (RT)M.ar[p], ai = (RT)M.ai[p];
    const int bit = (t >> k) & 1;
    const RT aim = bit ? ai : -ai;
    const int tp = t ^ (1 << k);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const G16<V> q = lds[(r << TT) + tp];
#pragma unroll
      for (int c = 0; c < EPG; ++c) vcfma(acc[r][c], ar, aim, q.v[c]);
    }
  }
  // bits TB .. TB+nfar-1: partner tile through the cache hierarchy
#pragma unroll RYD_FAR_UNROLL
  for (int kk = 0; kk < io.nfar; ++kk) {
    const int k = RYD_FAR_REV ? io.nfar - 1 - kk : kk;
    const int p = TB + k;
    const uint64_t pbase = (tile ^ (1ull << k)) << TB;
    const int bit = (int)((tile >> k) & 1ull);
    const RT ar = (RT)(M.ar[p] * s);
    const RT aim = (RT)((bit ? M.ai[p] : -M.ai[p]) * s);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const G16<V> q = *reinterpret_cast<const G16<V>*>(io.x + pbase + r * GS + t * EPG);
#pragma unroll
      for (int c = 0; c < EPG; ++c) vcfma(acc[r][c], ar, aim, q.v[c]);
    }
  }

  if (io.partial) {   // pass A: raw sum for the top-bit pass, streamed past L2
#pragma unroll
    for (int r = 0; r < R; ++r) {
      G16<V> o;
#pragma unroll
      for (int c = 0; c < EPG; ++c) o.v[c] = acc[r][c];
      nt_store16(io.y + base + r * GS + t * EPG, o);
    }
    return;
  }
  // combine, store, reduce
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  const RT chr = (RT)ch, cxr = (RT)cx, czr = (RT)cz;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint64_t g0 = base + r * GS + t * EPG;
    G16<V> zq;
    if (io.z) zq = *reinterpret_cast<const G16<V>*>(io.z + g0);
    uint32_t pw = 0;
    if (io.probe) pw = io.probe[g0 >> 5] >> (g0 & 31);
    G16<V> o;
#pragma unroll
    for (int c = 0; c < EPG; ++c) {
      V v = vscale(acc[r][c], chr);
      vfma(v, cxr, xs[r][c]);
      if (io.z) vfma(v, czr, zq.v[c]);
      o.v[c] = v;
      if (io.reduce) {
        s0 += rdot(xs[r][c], v);
        s1 += nrm2(v);
        s2 += nrm2(xs[r][c]);
        if (io.probe) s3 += zdot(((pw >> c) & 1u) ? 1.0 : -1.0, v);
      }
    }
    *reinterpret_cast<G16<V>*>(io.y + g0) = o;
  }
  if (io.reduce) block_reduce_store(s0, s1, s2, s3, rd.partials);
}

// Pass B of the split H·v for large n: the top TOPB bits (n-TOPB .. n-1).
// A block holds every combination of those bits for 2^CB consecutive codes
// (one 128- or 256-byte run per combination), so each top-bit flip is an LDS
// or register exchange.  It reads the pass-A partial sum from y and finishes
//   y = ch·(y_A + Σ_top flips of s·x) + cx·(s·x) + cz·z  with the reductions.
// Thread u (0..2^UB-1): c = u & (2^CB-1) selects the code inside the run,
// u >> CB gives the low (UB-CB) top bits, register slot j the remaining top
// bits.  Per element size (4 / 8 / 16 bytes): TOPB, UB and CB below.
// Measured on MI355X at n = 28, FP64 / FP32 ms per H·v, with 14-bit FP64 and
// 15-bit FP32 pass-A tiles (CB = 5 unless stated): 8 top bits, 512 threads
// 2.88 / 1.37; 8 bits, 1024 threads 2.80 / 1.44; 9 bits, 1024 threads
// 2.64 / 1.40; 9 bits, 512 threads 3.11 / 1.58 (one block per CU); FP64 with
// 128-byte runs (CB = 4): 9 bits 2.79 (1024 threads), 2.84 (512); FP32 with
// 256-byte runs (CB = 6): 8 bits 1.32 (1024 threads), 1.52 (512).  With 12-
// and 13-bit pass-A tiles: 8 bits, 512 threads 2.95 / 1.48; 9 bits with
// 128-byte runs 3.01 / 1.72; 7 or 6 top bits lose (pass A's far partners
// leave L2).
#ifndef RYD_TOP_PREFETCH
#define RYD_TOP_PREFETCH 1
#endif
// z of an FP32 step in the same burst: n = 28 Chebyshev step 1.64 -> 1.33 ms,
// n = 25 0.204 -> 0.171 ms; FP64 loses (2.82 -> 2.92 ms: the registers)
#ifndef RYD_TOP_ZPRE
#define RYD_TOP_ZPRE 1
#endif
#ifndef RYD_TOPB4
#define RYD_TOPB4 8
#endif
#ifndef RYD_TOPB8
#define RYD_TOPB8 9
#endif
#ifndef RYD_TOPB16
#define RYD_TOPB16 9
#endif
#ifndef RYD_TOP_UB4
#define RYD_TOP_UB4 10
#endif
#ifndef RYD_TOP_UB8
#define RYD_TOP_UB8 10
#endif
#ifndef RYD_TOP_UB16
#define RYD_TOP_UB16 10
#endif
#ifndef RYD_RUNB4
#define RYD_RUNB4 6
#endif
#ifndef RYD_RUNB8
#define RYD_RUNB8 5
#endif
#ifndef RYD_RUNB16
#define RYD_RUNB16 3
#endif
template <class V> constexpr int top_bits() {
  return sizeof(V) == 4 ? RYD_TOPB4 : (sizeof(V) == 16 ? RYD_TOPB16 : RYD_TOPB8);
}
template <class V> constexpr int top_ub() {
  return sizeof(V) == 4 ? RYD_TOP_UB4 : (sizeof(V) == 16 ? RYD_TOP_UB16 : RYD_TOP_UB8);
}
template <class V> constexpr int run_bits() {
  return sizeof(V) == 4 ? RYD_RUNB4 : (sizeof(V) == 16 ? RYD_RUNB16 : RYD_RUNB8);
}

// pass B reads x, the partial and z exactly once: streaming loads with
// RYD_TOP_NT
template <class V> __device__ __forceinline__ V top_load(const V* p) {
#if RYD_TOP_NT
  return nt_load(p);
#else
  return *p;
#endif
}

template <class V>
__global__ __launch_bounds__(1 << top_ub<V>()) void k_full_top(FullModel M, ApplyIO<V> io, Reduce rd) {
  using RT = typename VT<V>::R;
  constexpr int TOPB = top_bits<V>();
  constexpr int UB = top_ub<V>();
  constexpr int NU = 1 << UB;                // threads
  constexpr int CB = run_bits<V>();
  constexpr int TU = UB - CB;                // top bits carried by the thread index
  static_assert(TOPB >= 1 && TOPB <= 10 && UB >= 8 && UB <= 10, "top pass: 1..10 bits, 256..1024 threads");
  static_assert(TU >= 0 && TU <= TOPB, "top pass: the thread index holds at most TOPB top bits");
  constexpr int J = 1 << (TOPB - TU);        // register slots per thread
  static_assert((size_t)NU * J * sizeof(V) <= 160 * 1024, "top pass: LDS");
  __shared__ V lds[NU * J];
  if (io.stop && load_coherent(io.stop)) return;
  double s, ch, cx, cz;
  apply_scalars(io, s, ch, cx, cz);
  const RT sr = (RT)s;
  const int u = threadIdx.x;
  const int S = M.n - TOPB;                  // first top bit
  const uint64_t nm = 1ull << (S - CB);      // blocks: one per value of bits CB..S-1
  const uint64_t b = blockIdx.x;
  const uint64_t m = nm >= 8 ? ((b & 7ull) * (nm >> 3) + (b >> 3)) : b;
  const uint64_t c = (uint64_t)(u & ((1 << CB) - 1));
  const uint64_t tlo = (uint64_t)(u >> CB);
  const uint64_t ebase = (tlo << S) | (m << CB) | c;
  const uint64_t jstride = 1ull << (S + TU);
  // stage the scaled x of all top-bit combinations in LDS; the pass-A
  // partials are loaded in the same burst (RYD_TOP_PREFETCH), so that every
  // read of the block is in flight before the barrier (n = 28: FP64
  // 2.65 -> 2.53 ms, FP32 1.33 -> 1.28 ms per H·v)
  constexpr bool PRE = RYD_TOP_PREFETCH && J * sizeof(V) <= 128 && (!VT<V>::cplx || J <= 8);
  constexpr bool PREZ = PRE && RYD_TOP_ZPRE && sizeof(V) == 4;
  V pa[PRE ? J : 1];
  V zv[PREZ ? J : 1];
  {
    V xv[J];
#pragma unroll
    for (int j = 0; j < J; ++j) xv[j] = top_load(io.x + ebase + j * jstride);
    if constexpr (PRE) {
#pragma unroll
      for (int j = 0; j < J; ++j) pa[j] = top_load(io.y + ebase + j * jstride);
    }
    if constexpr (PREZ) {
      if (io.z) {
#pragma unroll
        for (int j = 0; j < J; ++j) zv[j] = top_load(io.z + ebase + j * jstride);
      }
    }
#pragma unroll
    for (int j = 0; j < J; ++j) lds[u + NU * j] = vscale(xv[j], sr);
  }
  __syncthreads();
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  const RT chr = (RT)ch, cxr = (RT)cx, czr = (RT)cz;
#pragma unroll
  for (int j = 0; j < J; ++j) {
    const uint64_t e = ebase + j * jstride;
    V a;
    if constexpr (PRE) a = pa[j];
    else a = io.y[e];                        // pass-A partial
#pragma unroll
    for (int k = 0; k < TU; ++k) {           // partner in another thread's column
      const int p = S + k;
      const RT ai = (RT)M.ai[p];
      const int bit = (int)((tlo >> k) & 1ull);
      vcfma(a, (RT)M.ar[p], bit ? ai : -ai, lds[(u ^ (1 << (CB + k))) + NU * j]);
    }
#pragma unroll
    for (int k = 0; k < TOPB - TU; ++k) {    // partner in another register slot
      const int p = S + TU + k;
      const RT ai = (RT)M.ai[p];
      const int bit = (j >> k) & 1;
      vcfma(a, (RT)M.ar[p], bit ? ai : -ai, lds[u + NU * (j ^ (1 << k))]);
    }
    const V xj = lds[u + NU * j];
    V v = vscale(a, chr);
    vfma(v, cxr, xj);
    if (io.z) {
      if constexpr (PREZ) vfma(v, czr, zv[j]);
      else vfma(v, czr, top_load(io.z + e));
    }
#if RYD_TOP_NT
    nt_store(io.y + e, v);
#else
    io.y[e] = v;
#endif
    if (io.reduce) {
      s0 += rdot(xj, v);
      s1 += nrm2(v);
      s2 += nrm2(xj);
      if (io.probe) s3 += zdot(((io.probe[e >> 5] >> (e & 31)) & 1u) ? 1.0 : -1.0, v);
    }
  }
  if (io.reduce) block_reduce_store(s0, s1, s2, s3, rd.partials);
}

// n below the tile size: one element per thread, partners from global memory.
template <class V>
__global__ __launch_bounds__(256) void k_full_small(FullModel M, ApplyIO<V> io, Reduce rd) {
  using RT = typename VT<V>::R;
  if (io.stop && load_coherent(io.stop)) return;
  double s, ch, cx, cz;
  apply_scalars(io, s, ch, cx, cz);
  const uint64_t D = 1ull << M.n;
  const uint64_t g = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  const bool ok = g < D;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  if (ok) {
    const V xe = vscale(io.x[g], (RT)s);
    const uint64_t lo = g & ((1ull << M.tb) - 1);
    const uint64_t hi = g >> M.tb;
    V acc = vscale(xe, (RT)(M.elo[lo] + M.ehi[hi]));
    // cross terms between low and high parts (only when tb < n)
    if (M.tb < M.n) {
      double dc = 0.0;
      for (int p = 0; p < M.tb; ++p) {
        if (!((g >> p) & 1ull)) continue;
        for (int q = M.tb; q < M.n; ++q)
          if ((g >> q) & 1ull) dc += M.vb[p * M.n + q];
      }
      vfma(acc, (RT)dc, xe);
    }
    for (int p = 0; p < M.n; ++p) {
      const int bit = (int)((g >> p) & 1ull);
      vcfma(acc, (RT)(M.ar[p] * s), (RT)((bit ? M.ai[p] : -M.ai[p]) * s), io.x[g ^ (1ull << p)]);
    }
    V v = vscale(acc, (RT)ch);
    vfma(v, (RT)cx, xe);
    if (io.z) vfma(v, (RT)cz, io.z[g]);
    io.y[g] = v;
    if (io.reduce) {
      s0 = rdot(xe, v);
      s1 = nrm2(v);
      s2 = nrm2(xe);
      if (io.probe) s3 = zdot(((io.probe[g >> 5] >> (g & 31)) & 1u) ? 1.0 : -1.0, v);
    }
  }
  if (io.reduce) block_reduce_store(s0, s1, s2, s3, rd.partials);
}

}  // namespace ryd
