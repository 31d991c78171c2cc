// Independent-set bases of a blockade graph, enumerated and ranked on the
// device.
//
// The basis is every code with no two excited sites joined by an edge, in
// ascending order: the order of tomisnot/energy-level
// src/core/manybody/basis.py:24-47 (site i is bit n-1-i; the depth-first
// search visits "site empty" before "site excited", which is ascending).
//
// Split enumeration.  A code is (h, l): its H high bits and L low bits.
//   * low sets  : the independent sets of the graph induced on the low bits,
//                 ascending (lowtab[T_L]) with lo_idx[l] = index or NONE;
//   * classes   : the boundary sites are the high sites with a neighbour
//                 among the low bits; a class is an independent pattern of
//                 boundary sites, and F_c is the set of low bits they block;
//   * high parts: h is valid when it is independent on the high bits; its
//                 class is its boundary pattern and its segment holds the
//                 low sets l with l & F_c = 0, in ascending order;
//   * offsets   : hpk[h] = (first rank of h) << 24 | class, class = NONE for
//                 an invalid h; hpk[2^H] carries the dimension.
// With the per-class prefix table pre[c][t] = #{t' < t : lowtab[t'] & F_c = 0}
// the rank of a code is hpk[h] >> 24 + pre[c][lo_idx[l]]: four table loads,
// no search.  The code list itself is written once (codes[D]) for the
// row-ordered kernels.
#pragma once
#include "common.hpp"

namespace ryd {

constexpr int GR_MAXB = 30;                    // bits per part (L, H <= 30)
constexpr uint32_t GR_NONE = 0xffffffffu;
constexpr uint64_t GR_CLS = 0xffffffull;       // class field of hpk; all ones = invalid h
constexpr int GR_ITEMS = 16;                   // scan: elements per thread
constexpr int64_t GR_TILE = 256 * GR_ITEMS;    // scan: elements per block

struct GraphParts {
  int L, H, k;                    // low bits, high bits, boundary sites
  uint32_t nb_lo[GR_MAXB];        // low bit q: its neighbours among the low bits
  uint32_t nb_hi[GR_MAXB];        // high bit p (high coordinates): neighbours among high bits
  uint32_t hb_pos[GR_MAXB];       // boundary site j: its high-coordinate bit
  uint32_t nb_cls[GR_MAXB];       // boundary site j: neighbours among boundary sites
  uint32_t hb_low[GR_MAXB];       // boundary site j: the low bits it blocks
};

__device__ __forceinline__ bool gr_indep(uint32_t x, const uint32_t* nb) {
  for (uint32_t y = x; y; y &= y - 1)
    if (x & nb[__builtin_ctz(y)]) return false;
  return true;
}

__device__ __forceinline__ uint32_t gr_pattern(const GraphParts& G, uint32_t h) {
  uint32_t p = 0;
  for (int j = 0; j < G.k; ++j) p |= ((h >> G.hb_pos[j]) & 1u) << j;
  return p;
}

// Exclusive prefix of v over a 256-thread block; total of the block in `total`.
__device__ inline int64_t gr_block_scan(int64_t v, int64_t* lds, int64_t& total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  long long x = (long long)v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const long long y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) lds[w] = x;
  __syncthreads();
  int64_t off = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int64_t c = lds[i];
    if (i < w) off += c;
    tot += c;
  }
  __syncthreads();
  total = tot;
  return off + (int64_t)x - v;
}

// Position of a set flag among the block's set flags (256 threads).
__device__ inline uint32_t gr_flag_scan(bool f, uint32_t* lds, uint32_t& total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const unsigned long long bal = __ballot(f);
  const uint32_t below = (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
  if (lane == 0) lds[w] = (uint32_t)__popcll(bal);
  __syncthreads();
  uint32_t off = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const uint32_t c = lds[i];
    if (i < w) off += c;
    tot += c;
  }
  __syncthreads();
  total = tot;
  return off + below;
}

// ---- three-phase exclusive scan over [0, count) of a value functor --------
// phase 1: per-block sums; phase 2: one block scans them (bsum[nb] = total);
// phase 3: each block rescans its tile and hands (i, prefix, value) to a writer.
template <class F>
__global__ __launch_bounds__(256) void k_gr_scan_sum(F f, int64_t count, int64_t* bsum) {
  __shared__ int64_t lds[4];
  const int64_t base = (int64_t)blockIdx.x * GR_TILE + (int64_t)threadIdx.x * GR_ITEMS;
  int64_t s = 0;
#pragma unroll
  for (int q = 0; q < GR_ITEMS; ++q)
    if (base + q < count) s += f(base + q);
  int64_t tot;
  gr_block_scan(s, lds, tot);
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ __launch_bounds__(256) void k_gr_scan_blocks(int64_t* bsum, int64_t nb) {
  __shared__ int64_t lds[4];
  int64_t carry = 0;
  for (int64_t c = 0; c < nb; c += 256) {
    const int64_t i = c + threadIdx.x;
    const int64_t v = i < nb ? bsum[i] : 0;
    int64_t tot;
    const int64_t ex = gr_block_scan(v, lds, tot);
    if (i < nb) bsum[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) bsum[nb] = carry;
}

template <class F, class W>
__global__ __launch_bounds__(256) void k_gr_scan_write(F f, W w, int64_t count,
                                                       const int64_t* __restrict__ bsum) {
  __shared__ int64_t lds[4];
  const int64_t base = (int64_t)blockIdx.x * GR_TILE + (int64_t)threadIdx.x * GR_ITEMS;
  int64_t v[GR_ITEMS];
  int64_t s = 0;
#pragma unroll
  for (int q = 0; q < GR_ITEMS; ++q) {
    v[q] = base + q < count ? f(base + q) : 0;
    s += v[q];
  }
  int64_t tot;
  int64_t pre = bsum[blockIdx.x] + gr_block_scan(s, lds, tot);
#pragma unroll
  for (int q = 0; q < GR_ITEMS; ++q) {
    if (base + q < count) w(base + q, pre, v[q]);
    pre += v[q];
  }
}

// ---- functors ----------------------------------------------------------------
struct GrLowFlag {             // independent on the low bits
  GraphParts G;
  __device__ int64_t operator()(int64_t i) const { return gr_indep((uint32_t)i, G.nb_lo) ? 1 : 0; }
};
struct GrClassFlag {           // independent boundary pattern
  GraphParts G;
  __device__ int64_t operator()(int64_t i) const { return gr_indep((uint32_t)i, G.nb_cls) ? 1 : 0; }
};
struct GrIndexWrite {          // idx[i] = position among set flags or NONE; list[pos] = i
  uint32_t* idx;
  uint32_t* list;              // may be null in the counting pass
  __device__ void operator()(int64_t i, int64_t pre, int64_t v) const {
    idx[i] = v ? (uint32_t)pre : GR_NONE;
    if (v && list) list[pre] = (uint32_t)i;
  }
};
struct GrHighCount {           // rows of high part h (0 when h is not independent)
  GraphParts G;
  const uint32_t* cls_idx;
  const int64_t* ccnt;
  __device__ int64_t operator()(int64_t h) const {
    if (!gr_indep((uint32_t)h, G.nb_hi)) return 0;
    return ccnt[cls_idx[gr_pattern(G, (uint32_t)h)]];
  }
};
struct GrHighWrite {
  GraphParts G;
  const uint32_t* cls_idx;
  uint64_t* hpk;
  __device__ void operator()(int64_t h, int64_t pre, int64_t v) const {
    // every valid h has at least one row (the empty low set)
    const uint64_t c = v ? (uint64_t)cls_idx[gr_pattern(G, (uint32_t)h)] : GR_CLS;
    hpk[h] = ((uint64_t)pre << 24) | c;
  }
};

// fc[c] = low bits blocked by class c
__global__ __launch_bounds__(256) void k_gr_class_mask(GraphParts G, const uint32_t* __restrict__ clslist,
                                                       int64_t nc, uint32_t* __restrict__ fc) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += stride) {
    uint32_t f = 0;
    for (uint32_t p = clslist[c]; p; p &= p - 1) f |= G.hb_low[__builtin_ctz(p)];
    fc[c] = f;
  }
}

// ccnt[c] = #{t : lowtab[t] & fc[c] = 0}, one block per class (grid-stride)
__global__ __launch_bounds__(256) void k_gr_class_count(const uint32_t* __restrict__ lowtab, int64_t tl,
                                                        const uint32_t* __restrict__ fc, int64_t nc,
                                                        int64_t* __restrict__ ccnt) {
  __shared__ int64_t lds[4];
  for (int64_t c = blockIdx.x; c < nc; c += gridDim.x) {
    const uint32_t F = fc[c];
    int64_t s = 0;
    for (int64_t t = threadIdx.x; t < tl; t += 256) s += (lowtab[t] & F) ? 0 : 1;
    int64_t tot;
    gr_block_scan(s, lds, tot);
    if (threadIdx.x == 0) ccnt[c] = tot;
  }
}

// pre[c][t] (row length tl + 1): compatible low sets before t; one block per class
__global__ __launch_bounds__(256) void k_gr_prefix(const uint32_t* __restrict__ lowtab, int64_t tl,
                                                   const uint32_t* __restrict__ fc, int64_t nc,
                                                   uint32_t* __restrict__ pre) {
  __shared__ uint32_t lds[4];
  for (int64_t c = blockIdx.x; c < nc; c += gridDim.x) {
    const uint32_t F = fc[c];
    uint32_t* row = pre + (size_t)c * (size_t)(tl + 1);
    uint32_t run = 0;
    for (int64_t t0 = 0; t0 < tl; t0 += 256) {
      const int64_t t = t0 + threadIdx.x;
      const bool ok = t < tl && !(lowtab[t] & F);
      uint32_t tot;
      const uint32_t pos = gr_flag_scan(ok, lds, tot);
      if (t < tl) row[t] = run + pos;
      run += tot;
    }
    if (threadIdx.x == 0) row[tl] = run;
  }
}

// codes of every valid high part, in rank order; one block per h (grid-stride)
__global__ __launch_bounds__(256) void k_gr_emit(int L, const uint64_t* __restrict__ hpk, int64_t nh,
                                                 const uint32_t* __restrict__ lowtab, int64_t tl,
                                                 const uint32_t* __restrict__ fc,
                                                 int64_t* __restrict__ codes) {
  __shared__ uint32_t lds[4];
  for (int64_t h = blockIdx.x; h < nh; h += gridDim.x) {
    const uint64_t e = hpk[h];
    const uint64_t c = e & GR_CLS;
    if (c == GR_CLS) continue;
    const uint32_t F = fc[c];
    int64_t* out = codes + (int64_t)(e >> 24);
    const int64_t hc = h << L;
    uint32_t run = 0;
    for (int64_t t0 = 0; t0 < tl; t0 += 256) {
      const int64_t t = t0 + threadIdx.x;
      const uint32_t l = t < tl ? lowtab[t] : 0u;
      const bool ok = t < tl && !(l & F);
      uint32_t tot;
      const uint32_t pos = gr_flag_scan(ok, lds, tot);
      if (ok) out[run + pos] = hc | (int64_t)l;
      run += tot;
    }
  }
}

__global__ void k_gr_seal(uint64_t* hpk, int64_t nh, const int64_t* total) {
  hpk[nh] = ((uint64_t)*total << 24) | GR_CLS;
}

}  // namespace ryd
