// Site occupations <n_p> = Σ_e |x_e|^2 b_p(e) on the full basis, per bit
// position p, plus the total weight Σ|x|^2 (reference
// src/core/manybody/eq_route.py:108-122, observables.py:22-47).
//
// A block sweeps tiles of 2^OB codes: bits below OB are summed per thread
// in registers, bits at or above OB are constant over a tile and receive the
// tile's total weight.  Per-block partials [nblocks][MAXN+1] are summed by a
// second one-block kernel in fixed order.
#pragma once
#include "common.hpp"

namespace ryd {

constexpr int OB = 10;         // tile bits of the occupation sweep (1024 codes, 4 per thread)
constexpr int MAXN_OCC = 64;   // partial-sum columns per block (bit positions), then the total
constexpr int OCC_BLOCKS = 2048;

template <class V>
__global__ __launch_bounds__(256) void k_site_occ(const V* __restrict__ x, int n, double* partials) {
  __shared__ double acc_hi[64];
  __shared__ double red[4][OB + 1];
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  const int ob = n < OB ? n : OB;
  const uint64_t D = 1ull << n;
  const uint64_t tsize = 1ull << ob;
  const uint64_t ntiles = D >> ob;
  if (t < 64) acc_hi[t] = 0.0;
  double lo[OB + 1];
#pragma unroll
  for (int p = 0; p <= OB; ++p) lo[p] = 0.0;
  __syncthreads();
  for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    double tot = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint64_t e = (uint64_t)t * 4 + k;
      if (e < tsize) {
        const double w = nrm2(x[(tile << ob) + e]);
        tot += w;
#pragma unroll
        for (int p = 0; p < OB; ++p)
          if (p < ob && ((e >> p) & 1ull)) lo[p] += w;
      }
    }
    lo[OB] += tot;
    // tile total for the high bits
    tot = wave_sum(tot);
    if (lane == 0) red[wid][OB] = tot;
    __syncthreads();
    if (t == 0) {
      const double T = red[0][OB] + red[1][OB] + red[2][OB] + red[3][OB];
      for (int q = ob; q < n; ++q)
        if ((tile >> (q - ob)) & 1ull) acc_hi[q] += T;
    }
    __syncthreads();
  }
  // reduce the per-thread low-bit sums
#pragma unroll
  for (int p = 0; p <= OB; ++p) {
    const double v = wave_sum(lo[p]);
    if (lane == 0) red[wid][p] = v;
  }
  __syncthreads();
  double* out = partials + (size_t)blockIdx.x * (MAXN_OCC + 1);
  if (t <= OB) {
    const double v = red[0][t] + red[1][t] + red[2][t] + red[3][t];
    if (t < OB) { if (t < ob) out[t] = v; }
    else out[MAXN_OCC] = v;
  }
  if (t >= ob && t < n) out[t] = acc_hi[t];
}

__global__ __launch_bounds__(256) void k_occ_finish(const double* partials, int nblocks, int n,
                                                    double* out) {
  // out[0..n-1] per site, out[n] = total weight
  __shared__ double red[4];
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  for (int p = 0; p <= n; ++p) {
    const int col = p < n ? p : MAXN_OCC;
    double s = 0.0;
    for (int b = t; b < nblocks; b += blockDim.x) s += partials[(size_t)b * (MAXN_OCC + 1) + col];
    s = wave_sum(s);
    if (lane == 0) red[wid] = s;
    __syncthreads();
    // bit position p is site n-1-p; the total goes last
    if (t == 0) out[p < n ? n - 1 - p : n] = red[0] + red[1] + red[2] + red[3];
    __syncthreads();
  }
}

}  // namespace ryd
