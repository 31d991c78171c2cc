// Observables of a state on any plan's basis: Pauli-string expectations and
// the site correlation sums Σ_s |x_s|^2 b_i(s) b_j(s).
//
// Pauli strings (tomisnot/energy-level src/core/manybody/observables.py:486-622,
// the basis path of PauliPlan): a string is X, Y, Z site sets as bit masks;
// the source row k with code c contributes
//     conj(x[rank(c ^ (X|Y))]) · i^{|Y|} · (-1)^{popcount(c & (Y|Z))} · x[k],
// (σ^z: -1 on |r>; σ^y: +i from |g>, -i from |r>; σ^x: 1), targets outside
// the basis contribute nothing.  Correlations (observables.py:40-47,
// corr_matrix): one wave takes 64 rows, ballots one 64-bit lane mask per
// site, and every lane sums the weights |x|^2 under mask_i & mask_j for its
// own pairs (i <= j); no atomics, fixed row chunks per block, so both sums
// are deterministic.
#pragma once
#include "common.hpp"
#include "channels.hpp"

namespace ryd {

constexpr int PAULI_BATCH = 8;                 // strings per pass
constexpr int CORR_PPL = 33;                   // pairs per lane (n <= 64: 2080 pairs)

struct ObsBasis {
  int kind;                    // 0 identity (full basis), 1 code tables, 2 chain sector
  int n, mode, m, gl;
  int64_t dim, tl;
  const int64_t* codes;
  const int64_t* binom;
  int rsL;
  const int64_t* rsplit;
  const uint64_t* hpk;
  const uint32_t* lo_idx;
  const uint32_t* fc;
  const uint32_t* pre;
  int64_t W[MAXN];             // chain sector: W[p] = F(p + 2)
};

__device__ inline uint64_t obs_code(const ObsBasis& B, int64_t k) {
  if (B.kind == 0) return (uint64_t)k;
  if (B.kind == 1) return chan_code(B, k);
  uint64_t code = 0;
  int prev = 0;
  for (int p = B.n - 1; p >= 0; --p) {
    if (!prev && k >= B.W[p]) {
      code |= 1ull << p;
      k -= B.W[p];
      prev = 1;
    } else {
      prev = 0;
    }
  }
  return code;
}

__device__ inline int64_t obs_rank(const ObsBasis& B, uint64_t s) {
  if (B.n < 64 && (s >> B.n)) return -1;
  if (B.kind == 0) return (int64_t)s;
  if (B.kind == 1) return chan_rank(B, s);
  if (s & (s >> 1)) return -1;
  int64_t r = 0;
  for (uint64_t b = s; b; b &= b - 1) r += B.W[__builtin_ctzll(b)];
  return r;
}

// conj(a) * b
__device__ __forceinline__ void cdot_acc(float a, float b, double& re, double& im) {
  re += (double)a * b;
}
__device__ __forceinline__ void cdot_acc(double a, double b, double& re, double& im) {
  re += a * b;
}
__device__ __forceinline__ void cdot_acc(c32 a, c32 b, double& re, double& im) {
  re += (double)a.re * b.re + (double)a.im * b.im;
  im += (double)a.re * b.im - (double)a.im * b.re;
}
__device__ __forceinline__ void cdot_acc(c64 a, c64 b, double& re, double& im) {
  re += a.re * b.re + a.im * b.im;
  im += a.re * b.im - a.im * b.re;
}

struct PauliBatch {
  int count;                              // strings in this pass (<= PAULI_BATCH)
  uint64_t flip[PAULI_BATCH];             // X | Y by bit position
  uint64_t sign[PAULI_BATCH];             // Y | Z by bit position
};

// partials[block][2 * PAULI_BATCH] = (re, im) of Σ conj(x[t]) (±1) x[k] per string
template <class V>
__global__ __launch_bounds__(256) void k_pauli(ObsBasis B, PauliBatch S, const V* __restrict__ x,
                                               double* __restrict__ partials) {
  __shared__ double red[4][2 * PAULI_BATCH];
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  const int64_t per = (B.dim + gridDim.x - 1) / gridDim.x;
  const int64_t k0 = (int64_t)blockIdx.x * per;
  const int64_t k1 = k0 + per < B.dim ? k0 + per : B.dim;
  double re[PAULI_BATCH], im[PAULI_BATCH];
#pragma unroll
  for (int q = 0; q < PAULI_BATCH; ++q) re[q] = im[q] = 0.0;
  for (int64_t k = k0 + t; k < k1; k += 256) {
    const uint64_t c = obs_code(B, k);
    const V xk = x[k];
#pragma unroll
    for (int q = 0; q < PAULI_BATCH; ++q) {
      if (q >= S.count) break;
      const int64_t r = obs_rank(B, c ^ S.flip[q]);
      if (r < 0) continue;
      double a = 0.0, b = 0.0;
      cdot_acc(x[r], xk, a, b);
      const double sg = (__popcll(c & S.sign[q]) & 1) ? -1.0 : 1.0;
      re[q] += sg * a;
      im[q] += sg * b;
    }
  }
#pragma unroll
  for (int q = 0; q < PAULI_BATCH; ++q) {
    const double a = wave_sum(re[q]), b = wave_sum(im[q]);
    if (lane == 0) {
      red[wid][2 * q] = a;
      red[wid][2 * q + 1] = b;
    }
  }
  __syncthreads();
  if (t < 2 * PAULI_BATCH)
    partials[(size_t)blockIdx.x * 2 * PAULI_BATCH + t] = red[0][t] + red[1][t] + red[2][t] + red[3][t];
}

// partials[block][npairs]: Σ |x|^2 b_i b_j over the block's rows, pairs
// (i <= j) in row-major upper-triangle order; partials[block][npairs] = Σ |x|^2
template <class V>
__global__ __launch_bounds__(256) void k_corr(ObsBasis B, const V* __restrict__ x,
                                              double* __restrict__ partials) {
  __shared__ double wl[4][64];
  __shared__ unsigned long long msk[4][MAXN];
  __shared__ double red[4][CORR_PPL * 64 + 1];
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  const int n = B.n;
  const int npairs = n * (n + 1) / 2;
  const int64_t per = (B.dim + gridDim.x - 1) / gridDim.x;
  const int64_t k0 = (int64_t)blockIdx.x * per;
  const int64_t k1 = k0 + per < B.dim ? k0 + per : B.dim;
  // this lane's pairs: q = lane + 64·u
  int pi[CORR_PPL], pj[CORR_PPL];
#pragma unroll
  for (int u = 0; u < CORR_PPL; ++u) {
    int q = lane + 64 * u, i = 0;
    if (q >= npairs) { pi[u] = -1; pj[u] = -1; continue; }
    while (q >= n - i) { q -= n - i; ++i; }
    pi[u] = i;
    pj[u] = i + q;
  }
  double acc[CORR_PPL];
#pragma unroll
  for (int u = 0; u < CORR_PPL; ++u) acc[u] = 0.0;
  double tot = 0.0;
  for (int64_t b0 = k0; b0 < k1; b0 += 256) {        // same trip count in every wave
    const int64_t k = b0 + wid * 64 + lane;
    const bool in = k < k1;
    const uint64_t c = in ? obs_code(B, k) : 0ull;
    const double w = in ? nrm2(x[k]) : 0.0;
    tot += w;
    wl[wid][lane] = w;
    for (int i = 0; i < n; ++i) {
      const unsigned long long b = __ballot((c >> (n - 1 - i)) & 1ull);
      if (lane == 0) msk[wid][i] = b;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < CORR_PPL; ++u) {
      if (pi[u] < 0) break;
      unsigned long long mm = msk[wid][pi[u]] & msk[wid][pj[u]];
      double s = 0.0;
      for (; mm; mm &= mm - 1) s += wl[wid][__builtin_ctzll(mm)];
      acc[u] += s;
    }
    __syncthreads();
  }
  // the 4 waves' sums in fixed order
#pragma unroll
  for (int u = 0; u < CORR_PPL; ++u) red[wid][lane + 64 * u] = acc[u];
  tot = wave_sum(tot);
  if (lane == 0) red[wid][CORR_PPL * 64] = tot;
  __syncthreads();
  double* out = partials + (size_t)blockIdx.x * (npairs + 1);
  for (int q = t; q < npairs; q += 256) out[q] = red[0][q] + red[1][q] + red[2][q] + red[3][q];
  if (t == 0)
    out[npairs] = red[0][CORR_PPL * 64] + red[1][CORR_PPL * 64] + red[2][CORR_PPL * 64] +
                  red[3][CORR_PPL * 64];
}

// out[c] = Σ_b partials[b][c] in block order (one thread per column)
__global__ __launch_bounds__(256) void k_sum_columns(const double* __restrict__ partials, int nblocks,
                                                     int width, double* __restrict__ out) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < width; c += gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int b = 0; b < nblocks; ++b) s += partials[(size_t)b * width + c];
    out[c] = s;
  }
}

// out[i*n + j] (both triangles) and out[n*n] from the upper-triangle sums
__global__ __launch_bounds__(256) void k_corr_square(const double* __restrict__ sums, int n,
                                                     double* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k == n * n) {
    out[k] = sums[n * (n + 1) / 2];
    return;
  }
  if (k > n * n) return;
  int i = k / n, j = k % n;
  if (i > j) { const int t = i; i = j; j = t; }
  // upper-triangle index of (i, j), rows i' < i hold n - i' pairs each
  const int q = i * n - i * (i - 1) / 2 + (j - i);
  out[k] = sums[q];
}

}  // namespace ryd
