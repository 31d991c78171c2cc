// Lanczos runs of a small full-basis plan (D = 2^n <= 8192, complex128
// 4096) in one
// workgroup: the recurrence of do_lanczos (apply with alpha fused, orth with
// the norm, the breakdown test, the replay that accumulates Ritz vectors),
// every step of a call in one launch.  u_j sits in LDS for the bit flips;
// each thread keeps its own elements of u_{j-1} and w in registers; alpha
// and |u_{j+1}| are block sums formed in the same fixed order on every
// thread, so no value goes through global memory inside the launch.  On
// return the rotating buffers hold the two live vectors as the multi-launch
// form leaves them (u_j in b[j % 3] for the last step j and u_{j+1}, or the
// residual that broke down, in b[(j+1) % 3]), so a later call (continuation
// or replay, any path) finds the same state; the third buffer is scratch.
// Below this size a step of the multi-launch form is four launches of a few
// microseconds each around ~50 KB of work.
#pragma once
#include "common.hpp"
#include "hv_full.hpp"

namespace ryd {

constexpr int LZS_THREADS = 1024;
constexpr int LZS_WAVES = LZS_THREADS / 64;
// elements per thread: 8 (D up to 2^13) for 4- and 8-byte elements, 4 (2^12)
// for complex128, whose 8 would spill at the 128 VGPRs of a 1024-thread block
template <class V> constexpr int lzs_maxe() { return sizeof(V) == 16 ? 4 : 8; }
template <class V> constexpr int lzs_maxn() { return sizeof(V) == 16 ? 12 : 13; }

// Block sum of v, the same value on every thread (waves summed in order).
// Callers alternate two `red` slots between consecutive sums: a slot is
// written again only after a barrier that every reader of its previous
// value has passed.
__device__ __forceinline__ double lzs_sum(double v, double* red) {
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int k = 0; k < LZS_WAVES; ++k) t += red[k];
  return t;
}

template <class V>
__device__ __forceinline__ V* lzs_buf(V* b0, V* b1, V* b2, int64_t j) {
  const int r = (int)(j % 3);
  return r == 0 ? b0 : (r == 1 ? b1 : b2);
}

// the live vectors after step j: u_j (uj) to b[j % 3], u_{j+1} (w) to b[(j+1) % 3]
template <class V, int E>
__device__ __forceinline__ void lzs_store(V* b0, V* b1, V* b2, int64_t j, const V (&uj)[E],
                                          const V (&w)[E], int D) {
  V* pu = lzs_buf(b0, b1, b2, j);
  V* pw = lzs_buf(b0, b1, b2, j + 1);
#pragma unroll
  for (int k = 0; k < E; ++k) {
    const int g = (int)threadIdx.x + k * LZS_THREADS;
    if (g < D) {
      pu[g] = uj[k];
      pw[g] = w[k];
    }
  }
}

// basis != nullptr: the stored form (do_lanczos_stored), u_j in row
// basis + (j - jbase)·ld and every u_{j+1} written to its row; b0..b2 unused.
template <class V>
__global__ __launch_bounds__(LZS_THREADS) void k_lz_small(FullModel M, V* b0, V* b1, V* b2,
                                                          int64_t j0, int64_t steps,
                                                          double* alpha, double* nrm, LzState* S,
                                                          double tiny, V* ritz0, V* ritz1,
                                                          const double* y0, const double* y1,
                                                          int replay, V* basis, int64_t ld,
                                                          int64_t jbase) {
  using RT = typename VT<V>::R;
  constexpr int LZS_MAXE = lzs_maxe<V>();
  __shared__ V xs[LZS_THREADS * LZS_MAXE];
  __shared__ double dgs[LZS_THREADS * LZS_MAXE];    // diagonal (LDS: registers hold w, u_{j-1})
  __shared__ double redA[LZS_WAVES], redB[LZS_WAVES];
  const int n = M.n;
  const int D = 1 << n;
  const int tid = threadIdx.x;
  const int64_t jend = j0 + steps;
  auto row = [&](int64_t j) -> V* { return basis + (j - jbase) * ld; };
  double nj, njm;
  if (!replay) {
    if (j0 == 0) {
      // start vector: nrm[0] = |u_0|, stop when it is not above tiny
      const V* u0 = basis ? row(0) : b0;
      double nn = 0.0;
#pragma unroll
      for (int k = 0; k < LZS_MAXE; ++k) {
        const int g = tid + k * LZS_THREADS;
        if (g < D) nn += nrm2(u0[g]);
      }
      nn = lzs_sum(nn, redB);
      const double nv = sqrt(nn > 0.0 ? nn : 0.0);
      const bool brk = !(nv > tiny);
      if (tid == 0) {
        S->tiny = tiny;
        S->stop = brk ? 1 : 0;
        S->steps = 0;
        nrm[0] = nv;
      }
      if (brk) return;
      nj = nv;
    } else {
      if (S->stop) return;
      tiny = S->tiny;
      nj = nrm[j0];
    }
  } else {
    nj = nrm[j0];
  }
  njm = j0 > 0 ? nrm[j0 - 1] : 1.0;
  if (steps <= 0) return;

  // diagonal of the owned elements (the tables and cross terms of k_full_small)
  V um[LZS_MAXE], w[LZS_MAXE];
  {
    const V* uj = basis ? row(j0) : lzs_buf(b0, b1, b2, j0);
    const V* ujm = basis ? (j0 > 0 ? row(j0 - 1) : nullptr) : lzs_buf(b0, b1, b2, j0 + 2);
    const uint64_t lmask = (1ull << M.tb) - 1;
#pragma unroll
    for (int k = 0; k < LZS_MAXE; ++k) {
      const int g = tid + k * LZS_THREADS;
      um[k] = vzero(V{});
      w[k] = vzero(V{});
      if (g < D) {
        double d = M.elo[(uint64_t)g & lmask] + M.ehi[(uint64_t)g >> M.tb];
        for (int p = 0; p < M.tb && p < n; ++p) {
          if (!((g >> p) & 1)) continue;
          for (int q = M.tb; q < n; ++q)
            if ((g >> q) & 1) d += M.vb[p * n + q];
        }
        dgs[g] = d;
        xs[g] = uj[g];
        if (j0 > 0) um[k] = ujm[g];
      }
    }
  }
  __syncthreads();

  for (int64_t j = j0; j < jend; ++j) {
    const double s = 1.0 / nj;
    const RT sr = (RT)s;
    const RT cz = (RT)(j > 0 ? -nj / njm : 0.0);
    // w = H (u_j / nrm_j) - (nrm_j / nrm_{j-1}) u_{j-1};  alpha_j = <u_j/nrm_j, w>
    // (flips summed unscaled, every LDS read of an element issued together)
    double sa = 0.0;
#pragma unroll
    for (int k = 0; k < LZS_MAXE; ++k) {
      const int g = tid + k * LZS_THREADS;
      if (g >= D) continue;
      const V xe = vscale(xs[g], sr);
      V fl = vzero(V{});
#pragma unroll
      for (int p = 0; p < lzs_maxn<V>(); ++p) {
        if (p >= n) break;
        const int bit = (g >> p) & 1;
        vcfma(fl, (RT)M.ar[p], (RT)(bit ? M.ai[p] : -M.ai[p]), xs[g ^ (1 << p)]);
      }
      V acc = vscale(xe, (RT)dgs[g]);
      vfma(acc, sr, fl);
      if (j > 0) vfma(acc, cz, um[k]);
      w[k] = acc;
      sa += rdot(xe, acc);
    }
    double a;
    if (!replay) {
      a = lzs_sum(sa, redA);                       // also: every flip read of xs is done
      if (tid == 0) alpha[j] = a;
    } else {
      __syncthreads();
      a = alpha[j];
    }
    // u_{j+1} = w - (alpha_j / nrm_j) u_j;  Ritz vectors += (y[j] / nrm_j) u_j
    const RT coef = (RT)(-a * s);
    const RT r0c = ritz0 ? (RT)(y0[j] * s) : (RT)0;
    const RT r1c = ritz1 ? (RT)(y1[j] * s) : (RT)0;
    V* wrow = basis ? row(j + 1) : nullptr;
    double sn = 0.0;
#pragma unroll
    for (int k = 0; k < LZS_MAXE; ++k) {
      const int g = tid + k * LZS_THREADS;
      if (g >= D) continue;
      const V u = xs[g];
      vfma(w[k], coef, u);
      sn += nrm2(w[k]);
      if (ritz0) { V r = ritz0[g]; vfma(r, r0c, u); ritz0[g] = r; }
      if (ritz1) { V r = ritz1[g]; vfma(r, r1c, u); ritz1[g] = r; }
      if (wrow) wrow[g] = w[k];
      um[k] = u;
      xs[g] = w[k];                                // only this thread reads xs[g] now
    }
    njm = nj;
    if (!replay) {
      const double t = lzs_sum(sn, redB);          // also publishes xs for the next step
      const double nv = sqrt(t > 0.0 ? t : 0.0);
      const bool brk = !(nv > tiny);               // also catches NaN
      if (tid == 0) {
        nrm[j + 1] = nv;
        if (brk) {
          S->stop = 1;
          S->steps = (int)(j + 1);
        }
      }
      if (brk || j + 1 == jend) {
        if (!basis) lzs_store(b0, b1, b2, j, um, w, D);
        return;
      }
      nj = nv;
    } else {
      __syncthreads();
      nj = nrm[j + 1];
      if (j + 1 == jend && !basis) lzs_store(b0, b1, b2, j, um, w, D);
    }
  }
}

}  // namespace ryd
