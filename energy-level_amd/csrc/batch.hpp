// Batched Lanczos over many independent full-basis instances of uneven n.
//
// All instances' vectors are concatenated; every launch covers the tiles of
// all instances (a tile is 2^tb consecutive codes of one instance, or the
// whole instance when n < tb), so one Lanczos half-step of the whole batch is
// one launch plus a per-instance finishing launch for the FP64 sums.
//   apply : w = H (u_j / nrm_j) − (nrm_j / nrm_{j−1}) u_{j−1},   alpha_j = <q_j, w>
//   orth  : u_{j+1} = w − (alpha_j / nrm_j) u_j,                  nrm_{j+1} = |u_{j+1}|
// Instances whose recurrence broke down (nrm <= tiny) stop updating.
#pragma once
#include "common.hpp"
#include "hv_full.hpp"

namespace ryd {

constexpr int BMAXN = 24;

struct BatchInst {
  int n;
  int tb;                 // tile bits of this instance (min(n, batch tile bits))
  int ntiles;
  int pad;
  int64_t off;            // offset of the instance's vector in the concatenation
  int64_t tile0;          // first global tile
  const double* elo;      // [2^tb]
  const double* ehi;      // [2^(n-tb)]
  const double* vb;       // [n*n]
  double ar[BMAXN];       // Ω/2 by bit position (real drive)
  int64_t dim;            // codes (rows) of the instance
};

struct BatchLz {
  const BatchInst* inst;
  const int* tile_inst;   // global tile -> instance
  double* alpha;          // [count][ldk]
  double* nrm;            // [count][ldk+1]
  int* stop;              // [count]
  double* partial;        // [total tiles]
  int ldk;
  int j;
  double tiny;
};

// Per-instance reduction: each tile stores its block sum; k_batch_finish (one
// block per instance, next launch) sums the instance's tiles in fixed order and
// stores alpha_j (post 0) or nrm_{j+1} with the breakdown flag (post 1).
__device__ inline void batch_store(const BatchLz& L, const BatchInst& I, int ltile, double v) {
  __shared__ double red[4];
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  v = wave_sum(v);
  if (lane == 0) red[wid] = v;
  __syncthreads();
  if (t == 0) L.partial[I.tile0 + ltile] = red[0] + red[1] + red[2] + red[3];
}

__global__ __launch_bounds__(256) void k_batch_finish(BatchLz L, int post) {
  __shared__ double red[4];
  const int b = blockIdx.x;
  if (L.stop[b]) return;
  const BatchInst& I = L.inst[b];
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  double s = 0.0;
  for (int k = t; k < I.ntiles; k += blockDim.x) s += L.partial[I.tile0 + k];
  s = wave_sum(s);
  if (lane == 0) red[wid] = s;
  __syncthreads();
  if (t == 0) {
    const double tot = red[0] + red[1] + red[2] + red[3];
    if (post == 0) {
      L.alpha[(size_t)b * L.ldk + L.j] = tot;
    } else {
      const double nv = sqrt(tot > 0.0 ? tot : 0.0);
      L.nrm[(size_t)b * (L.ldk + 1) + L.j + 1] = nv;
      if (!(nv > L.tiny)) L.stop[b] = 1;
    }
  }
}

// Blocks go round-robin to the eight XCDs (block b runs on XCD b % 8).  XCD x
// takes the x-th eighth of the global tiles in order, so the partner tiles of
// an instance (bits >= tb) are fetched into one XCD's L2 instead of all
// eight; the last G % 8 tiles keep their own block.
__device__ inline int batch_tile_of(int b, int G) {
  const int per = G >> 3;
  return b < (per << 3) ? (b & 7) * per + (b >> 3) : b;
}

// apply half-step (real element types).  Instances with tb >= 10 use the
// full-basis tile layout (element e = r·1024 + t·4 + c, R = 2^(tb-10) groups):
// bits 0..1 and 10..tb-1 flip in registers, bits 2..9 through LDS, bits >= tb
// from the partner tile in global memory with 4-wide loads.  Smaller
// instances (one tile of 2^n < 1024 codes) take the generic LDS path.
template <class V>
__global__ __launch_bounds__(256) void k_batch_apply(BatchLz L, const V* __restrict__ xall,
                                                     const V* __restrict__ zall, V* __restrict__ yall) {
  __shared__ V lds[4096];
  __shared__ double Fx[16];
  __shared__ double T0[256];
  __shared__ double T1[16];
  const int gt = batch_tile_of(blockIdx.x, gridDim.x);
  const int b = L.tile_inst[gt];
  const BatchInst& I = L.inst[b];
  if (load_coherent(&L.stop[b])) return;
  const int ltile = (int)(gt - I.tile0);
  const int t = threadIdx.x;
  const int n = I.n, tb = I.tb;
  const int64_t base = (int64_t)ltile << tb;
  const V* x = xall + I.off;
  const V* z = zall + I.off;
  V* y = yall + I.off;
  const double* nrm = L.nrm + (size_t)b * (L.ldk + 1);
  const double nj = nrm[L.j];
  const double s = 1.0 / nj;
  const double cz = L.j > 0 ? -nj / nrm[L.j - 1] : 0.0;
  if (t < tb) {
    double f = 0.0;
    const double* row = I.vb + (size_t)t * n;
    for (int q = tb; q < n; ++q)
      if ((ltile >> (q - tb)) & 1) f += row[q];
    Fx[t] = f;
  }
  double sa = 0.0;
  if (tb >= 10) {
    const int R = 1 << (tb - 10);
    V xs[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      if (r < R) {
        V4<V> q = *reinterpret_cast<const V4<V>*>(x + base + r * 1024 + t * 4);
#pragma unroll
        for (int c = 0; c < 4; ++c) xs[r][c] = (V)(q.v[c] * s);
        V4<V> w;
#pragma unroll
        for (int c = 0; c < 4; ++c) w.v[c] = xs[r][c];
        *reinterpret_cast<V4<V>*>(lds + r * 1024 + t * 4) = w;
      } else {
#pragma unroll
        for (int c = 0; c < 4; ++c) xs[r][c] = (V)0;
      }
    }
    __syncthreads();
    {
      double v = 0.0;
#pragma unroll
      for (int p = 0; p < 8; ++p)
        if ((t >> p) & 1) v += Fx[p];
      T0[t] = v;
      if (t < 16) {
        double w = 0.0;
        for (int p = 8; p < tb; ++p)
          if ((t >> (p - 8)) & 1) w += Fx[p];
        T1[t] = w;
      }
    }
    const double eh = I.ehi[ltile];
    __syncthreads();
    V acc[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int e = r * 1024 + t * 4 + c;
        acc[r][c] = r < R ? (V)((eh + I.elo[e] + T0[e & 255] + T1[e >> 8]) * xs[r][c]) : (V)0;
      }
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      const V a = (V)I.ar[p];
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[r][c] += a * xs[r][c ^ (1 << p)];
    }
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      if (10 + k >= tb) break;
      const V a = (V)I.ar[10 + k];
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[r][c] += a * xs[(r ^ (1 << k)) & 3][c];
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const V a = (V)I.ar[2 + k];
      const int tp = t ^ (1 << k);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        if (r >= R) break;
        V4<V> q = *reinterpret_cast<const V4<V>*>(lds + r * 1024 + tp * 4);
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[r][c] += a * q.v[c];
      }
    }
    for (int p = tb; p < n; ++p) {
      const V a = (V)(I.ar[p] * s);
      const int64_t pbase = (int64_t)(ltile ^ (1 << (p - tb))) << tb;
      V4<V> q[4];
#pragma unroll
      for (int r = 0; r < 4; ++r)
        if (r < R) q[r] = *reinterpret_cast<const V4<V>*>(x + pbase + r * 1024 + t * 4);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        if (r >= R) break;
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[r][c] += a * q[r].v[c];
      }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      if (r >= R) break;
      const int64_t g0 = base + r * 1024 + t * 4;
      V4<V> zq;
      if (L.j > 0) zq = *reinterpret_cast<const V4<V>*>(z + g0);
      V4<V> o;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        V v = acc[r][c];
        if (L.j > 0) v += (V)cz * zq.v[c];
        o.v[c] = v;
        sa += (double)xs[r][c] * (double)v;
      }
      *reinterpret_cast<V4<V>*>(y + g0) = o;
    }
  } else {
    // generic path: one tile of 2^tb < 1024 codes, all flips through LDS
    const int D = 1 << tb;
    V xs[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int e = t * 4 + c;
      xs[c] = e < D ? (V)(x[base + e] * s) : (V)0;
      if (e < D) lds[e] = xs[c];
    }
    __syncthreads();
    {
      double v = 0.0;
      for (int p = 0; p < 8 && p < tb; ++p)
        if ((t >> p) & 1) v += Fx[p];
      T0[t] = v;
      if (t < 16) T1[t] = 0.0;
    }
    const double eh = I.ehi[ltile];
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int e = t * 4 + c;
      if (e >= D) continue;
      V acc = (V)((eh + I.elo[e] + T0[e & 255] + T1[e >> 8]) * xs[c]);
      for (int p = 0; p < tb; ++p) acc += (V)I.ar[p] * lds[e ^ (1 << p)];
      for (int p = tb; p < n; ++p) {
        const int64_t pe = ((int64_t)(ltile ^ (1 << (p - tb))) << tb) + e;
        acc += (V)(I.ar[p] * s) * x[pe];
      }
      V v = acc;
      if (L.j > 0) v += (V)cz * z[base + e];
      y[base + e] = v;
      sa += (double)xs[c] * (double)v;
    }
  }
  batch_store(L, I, ltile, sa);
}

// orth half-step; with ritz != nullptr (replay) it also accumulates the Ritz
// vector x_b += (y_b[j] / nrm_b[j]) u_j with y_b = ycoef[b][0..]
template <class V>
__global__ __launch_bounds__(256) void k_batch_orth(BatchLz L, const V* __restrict__ uall,
                                                    V* __restrict__ wall, V* __restrict__ ritz,
                                                    const double* __restrict__ ycoef) {
  const int gt = blockIdx.x;
  const int b = L.tile_inst[gt];
  const BatchInst& I = L.inst[b];
  if (load_coherent(&L.stop[b])) return;
  const int ltile = (int)(gt - I.tile0);
  const int D = 1 << I.tb;
  const int64_t base = I.off + ((int64_t)ltile << I.tb);
  const double* nrm = L.nrm + (size_t)b * (L.ldk + 1);
  const V coef = (V)(-L.alpha[(size_t)b * L.ldk + L.j] / nrm[L.j]);
  const V rc = ritz ? (V)(ycoef[(size_t)b * L.ldk + L.j] / nrm[L.j]) : (V)0;
  double acc = 0.0;
  for (int e = threadIdx.x; e < D; e += blockDim.x) {
    const V u = uall[base + e];
    const V v = wall[base + e] + coef * u;
    wall[base + e] = v;
    acc += (double)v * (double)v;
    if (ritz) ritz[base + e] += rc * u;
  }
  batch_store(L, I, ltile, acc);
}

// per-tile site occupations of the Ritz vectors: partial[tile][p] = Σ_e x_e^2
// over codes with bit p set (p < n), partial[tile][BMAXN] = Σ_e x_e^2
template <class V>
__global__ __launch_bounds__(256) void k_batch_occ(BatchLz L, const V* __restrict__ xall,
                                                   double* __restrict__ partial) {
  __shared__ double sw[256];
  __shared__ double sh[4][256];
  const int gt = blockIdx.x;
  const int b = L.tile_inst[gt];
  const BatchInst& I = L.inst[b];
  const int ltile = (int)(gt - I.tile0);
  const int D = 1 << I.tb;
  const int t = threadIdx.x;
  const V* x = xall + I.off + ((int64_t)ltile << I.tb);
  double w = 0.0, hi[4] = {0.0, 0.0, 0.0, 0.0};
  for (int e = t; e < D; e += 256) {       // bits 0..7 of e are those of t
    const double v = (double)x[e] * (double)x[e];
    w += v;
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (((e >> 8) >> q) & 1) hi[q] += v;
  }
  sw[t] = w;
#pragma unroll
  for (int q = 0; q < 4; ++q) sh[q][t] = hi[q];
  __syncthreads();
  double* out = partial + (size_t)gt * (BMAXN + 1);
  if (t <= BMAXN) {
    double s = 0.0;
    const int p = t;
    if (p == BMAXN) {
      for (int u = 0; u < 256; ++u) s += sw[u];
    } else if (p < I.n) {
      if (p < 8 && p < I.tb) {
        for (int u = 0; u < 256; ++u) if ((u >> p) & 1) s += sw[u];
      } else if (p < I.tb) {
        for (int u = 0; u < 256; ++u) s += sh[p - 8][u];
      } else if ((ltile >> (p - I.tb)) & 1) {
        for (int u = 0; u < 256; ++u) s += sw[u];
      }
    }
    out[p] = s;
  }
}

// per instance: occ[b][i] = Σ over its tiles (fixed order) for site i = n-1-p,
// occ[b][BMAXN] = |x_b|^2
__global__ __launch_bounds__(64) void k_batch_occ_finish(BatchLz L, const double* __restrict__ partial,
                                                         double* __restrict__ occ) {
  const int b = blockIdx.x;
  const BatchInst& I = L.inst[b];
  const int p = threadIdx.x;
  if (p > BMAXN) return;
  double s = 0.0;
  for (int k = 0; k < I.ntiles; ++k) s += partial[(size_t)(I.tile0 + k) * (BMAXN + 1) + p];
  double* o = occ + (size_t)b * (BMAXN + 1);
  if (p == BMAXN) o[BMAXN] = s;
  else if (p < I.n) o[I.n - 1 - p] = s;
  else o[p] = 0.0;
}

template <class V>
__global__ __launch_bounds__(256) void k_batch_zero(V* __restrict__ x, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) x[i] = vzero(V{});
}

// Ritz vectors from a stored Krylov basis (row j = u_j of every instance,
// rows total elements apart): ritz_b = Σ_j (ycoef[b][j] / nrm_b[j]) u_j, one
// streaming pass over the rows instead of a replay of the recurrence.  Rows
// whose coefficient is zero are skipped (the host leaves the rows past an
// instance's breakdown at zero; those rows were never written and may hold
// anything).  At most RYD_BATCH_COMBINE_MAX rows (one LDS table of FP64
// coefficients).
#define RYD_BATCH_COMBINE_MAX 2048
template <class V>
__global__ __launch_bounds__(256) void k_batch_ritz_stored(BatchLz L, const V* __restrict__ rows,
                                                           int64_t total, int nrows,
                                                           const double* __restrict__ ycoef,
                                                           V* __restrict__ ritz) {
  __shared__ double cf[RYD_BATCH_COMBINE_MAX];
  const int gt = blockIdx.x;
  const int b = L.tile_inst[gt];
  const BatchInst& I = L.inst[b];
  const int ltile = (int)(gt - I.tile0);
  const int D = 1 << I.tb;
  const int64_t base = I.off + ((int64_t)ltile << I.tb);
  const double* nrm = L.nrm + (size_t)b * (L.ldk + 1);
  for (int j = threadIdx.x; j < nrows; j += blockDim.x) {
    const double y = ycoef[(size_t)b * L.ldk + j];
    cf[j] = y != 0.0 ? y / nrm[j] : 0.0;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < D; e += blockDim.x) {
    double acc = 0.0;
    for (int j = 0; j < nrows; ++j) {
      const double c = cf[j];                 // block-uniform
      if (c != 0.0) acc += c * (double)rows[(size_t)j * total + base + e];
    }
    ritz[base + e] = (V)acc;
  }
}

// nrm[b][0] = |start_b|
template <class V>
__global__ __launch_bounds__(256) void k_batch_norm0(BatchLz L, const V* __restrict__ vall) {
  const int gt = blockIdx.x;
  const int b = L.tile_inst[gt];
  const BatchInst& I = L.inst[b];
  const int ltile = (int)(gt - I.tile0);
  const int D = 1 << I.tb;
  const int64_t base = I.off + ((int64_t)ltile << I.tb);
  double acc = 0.0;
  for (int e = threadIdx.x; e < D; e += blockDim.x) acc += (double)vall[base + e] * (double)vall[base + e];
  batch_store(L, I, ltile, acc);
}

}  // namespace ryd
