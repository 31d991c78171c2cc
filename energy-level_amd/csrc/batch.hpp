// Batched Lanczos over many independent full-basis instances of uneven n.
//
// All instances' vectors are concatenated; every launch covers the tiles of
// all instances (a tile is 2^tb consecutive codes of one instance, or the
// whole instance when n < tb), so one Lanczos half-step of the whole batch is
// one launch.  Per-instance FP64 sums close through a per-instance arrival
// counter: the last tile of an instance to finish writes alpha/nrm for it.
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
};

struct BatchLz {
  const BatchInst* inst;
  const int* tile_inst;   // global tile -> instance
  double* alpha;          // [count][ldk]
  double* nrm;            // [count][ldk+1]
  int* stop;              // [count]
  unsigned* counter;      // [count]
  double* partial;        // [total tiles]
  int ldk;
  int j;
  double tiny;
};

// Per-instance reduction: each tile stores its partial, the instance's last
// tile sums them in fixed order and runs the post-operation (0: alpha, 1: norm).
__device__ inline void batch_finish(const BatchLz& L, const BatchInst& I, int b, int ltile,
                                    double v, int post) {
  __shared__ double red[4];
  __shared__ int last;
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  v = wave_sum(v);
  if (lane == 0) red[wid] = v;
  __syncthreads();
  if (t == 0) {
    L.partial[I.tile0 + ltile] = red[0] + red[1] + red[2] + red[3];
    __threadfence();
    const unsigned prev = atomicAdd(&L.counter[b], 1u);
    last = prev == (unsigned)(I.ntiles - 1);
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
  double s = 0.0;
  for (int k = t; k < I.ntiles; k += blockDim.x) s += load_coherent(L.partial + I.tile0 + k);
  s = wave_sum(s);
  __syncthreads();
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
    __threadfence();
    L.counter[b] = 0u;
  }
}

// apply half-step (real element types)
template <class V>
__global__ __launch_bounds__(256) void k_batch_apply(BatchLz L, const V* __restrict__ xall,
                                                     const V* __restrict__ zall, V* __restrict__ yall) {
  __shared__ V lds[4096];
  __shared__ double Fx[16];
  __shared__ double T0[256];
  __shared__ double T1[16];
  const int gt = blockIdx.x;
  const int b = L.tile_inst[gt];
  const BatchInst& I = L.inst[b];
  if (load_coherent(&L.stop[b])) return;
  const int ltile = (int)(gt - I.tile0);
  const int t = threadIdx.x;
  const int n = I.n, tb = I.tb;
  const int R = tb > 10 ? 1 << (tb - 10) : 1;
  const int64_t base = (int64_t)ltile << tb;
  const V* x = xall + I.off;
  const V* z = zall + I.off;
  V* y = yall + I.off;
  const double* nrm = L.nrm + (size_t)b * (L.ldk + 1);
  const double nj = nrm[L.j];
  const double s = 1.0 / nj;
  const double cz = L.j > 0 ? -nj / nrm[L.j - 1] : 0.0;
  const int D = 1 << tb;                 // codes in this tile
  // stage
  V xs[4][4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    if (r >= R) break;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int e = r * 1024 + t * 4 + c;
      xs[r][c] = e < D ? (V)(x[base + e] * s) : (V)0;
      if (e < D) lds[e] = xs[r][c];
    }
  }
  // cross field of the high bits on the low bits, per tile
  if (t < tb) {
    double f = 0.0;
    const double* row = I.vb + (size_t)t * n;
    for (int q = tb; q < n; ++q)
      if ((ltile >> (q - tb)) & 1) f += row[q];
    Fx[t] = f;
  }
  __syncthreads();
  {
    double v = 0.0;
    for (int p = 0; p < 8 && p < tb; ++p)
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
  double sa = 0.0;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    if (r >= R) break;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int e = r * 1024 + t * 4 + c;
      if (e >= D) continue;
      const double d = eh + I.elo[e] + T0[e & 255] + T1[e >> 8];
      V acc = (V)(d * xs[r][c]);
      // flips inside the tile (LDS / registers)
      for (int p = 0; p < tb; ++p) acc += (V)I.ar[p] * lds[e ^ (1 << p)];
      // flips to other tiles of the instance
      for (int p = tb; p < n; ++p) {
        const int64_t pe = ((int64_t)(ltile ^ (1 << (p - tb))) << tb) + e;
        acc += (V)(I.ar[p] * s) * x[pe];
      }
      V v = acc;
      if (L.j > 0) v += (V)cz * z[base + e];
      y[base + e] = v;
      sa += (double)xs[r][c] * (double)v;
    }
  }
  batch_finish(L, I, b, ltile, sa, 0);
}

template <class V>
__global__ __launch_bounds__(256) void k_batch_orth(BatchLz L, const V* __restrict__ uall,
                                                    V* __restrict__ wall) {
  const int gt = blockIdx.x;
  const int b = L.tile_inst[gt];
  const BatchInst& I = L.inst[b];
  if (load_coherent(&L.stop[b])) return;
  const int ltile = (int)(gt - I.tile0);
  const int D = 1 << I.tb;
  const int64_t base = I.off + ((int64_t)ltile << I.tb);
  const double* nrm = L.nrm + (size_t)b * (L.ldk + 1);
  const V coef = (V)(-L.alpha[(size_t)b * L.ldk + L.j] / nrm[L.j]);
  double acc = 0.0;
  for (int e = threadIdx.x; e < D; e += blockDim.x) {
    const V v = wall[base + e] + coef * uall[base + e];
    wall[base + e] = v;
    acc += (double)v * (double)v;
  }
  batch_finish(L, I, b, ltile, acc, 1);
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
  BatchLz L0 = L;
  L0.j = -1;
  batch_finish(L0, I, b, ltile, acc, 1);
}

}  // namespace ryd
