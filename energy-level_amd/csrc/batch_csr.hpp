// Batched Lanczos over caller-supplied sparse Hermitian instances (CSR).
//
// The instances' vectors are concatenated as in the full-basis batch, and a
// tile is a run of `tr` consecutive rows of one instance (the last tile of
// an instance may be shorter), so every Lanczos half-step of the whole batch
// is one launch over all tiles plus a per-instance finishing launch:
//   apply : w = H (u_j / nrm_j) − (nrm_j / nrm_{j−1}) u_{j−1},  alpha_j = Re<q_j, w>
//   orth  : u_{j+1} = w − (alpha_j / nrm_j) u_j [− γ g],       nrm_{j+1} = |u_{j+1}|
// With a lock vector g per instance (unit, e.g. a converged ground vector)
// the orth removes γ g with γ = <g, w> − (alpha_j / nrm_j)<g, u_j>, the g
// component of the finished residual (the apply sums <g, w>, the previous
// orth <g, u_j>), so <g, u_{j+1}> stays at rounding level: the recurrence
// runs in the complement of g and its lowest Ritz value is the next level
// counted with multiplicity (a second ground vector gives E1 = E0).
// Removing only <g, w> would leave −alpha_j <g, q_j>, which grows by a
// factor ~|alpha| per step.
// Values are FP64 (real, or real and imaginary parts); the vectors are
// float, double, complex64 or complex128.
#pragma once
#include "common.hpp"
#include "batch.hpp"
#include "reortho.hpp"   // csub, and cdot_acc: conj(a)·b accumulated in FP64
#include "lz_small.hpp"  // LZS_THREADS, LZS_WAVES

namespace ryd {

struct BcLz {
  const BatchInst* inst;  // off, dim, tile0, ntiles, n (occupation bits, 0 if dim is not 2^n)
  const int* tile_inst;   // global tile -> instance
  double* alpha;          // [count][ldk]
  double* nrm;            // [count][ldk+1]
  int* stop;              // [count]
  double* partial;        // [tiles][3]
  double* gam;            // [count][4]: γ of the current step, then <g_b, u_j>
  int ldk;
  int j;
  int tr;                 // rows per tile
  double tiny;
};

struct BcMat {
  const int64_t* ptr;     // [total + 1] row pointers of the concatenated rows
  const int32_t* col;     // [nnz] column inside the row's instance
  const double* vre;      // [nnz]
  const double* vim;      // [nnz], null for real values
};

// the block's three sums to partial[gt][0..2] (256 threads)
__device__ inline void bc_store3(const BcLz& L, int gt, double a, double b, double c) {
  __shared__ double red[3][4];
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  a = wave_sum(a);
  b = wave_sum(b);
  c = wave_sum(c);
  if (lane == 0) {
    red[0][wid] = a;
    red[1][wid] = b;
    red[2][wid] = c;
  }
  __syncthreads();
  if (t < 3) L.partial[(size_t)gt * 3 + t] = red[t][0] + red[t][1] + red[t][2] + red[t][3];
}

// Per instance, tiles summed in fixed order: post 0 stores alpha_j (and, in a
// locked run, γ), post 1 stores nrm_{j+1}, the breakdown flag and, locked,
// <g, u_{j+1}>.
__global__ __launch_bounds__(256) void k_bc_finish(BcLz L, int post, int locked) {
  __shared__ double red[3][4];
  const int b = blockIdx.x;
  if (L.stop[b]) return;
  const BatchInst& I = L.inst[b];
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (int k = t; k < I.ntiles; k += blockDim.x) {
    const double* p = L.partial + (size_t)(I.tile0 + k) * 3;
    s0 += p[0];
    s1 += p[1];
    s2 += p[2];
  }
  s0 = wave_sum(s0);
  s1 = wave_sum(s1);
  s2 = wave_sum(s2);
  if (lane == 0) {
    red[0][wid] = s0;
    red[1][wid] = s1;
    red[2][wid] = s2;
  }
  __syncthreads();
  if (t == 0) {
    const double a = red[0][0] + red[0][1] + red[0][2] + red[0][3];
    const double gr = red[1][0] + red[1][1] + red[1][2] + red[1][3];
    const double gi = red[2][0] + red[2][1] + red[2][2] + red[2][3];
    double* gm = L.gam + 4 * (size_t)b;
    if (post == 0) {
      L.alpha[(size_t)b * L.ldk + L.j] = a;
      if (locked) {
        const double c = a / L.nrm[(size_t)b * (L.ldk + 1) + L.j];
        gm[0] = gr - c * gm[2];
        gm[1] = gi - c * gm[3];
      }
    } else {
      const double nv = sqrt(a > 0.0 ? a : 0.0);
      L.nrm[(size_t)b * (L.ldk + 1) + L.j + 1] = nv;
      if (!(nv > L.tiny)) L.stop[b] = 1;
      if (locked) {
        gm[2] = gr;
        gm[3] = gi;
      }
    }
  }
}

// partial[gt] = {|u_0|^2, <g, u_0>} over the tile
template <class V>
__global__ __launch_bounds__(256) void k_bc_norm0(BcLz L, const V* __restrict__ uall,
                                                  const V* __restrict__ lock) {
  const int gt = blockIdx.x;
  const BatchInst& I = L.inst[L.tile_inst[gt]];
  const int64_t r0 = (int64_t)(gt - I.tile0) * L.tr;
  const int64_t r1 = r0 + L.tr < I.dim ? r0 + L.tr : I.dim;
  const V* u = uall + I.off;
  const V* g = lock ? lock + I.off : nullptr;
  double acc = 0.0, gre = 0.0, gim = 0.0;
  for (int64_t r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
    acc += nrm2(u[r]);
    if (g) cdot_acc(g[r], u[r], gre, gim);
  }
  bc_store3(L, gt, acc, gre, gim);
}

// apply half-step.  WAVE = false: one row per thread (short rows); WAVE =
// true: one row per wavefront, its entries split over the 64 lanes (rows of
// a dense or nearly dense matrix), the row sum then a wavefront reduction.
template <class V, bool WAVE>
__global__ __launch_bounds__(256) void k_bc_apply(BcLz L, BcMat M, const V* __restrict__ xall,
                                                  const V* __restrict__ zall, V* __restrict__ yall,
                                                  const V* __restrict__ lock) {
  using RT = typename VT<V>::R;
  const int gt = blockIdx.x;
  const int b = L.tile_inst[gt];
  const BatchInst& I = L.inst[b];
  if (load_coherent(&L.stop[b])) return;
  const int64_t r0 = (int64_t)(gt - I.tile0) * L.tr;
  const int64_t r1 = r0 + L.tr < I.dim ? r0 + L.tr : I.dim;
  const V* x = xall + I.off;
  const V* z = zall + I.off;
  V* y = yall + I.off;
  const V* g = lock ? lock + I.off : nullptr;
  const int64_t* ptr = M.ptr + I.off;
  const double* nrm = L.nrm + (size_t)b * (L.ldk + 1);
  const double nj = nrm[L.j];
  const double s = 1.0 / nj;
  const RT sr = (RT)s;
  const RT czr = (RT)(L.j > 0 ? -nj / nrm[L.j - 1] : 0.0);
  double sa = 0.0, gre = 0.0, gim = 0.0;
  if constexpr (!WAVE) {
    for (int64_t r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
      V acc = vzero(V{});
      const int64_t qa = ptr[r], qb = ptr[r + 1];
      for (int64_t q = qa; q < qb; ++q)
        vcfma(acc, (RT)(M.vre[q] * s), (RT)(M.vim ? M.vim[q] * s : 0.0), x[M.col[q]]);
      if (L.j > 0) vfma(acc, czr, z[r]);
      y[r] = acc;
      sa += rdot(vscale(x[r], sr), acc);
      if (g) cdot_acc(g[r], acc, gre, gim);
    }
  } else {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int64_t r = r0 + wid; r < r1; r += 4) {
      double are = 0.0, aim = 0.0;
      const int64_t qa = ptr[r], qb = ptr[r + 1];
      for (int64_t q = qa + lane; q < qb; q += 64) {
        V t = vzero(V{});
        vcfma(t, (RT)M.vre[q], (RT)(M.vim ? M.vim[q] : 0.0), x[M.col[q]]);
        cdot_acc(vfrom_real<V>(1.0), t, are, aim);
      }
      are = wave_sum(are);
      aim = wave_sum(aim);
      if (lane == 0) {
        V acc;
        if constexpr (VT<V>::cplx) {
          acc.re = (RT)(are * s);
          acc.im = (RT)(aim * s);
        } else {
          acc = (RT)(are * s);
        }
        if (L.j > 0) vfma(acc, czr, z[r]);
        y[r] = acc;
        sa += rdot(vscale(x[r], sr), acc);
        if (g) cdot_acc(g[r], acc, gre, gim);
      }
    }
  }
  bc_store3(L, gt, sa, gre, gim);
}

// orth half-step; in a locked run it also removes γ g; with ritz != nullptr
// (replay) it accumulates x_b += (y_b[j] / nrm_b[j]) u_j
template <class V>
__global__ __launch_bounds__(256) void k_bc_orth(BcLz L, const V* __restrict__ uall,
                                                 V* __restrict__ wall, V* __restrict__ ritz,
                                                 const double* __restrict__ ycoef,
                                                 const V* __restrict__ lock) {
  using RT = typename VT<V>::R;
  const int gt = blockIdx.x;
  const int b = L.tile_inst[gt];
  const BatchInst& I = L.inst[b];
  if (load_coherent(&L.stop[b])) return;
  const int64_t r0 = (int64_t)(gt - I.tile0) * L.tr;
  const int64_t r1 = r0 + L.tr < I.dim ? r0 + L.tr : I.dim;
  const V* u = uall + I.off;
  V* w = wall + I.off;
  V* xr = ritz ? ritz + I.off : nullptr;
  const V* g = lock ? lock + I.off : nullptr;
  const double* nrm = L.nrm + (size_t)b * (L.ldk + 1);
  const RT coef = (RT)(-L.alpha[(size_t)b * L.ldk + L.j] / nrm[L.j]);
  const RT rc = ritz ? (RT)(ycoef[(size_t)b * L.ldk + L.j] / nrm[L.j]) : (RT)0;
  const double gr = g ? L.gam[4 * (size_t)b] : 0.0, gi = g ? L.gam[4 * (size_t)b + 1] : 0.0;
  double acc = 0.0, gre = 0.0, gim = 0.0;
  for (int64_t r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
    const V ue = u[r];
    V v = w[r];
    vfma(v, coef, ue);
    if (g) {
      const V ge = g[r];
      csub(v, gr, gi, ge);
      cdot_acc(ge, v, gre, gim);
    }
    w[r] = v;
    acc += nrm2(v);
    if (xr) {
      V a = xr[r];
      vfma(a, rc, ue);
      xr[r] = a;
    }
  }
  bc_store3(L, gt, acc, gre, gim);
}

// Ritz vectors from the stored rows of the last run: ritz_b = Σ_j
// (ycoef[b][j] / nrm_b[j]) u_j over the tile's rows; zero coefficients skip
// their rows (never written past an instance's breakdown).
template <class V>
__global__ __launch_bounds__(256) void k_bc_ritz_stored(BcLz L, const V* __restrict__ rows,
                                                        int64_t total, int nrows,
                                                        const double* __restrict__ ycoef,
                                                        V* __restrict__ ritz) {
  using RT = typename VT<V>::R;
  __shared__ double cf[RYD_BATCH_COMBINE_MAX];
  const int gt = blockIdx.x;
  const int b = L.tile_inst[gt];
  const BatchInst& I = L.inst[b];
  const int64_t r0 = (int64_t)(gt - I.tile0) * L.tr;
  const int64_t r1 = r0 + L.tr < I.dim ? r0 + L.tr : I.dim;
  const double* nrm = L.nrm + (size_t)b * (L.ldk + 1);
  for (int j = threadIdx.x; j < nrows; j += blockDim.x) {
    const double y = ycoef[(size_t)b * L.ldk + j];
    cf[j] = y != 0.0 ? y / nrm[j] : 0.0;
  }
  __syncthreads();
  for (int64_t r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
    double are = 0.0, aim = 0.0;
    for (int j = 0; j < nrows; ++j) {
      const double c = cf[j];                 // block-uniform
      if (c != 0.0) {
        const V e = rows[(size_t)j * total + I.off + r];
        if constexpr (VT<V>::cplx) {
          are += c * (double)e.re;
          aim += c * (double)e.im;
        } else {
          are += c * (double)e;
        }
      }
    }
    V o;
    if constexpr (VT<V>::cplx) {
      o.re = (RT)are;
      o.im = (RT)aim;
    } else {
      o = (RT)are;
    }
    ritz[I.off + r] = o;
  }
}

// per-tile weights of the Ritz vectors: partial[tile][p] = Σ |x_r|^2 over
// local rows r with bit p set (p < n, the instance's qubit count when its
// dim is 2^n), partial[tile][BMAXN] = Σ |x_r|^2 (finished by k_batch_occ_finish)
template <class V>
__global__ __launch_bounds__(256) void k_bc_occ(BcLz L, const V* __restrict__ xall,
                                                double* __restrict__ partial) {
  __shared__ double red[BMAXN + 1][4];
  const int gt = blockIdx.x;
  const BatchInst& I = L.inst[L.tile_inst[gt]];
  const int64_t r0 = (int64_t)(gt - I.tile0) * L.tr;
  const int64_t r1 = r0 + L.tr < I.dim ? r0 + L.tr : I.dim;
  const V* x = xall + I.off;
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  double acc[BMAXN + 1];
#pragma unroll
  for (int p = 0; p <= BMAXN; ++p) acc[p] = 0.0;
  for (int64_t r = r0 + t; r < r1; r += blockDim.x) {
    const double w = nrm2(x[r]);
    acc[BMAXN] += w;
#pragma unroll
    for (int p = 0; p < BMAXN; ++p)
      if ((r >> p) & 1) acc[p] += w;
  }
#pragma unroll
  for (int p = 0; p <= BMAXN; ++p) {
    const double v = wave_sum(acc[p]);
    if (lane == 0) red[p][wid] = v;
  }
  __syncthreads();
  if (t <= BMAXN) {
    const bool used = t == BMAXN || t < I.n;
    partial[(size_t)gt * (BMAXN + 1) + t] = used ? red[t][0] + red[t][1] + red[t][2] + red[t][3] : 0.0;
  }
}

// ---- one launch per run for small instances -----------------------------------
// Instances of at most LZS_THREADS·bcc_maxe<V>() rows with short rows: one
// 1024-thread workgroup per instance runs the whole sequence of a run
// (stop reset, |u_0| and <g, u_0>, then per step the apply with alpha and
// <g, w>, the finishing formulas, the orth with γ g and the Ritz
// accumulation, the norm and breakdown test), u_j in LDS for the column
// gathers, each thread's rows of u_{j-1} and w in registers, the sums formed
// in the same order on every thread.  alpha, nrm, stop and gam receive what
// the multi-launch form writes; the stored form writes every row, the
// rotating form the two live vectors on exit (as lz_small.hpp).
template <class V> constexpr int bcc_maxe() { return sizeof(V) == 16 ? 4 : 8; }

// three block sums, the same on every thread; callers alternate two slots
__device__ __forceinline__ void bcc_sum3(double& a, double& b, double& c,
                                         double (*red)[LZS_WAVES]) {
  a = wave_sum(a);
  b = wave_sum(b);
  c = wave_sum(c);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[0][w] = a;
    red[1][w] = b;
    red[2][w] = c;
  }
  __syncthreads();
  double x = 0.0, y = 0.0, z = 0.0;
#pragma unroll
  for (int k = 0; k < LZS_WAVES; ++k) {
    x += red[0][k];
    y += red[1][k];
    z += red[2][k];
  }
  a = x;
  b = y;
  c = z;
}

template <class V>
__global__ __launch_bounds__(LZS_THREADS) void k_bc_cycle(BcLz L, BcMat M, V* base, int64_t total,
                                                          int stored, const V* __restrict__ lock,
                                                          V* __restrict__ ritz,
                                                          const double* __restrict__ ycoef,
                                                          int steps) {
  using RT = typename VT<V>::R;
  constexpr int E = bcc_maxe<V>();
  __shared__ V xs[LZS_THREADS * E];
  __shared__ double redA[3][LZS_WAVES], redB[3][LZS_WAVES];
  const int b = blockIdx.x;
  const BatchInst& I = L.inst[b];
  const int dim = (int)I.dim;
  const int tid = threadIdx.x;
  auto U = [&](int j) -> V* {
    const int r = stored ? j : ((j % 3) + 3) % 3;
    return base + (size_t)r * (size_t)total + I.off;
  };
  const V* g = lock ? lock + I.off : nullptr;
  const int64_t* ptr = M.ptr + I.off;
  V* xr = ritz ? ritz + I.off : nullptr;
  double* alpha = L.alpha + (size_t)b * L.ldk;
  double* nrm = L.nrm + (size_t)b * (L.ldk + 1);
  double* gm = L.gam + 4 * (size_t)b;
  const double* yc = ritz ? ycoef + (size_t)b * L.ldk : nullptr;

  // |u_0|, <g, u_0>
  double nn = 0.0, g2 = 0.0, g3 = 0.0;
  {
    const V* u0 = U(0);
#pragma unroll
    for (int k = 0; k < E; ++k) {
      const int r = tid + k * LZS_THREADS;
      if (r >= dim) continue;
      const V u = u0[r];
      xs[r] = u;
      nn += nrm2(u);
      if (g) cdot_acc(g[r], u, g2, g3);
    }
  }
  bcc_sum3(nn, g2, g3, redB);                      // also publishes xs
  double nj = sqrt(nn > 0.0 ? nn : 0.0);
  const bool brk0 = !(nj > L.tiny);
  if (tid == 0) {
    L.stop[b] = brk0 ? 1 : 0;
    nrm[0] = nj;
    if (g) {
      gm[2] = g2;
      gm[3] = g3;
    }
  }
  if (brk0) return;

  V um[E], w[E];
#pragma unroll
  for (int k = 0; k < E; ++k) {
    um[k] = vzero(V{});
    w[k] = vzero(V{});
  }
  double njm = 1.0;
  for (int j = 0; j < steps; ++j) {
    const double s = 1.0 / nj;
    const RT sr = (RT)s;
    const RT czr = (RT)(j > 0 ? -nj / njm : 0.0);
    // apply: w = H (u_j / nrm_j) - (nrm_j / nrm_{j-1}) u_{j-1}
    double sa = 0.0, gre = 0.0, gim = 0.0;
#pragma unroll
    for (int k = 0; k < E; ++k) {
      const int r = tid + k * LZS_THREADS;
      if (r >= dim) continue;
      V acc = vzero(V{});
      const int64_t qa = ptr[r], qb = ptr[r + 1];
      // four entries' loads in flight at once (they come from L2)
#pragma unroll 4
      for (int64_t q = qa; q < qb; ++q)
        vcfma(acc, (RT)(M.vre[q] * s), (RT)(M.vim ? M.vim[q] * s : 0.0), xs[M.col[q]]);
      if (j > 0) vfma(acc, czr, um[k]);
      w[k] = acc;
      sa += rdot(vscale(xs[r], sr), acc);
      if (g) cdot_acc(g[r], acc, gre, gim);
    }
    bcc_sum3(sa, gre, gim, redA);                  // also: every gather of xs is done
    const double a = sa;
    double gm0 = 0.0, gm1 = 0.0;
    if (g) {
      const double c = a / nj;
      gm0 = gre - c * g2;
      gm1 = gim - c * g3;
    }
    if (tid == 0) {
      alpha[j] = a;
      if (g) {
        gm[0] = gm0;
        gm[1] = gm1;
      }
    }
    // orth: u_{j+1} = w - (alpha_j / nrm_j) u_j [- γ g]
    const RT coef = (RT)(-a / nj);
    const RT rc = xr ? (RT)(yc[j] / nj) : (RT)0;
    V* wrow = stored ? U(j + 1) : nullptr;
    double acc2 = 0.0, h0 = 0.0, h1 = 0.0;
#pragma unroll
    for (int k = 0; k < E; ++k) {
      const int r = tid + k * LZS_THREADS;
      if (r >= dim) continue;
      const V ue = xs[r];
      V v = w[k];
      vfma(v, coef, ue);
      if (g) {
        const V ge = g[r];
        csub(v, gm0, gm1, ge);
        cdot_acc(ge, v, h0, h1);
      }
      acc2 += nrm2(v);
      if (xr) {
        V q = xr[r];
        vfma(q, rc, ue);
        xr[r] = q;
      }
      if (wrow) wrow[r] = v;
      w[k] = v;
      um[k] = ue;
      xs[r] = v;                                   // only this thread reads xs[r] now
    }
    bcc_sum3(acc2, h0, h1, redB);                  // also publishes xs
    const double nv = sqrt(acc2 > 0.0 ? acc2 : 0.0);
    const bool brk = !(nv > L.tiny);               // also catches NaN
    if (tid == 0) {
      nrm[j + 1] = nv;
      if (brk) L.stop[b] = 1;
      if (g) {
        gm[2] = h0;
        gm[3] = h1;
      }
    }
    g2 = h0;
    g3 = h1;
    if (brk || j + 1 == steps) {
      if (!stored) {
        V* pu = U(j);
        V* pw = U(j + 1);
#pragma unroll
        for (int k = 0; k < E; ++k) {
          const int r = tid + k * LZS_THREADS;
          if (r < dim) {
            pu[r] = um[k];
            pw[r] = w[k];
          }
        }
      }
      return;
    }
    njm = nj;
    nj = nv;
  }
}

}  // namespace ryd
