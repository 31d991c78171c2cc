// H·v on an independent-set basis of a blockade graph (graph.hpp tables),
// one block per valid high part h.
//
// The rows of h form one contiguous segment [off, off + cnt); row off + j
// holds low set t = post_c[j] (c = class of h), so x and y move in
// coalesced runs.  Every high part of class c has the same row layout, so
// the per-class row tables (indexed by base_c + j, coalesced across lanes)
// hold what depends on the low set only: its bits rcode, its diagonal rdiag
// and, for each low bit q, the segment position lrow[q] of the row with q
// flipped (GR_NONE when that set is not independent or h blocks it).  A flip
// of high bit p moves to the segment of h ^ 2^p with the same low set:
// off2 + pre_c2[t], block-uniform offsets held in LDS.  The diagonal is
// dhi(h) (block-uniform) + rdiag + Σ_{q in l} w_q(h), with
// w_q(h) = Σ_{p in h} V_{p,q} the high-low cross terms in LDS.
#pragma once
#include "common.hpp"
#include "graph.hpp"
#include "hv_full.hpp"

namespace ryd {

constexpr int GR_FLIPS = 8;      // flips whose loads are issued together

struct GraphModel {
  int n, L, H;
  int64_t tl;                   // low sets
  int64_t nv;                   // valid high parts
  const uint32_t* hlist;        // [nv] valid high parts, ascending
  const uint64_t* hpk;          // [2^H + 1]
  int64_t rows;                 // rows of all classes (class c from post_off[c])
  const uint32_t* fc;           // [nc]
  const uint32_t* pre;          // [nc][tl + 1]
  const uint32_t* post;         // [rows] low-set index of each class row
  const int64_t* post_off;      // [nc] start of class c in the row tables
  const uint32_t* rcode;        // [rows] low bits of the row
  const double* rdiag;          // [rows] diagonal of those bits alone
  const void* lrow;             // [L][rows] segment position of the row with bit q flipped
                                // (uint16_t when every segment is shorter than 65535 rows)
  const double* vb;             // [n*n] V by bit position
  int hfar;                     // high bits 0 .. hfar-1 are applied here (H: all)
  double dl[MAXN];              // delta by bit position
  double ar[MAXN], ai[MAXN];    // Ω/2 by bit position
};

template <class V, class IT>
__global__ __launch_bounds__(256) void k_graph_apply(GraphModel M, ApplyIO<V> io, Reduce rd) {
  using RT = typename VT<V>::R;
  const IT* lrow = static_cast<const IT*>(M.lrow);
  constexpr IT INONE = (IT)~(IT)0;
  __shared__ double wq[GR_MAXB];
  __shared__ double dhi_s;
  __shared__ int64_t off2[GR_MAXB];     // first rank of h ^ 2^p, -1 when not independent
  __shared__ int64_t prow2[GR_MAXB];    // its class row in pre
  __shared__ uint32_t f2[GR_MAXB];      // the low bits its class blocks
  if (io.stop && load_coherent(io.stop)) return;
  double s, ch, cx, cz;
  apply_scalars(io, s, ch, cx, cz);
  const RT sr = (RT)s, chr = (RT)ch, cxr = (RT)cx, czr = (RT)cz;
  const int tid = threadIdx.x;
  const int L = M.L, H = M.H;
  const int64_t tl = M.tl;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  // contiguous runs of high parts per XCD (blocks go round-robin to the 8 XCDs)
  int64_t v0, v1, vstep;
  if (gridDim.x >= 8 && (gridDim.x & 7) == 0) {
    const int64_t chunk = (M.nv + 7) / 8;
    v0 = (int64_t)(blockIdx.x & 7) * chunk + (blockIdx.x >> 3);
    v1 = ((int64_t)(blockIdx.x & 7) + 1) * chunk;
    if (v1 > M.nv) v1 = M.nv;
    vstep = gridDim.x >> 3;
  } else {
    v0 = blockIdx.x;
    v1 = M.nv;
    vstep = gridDim.x;
  }
  for (int64_t vi = v0; vi < v1; vi += vstep) {
    const uint32_t h = M.hlist[vi];
    const uint64_t e = M.hpk[h];
    const uint64_t c = e & GR_CLS;
    const int64_t off = (int64_t)(e >> 24);
    const int64_t cnt = (int64_t)(M.hpk[h + 1] >> 24) - off;
    __syncthreads();                            // LDS of the previous high part is free
    if (tid < L) {
      double w = 0.0;
      for (uint32_t y = h; y; y &= y - 1) w += M.vb[(size_t)(__builtin_ctz(y) + L) * M.n + tid];
      wq[tid] = w;
    } else if (tid >= 64 && tid < 64 + H) {
      const int p = tid - 64;
      const uint64_t e2 = M.hpk[h ^ (1u << p)];
      const uint64_t c2 = e2 & GR_CLS;
      if (c2 == GR_CLS) {
        off2[p] = -1;
      } else {
        off2[p] = (int64_t)(e2 >> 24);
        prow2[p] = (int64_t)c2 * (tl + 1);
        f2[p] = M.fc[c2];
      }
    } else if (tid == 128) {
      double d = 0.0;
      for (uint32_t y = h; y; y &= y - 1) {
        const int p = __builtin_ctz(y) + L;
        d += M.dl[p];
        for (uint32_t z = y & (y - 1); z; z &= z - 1) d += M.vb[(size_t)p * M.n + __builtin_ctz(z) + L];
      }
      dhi_s = d;
    }
    __syncthreads();
    const int64_t base = M.post_off[c];
    const double dh = dhi_s;
    for (int64_t j = tid; j < cnt; j += 256) {
      const int64_t row = off + j;
      const int64_t cr = base + j;
      const uint32_t t = M.post[cr];
      const uint32_t l = M.rcode[cr];
      double dg = dh + M.rdiag[cr];
      for (uint32_t y = l; y; y &= y - 1) dg += wq[__builtin_ctz(y)];
      const V xe = vscale(io.x[row], sr);
      V acc = vscale(xe, (RT)dg);
      // flips in groups of GR_FLIPS: every load of a stage is issued before
      // the next stage needs it (position, then x)
      for (int q0 = 0; q0 < L; q0 += GR_FLIPS) {
        int64_t r2[GR_FLIPS];
#pragma unroll
        for (int u = 0; u < GR_FLIPS; ++u) {
          const int q = q0 + u;
          const IT pos = q < L ? lrow[(size_t)q * (size_t)M.rows + cr] : INONE;
          r2[u] = (pos != INONE && (M.ar[q] != 0.0 || M.ai[q] != 0.0)) ? off + (int64_t)pos : -1;
        }
        V xv[GR_FLIPS];
#pragma unroll
        for (int u = 0; u < GR_FLIPS; ++u) xv[u] = r2[u] >= 0 ? io.x[r2[u]] : vzero(V{});
#pragma unroll
        for (int u = 0; u < GR_FLIPS; ++u) {
          if (r2[u] < 0) continue;
          const int q = q0 + u;
          const double aim = M.ai[q];
          vcfma(acc, (RT)(M.ar[q] * s), (RT)((((l >> q) & 1u) ? aim : -aim) * s), xv[u]);
        }
      }
      const int hf = M.hfar;
      for (int p0 = 0; p0 < hf; p0 += GR_FLIPS) {
        int64_t r2[GR_FLIPS];
#pragma unroll
        for (int u = 0; u < GR_FLIPS; ++u) {
          const int p = p0 + u;
          bool ok = p < hf;
          if (ok) {
            const int64_t o2 = off2[p];
            ok = o2 >= 0 && (((h >> p) & 1u) || !(l & f2[p])) &&
                 (M.ar[p + L] != 0.0 || M.ai[p + L] != 0.0);
          }
          r2[u] = ok ? (int64_t)M.pre[prow2[p] + t] : -1;
        }
        V xv[GR_FLIPS];
#pragma unroll
        for (int u = 0; u < GR_FLIPS; ++u)
          xv[u] = r2[u] >= 0 ? io.x[off2[p0 + u] + r2[u]] : vzero(V{});
#pragma unroll
        for (int u = 0; u < GR_FLIPS; ++u) {
          if (r2[u] < 0) continue;
          const int p = p0 + u;
          const double aim = M.ai[p + L];
          vcfma(acc, (RT)(M.ar[p + L] * s), (RT)((((h >> p) & 1u) ? aim : -aim) * s), xv[u]);
        }
      }
      if (io.partial) {                 // pass A of a split H·v: the raw sum
        nt_store(io.y + row, acc);
        continue;
      }
      V v = vscale(acc, chr);
      vfma(v, cxr, xe);
      if (io.z) vfma(v, czr, io.z[row]);
      io.y[row] = v;
      if (io.reduce) {
        s0 += rdot(xe, v);
        s1 += nrm2(v);
        s2 += nrm2(xe);
        if (io.probe) s3 += zdot(((io.probe[row >> 5] >> (row & 31)) & 1u) ? 1.0 : -1.0, v);
      }
    }
  }
  if (io.reduce) block_reduce_store(s0, s1, s2, s3, rd.partials);
}

// ---- pass B of a split graph H·v: the top GT bits of the high part ---------------
// When the top K high bits are sites with no neighbour among the low bits,
// every valid top pattern g over a body high part b (the other high bits)
// has the class of b, so row j of the segment of h = g·2^(H-K) | b holds the
// same low set for every g.  A block takes one body and a run of GR_RUN
// positions j, stages s·x of that run for every valid g in LDS, and adds the
// top flips (g ^ 2^k, an LDS read) to the pass-A partial, then finishes
//   y = ch·(y_A + Σ top flips) + cx·(s·x) + cz·z  with the reductions.
constexpr int GR_TOPMAX = 128;      // valid top patterns
constexpr int GR_KMAX = 10;         // top bits
constexpr int GR_HNEAR = 12;        // split only when pass A keeps at most this many high bits
// segment positions per block: 64 KB of LDS for 128 patterns
template <class V> constexpr int gr_run() { return 32; }

struct GraphTop {
  int H, K, nt;                 // high bits, top bits, valid top patterns
  int L;
  int64_t nbody;                // valid high parts with the top bits clear (hlist prefix)
  int64_t tl;                   // low sets (the longest segment)
  const uint32_t* hlist;
  const uint64_t* hpk;
  const uint32_t* tp;           // [nt] valid top patterns, ascending (top-bit coordinates)
  const int16_t* tpk;           // [nt][K] index of tp ^ 2^k in tp, -1 if not independent
  double ar[GR_KMAX], ai[GR_KMAX];   // Ω/2 of the top bits
};

template <class V>
__global__ __launch_bounds__(256) void k_graph_top(GraphTop T, ApplyIO<V> io, Reduce rd) {
  using RT = typename VT<V>::R;
  constexpr int GR_RUN = gr_run<V>();
  __shared__ V xs[GR_TOPMAX * GR_RUN];
  __shared__ int64_t voff[GR_TOPMAX];
  __shared__ int16_t pk[GR_TOPMAX * GR_KMAX];     // partner index, -1 when not valid for the body
  __shared__ uint32_t tpat[GR_TOPMAX];
  const int64_t nruns = (T.tl + GR_RUN - 1) / GR_RUN;
  for (int i = threadIdx.x; i < T.nt; i += 256) tpat[i] = T.tp[i];
  if (io.stop && load_coherent(io.stop)) return;
  double s, ch, cx, cz;
  apply_scalars(io, s, ch, cx, cz);
  const RT sr = (RT)s, chr = (RT)ch, cxr = (RT)cx, czr = (RT)cz;
  const int tid = threadIdx.x;
  const int HB = T.H - T.K;
  const int items = T.nt * GR_RUN;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  const int64_t units = T.nbody * nruns;
  // contiguous runs of units per XCD
  int64_t u0, u1, ustep;
  if (gridDim.x >= 8 && (gridDim.x & 7) == 0) {
    const int64_t chunk = (units + 7) / 8;
    u0 = (int64_t)(blockIdx.x & 7) * chunk + (blockIdx.x >> 3);
    u1 = ((int64_t)(blockIdx.x & 7) + 1) * chunk;
    if (u1 > units) u1 = units;
    ustep = gridDim.x >> 3;
  } else {
    u0 = blockIdx.x;
    u1 = units;
    ustep = gridDim.x;
  }
  for (int64_t u = u0; u < u1; u += ustep) {
    const int64_t bi = u / nruns;
    const int64_t j0 = (u - bi * nruns) * GR_RUN;
    const uint32_t hb = T.hlist[bi];
    const int64_t cnt = (int64_t)(T.hpk[hb + 1] >> 24) - (int64_t)(T.hpk[hb] >> 24);
    if (j0 >= cnt) continue;                       // block-uniform
    __syncthreads();                               // LDS of the previous unit is free
    if (tid < T.nt) {
      const uint64_t e = T.hpk[((uint32_t)tpat[tid] << HB) | hb];
      voff[tid] = (e & GR_CLS) == GR_CLS ? -1 : (int64_t)(e >> 24);
    }
    __syncthreads();
    for (int i = tid; i < T.nt * T.K; i += 256) {
      const int pv = T.tpk[i];
      pk[i] = (int16_t)((pv >= 0 && voff[pv] >= 0) ? pv : -1);
    }
    for (int i = tid; i < items; i += 256) {
      const int v = i / GR_RUN, jj = i - v * GR_RUN;
      const int64_t o = voff[v];
      xs[i] = (o >= 0 && j0 + jj < cnt) ? vscale(io.x[o + j0 + jj], sr) : vzero(xs[0]);
    }
    __syncthreads();
    for (int i = tid; i < items; i += 256) {
      const int v = i / GR_RUN, jj = i - v * GR_RUN;
      const int64_t o = voff[v];
      if (o < 0 || j0 + jj >= cnt) continue;
      const int64_t row = o + j0 + jj;
      V acc = io.y[row];                           // pass-A partial
      const uint32_t g = tpat[v];
      for (int k = 0; k < T.K; ++k) {
        const int pv = pk[v * T.K + k];
        if (pv < 0) continue;
        const int bit = (int)((g >> k) & 1u);
        vcfma(acc, (RT)T.ar[k], (RT)(bit ? T.ai[k] : -T.ai[k]), xs[pv * GR_RUN + jj]);
      }
      const V xe = xs[i];
      V y = vscale(acc, chr);
      vfma(y, cxr, xe);
      if (io.z) vfma(y, czr, io.z[row]);
      io.y[row] = y;
      if (io.reduce) {
        s0 += rdot(xe, y);
        s1 += nrm2(y);
        s2 += nrm2(xe);
        if (io.probe) s3 += zdot(((io.probe[row >> 5] >> (row & 31)) & 1u) ? 1.0 : -1.0, y);
      }
    }
  }
  if (io.reduce) block_reduce_store(s0, s1, s2, s3, rd.partials);
}

// ---- tables for k_graph_apply -------------------------------------------------
// Per class c and segment position j (row cr = post_off[c] + j, t = post[cr]):
// rcode = lowtab[t], rdiag = diagonal of those low bits, lrow[q][cr] =
// pre_c[index of lowtab[t] ^ 2^q] or GR_NONE.  One block per class.
template <class IT>
__global__ __launch_bounds__(256) void k_gr_rows(int n, int L, const uint32_t* __restrict__ lowtab,
                                                 int64_t tl, const uint32_t* __restrict__ lo_idx,
                                                 const uint32_t* __restrict__ fc, int64_t nc,
                                                 const uint32_t* __restrict__ pre,
                                                 const uint32_t* __restrict__ post,
                                                 const int64_t* __restrict__ post_off, int64_t rows,
                                                 const double* __restrict__ dl,
                                                 const double* __restrict__ vb,
                                                 uint32_t* __restrict__ rcode,
                                                 double* __restrict__ rdiag,
                                                 IT* __restrict__ lrow) {
  for (int64_t c = blockIdx.x; c < nc; c += gridDim.x) {
    const uint32_t F = fc[c];
    const uint32_t* prec = pre + (size_t)c * (size_t)(tl + 1);
    const int64_t b0 = post_off[c];
    const int64_t b1 = c + 1 < nc ? post_off[c + 1] : rows;
    for (int64_t cr = b0 + threadIdx.x; cr < b1; cr += blockDim.x) {
      const uint32_t l = lowtab[post[cr]];
      rcode[cr] = l;
      double d = 0.0;
      for (uint32_t y = l; y; y &= y - 1) {
        const int q = __builtin_ctz(y);
        d += dl[q];
        for (uint32_t z = y & (y - 1); z; z &= z - 1) d += vb[(size_t)q * n + __builtin_ctz(z)];
      }
      rdiag[cr] = d;
      for (int q = 0; q < L; ++q) {
        const uint32_t l2 = l ^ (1u << q);
        const uint32_t t2 = lo_idx[l2];
        const bool ok = t2 != GR_NONE && !(l2 & F);
        lrow[(size_t)q * (size_t)rows + cr] = ok ? (IT)prec[t2] : (IT)~(IT)0;
      }
    }
  }
}

// post[post_off[c] + j] = t of the j-th low set compatible with class c
__global__ __launch_bounds__(256) void k_gr_post(const uint32_t* __restrict__ lowtab, int64_t tl,
                                                 const uint32_t* __restrict__ fc, int64_t nc,
                                                 const int64_t* __restrict__ post_off,
                                                 uint32_t* __restrict__ post) {
  __shared__ uint32_t lds[4];
  for (int64_t c = blockIdx.x; c < nc; c += gridDim.x) {
    const uint32_t F = fc[c];
    uint32_t* out = post + post_off[c];
    uint32_t run = 0;
    for (int64_t t0 = 0; t0 < tl; t0 += 256) {
      const int64_t t = t0 + threadIdx.x;
      const bool ok = t < tl && !(lowtab[t] & F);
      uint32_t tot;
      const uint32_t pos = gr_flag_scan(ok, lds, tot);
      if (ok) out[run + pos] = (uint32_t)t;
      run += tot;
    }
  }
}

struct GrValidFlag {           // h is a valid high part
  const uint64_t* hpk;
  __device__ int64_t operator()(int64_t h) const { return (hpk[h] & GR_CLS) != GR_CLS ? 1 : 0; }
};
struct GrListWrite {           // list[pos] = i for set flags
  uint32_t* list;
  __device__ void operator()(int64_t i, int64_t pre, int64_t v) const {
    if (v) list[pre] = (uint32_t)i;
  }
};

}  // namespace ryd
