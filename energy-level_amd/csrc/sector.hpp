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
// Tiling: a code splits into its high part h (bits >= L) and its low part
// (bits < L).  All codes sharing h are the contiguous rank range
// [R(h), R(h) + cnt(h)), R(h) = Σ_{p >= L, b_p = 1} W_p, cnt(h) = F(L+2), or
// F(L+1) when bit L of h is set (bit L-1 must then stay clear); inside it the
// offset is the Zeckendorf rank of the low part.  One block owns one high
// part: every flip of a low bit stays inside the block (LDS), every flip of a
// high bit reads one contiguous partner range R(h') + offset, and whether that
// flip is allowed depends on h alone (bit L adds the offset < F(L+1) test).
// The diagonal is E_hi(h) + E_lo[offset] + cross(h, low bits), with the cross
// field of h on the low bits held in per-block byte tables (LDS).
#pragma once
#include "common.hpp"
#include "hv_full.hpp"
#include "observables.hpp"

namespace ryd {

constexpr int FIB_L = 18;          // low bits per block
constexpr int FIB_TMAX = 6765;     // F(FIB_L + 2): codes of the largest block
constexpr int FIB_EPT = (FIB_TMAX + 255) / 256;   // offsets per thread (27)

struct FibHi {
  uint64_t code;    // high part (bits >= L, in place)
  int64_t R;        // rank of its first code
  double ehi;       // diagonal of the high bits alone (delta and V among them)
  double pad;
};

struct FibTop {
  uint64_t g;       // top pattern (bits >= S, in place)
  int64_t R;        // rank offset of the pattern
};

struct FibModel {
  int n;
  int L;                      // min(n, FIB_L)
  int far_end;                // high bits L .. far_end-1 are applied (n: all)
  int pad;
  int64_t dim;                // F(n+2)
  int64_t nh;                 // high parts, F(n-L+2)
  const FibHi* hi;            // [nh] in ascending order
  const uint32_t* lo_code;    // [F(L+2)] L-bit sector codes, ascending
  const double* lo_diag;      // [F(L+2)] diagonal from bits < L
  const double* vb;           // [n*n] V by bit position
  int64_t W[MAXN];            // W_p = F(p+2)
  int64_t cnt_lo[2];          // F(L+2), F(L+1)
  double ar[MAXN], ai[MAXN];  // Ω/2 by bit position
  // top pass (bits S .. n-1, S = n - FIB_K) of a split H·v
  int S;
  int ng;                     // valid FIB_K-bit top patterns, F(FIB_K + 2)
  const struct FibTop* top;   // [ng] ascending
  int64_t nbody[2];           // F(S+2), F(S+1): body ranks, all / with bit S-1 clear
};

// Block order over the high parts.  RYD_FIB_XCD = 1 lets XCD b%8 sweep its
// own contiguous eighth; the default sweeps them in rank order across all
// XCDs, which keeps one compact front whose recent high parts the MALL still
// holds when a far high-bit flip reads them (N=40 FP32 2.53 -> 2.41 ms).
#ifndef RYD_FIB_XCD
#define RYD_FIB_XCD 0
#endif
// Runs of RYD_FIB_CHUNK consecutive high parts per XCD inside the global
// rank-order front: N=40 FP32 2.41 -> 2.38 ms, FP64 4.71 -> 4.66 ms per H·v
// (runs of 64 or 256: 2.40 ms).
#ifndef RYD_FIB_CHUNK
#define RYD_FIB_CHUNK 16
#endif
__device__ __forceinline__ int64_t fib_block_hi(const FibModel& M, int64_t b) {
  const int64_t per = (M.nh + 7) / 8;
  if (RYD_FIB_CHUNK > 0 && M.nh >= 8 * RYD_FIB_CHUNK) {
    // XCD b % 8 takes runs of RYD_FIB_CHUNK consecutive high parts; the
    // eight XCDs work on adjacent runs (one compact front)
    const int64_t xcd = b & 7, k = b >> 3;
    const int64_t hb = ((k / RYD_FIB_CHUNK) * 8 + xcd) * RYD_FIB_CHUNK + k % RYD_FIB_CHUNK;
    return hb;
  }
  return (RYD_FIB_XCD && M.nh >= 64) ? ((b & 7) * per + (b >> 3)) : b;
}

// per-block cross field of the high part on the low bits, as byte tables
__device__ inline void fib_cross_tables(const FibModel& M, uint64_t h, double* Fx, double* T0,
                                        double* T1, double* T2) {
  const int t = threadIdx.x;
  if (t < M.L) {
    double f = 0.0;
    const double* row = M.vb + (size_t)t * M.n;
    for (uint64_t c = h; c; c &= c - 1) f += row[__builtin_ctzll(c)];
    Fx[t] = f;
  }
  __syncthreads();
  if (t < 256) {
    double v0 = 0.0, v1 = 0.0;
#pragma unroll
    for (int p = 0; p < 8; ++p) {
      if ((t >> p) & 1) {
        if (p < M.L) v0 += Fx[p];
        if (p + 8 < M.L) v1 += Fx[p + 8];
      }
    }
    T0[t] = v0;
    T1[t] = v1;
  }
  if (t < 4) {
    double v2 = 0.0;
    for (int p = 0; p < 2; ++p)
      if (((t >> p) & 1) && p + 16 < M.L) v2 += Fx[p + 16];
    T2[t] = v2;
  }
  __syncthreads();
}

// Lane layout inside a block: the low part splits into inner bits (< FIB_IB)
// and outer bits (FIB_IB .. L-1).  Codes sharing the outer part o form a
// group of F(IB+2) consecutive offsets starting at B(o) = Σ_{outer b_p} W_p
// (F(IB+1) when bit IB is set), and lane ℓ of a wavefront owns inner rank ℓ
// of every group its wave handles.  So an inner flip is a fixed per-lane
// partner offset and coefficient, an outer or high flip is allowed or not for
// the whole group (bit IB adds the lane's inner bit IB-1), and all LDS and
// global accesses of a group are contiguous across lanes.
#ifndef RYD_FIB_BATCH
#define RYD_FIB_BATCH 1
#endif
#ifndef RYD_FIB_FL
#define RYD_FIB_FL 1
#endif
#ifndef RYD_FIB_FL_WAVES
#define RYD_FIB_FL_WAVES(V) (sizeof(V) >= 8 ? 2 : 4)
#endif
#ifndef RYD_FIB_FL_NW4
#define RYD_FIB_FL_NW4 8
#endif
#ifndef RYD_FIB_FL_NW8
#define RYD_FIB_FL_NW8 8
#endif
constexpr int FIB_IB = 8;          // inner bits: groups of F(10) = 55 or F(9) = 34 offsets
constexpr int FIB_THREADS = 512;   // 8 wavefronts per block
constexpr int FIB_WAVES = FIB_THREADS / 64;
constexpr int FIB_GPW = 18;        // groups per wavefront: F(12) = 144 groups over 8 waves
constexpr int FIB_SEPT = (FIB_TMAX + FIB_THREADS - 1) / FIB_THREADS;   // staged offsets per thread
constexpr int FIB_PAD = 128;       // LDS slack read (never used) by lanes past a group's end

template <class V>
__global__ __launch_bounds__(FIB_THREADS, 4) void k_fib_apply(FibModel M, ApplyIO<V> io, Reduce rd) {
  using RT = typename VT<V>::R;
  __shared__ V xt[FIB_TMAX + FIB_PAD];
  __shared__ double Fx[FIB_L];
  __shared__ double T0[256], T1[256], T2[4];
  __shared__ int gbase[FIB_WAVES * FIB_GPW];
  __shared__ uint32_t gcode[FIB_WAVES * FIB_GPW];
  __shared__ uint32_t gmask[FIB_WAVES * FIB_GPW];   // outer flips allowed for the group
  if (io.stop && load_coherent(io.stop)) return;
  double s, ch, cx, cz;
  apply_scalars(io, s, ch, cx, cz);
  const RT sr = (RT)s;
  const int t = threadIdx.x, lane = t & 63;
  const int wv = __builtin_amdgcn_readfirstlane(t >> 6);   // wave-uniform (SGPR)
  const int64_t hb = fib_block_hi(M, blockIdx.x);
  if (hb >= M.nh) {   // tail of the XCD remap: idle block still joins the reduction
    if (io.reduce) block_reduce_store(0.0, 0.0, 0.0, 0.0, rd.partials);
    return;
  }
  const FibHi H = M.hi[hb];
  const uint64_t h = H.code;
  const int hL = (int)((h >> M.L) & 1ull);
  const int cnt = (int)M.cnt_lo[hL];
  const int64_t R = H.R;
  const int IB = M.L < FIB_IB ? M.L : FIB_IB;
  const int nout = M.L - IB;
  const int Fbig = (int)M.W[IB];               // F(IB+2): group size, bit IB clear
  const int Fsmall = (int)(IB > 0 ? M.W[IB - 1] : 1);   // F(IB+1): bit IB set
  // groups (outer parts) of this block, ascending
  const int ng = nout > 0 ? (int)(hL ? M.W[nout - 1] : M.W[nout]) : 1;
  // stage s·x with every load in flight before the first store
  {
    V v[FIB_SEPT];
#pragma unroll
    for (int j = 0; j < FIB_SEPT; ++j) {
      const int k = j * FIB_THREADS + t;
      v[j] = k < cnt ? io.x[R + k] : vzero(v[0]);
    }
#pragma unroll
    for (int j = 0; j < FIB_SEPT; ++j) {
      const int k = j * FIB_THREADS + t;
      if (k < cnt) xt[k] = vscale(v[j], sr);
    }
  }
  for (int g = t; g < ng; g += FIB_THREADS) {
    uint32_t o = 0;
    int r = g, prev = 0, base = 0;
    for (int q = nout - 1; q >= 0; --q) {
      if (!prev && r >= (int)M.W[q]) {
        o |= 1u << q;
        r -= (int)M.W[q];
        base += (int)M.W[IB + q];
        prev = 1;
      } else {
        prev = 0;
      }
    }
    uint32_t mk = 0;
    for (int q = 0; q < nout; ++q) {
      const bool bit = (o >> q) & 1u;
      const bool up = q + 1 < nout ? ((o >> (q + 1)) & 1u) != 0 : hL != 0;
      const bool dn = q > 0 && ((o >> (q - 1)) & 1u);
      if (bit || !(up || dn)) mk |= 1u << q;
    }
    gcode[g] = o;
    gbase[g] = base;
    gmask[g] = mk;
  }
  fib_cross_tables(M, h, Fx, T0, T1, T2);     // its barriers also publish xt and the groups
  // per-lane inner code, flip offsets and coefficients (upper neighbour of
  // bit IB-1 taken clear; a set bit IB is handled per group below)
  uint32_t cin = 0;
  {
    int r = lane, prev = 0;
    for (int p = IB - 1; p >= 0; --p) {
      if (!prev && r >= (int)M.W[p]) { cin |= 1u << p; r -= (int)M.W[p]; prev = 1; }
      else prev = 0;
    }
  }
  RT cfr[FIB_IB], cfi[FIB_IB];
  int off[FIB_IB];
#pragma unroll
  for (int p = 0; p < FIB_IB; ++p) {
    const int bit = (int)((cin >> p) & 1u);
    const bool fr = !((cin >> (p + 1)) & 1u) && !(p > 0 && ((cin >> (p - 1)) & 1u));
    const bool ok = p < IB && (bit || fr);
    off[p] = ok ? (bit ? -(int)M.W[p] : (int)M.W[p]) : 0;   // blocked: own slot, coefficient 0
    cfr[p] = ok ? (RT)M.ar[p] : (RT)0;
    cfi[p] = ok ? (RT)(bit ? M.ai[p] : -M.ai[p]) : (RT)0;
  }
  const bool top_in = IB > 0 && ((cin >> (IB - 1)) & 1u);   // inner bit IB-1 set
  // outer-bit coefficients: FP32 converts them once into registers, FP64
  // reads them from the (scalar) kernel arguments
  constexpr int NOUT = FIB_L - FIB_IB;
  constexpr bool CONV = sizeof(RT) == 4;
  RT oar[CONV ? NOUT : 1], oai[CONV ? NOUT : 1];
  if constexpr (CONV) {
#pragma unroll
    for (int q = 0; q < NOUT; ++q) {
      const int p = q + IB < MAXN ? q + IB : 0;
      oar[q] = (RT)M.ar[p];
      oai[q] = (RT)M.ai[p];
    }
  }
  auto car = [&](int q) -> RT { if constexpr (CONV) return oar[q]; else return (RT)M.ar[IB + q]; };
  auto cai = [&](int q) -> RT { if constexpr (CONV) return oai[q]; else return (RT)M.ai[IB + q]; };
  // FP32: every flip of a group predicated and its LDS reads batched (wider
  // types would spill at 128 registers and keep the branch form)
  constexpr bool BATCH = RYD_FIB_BATCH && sizeof(V) == 4;
  int Wq[BATCH ? NOUT : 1];      // outer-bit weights held for the whole block
  if constexpr (BATCH) {
#pragma unroll
    for (int q = 0; q < NOUT; ++q) Wq[q] = (int)M.W[IB + q < MAXN ? IB + q : 0];
  }
  const double eh = H.ehi;
  V acc[FIB_GPW];
  int koff[FIB_GPW];   // this lane's offset in the block for each group of the wave
  bool act[FIB_GPW];
#pragma unroll
  for (int i = 0; i < FIB_GPW; ++i) {
    const int g = wv + FIB_WAVES * i;
    const bool gv = g < ng;
    // group code, base and mask are the same for the whole wave: keep them
    // scalar so every group test below is a uniform branch
    const uint32_t o = (uint32_t)__builtin_amdgcn_readfirstlane(gv ? (int)gcode[g] : 0);
    const uint32_t om = (uint32_t)__builtin_amdgcn_readfirstlane(gv ? (int)gmask[g] : 0);
    const int gbi = __builtin_amdgcn_readfirstlane(gv ? gbase[g] : 0);
    const bool small = nout > 0 ? (o & 1u) : (hL != 0);
    const int k = gbi + lane;
    koff[i] = k;
    act[i] = gv && lane < (small ? Fsmall : Fbig);
    const uint32_t lo = (o << IB) | cin;
    const double d = act[i] ? eh + M.lo_diag[k] + T0[lo & 255u] + T1[(lo >> 8) & 255u] +
                                  T2[(lo >> 16) & 3u]
                            : 0.0;
    acc[i] = vscale(xt[k], (RT)d);
    if (!gv) continue;
    if constexpr (BATCH) {
      // every flip of the group predicated (a blocked flip reads the lane's
      // own slot with coefficient 0); all LDS reads issued before the FMAs
      constexpr int NF = FIB_IB + NOUT;
      int fo[NF];
      RT fr[NF], fi[NF];
#pragma unroll
      for (int p = 0; p < FIB_IB; ++p) {
        const bool skip = p == IB - 1 && small;          // uniform
        fo[p] = skip ? 0 : off[p];
        fr[p] = skip ? (RT)0 : cfr[p];
        fi[p] = skip ? (RT)0 : cfi[p];
      }
#pragma unroll
      for (int q = 0; q < NOUT; ++q) {
        const bool bit = (o >> q) & 1u;
        bool use = q < nout && ((om >> q) & 1u);         // uniform
        if (q == 0 && !bit) use = use && !top_in;        // setting bit IB: per lane
        const int W = Wq[BATCH ? q : 0];
        fo[FIB_IB + q] = use ? (bit ? -W : W) : 0;
        fr[FIB_IB + q] = use ? car(q) : (RT)0;
        fi[FIB_IB + q] = use ? (bit ? cai(q) : -cai(q)) : (RT)0;
      }
      V xv[NF];
#pragma unroll
      for (int f = 0; f < NF; ++f) xv[f] = xt[k + fo[f]];
#pragma unroll
      for (int f = 0; f < NF; ++f) vcfma(acc[i], fr[f], fi[f], xv[f]);
      continue;
    }
    // inner flips: fixed per lane; in a small group bit IB-1 cannot be set
#pragma unroll
    for (int p = 0; p < FIB_IB; ++p) {
      if (p == IB - 1 && small) continue;
      vcfma(acc[i], cfr[p], cfi[p], xt[k + off[p]]);
    }
    // outer flips: allowed per group (mask); setting bit IB also needs the
    // lane's inner bit IB-1 clear (a blocked lane reads its own finite slot)
#pragma unroll
    for (int q = 0; q < NOUT; ++q) {
      if (q >= nout || !((om >> q) & 1u)) continue;
      const bool bit = (o >> q) & 1u;
      const int W = (int)M.W[IB + q];
      if (q == 0 && !bit) {
        const RT ar = top_in ? (RT)0 : car(0), ai = top_in ? (RT)0 : -cai(0);
        vcfma(acc[i], ar, ai, xt[top_in ? k : k + W]);
      } else {
        vcfma(acc[i], car(q), bit ? cai(q) : -cai(q), xt[k + (bit ? -W : W)]);
      }
    }
  }
  // high bits: one contiguous partner range per allowed bit; the loads of
  // every group of the wave are in flight together (uniform base pointer,
  // per-lane 32-bit offsets)
  for (int p = M.L; p < M.far_end; ++p) {
    const int bit = (int)((h >> p) & 1ull);
    const bool set_ok = !((h >> (p + 1)) & 1ull) && (p == M.L || !((h >> (p - 1)) & 1ull));
    if (!bit && !set_ok) continue;
    const bool need_lo = !bit && p == M.L;     // setting bit L: low bit L-1 must be clear
    const int lim = need_lo ? (int)M.cnt_lo[1] : cnt;
    const V* px = io.x + (R + (bit ? -M.W[p] : M.W[p]));
    const RT ar = (RT)(M.ar[p] * s), ai = (RT)((bit ? M.ai[p] : -M.ai[p]) * s);
    V q[FIB_GPW];
#pragma unroll
    for (int i = 0; i < FIB_GPW; ++i) q[i] = act[i] && koff[i] < lim ? px[koff[i]] : vzero(acc[0]);
#pragma unroll
    for (int i = 0; i < FIB_GPW; ++i) vcfma(acc[i], ar, ai, q[i]);
  }
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  const RT chr = (RT)ch, cxr = (RT)cx, czr = (RT)cz;
  V zr[FIB_GPW];
  V* py = io.y + R;
  if (io.z && !io.partial) {
    const V* pz = io.z + R;
#pragma unroll
    for (int i = 0; i < FIB_GPW; ++i) zr[i] = act[i] ? pz[koff[i]] : vzero(acc[0]);
  }
#pragma unroll
  for (int i = 0; i < FIB_GPW; ++i) {
    if (!act[i]) continue;
    if (io.partial) {
      py[koff[i]] = acc[i];
      continue;
    }
    const V xe = xt[koff[i]];
    V v = vscale(acc[i], chr);
    vfma(v, cxr, xe);
    if (io.z) vfma(v, czr, zr[i]);
    py[koff[i]] = v;
    if (io.reduce) {
      s0 += rdot(xe, v);
      s1 += nrm2(v);
      s2 += nrm2(xe);
      if (io.probe) {
        const int64_t r = R + koff[i];
        s3 += zdot(((io.probe[r >> 5] >> (r & 31)) & 1u) ? 1.0 : -1.0, v);
      }
    }
  }
  if (io.reduce) block_reduce_store(s0, s1, s2, s3, rd.partials);
}

// The same H·v for blocks with the full FIB_L low bits (n >= FIB_L), written
// so that a group costs no scalar instructions.  The setup tabulates, per
// group, the LDS base of every outer-bit partner run, or the base of a
// 64-slot zero stripe when the group may not flip that bit; a blocked flip
// then reads zeros with the flip's own coefficient, so the loop over a
// wave's groups is straight-line vector and LDS code.  The diagonal splits
// as lo_diag[k] + T0[inner] (per lane) + T1, T2 of the outer code (per
// group).  High-bit partner runs are read through a buffer descriptor of the
// run's length, whose range check zeroes the lanes past it.
// Measured at N=40 (FP32 / FP64 pass A + top pass, ms per H·v): the
// branch-per-group kernel above 4.28 / ~6.0; this form 2.70 / 5.02; with
// buffer loads 2.58 / 4.78; global block order 2.41 / 4.70; 16 waves per
// block 3.22 / 4.79; an FP32 top pass 3.04.  Its FP32 form fetches about
// 10× the vector from beyond L2 (high-bit partners), so it is bound there.
constexpr int FIB_NOUT = FIB_L - FIB_IB;          // 10 outer bits
constexpr int FIB_NGMAX = 144;                    // F(12): groups of a block
constexpr int FIB_ZB = FIB_TMAX + FIB_PAD;        // first slot of the zero stripe
constexpr int FIB_PBW = 12;                       // ints per group row: 10 bases, base, code

// NW wavefronts per block (8 or 16): 16 halves the groups a wave holds, and
// with them the registers of 8-byte types.
template <class V> constexpr int fib_fl_waves() {
  return sizeof(V) == 4 ? RYD_FIB_FL_NW4 : (sizeof(V) == 16 ? 8 : RYD_FIB_FL_NW8);
}

template <class V, int NW = fib_fl_waves<V>()>
__global__ __launch_bounds__(NW * 64, NW >= 16 ? 4 : RYD_FIB_FL_WAVES(V)) void k_fib_apply_fl(FibModel M, ApplyIO<V> io, Reduce rd) {
  using RT = typename VT<V>::R;
  constexpr int NT = NW * 64;                       // threads
  constexpr int GPW = (FIB_NGMAX + NW - 1) / NW;    // groups per wave
  constexpr int SEPT = (FIB_TMAX + NT - 1) / NT;    // staged codes per thread
  constexpr bool CPLX = VT<V>::cplx;
  constexpr int IB = FIB_IB;
  __shared__ V xt[FIB_ZB + 64];
  __shared__ double Fx[FIB_L];
  __shared__ double T0[256], T1[256], T2[4];
  __shared__ __align__(16) int gpb[FIB_NGMAX * FIB_PBW];
  if (io.stop && load_coherent(io.stop)) return;
  double s, ch, cx, cz;
  apply_scalars(io, s, ch, cx, cz);
  const RT sr = (RT)s;
  const int t = threadIdx.x, lane = t & 63;
  const int wv = __builtin_amdgcn_readfirstlane(t >> 6);
  const int64_t hb = fib_block_hi(M, blockIdx.x);
  if (hb >= M.nh) {
    if (io.reduce) block_reduce_store(0.0, 0.0, 0.0, 0.0, rd.partials);
    return;
  }
  const FibHi H = M.hi[hb];
  const uint64_t h = H.code;
  const int hL = (int)((h >> FIB_L) & 1ull);
  const int cnt = (int)M.cnt_lo[hL];
  const int64_t R = H.R;
  const int ng = (int)(hL ? M.W[FIB_NOUT - 1] : M.W[FIB_NOUT]);   // F(11) or F(12)
  // the block's codes are requested first; the group tables below are
  // computed while the loads are in flight, and only then staged in LDS
  V xstage[SEPT];
#pragma unroll
  for (int j = 0; j < SEPT; ++j) {
    const int k = j * NT + t;
    xstage[j] = k < cnt ? io.x[R + k] : vzero(xstage[0]);
  }
  if (t < 64) xt[FIB_ZB + t] = vzero(xt[0]);
  for (int g = t; g < ng; g += NT) {
    uint32_t o = 0;
    int r = g, prev = 0, base = 0;
    for (int q = FIB_NOUT - 1; q >= 0; --q) {
      if (!prev && r >= (int)M.W[q]) {
        o |= 1u << q;
        r -= (int)M.W[q];
        base += (int)M.W[IB + q];
        prev = 1;
      } else {
        prev = 0;
      }
    }
    int* row = gpb + g * FIB_PBW;
    for (int q = 0; q < FIB_NOUT; ++q) {
      const bool bit = (o >> q) & 1u;
      const bool up = q + 1 < FIB_NOUT ? ((o >> (q + 1)) & 1u) != 0 : hL != 0;
      const bool dn = q > 0 && ((o >> (q - 1)) & 1u);
      const int W = (int)M.W[IB + q];
      row[q] = bit ? base - W : (!(up || dn) ? base + W : FIB_ZB);
    }
    row[FIB_NOUT] = base;
    row[FIB_NOUT + 1] = (int)o;
  }
#pragma unroll
  for (int j = 0; j < SEPT; ++j) {
    const int k = j * NT + t;
    if (k < cnt) xt[k] = vscale(xstage[j], sr);
  }
  fib_cross_tables(M, h, Fx, T0, T1, T2);     // its barriers publish xt, the stripe and gpb
  // per-lane inner code, flip offsets and coefficients (upper neighbour of
  // bit IB-1 taken clear; a set bit IB is the group's "small" flag)
  uint32_t cin = 0;
  {
    int r = lane, prev = 0;
    for (int p = IB - 1; p >= 0; --p) {
      if (!prev && r >= (int)M.W[p]) { cin |= 1u << p; r -= (int)M.W[p]; prev = 1; }
      else prev = 0;
    }
  }
  RT cfr[IB], cfi[IB];
  int off[IB];
#pragma unroll
  for (int p = 0; p < IB; ++p) {
    const int bit = (int)((cin >> p) & 1u);
    const bool fr = !((cin >> (p + 1)) & 1u) && !(p > 0 && ((cin >> (p - 1)) & 1u));
    const bool ok = bit || fr;
    off[p] = ok ? (bit ? -(int)M.W[p] : (int)M.W[p]) : 0;   // blocked: own slot, coefficient 0
    cfr[p] = ok ? (RT)M.ar[p] : (RT)0;
    cfi[p] = ok ? (RT)(bit ? M.ai[p] : -M.ai[p]) : (RT)0;
  }
  const int zin = ((cin >> (IB - 1)) & 1u) ? FIB_ZB : 0;   // setting bit IB: blocked for this lane
  const double t0c = T0[cin];
  RT oar[FIB_NOUT], oai[FIB_NOUT];
#pragma unroll
  for (int q = 0; q < FIB_NOUT; ++q) {
    oar[q] = (RT)M.ar[IB + q];
    oai[q] = (RT)M.ai[IB + q];
  }
  const double eh = H.ehi;
  V acc[GPW];
  uint32_t kb[GPW];     // byte offset of this lane's code in the block, per group of the wave
  uint32_t actm = 0;        // bit i: group i of the wave holds this lane
#pragma unroll
  for (int i = 0; i < GPW; ++i) {
    const int g = wv + NW * i;
    if (g >= ng) {                                 // uniform
      kb[i] = (uint32_t)lane * (uint32_t)sizeof(V);
      acc[i] = vzero(acc[0]);
      continue;
    }
    const int4* row = reinterpret_cast<const int4*>(gpb + g * FIB_PBW);
    const int4 r0 = row[0], r1 = row[1], r2 = row[2];
    const int pb[FIB_NOUT] = {r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w, r2.x, r2.y};
    const int base = r2.z;
    const uint32_t o = (uint32_t)r2.w;
    const bool small = o & 1u;
    const int k = base + lane;
    kb[i] = (uint32_t)k * (uint32_t)sizeof(V);
    if (lane < (small ? (int)M.W[IB - 1] : (int)M.W[IB])) actm |= 1u << i;
    const int kd = k < FIB_TMAX ? k : FIB_TMAX - 1;
    const double d = eh + M.lo_diag[kd] + t0c + T1[o & 255u] + T2[(o >> 8) & 3u];
    V a = vscale(xt[k], (RT)d);
    {
      V xv[IB];
#pragma unroll
      for (int p = 0; p < IB - 1; ++p) xv[p] = xt[k + off[p]];
      xv[IB - 1] = xt[small ? FIB_ZB + lane : k + off[IB - 1]];
#pragma unroll
      for (int p = 0; p < IB; ++p) vcfma(a, cfr[p], cfi[p], xv[p]);
    }
    {
      V xv[FIB_NOUT];
      xv[0] = xt[((o & 1u) ? pb[0] : (zin ? FIB_ZB : pb[0])) + lane];
#pragma unroll
      for (int q = 1; q < FIB_NOUT; ++q) xv[q] = xt[pb[q] + lane];
#pragma unroll
      for (int q = 0; q < FIB_NOUT; ++q) {
        RT ci = (RT)0;
        if constexpr (CPLX) ci = ((o >> q) & 1u) ? oai[q] : -oai[q];
        vcfma(a, oar[q], ci, xv[q]);
      }
    }
    acc[i] = a;
  }
  // high bits: one contiguous partner range per allowed bit, read through a
  // buffer descriptor of the range's length (lanes past it read zero), every
  // load of the wave in flight together
  for (int p = FIB_L; p < M.far_end; ++p) {
    const int bit = (int)((h >> p) & 1ull);
    const bool set_ok = !((h >> (p + 1)) & 1ull) && (p == FIB_L || !((h >> (p - 1)) & 1ull));
    if (!bit && !set_ok) continue;
    const bool need_lo = !bit && p == FIB_L;     // setting bit L: low bit L-1 must be clear
    const uint32_t lim = need_lo ? (uint32_t)M.cnt_lo[1] : (uint32_t)cnt;
    const auto rs = buf_rsrc(io.x + (R + (bit ? -M.W[p] : M.W[p])), lim * (uint32_t)sizeof(V));
    const RT ar = (RT)(M.ar[p] * s), ai = (RT)((bit ? M.ai[p] : -M.ai[p]) * s);
    // in two halves of the wave's groups (register pressure of 8-byte types)
    constexpr int HG = (GPW + 1) / 2;
#pragma unroll
    for (int hf = 0; hf < 2; ++hf) {
      V q[HG];
#pragma unroll
      for (int j = 0; j < HG; ++j)
        if (hf * HG + j < GPW) q[j] = buf_load<V>(rs, kb[hf * HG + j]);
#pragma unroll
      for (int j = 0; j < HG; ++j)
        if (hf * HG + j < GPW) vcfma(acc[hf * HG + j], ar, ai, q[j]);
    }
  }
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  const RT chr = (RT)ch, cxr = (RT)cx, czr = (RT)cz;
  char* py = reinterpret_cast<char*>(io.y + R);
  const char* pz = io.z ? reinterpret_cast<const char*>(io.z + R) : nullptr;
  const char* pxt = reinterpret_cast<const char*>(xt);
#pragma unroll
  for (int i = 0; i < GPW; ++i) {
    if (!((actm >> i) & 1u)) continue;
    V* yi = reinterpret_cast<V*>(py + kb[i]);
    if (io.partial) {
      *yi = acc[i];
      continue;
    }
    const V xe = *reinterpret_cast<const V*>(pxt + kb[i]);
    V v = vscale(acc[i], chr);
    vfma(v, cxr, xe);
    if (io.z) vfma(v, czr, *reinterpret_cast<const V*>(pz + kb[i]));
    *yi = v;
    if (io.reduce) {
      s0 += rdot(xe, v);
      s1 += nrm2(v);
      s2 += nrm2(xe);
      if (io.probe) {
        const int64_t r = R + kb[i] / (uint32_t)sizeof(V);
        s3 += zdot(((io.probe[r >> 5] >> (r & 31)) & 1u) ? 1.0 : -1.0, v);
      }
    }
  }
  if (io.reduce) block_reduce_store(s0, s1, s2, s3, rd.partials);
}

// Pass B of a split sector H·v: the top FIB_K bits.  A code is a top pattern
// g (bits >= S) over a body of S bits with body rank ρ; its rank is
// R(g) + ρ, and bit S of g set needs the body's top bit clear (ρ < F(S+1)).
// A block holds one run of FIB_RUN consecutive body ranks (256 bytes) for
// every valid pattern, so each top-bit flip (pattern rank ± F(q+2)) is an LDS
// read.  It reads the pass-A partial from y and finishes
//   y = ch·(y_A + Σ_top flips of s·x) + cx·(s·x) + cz·z  with the reductions.
// N=40 FP64 H·v (pass A + this pass): 8 top bits 4.82 ms, 9 bits 4.69 ms,
// 10 bits 5.25 ms.
#ifndef RYD_FIB_K
#define RYD_FIB_K 9
#endif
constexpr int FIB_K = RYD_FIB_K;
constexpr int fib_const(int k) { return k < 2 ? k : fib_const(k - 1) + fib_const(k - 2); }
constexpr int FIB_NG = fib_const(FIB_K + 2);    // valid top patterns
template <class V> constexpr int fib_run() { return 256 / (int)sizeof(V); }

template <class V>
__global__ __launch_bounds__(256) void k_fib_top(FibModel M, ApplyIO<V> io, Reduce rd) {
  using RT = typename VT<V>::R;
  constexpr int RUN = fib_run<V>();
  constexpr int GR = 256 / RUN;                 // patterns per pass of the block
  constexpr int ROWS = (FIB_NG + GR - 1) / GR;
  __shared__ V xt[FIB_NG][RUN];
  __shared__ FibTop tp[FIB_NG];
  if (io.stop && load_coherent(io.stop)) return;
  double s, ch, cx, cz;
  apply_scalars(io, s, ch, cx, cz);
  const RT sr = (RT)s;
  const int t = threadIdx.x;
  const int rl = t % RUN, gr = t / RUN;
  const int64_t nruns = (M.nbody[0] + RUN - 1) / RUN;
  const int64_t per = (nruns + 7) / 8;
  const int64_t m = nruns >= 64 ? ((int64_t)(blockIdx.x & 7) * per + (blockIdx.x >> 3)) : blockIdx.x;
  if (m >= nruns) {
    if (io.reduce) block_reduce_store(0.0, 0.0, 0.0, 0.0, rd.partials);
    return;
  }
  if (t < FIB_NG) tp[t] = M.top[t];
  __syncthreads();
  const int64_t rho = m * RUN + rl;
  V acc[ROWS], xs[ROWS], zr[ROWS];
  bool ok[ROWS];
#pragma unroll
  for (int j = 0; j < ROWS; ++j) {
    const int gi = gr + GR * j;
    bool v = gi < FIB_NG && rho < M.nbody[0];
    if (v) v = !((tp[gi].g >> M.S) & 1ull) || rho < M.nbody[1];
    ok[j] = v;
    const int64_t r = v ? tp[gi].R + rho : 0;
    xs[j] = v ? io.x[r] : vzero(acc[0]);
    acc[j] = v ? io.y[r] : vzero(acc[0]);
    if (io.z) zr[j] = v ? io.z[r] : vzero(acc[0]);
  }
#pragma unroll
  for (int j = 0; j < ROWS; ++j) {
    const int gi = gr + GR * j;
    xs[j] = vscale(xs[j], sr);
    if (gi < FIB_NG) xt[gi][rl] = xs[j];
  }
  __syncthreads();
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  const RT chr = (RT)ch, cxr = (RT)cx, czr = (RT)cz;
#pragma unroll
  for (int j = 0; j < ROWS; ++j) {
    if (!ok[j]) continue;
    const int gi = gr + GR * j;
    const uint64_t g = tp[gi].g;
    V a = acc[j];
    int w = 1, wprev = 1;                       // F(q+2), F(q+1)
    for (int q = 0; q < FIB_K; ++q) {
      const int p = M.S + q;
      const int bit = (int)((g >> p) & 1ull);
      const bool up = (g >> (p + 1)) & 1ull;
      const bool dn = q == 0 ? rho >= M.nbody[1] : (((g >> (p - 1)) & 1ull) != 0);
      if (bit || !(up || dn)) {
        const RT ai = (RT)M.ai[p];
        vcfma(a, (RT)M.ar[p], bit ? ai : -ai, xt[gi + (bit ? -w : w)][rl]);
      }
      const int wn = w + wprev;
      wprev = w;
      w = wn;
    }
    const int64_t r = tp[gi].R + rho;
    V v = vscale(a, chr);
    vfma(v, cxr, xs[j]);
    if (io.z) vfma(v, czr, zr[j]);
    io.y[r] = v;
    if (io.reduce) {
      s0 += rdot(xs[j], v);
      s1 += nrm2(v);
      s2 += nrm2(xs[j]);
      if (io.probe) s3 += zdot(((io.probe[r >> 5] >> (r & 31)) & 1u) ? 1.0 : -1.0, v);
    }
  }
  if (io.reduce) block_reduce_store(s0, s1, s2, s3, rd.partials);
}

// Site occupations over the sector.  Low bits accumulate per thread; the high
// part is constant over a block's range, so its total weight goes to the set
// high bits.  partials: [grid][MAXN_OCC + 1], column MAXN_OCC holds the total.
template <class V>
__global__ __launch_bounds__(256) void k_fib_occ(FibModel M, const V* __restrict__ x,
                                                 double* partials) {
  __shared__ double acc[MAXN_OCC + 1];
  __shared__ double red[4][FIB_L + 1];
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  if (t <= MAXN_OCC) acc[t] = 0.0;
  double lo_acc[FIB_L];
#pragma unroll
  for (int p = 0; p < FIB_L; ++p) lo_acc[p] = 0.0;
  for (int64_t hb = blockIdx.x; hb < M.nh; hb += gridDim.x) {
    const FibHi H = M.hi[hb];
    const int cnt = (int)M.cnt_lo[(H.code >> M.L) & 1ull];
    double tot = 0.0;
    for (int k = t; k < cnt; k += 256) {
      const double pw = nrm2(x[H.R + k]);
      const uint32_t c = M.lo_code[k];
      tot += pw;
#pragma unroll
      for (int p = 0; p < FIB_L; ++p)
        if ((c >> p) & 1u) lo_acc[p] += pw;
    }
    tot = wave_sum(tot);
    __syncthreads();
    if (lane == 0) red[wid][FIB_L] = tot;
    __syncthreads();
    if (t == 0) {
      const double T = red[0][FIB_L] + red[1][FIB_L] + red[2][FIB_L] + red[3][FIB_L];
      acc[MAXN_OCC] += T;
      for (uint64_t c = H.code; c; c &= c - 1) acc[__builtin_ctzll(c)] += T;
    }
  }
#pragma unroll
  for (int p = 0; p < FIB_L; ++p) {
    const double v = wave_sum(lo_acc[p]);
    if (lane == 0) red[wid][p] = v;
  }
  __syncthreads();
  if (t < M.L) acc[t] += red[0][t] + red[1][t] + red[2][t] + red[3][t];
  __syncthreads();
  if (t <= MAXN_OCC) partials[(size_t)blockIdx.x * (MAXN_OCC + 1) + t] = acc[t];
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
