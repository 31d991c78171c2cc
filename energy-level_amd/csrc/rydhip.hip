// C ABI of the Rydberg H·v engine: plan construction (host tables), launch
// logic and the device-resident Krylov drivers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/rydhip.h"
#include "common.hpp"
#include "hv_full.hpp"
#include "stream_ops.hpp"
#include "lz_small.hpp"
#include "channels.hpp"
#include "graph_apply.hpp"
#include "measure.hpp"
#include "observables.hpp"
#include "sector.hpp"
#include "evolve.hpp"
#include "batch.hpp"
#include "batch_csr.hpp"
#include "tridiag.hpp"
#include "reortho.hpp"

using namespace ryd;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(expr)                                                            \
  do {                                                                           \
    hipError_t _e = (expr);                                                      \
    if (_e != hipSuccess)                                                        \
      return fail(_e == hipErrorOutOfMemory ? RYD_ERR_MEMORY : RYD_ERR_DEVICE,   \
                  std::string(#expr) + ": " + hipGetErrorString(_e));            \
  } while (0)

#define LAUNCH_CHECK()                                                           \
  do {                                                                           \
    hipError_t _e = hipGetLastError();                                           \
    if (_e != hipSuccess)                                                        \
      return fail(RYD_ERR_DEVICE, std::string("kernel launch: ") + hipGetErrorString(_e)); \
  } while (0)

}  // namespace

struct ryd_plan {
  int n = 0;
  int basis = RYD_BASIS_FULL;
  int64_t dim = 0;
  int device = 0;
  bool cplx = false;                 // Ω has an imaginary part
  double ar[MAXN] = {0}, ai[MAXN] = {0};   // per bit position, already halved
  double* d_vb = nullptr;            // [n*n] V by bit position
  // diagonal tables per tile width tb = 10 + k (built for the widths the
  // element types and split passes use; tbits[k] = min(n, 10 + k))
  double* d_elo[NTBW] = {};
  double* d_ehi[NTBW] = {};
  int tbits[NTBW] = {};
  size_t ws_bytes = 0;
  // chain-blockade sector
  uint32_t* d_lo_code = nullptr;     // [F(L+2)] low-part codes
  double* d_lo_diag = nullptr;       // [F(L+2)] low-part diagonal
  FibHi* d_hi = nullptr;             // [nh] high parts in ascending order
  int64_t fib_nh = 0;
  FibTop* d_top = nullptr;           // [FIB_NG] top patterns (split H·v only)
  bool fib_split = false;
  // generic channel path: code bases and pair flips (J, A)
  bool generic = false;
  int rank_mode = 0;                 // RankMode
  int popcount = 0;                  // number sector
  int64_t* d_codes = nullptr;        // [dim] ascending
  int64_t* d_csr_ptr = nullptr;      // RYD_BASIS_CSR: [dim + 1] row pointers
  int32_t* d_csr_col = nullptr;      // [nnz] column indices
  double* d_csr_re = nullptr;        // [nnz] values (real part)
  double* d_csr_im = nullptr;        // [nnz] imaginary part, null for a real matrix
  int64_t* d_binom = nullptr;        // [64][BINOM_LD]
  int rs_L = 0;                      // number-sector split rank tables (RANK_NUMBER, n <= RS_MAXN)
  int64_t* d_rsplit = nullptr;       // [2^rs_L] + [2^(n-rs_L)][rs_L+1]
  PairChan* d_pairs = nullptr;       // pair flip channels
  int npairs = 0;
  double dl[MAXN] = {0};             // delta by bit position
  // independent-set basis of a blockade graph (graph.hpp)
  int gr_L = 0;                      // low bits
  int64_t gr_tl = 0;                 // low sets
  int64_t gr_nc = 0;                 // boundary classes
  uint64_t* d_hpk = nullptr;         // [2^H + 1]
  uint32_t* d_lo_idx = nullptr;      // [2^L]
  uint32_t* d_lowtab = nullptr;      // [tl]
  uint32_t* d_fc = nullptr;          // [nc]
  uint32_t* d_pre = nullptr;         // [nc][tl + 1] (RANK_GRAPH only)
  // k_graph_apply tables (RANK_GRAPH only)
  int64_t gr_nv = 0;                 // valid high parts
  uint32_t* d_hlist = nullptr;       // [nv]
  int64_t gr_rows = 0;               // rows of all classes
  uint32_t* d_post = nullptr;        // [rows]
  int64_t* d_post_off = nullptr;     // [nc]
  uint32_t* d_rcode = nullptr;       // [rows]
  double* d_rdiag = nullptr;         // [rows]
  void* d_lrow = nullptr;            // [L][rows], uint16_t entries when gr_l16
  bool gr_l16 = false;               // every segment shorter than 65535 rows
  // top pass of a split graph H·v (gr_K = 0: one pass)
  int gr_K = 0;
  int gr_nt = 0;
  int64_t gr_nbody = 0;
  uint32_t* d_tp = nullptr;          // [nt] valid top patterns
  int16_t* d_tpk = nullptr;          // [nt][K] partner indices
  int fib_L = 0;
  int64_t fib_cnt[2] = {0, 0};
};

namespace {

// tile bits of the single-pass kernel and of pass A of a split H·v
int tile_bits_of(int dtype) {
  return dtype == RYD_F32 ? tile_bits<float>()
         : dtype == RYD_F64 ? tile_bits<double>()
         : dtype == RYD_C64 ? tile_bits<c32>()
                            : tile_bits<c64>();
}

int passa_bits_of(int dtype) {
  return dtype == RYD_F32 ? tile_bits<float, passa_tt<float>()>()
         : dtype == RYD_F64 ? tile_bits<double, passa_tt<double>()>()
         : dtype == RYD_C64 ? tile_bits<c32, passa_tt<c32>()>()
                            : tile_bits<c64, passa_tt<c64>()>();
}

size_t elem_bytes(int dtype) {
  return dtype == RYD_F32 ? 4 : (dtype == RYD_C128 ? 16 : 8);
}

// diagonal tables of tile width `width`
FullModel full_model(const ryd_plan* P, int width) {
  FullModel M;
  const int k = width - 10;
  M.n = P->n;
  M.tb = P->tbits[k];
  M.elo = P->d_elo[k];
  M.ehi = P->d_ehi[k];
  M.vb = P->d_vb;
  for (int p = 0; p < MAXN; ++p) {
    M.ar[p] = P->ar[p];
    M.ai[p] = P->ai[p];
  }
  return M;
}

// Diagonal tables for tile bits tb: elo over bit positions < tb, ehi over >= tb.
void build_tables(int n, int tb, const std::vector<double>& dl, const std::vector<double>& vb,
                  std::vector<double>& elo, std::vector<double>& ehi) {
  const int64_t nl = 1ll << tb, nh = 1ll << (n - tb);
  elo.assign(nl, 0.0);
  ehi.assign(nh, 0.0);
  for (int64_t e = 1; e < nl; ++e) {
    const int p = __builtin_ctzll(e);
    const int64_t rest = e & (e - 1);
    double v = elo[rest] + dl[p];
    for (int64_t r = rest; r; r &= r - 1) v += vb[p * n + __builtin_ctzll(r)];
    elo[e] = v;
  }
  for (int64_t h = 1; h < nh; ++h) {
    const int q = __builtin_ctzll(h);
    const int64_t rest = h & (h - 1);
    double v = ehi[rest] + dl[tb + q];
    for (int64_t r = rest; r; r &= r - 1) v += vb[(tb + q) * n + tb + __builtin_ctzll(r)];
    ehi[h] = v;
  }
}

// Split threshold: from this n on, the top bits run in a second pass
// (k_full_top) so that pass A's partner tiles stay within the caches.  The
// split costs a write and a read of the partial: it pays from n = 23 for
// 8- and 16-byte elements and for an FP32 H·v, from n = 25 for an FP32 step
// that also streams z.  Measured on MI355X, FP32 ms per H·v / per Chebyshev
// step (H·v with z): n = 23 split 0.035 / 0.050, one pass 0.040 / 0.045;
// n = 24 0.076 / 0.107 against 0.085 / 0.093; n = 26 0.27 / 0.40 against
// 0.27 / 0.44; n = 28 1.18 / 1.64 against 1.18 (split) / 2.12.
#ifndef RYD_SPLIT_MIN_N
#define RYD_SPLIT_MIN_N 23
#endif
// FP32 thresholds above SPLIT_MIN_N: an H·v, and a step that streams z
#ifndef RYD_SPLIT4_DN
#define RYD_SPLIT4_DN 0
#endif
#ifndef RYD_SPLIT4Z_DN
#define RYD_SPLIT4Z_DN 2
#endif
constexpr int SPLIT_MIN_N = RYD_SPLIT_MIN_N;
constexpr int split_min_n(size_t elem, bool with_z) {
  if (elem >= 8) return SPLIT_MIN_N;
  return with_z ? SPLIT_MIN_N + RYD_SPLIT4Z_DN : SPLIT_MIN_N + RYD_SPLIT4_DN;
}

// Chain sector: split the top FIB_K bits into a second pass from this n on.
#ifndef RYD_FIB_SPLIT_MIN_N
#define RYD_FIB_SPLIT_MIN_N 38
#endif
constexpr int FIB_SPLIT_MIN_N = RYD_FIB_SPLIT_MIN_N;
#ifndef RYD_FIB_SPLIT4
#define RYD_FIB_SPLIT4 0
#endif

int64_t fib(int k) {
  int64_t a = 0, b = 1;
  for (int i = 0; i < k; ++i) { const int64_t c = a + b; a = b; b = c; }
  return a;
}

FibModel fib_model(const ryd_plan* P) {
  FibModel M;
  M.n = P->n;
  M.L = P->fib_L;
  M.far_end = P->n;
  M.pad = 0;
  M.dim = P->dim;
  M.nh = P->fib_nh;
  M.hi = P->d_hi;
  M.lo_code = P->d_lo_code;
  M.lo_diag = P->d_lo_diag;
  M.vb = P->d_vb;
  for (int p = 0; p < MAXN; ++p) {
    M.W[p] = fib(p + 2);              // all positions: group sizes read W[IB] with IB = n
    M.ar[p] = P->ar[p];
    M.ai[p] = P->ai[p];
  }
  M.cnt_lo[0] = P->fib_cnt[0];
  M.cnt_lo[1] = P->fib_cnt[1];
  M.S = P->fib_split ? P->n - FIB_K : P->n;
  M.ng = FIB_NG;
  M.top = P->d_top;
  M.nbody[0] = fib(M.S + 2);
  M.nbody[1] = fib(M.S + 1);
  if (P->fib_split) M.far_end = M.S;
  return M;
}

ChanModel chan_model(const ryd_plan* P) {
  ChanModel M;
  M.n = P->n;
  M.mode = P->rank_mode;
  M.m = P->popcount;
  M.npairs = P->npairs;
  M.dim = P->dim;
  M.codes = P->d_codes;
  M.binom = P->d_binom;
  M.rsL = P->rs_L;
  M.rsplit = P->d_rsplit;
  M.pairs = P->d_pairs;
  M.vb = P->d_vb;
  M.gl = P->gr_L;
  M.hpk = P->d_hpk;
  M.lo_idx = P->d_lo_idx;
  M.fc = P->d_fc;
  M.pre = P->d_pre;
  M.tl = P->gr_tl;
  for (int p = 0; p < MAXN; ++p) {
    M.dl[p] = P->dl[p];
    M.ar[p] = P->ar[p];
    M.ai[p] = P->ai[p];
  }
  return M;
}

// Full-basis pair-channel plans run the tiled kernel when the diagonal tables
// of the element type's tile width exist (n >= that width).
#ifndef RYD_CHAN_TILED
#define RYD_CHAN_TILED 1
#endif
bool chan_full_tiled(const ryd_plan* P, int width) {
  return RYD_CHAN_TILED && P->basis == RYD_BASIS_FULL && P->rank_mode == RANK_FULL &&
         P->n >= width && P->d_elo[width - 10] && P->tbits[width - 10] == width;
}

ChanFullModel chan_full_model(const ryd_plan* P, int width) {
  ChanFullModel M;
  M.n = P->n;
  M.npairs = P->npairs;
  M.elo = P->d_elo[width - 10];
  M.ehi = P->d_ehi[width - 10];
  M.vb = P->d_vb;
  M.pairs = P->d_pairs;
  for (int p = 0; p < MAXN; ++p) {
    M.ar[p] = P->ar[p];
    M.ai[p] = P->ai[p];
  }
  return M;
}

GraphModel graph_model(const ryd_plan* P) {
  GraphModel M;
  M.n = P->n;
  M.L = P->gr_L;
  M.H = P->n - P->gr_L;
  M.tl = P->gr_tl;
  M.nv = P->gr_nv;
  M.hlist = P->d_hlist;
  M.hpk = P->d_hpk;
  M.rows = P->gr_rows;
  M.fc = P->d_fc;
  M.pre = P->d_pre;
  M.post = P->d_post;
  M.post_off = P->d_post_off;
  M.rcode = P->d_rcode;
  M.rdiag = P->d_rdiag;
  M.lrow = P->d_lrow;
  M.vb = P->d_vb;
  M.hfar = M.H - P->gr_K;
  for (int p = 0; p < MAXN; ++p) {
    M.dl[p] = P->dl[p];
    M.ar[p] = P->ar[p];
    M.ai[p] = P->ai[p];
  }
  return M;
}

GraphTop graph_top(const ryd_plan* P) {
  GraphTop T;
  std::memset(&T, 0, sizeof(T));
  T.L = P->gr_L;
  T.H = P->n - P->gr_L;
  T.K = P->gr_K;
  T.nt = P->gr_nt;
  T.nbody = P->gr_nbody;
  T.tl = P->gr_tl;
  T.hlist = P->d_hlist;
  T.hpk = P->d_hpk;
  T.tp = P->d_tp;
  T.tpk = P->d_tpk;
  for (int k = 0; k < T.K; ++k) {
    T.ar[k] = P->ar[T.L + T.H - T.K + k];
    T.ai[k] = P->ai[T.L + T.H - T.K + k];
  }
  return T;
}

template <class V>
unsigned graph_top_grid(const ryd_plan* P) {
  const int64_t units = P->gr_nbody * ((P->gr_tl + gr_run<V>() - 1) / gr_run<V>());
  if (units < 64) return (unsigned)(units < 1 ? 1 : units);
  const int64_t g = 8 * ((units + 7) / 8);
  return (unsigned)(g < CHAN_MAX_BLOCKS ? g : CHAN_MAX_BLOCKS);
}

// k_graph_apply: whole XCD groups of blocks, at most CHAN_MAX_BLOCKS
unsigned graph_grid(const ryd_plan* P) {
  const int64_t nv = P->gr_nv;
  if (nv < 64) return (unsigned)(nv < 1 ? 1 : nv);
  const int64_t g = 8 * ((nv + 7) / 8);
  return (unsigned)(g < CHAN_MAX_BLOCKS ? g : CHAN_MAX_BLOCKS);
}

bool graph_kernel(const ryd_plan* P) {
  return P->basis == RYD_BASIS_GRAPH && P->rank_mode == RANK_GRAPH && P->npairs == 0 && P->d_hlist;
}

unsigned chan_grid(const ryd_plan* P) {
  const int64_t g = (P->dim + 255) / 256;
  return (unsigned)(g < CHAN_MAX_BLOCKS ? (g < 1 ? 1 : g) : CHAN_MAX_BLOCKS);
}

unsigned fib_grid(const ryd_plan* P) {
  const int64_t nh = P->fib_nh;
  if (RYD_FIB_CHUNK > 0 && nh >= 8 * RYD_FIB_CHUNK) {
    // whole runs of RYD_FIB_CHUNK per XCD (fib_block_hi); the tail blocks idle
    const int64_t run = 8 * (int64_t)RYD_FIB_CHUNK;
    return (unsigned)(run * ((nh + run - 1) / run));
  }
  return (unsigned)(nh >= 64 ? 8 * ((nh + 7) / 8) : nh);
}

unsigned fib_top_grid(const ryd_plan* P, int run) {
  const int64_t nruns = (fib(P->n - FIB_K + 2) + run - 1) / run;
  return (unsigned)(nruns >= 64 ? 8 * ((nruns + 7) / 8) : nruns);
}

template <class V>
int launch_apply(const ryd_plan* P, int dtype, ApplyIO<V> io, Reduce rd, hipStream_t st) {
  if (P->basis == RYD_BASIS_CSR) {
    io.nfar = 0;
    io.partial = 0;
    CsrModel M{P->dim, P->d_csr_ptr, P->d_csr_col, P->d_csr_re, P->d_csr_im};
    const unsigned grid = chan_grid(P);
    k_csr_apply<V><<<dim3(grid), dim3(256), 0, st>>>(M, io, rd);
    LAUNCH_CHECK();
    if (io.reduce) k_reduce_finish<<<1, 256, 0, st>>>(rd, grid);
    LAUNCH_CHECK();
    return RYD_OK;
  }
  if (graph_kernel(P)) {
    io.nfar = 0;
    const bool split = P->gr_K > 0;
    io.partial = split ? 1 : 0;
    unsigned grid = graph_grid(P);
    if (P->gr_l16)
      k_graph_apply<V, uint16_t><<<dim3(grid), dim3(256), 0, st>>>(graph_model(P), io, rd);
    else
      k_graph_apply<V, uint32_t><<<dim3(grid), dim3(256), 0, st>>>(graph_model(P), io, rd);
    LAUNCH_CHECK();
    if (split) {
      io.partial = 0;
      grid = graph_top_grid<V>(P);
      k_graph_top<V><<<dim3(grid), dim3(256), 0, st>>>(graph_top(P), io, rd);
      LAUNCH_CHECK();
    }
    if (io.reduce) k_reduce_finish<<<1, 256, 0, st>>>(rd, grid);
    LAUNCH_CHECK();
    return RYD_OK;
  }
  if (P->generic) {
    io.nfar = 0;
    io.partial = 0;
    constexpr int TB = tile_bits<V>();
    if (chan_full_tiled(P, TB)) {
      const unsigned grid = (unsigned)(1ull << (P->n - TB));
      k_chan_full_tiled<V><<<dim3(grid), dim3(256), 0, st>>>(chan_full_model(P, TB), io, rd);
      LAUNCH_CHECK();
      if (io.reduce) k_reduce_finish<<<1, 256, 0, st>>>(rd, grid);
      LAUNCH_CHECK();
      return RYD_OK;
    }
    const unsigned grid = chan_grid(P);
    k_chan_apply<V><<<dim3(grid), dim3(256), 0, st>>>(chan_model(P), io, rd);
    LAUNCH_CHECK();
    if (io.reduce) k_reduce_finish<<<1, 256, 0, st>>>(rd, grid);
    LAUNCH_CHECK();
    return RYD_OK;
  }
  if (P->basis == RYD_BASIS_CHAIN_FIB) {
    io.nfar = 0;
    // the top-bit pass pays for 8- and 16-byte elements only (FP32 blocks
    // read their far partners fast enough in one pass)
    const bool split = P->fib_split && (sizeof(V) >= 8 || RYD_FIB_SPLIT4);
    FibModel M = fib_model(P);
    if (!split) M.far_end = P->n;
    io.partial = split ? 1 : 0;
    if (P->fib_L == FIB_L && RYD_FIB_FL)
      k_fib_apply_fl<V><<<dim3(fib_grid(P)), dim3(64 * fib_fl_waves<V>()), 0, st>>>(M, io, rd);
    else
      k_fib_apply<V><<<dim3(fib_grid(P)), dim3(FIB_THREADS), 0, st>>>(M, io, rd);
    LAUNCH_CHECK();
    unsigned grid = fib_grid(P);
    if (split) {
      io.partial = 0;
      grid = fib_top_grid(P, fib_run<V>());
      k_fib_top<V><<<dim3(grid), dim3(256), 0, st>>>(M, io, rd);
      LAUNCH_CHECK();
    }
    if (io.reduce) k_reduce_finish<<<1, 256, 0, st>>>(rd, grid);
    LAUNCH_CHECK();
    return RYD_OK;
  }
  constexpr int TB = tile_bits<V>();
  if (P->n >= TB) {
    const bool split = P->n >= split_min_n(sizeof(V), io.z != nullptr);
    unsigned grid;
    if (split) {
      // pass A: tiles of 2^PTB codes, far bits up to the top pass
      constexpr int PTT = passa_tt<V>();
      constexpr int PTB = tile_bits<V, PTT>();
      static_assert(PTB + top_bits<V>() <= split_min_n(sizeof(V), false),
                    "split: pass-A tile and top bits fit the smallest split n");
      FullModel M = full_model(P, PTB);
      io.nfar = P->n - top_bits<V>() - PTB;
      io.partial = 1;
      k_full_tiled<V, PTT><<<dim3((unsigned)(1ull << (P->n - PTB))), dim3(1 << PTT), 0, st>>>(M, io, rd);
      LAUNCH_CHECK();
    } else {
      FullModel M = full_model(P, TB);
      io.nfar = P->n - TB;
      io.partial = 0;
      grid = (unsigned)(1ull << (P->n - TB));
      k_full_tiled<V><<<dim3(grid), dim3(256), 0, st>>>(M, io, rd);
      LAUNCH_CHECK();
    }
    if (split) {
      FullModel M = full_model(P, TB);
      io.partial = 0;
      const uint64_t nb = 1ull << (P->n - top_bits<V>() - run_bits<V>());
      k_full_top<V><<<dim3((unsigned)nb), dim3(1 << top_ub<V>()), 0, st>>>(M, io, rd);
      LAUNCH_CHECK();
      grid = (unsigned)nb;
    }
    if (io.reduce) k_reduce_finish<<<1, 256, 0, st>>>(rd, grid);
    LAUNCH_CHECK();
  } else {
    FullModel M = full_model(P, TB);
    const uint64_t D = 1ull << P->n;
    const unsigned grid = (unsigned)((D + 255) / 256);
    io.nfar = 0;
    io.partial = 0;
    k_full_small<V><<<dim3(grid), dim3(256), 0, st>>>(M, io, rd);
    LAUNCH_CHECK();
    if (io.reduce) k_reduce_finish<<<1, 256, 0, st>>>(rd, grid);
    LAUNCH_CHECK();
  }
  return RYD_OK;
}

unsigned apply_grid(const ryd_plan* P, int dtype) {
  const int TB = tile_bits_of(dtype);
  if (P->n >= TB) return (unsigned)(1ull << (P->n - TB));
  return (unsigned)(((1ull << P->n) + 255) / 256);
}

unsigned stream_grid(int64_t D) {
  int64_t g = (D + 1023) / 1024;
  if (g < 1) g = 1;
  if (g > 8192) g = 8192;
  return (unsigned)g;
}

Reduce make_reduce(void* ws, int post, double* out) {
  Reduce rd;
  rd.partials = reinterpret_cast<double*>(reinterpret_cast<char*>(ws) + WS_HEADER);
  rd.post = post;
  rd.out = out;
  rd.alpha = nullptr;
  rd.nrm = nullptr;
  rd.st = nullptr;
  rd.j = 0;
  return rd;
}

template <class V>
int do_apply(const ryd_plan* P, int dtype, const void* x, const void* z, void* y, double s,
             double ch, double cx, double cz, const uint32_t* probe, double* sums, int accumulate,
             void* ws, hipStream_t st) {
  ApplyIO<V> io;
  io.x = static_cast<const V*>(x);
  io.z = static_cast<const V*>(z);
  io.y = static_cast<V*>(y);
  io.probe = probe;
  io.smode = SM_CONST;
  io.s = s; io.ch = ch; io.cx = cx; io.cz = cz;
  io.nrm = nullptr;
  io.j = 0;
  io.stop = nullptr;
  io.reduce = sums != nullptr;
  Reduce rd = make_reduce(ws, accumulate ? POST_ACCUM : POST_STORE, sums);
  return launch_apply<V>(P, dtype, io, rd, st);
}

__global__ void k_lz_init(LzState* s, double tiny) {
  s->stop = 0;
  s->steps = 0;
  s->tiny = tiny;
}

// One-launch Lanczos (k_lz_small) where launch_apply would run the standard
// full-basis kernels and the vector fits one workgroup's LDS.  Measured at
// n = 12 FP64 (C1, 200 steps and the replay): see lz_small.hpp.
#ifndef RYD_LZ_SMALL
#define RYD_LZ_SMALL 1
#endif
template <class V>
bool lz_small(const ryd_plan* P) {
  return RYD_LZ_SMALL && P->basis == RYD_BASIS_FULL && !P->generic && P->n >= 1 &&
         P->n <= lzs_maxn<V>();
}

template <class V>
int do_lanczos(const ryd_plan* P, int dtype, void* b0, void* b1, void* b2, int64_t j0,
               int64_t steps, double* alpha, double* nrm, void* state, double tiny, void* ritz0,
               void* ritz1, const double* y0, const double* y1, int replay, void* ws,
               hipStream_t st) {
  V* B[3] = {static_cast<V*>(b0), static_cast<V*>(b1), static_cast<V*>(b2)};
  LzState* S = static_cast<LzState*>(state);
  const int64_t D = P->dim;
  if (lz_small<V>(P)) {
    const FullModel M = full_model(P, tile_bits<V>());
    k_lz_small<V><<<1, LZS_THREADS, 0, st>>>(M, B[0], B[1], B[2], j0, steps, alpha, nrm, S, tiny,
                                             static_cast<V*>(ritz0), static_cast<V*>(ritz1), y0,
                                             y1, replay, nullptr, 0, 0);
    LAUNCH_CHECK();
    return RYD_OK;
  }
  const unsigned sg = stream_grid(D);
  if (j0 == 0 && !replay) {
    k_lz_init<<<1, 1, 0, st>>>(S, tiny);
    LAUNCH_CHECK();
    Reduce rd = make_reduce(ws, POST_LZ_NORM, nullptr);
    rd.nrm = nrm;
    rd.st = S;
    rd.j = -1;
    k_dots<V><<<sg, 256, 0, st>>>(nullptr, B[0], D, rd);
    LAUNCH_CHECK();
    k_reduce_finish<<<1, 256, 0, st>>>(rd, sg);
    LAUNCH_CHECK();
  }
  const int* stop = replay ? nullptr : &S->stop;
  for (int64_t j = j0; j < j0 + steps; ++j) {
    V* uj = B[j % 3];
    V* ujm = B[(j + 2) % 3];
    V* w = B[(j + 1) % 3];
    ApplyIO<V> io;
    io.x = uj;
    io.z = j > 0 ? ujm : nullptr;
    io.y = w;
    io.probe = nullptr;
    io.smode = SM_LANCZOS;
    io.s = io.ch = io.cx = io.cz = 0.0;
    io.nrm = nrm;
    io.j = (int)j;
    io.stop = stop;
    io.reduce = !replay;
    Reduce ra = make_reduce(ws, POST_LZ_ALPHA, nullptr);
    ra.alpha = alpha;
    ra.st = S;
    ra.j = (int)j;
    int rc = launch_apply<V>(P, dtype, io, ra, st);
    if (rc) return rc;
    Reduce rb = make_reduce(ws, POST_LZ_NORM, nullptr);
    rb.nrm = nrm;
    rb.st = S;
    rb.j = (int)j;
    k_lz_orth<V><<<sg, 256, 0, st>>>(uj, w, D, alpha, nrm, (int)j, static_cast<V*>(ritz0),
                                     static_cast<V*>(ritz1), y0, y1, stop, !replay, rb);
    LAUNCH_CHECK();
    if (!replay) k_reduce_finish<<<1, 256, 0, st>>>(rb, sg);
    LAUNCH_CHECK();
  }
  return RYD_OK;
}

// Recurrence on a stored basis chunk: u_j at basis + (j - jbase)·ld for
// j in [jbase, jbase + rows); steps j0 .. j0+steps-1 write u_{j+1} in place.
template <class V>
int do_lanczos_stored(const ryd_plan* P, int dtype, void* basis_v, int64_t ld, int64_t jbase,
                      int64_t j0, int64_t steps, double* alpha, double* nrm, void* state,
                      double tiny, void* ws, hipStream_t st) {
  V* basis = static_cast<V*>(basis_v);
  auto row = [&](int64_t j) { return basis + (j - jbase) * ld; };
  LzState* S = static_cast<LzState*>(state);
  const int64_t D = P->dim;
  const unsigned sg = stream_grid(D);
  if (lz_small<V>(P)) {
    const FullModel M = full_model(P, tile_bits<V>());
    k_lz_small<V><<<1, LZS_THREADS, 0, st>>>(M, nullptr, nullptr, nullptr, j0, steps, alpha, nrm,
                                             S, tiny, nullptr, nullptr, nullptr, nullptr, 0,
                                             basis, ld, jbase);
    LAUNCH_CHECK();
    return RYD_OK;
  }
  if (j0 == 0) {
    k_lz_init<<<1, 1, 0, st>>>(S, tiny);
    LAUNCH_CHECK();
    Reduce rd = make_reduce(ws, POST_LZ_NORM, nullptr);
    rd.nrm = nrm;
    rd.st = S;
    rd.j = -1;
    k_dots<V><<<sg, 256, 0, st>>>(nullptr, row(0), D, rd);
    LAUNCH_CHECK();
    k_reduce_finish<<<1, 256, 0, st>>>(rd, sg);
    LAUNCH_CHECK();
  }
  for (int64_t j = j0; j < j0 + steps; ++j) {
    ApplyIO<V> io;
    io.x = row(j);
    io.z = j > 0 ? row(j - 1) : nullptr;
    io.y = row(j + 1);
    io.probe = nullptr;
    io.smode = SM_LANCZOS;
    io.s = io.ch = io.cx = io.cz = 0.0;
    io.nrm = nrm;
    io.j = (int)j;
    io.stop = &S->stop;
    io.reduce = 1;
    Reduce ra = make_reduce(ws, POST_LZ_ALPHA, nullptr);
    ra.alpha = alpha;
    ra.st = S;
    ra.j = (int)j;
    if (int rc = launch_apply<V>(P, dtype, io, ra, st)) return rc;
    Reduce rb = make_reduce(ws, POST_LZ_NORM, nullptr);
    rb.nrm = nrm;
    rb.st = S;
    rb.j = (int)j;
    k_lz_orth<V><<<sg, 256, 0, st>>>(row(j), row(j + 1), D, alpha, nrm, (int)j, nullptr, nullptr,
                                     nullptr, nullptr, &S->stop, 1, rb);
    LAUNCH_CHECK();
    k_reduce_finish<<<1, 256, 0, st>>>(rb, sg);
    LAUNCH_CHECK();
  }
  return RYD_OK;
}

template <class V>
int do_cheb(const ryd_plan* P, int dtype, const uint32_t* probe, void* t0, void* t1, void* t2,
            int64_t nsteps, double c, double r, double* sums, void* ws, hipStream_t st) {
  V* B[3] = {static_cast<V*>(t0), static_cast<V*>(t1), static_cast<V*>(t2)};
  const int64_t D = P->dim;
  k_probe_expand<V><<<stream_grid(D), 256, 0, st>>>(probe, B[0], D);
  LAUNCH_CHECK();
  for (int64_t k = 1; k <= nsteps; ++k) {
    int rc;
    if (k == 1)
      rc = do_apply<V>(P, dtype, B[0], nullptr, B[1], 1.0, 1.0 / r, -c / r, 0.0, probe,
                       sums + 4 * k, 1, ws, st);
    else
      rc = do_apply<V>(P, dtype, B[(k - 1) % 3], B[(k - 2) % 3], B[k % 3], 1.0, 2.0 / r,
                       -2.0 * c / r, -1.0, probe, sums + 4 * k, 1, ws, st);
    if (rc) return rc;
  }
  return RYD_OK;
}

// Lanczos with two classical Gram-Schmidt passes against every stored row
// (reortho.hpp); basis rows are unit vectors, row 0 is phi·phi_scale on exit.
template <class V>
int do_lanczos_reortho(const ryd_plan* P, int dtype, void* basis_v, int64_t ld, int64_t nlock,
                       void* w_v, int64_t k, double phi_scale, double* alpha, double* nrm,
                       void* state, double tiny, int passes, double* gs_ws, void* ws,
                       hipStream_t st) {
  // rows [0, nlock): locked unit vectors (swept by every Gram-Schmidt pass,
  // never part of the recurrence); row nlock: the start vector
  V* basis = static_cast<V*>(basis_v);
  V* w = static_cast<V*>(w_v);
  LzState* S = static_cast<LzState*>(state);
  const int64_t D = P->dim;
  const unsigned sg = stream_grid(D);
  const int64_t nchunk = (D + GS_CHUNK - 1) / GS_CHUNK;
  double* partials = gs_ws;
  double* coef = gs_ws + (size_t)2 * (k + nlock) * nchunk;
  V* q0 = basis + nlock * ld;
  k_lz_init<<<1, 1, 0, st>>>(S, tiny);
  LAUNCH_CHECK();
  k_axpby<V><<<sg, 256, 0, st>>>(nullptr, q0, D, 0.0, phi_scale);
  LAUNCH_CHECK();
  if (nlock > 0) {
    // the start vector orthogonal to the locked rows, then unit again
    for (int pass = 0; pass < 2; ++pass) {
      k_gs_dots<V><<<(unsigned)nchunk, 256, 0, st>>>(basis, ld, D, (int)nlock, q0, partials,
                                                     &S->stop);
      LAUNCH_CHECK();
      k_gs_finish<<<(unsigned)nlock, 256, 0, st>>>(partials, nchunk, coef, &S->stop);
      LAUNCH_CHECK();
      for (int r0 = 0; r0 < nlock; r0 += GS_MAX_ROWS) {
        const int cnt = (int)std::min<int64_t>(GS_MAX_ROWS, nlock - r0);
        k_gs_axpy<V><<<sg, 256, 0, st>>>(basis + (int64_t)r0 * ld, ld, D, cnt, coef + 2 * r0, q0,
                                         &S->stop);
        LAUNCH_CHECK();
      }
    }
    Reduce rn = make_reduce(ws, POST_STORE, coef);
    k_dots<V><<<sg, 256, 0, st>>>(nullptr, q0, D, rn);
    LAUNCH_CHECK();
    k_reduce_finish<<<1, 256, 0, st>>>(rn, sg);
    LAUNCH_CHECK();
    k_scale_rsqrt<V><<<sg, 256, 0, st>>>(q0, D, coef);
    LAUNCH_CHECK();
  }
  for (int64_t j = 0; j < k; ++j) {
    V* qj = q0 + j * ld;
    ApplyIO<V> io;
    io.x = qj;
    io.z = nullptr;
    io.y = w;
    io.probe = nullptr;
    io.smode = SM_CONST;
    io.s = 1.0; io.ch = 1.0; io.cx = 0.0; io.cz = 0.0;
    io.nrm = nullptr;
    io.j = 0;
    io.stop = &S->stop;
    io.reduce = 1;
    Reduce ra = make_reduce(ws, POST_LZ_ALPHA, nullptr);
    ra.alpha = alpha;
    ra.st = S;
    ra.j = (int)j;
    if (int rc = launch_apply<V>(P, dtype, io, ra, st)) return rc;
    k_rl_three_term<V><<<sg, 256, 0, st>>>(w, qj, j > 0 ? qj - ld : qj, D, alpha, nrm, (int)j,
                                           &S->stop);
    LAUNCH_CHECK();
    const int rows = (int)(nlock + j + 1);
    for (int pass = 0; pass < passes; ++pass) {
      k_gs_dots<V><<<(unsigned)nchunk, 256, 0, st>>>(basis, ld, D, rows, w, partials, &S->stop);
      LAUNCH_CHECK();
      k_gs_finish<<<(unsigned)rows, 256, 0, st>>>(partials, nchunk, coef, &S->stop);
      LAUNCH_CHECK();
      for (int r0 = 0; r0 < rows; r0 += GS_MAX_ROWS) {
        const int cnt = std::min(GS_MAX_ROWS, rows - r0);
        k_gs_axpy<V><<<sg, 256, 0, st>>>(basis + (int64_t)r0 * ld, ld, D, cnt, coef + 2 * r0, w,
                                         &S->stop);
        LAUNCH_CHECK();
      }
    }
    Reduce rb = make_reduce(ws, POST_LZ_NORM, nullptr);
    rb.nrm = nrm;
    rb.st = S;
    rb.j = (int)j;
    k_dots<V><<<sg, 256, 0, st>>>(nullptr, w, D, rb);
    LAUNCH_CHECK();
    k_reduce_finish<<<1, 256, 0, st>>>(rb, sg);
    LAUNCH_CHECK();
    if (j + 1 < k) {
      k_rl_scale<V><<<sg, 256, 0, st>>>(qj + ld, w, D, nrm, (int)j, &S->stop);
      LAUNCH_CHECK();
    }
  }
  return RYD_OK;
}

int check_dtype(const ryd_plan* P, int dtype) {
  if (dtype < RYD_F32 || dtype > RYD_C128) return fail(RYD_ERR_INVALID, "unknown dtype");
  if (P->cplx && (dtype == RYD_F32 || dtype == RYD_F64))
    return fail(RYD_ERR_INVALID, "complex Rabi phases need a complex dtype (complex64/complex128)");
  return RYD_OK;
}

}  // namespace

namespace {

// Phases 1-2 of the graph scans: block sums, their prefix, and the total.
template <class F>
int gr_total(F f, int64_t count, int64_t* bsum, int64_t& total) {
  const int64_t nb = (count + GR_TILE - 1) / GR_TILE;
  k_gr_scan_sum<F><<<dim3((unsigned)nb), dim3(256)>>>(f, count, bsum);
  LAUNCH_CHECK();
  k_gr_scan_blocks<<<1, 256>>>(bsum, nb);
  LAUNCH_CHECK();
  HIP_TRY(hipMemcpy(&total, bsum + nb, sizeof(int64_t), hipMemcpyDeviceToHost));
  return RYD_OK;
}

template <class F, class W>
int gr_write(F f, W w, int64_t count, const int64_t* bsum) {
  const int64_t nb = (count + GR_TILE - 1) / GR_TILE;
  k_gr_scan_write<F, W><<<dim3((unsigned)nb), dim3(256)>>>(f, w, count, bsum);
  LAUNCH_CHECK();
  return RYD_OK;
}

// Blockade-graph bases from this dimension on split H·v into two passes (the
// top high bits in the second); RYD_GRAPH_SPLIT_MIN_DIM overrides it.
int64_t gr_split_min_dim() {
  if (const char* env = std::getenv("RYD_GRAPH_SPLIT_MIN_DIM")) {
    char* end = nullptr;
    const long long v = std::strtoll(env, &end, 10);
    if (end != env) return (int64_t)v;
  }
  return 1ll << 22;
}

// 16-bit row tables for k_graph_apply when every segment is shorter than
// 65535 rows; RYD_GRAPH_WIDE_ROWS=1 keeps 32-bit entries.
bool gr_row16(int64_t tl) {
  if (const char* env = std::getenv("RYD_GRAPH_WIDE_ROWS"))
    if (env[0] == '1') return false;
  return tl < 65535;
}

// Bytes allowed for the per-class prefix table: beyond this the ranks fall
// back to a binary search over the code list.  Default: twice the code list,
// at least 256 MiB; RYD_GRAPH_PREFIX_MAX_BYTES overrides it (0 = always search).
size_t gr_prefix_budget(int64_t dim) {
  if (const char* env = std::getenv("RYD_GRAPH_PREFIX_MAX_BYTES")) {
    char* end = nullptr;
    const unsigned long long v = std::strtoull(env, &end, 10);
    if (end != env) return (size_t)v;
  }
  const size_t codes = (size_t)dim * sizeof(int64_t);
  return codes * 2 > ((size_t)256 << 20) ? codes * 2 : ((size_t)256 << 20);
}

// Independent-set basis of the blockade graph in m (graph.hpp): device tables,
// the dimension and the ascending code list.
int build_graph(ryd_plan* P, const ryd_model* m) {
  const int n = P->n;
  std::vector<uint64_t> nbb(n, 0);          // neighbours by bit position
  for (int k = 0; k < m->n_edges; ++k) {
    const int p = n - 1 - m->edge_i[k], q = n - 1 - m->edge_j[k];
    nbb[p] |= 1ull << q;
    nbb[q] |= 1ull << p;
  }
  auto boundary = [&](int L) {
    const uint64_t low = (1ull << L) - 1ull;
    int c = 0;
    for (int p = L; p < n; ++p) c += (nbb[p] & low) ? 1 : 0;
    return c;
  };
  // low part: all bits up to 16 sites; else the cut near n/2 with the fewest
  // boundary sites (fewest classes), both parts at most GR_MAXB bits
  int L = n;
  if (n > 16) {
    int best = -1, bscore = 1 << 30;
    for (int c = n / 2 - 4; c <= n / 2 + 4; ++c) {
      if (c < 1 || c >= n || c > GR_MAXB || n - c > GR_MAXB) continue;
      const int score = 4 * boundary(c) + (c > n / 2 ? c - n / 2 : n / 2 - c);
      if (score < bscore) { bscore = score; best = c; }
    }
    if (best < 0) return fail(RYD_ERR_INVALID, "blockade-graph basis needs n <= 60");
    L = best;
  }
  const int H = n - L;
  const uint64_t low = L == 64 ? ~0ull : (1ull << L) - 1ull;
  GraphParts G;
  std::memset(&G, 0, sizeof(G));
  G.L = L;
  G.H = H;
  for (int q = 0; q < L; ++q) G.nb_lo[q] = (uint32_t)(nbb[q] & low);
  for (int p = 0; p < H; ++p) G.nb_hi[p] = (uint32_t)(nbb[p + L] >> L);
  int k = 0;
  for (int p = 0; p < H; ++p)
    if (nbb[p + L] & low) {
      G.hb_pos[k] = (uint32_t)p;
      G.hb_low[k] = (uint32_t)(nbb[p + L] & low);
      ++k;
    }
  G.k = k;
  for (int j = 0; j < k; ++j)
    for (int j2 = 0; j2 < k; ++j2)
      if ((G.nb_hi[G.hb_pos[j]] >> G.hb_pos[j2]) & 1u) G.nb_cls[j] |= 1u << j2;
  P->gr_L = L;

  const int64_t nlo = 1ll << L, ncls = 1ll << k, nhi = 1ll << H;
  int64_t most = nlo > ncls ? nlo : ncls;
  if (nhi > most) most = nhi;
  int64_t* bsum = nullptr;
  uint32_t* cls_idx = nullptr;
  uint32_t* clslist = nullptr;
  int64_t* ccnt = nullptr;
  double* dl_dev = nullptr;
  auto release = [&]() {
    if (dl_dev) (void)hipFree(dl_dev);
    if (bsum) (void)hipFree(bsum);
    if (cls_idx) (void)hipFree(cls_idx);
    if (clslist) (void)hipFree(clslist);
    if (ccnt) (void)hipFree(ccnt);
  };
  auto run = [&]() -> int {
    HIP_TRY(hipMalloc(&bsum, sizeof(int64_t) * (size_t)((most + GR_TILE - 1) / GR_TILE + 1)));
    // low sets
    int64_t tl = 0;
    if (int rc = gr_total(GrLowFlag{G}, nlo, bsum, tl)) return rc;
    HIP_TRY(hipMalloc(&P->d_lo_idx, sizeof(uint32_t) * (size_t)nlo));
    HIP_TRY(hipMalloc(&P->d_lowtab, sizeof(uint32_t) * (size_t)tl));
    if (int rc = gr_write(GrLowFlag{G}, GrIndexWrite{P->d_lo_idx, P->d_lowtab}, nlo, bsum)) return rc;
    // boundary classes and the low bits each blocks
    int64_t nc = 0;
    if (int rc = gr_total(GrClassFlag{G}, ncls, bsum, nc)) return rc;
    HIP_TRY(hipMalloc(&cls_idx, sizeof(uint32_t) * (size_t)ncls));
    HIP_TRY(hipMalloc(&clslist, sizeof(uint32_t) * (size_t)nc));
    if (int rc = gr_write(GrClassFlag{G}, GrIndexWrite{cls_idx, clslist}, ncls, bsum)) return rc;
    if (nc >= (int64_t)GR_CLS)
      return fail(RYD_ERR_UNSUPPORTED, "blockade graph: 2^24 or more boundary classes");
    if ((double)nc * (double)tl > (double)(1ll << 38))
      return fail(RYD_ERR_UNSUPPORTED, "blockade graph: boundary classes x low sets exceeds 2^38; "
                                       "order the sites so that the cut crosses fewer edges");
    HIP_TRY(hipMalloc(&P->d_fc, sizeof(uint32_t) * (size_t)nc));
    const unsigned cg = (unsigned)(nc < 65536 ? nc : 65536);
    k_gr_class_mask<<<dim3((unsigned)((nc + 255) / 256)), dim3(256)>>>(G, clslist, nc, P->d_fc);
    LAUNCH_CHECK();
    HIP_TRY(hipMalloc(&ccnt, sizeof(int64_t) * (size_t)nc));
    k_gr_class_count<<<dim3(cg), dim3(256)>>>(P->d_lowtab, tl, P->d_fc, nc, ccnt);
    LAUNCH_CHECK();
    // high parts: first ranks and classes
    int64_t dim = 0;
    if (int rc = gr_total(GrHighCount{G, cls_idx, ccnt}, nhi, bsum, dim)) return rc;
    if (dim >= (1ll << 40)) return fail(RYD_ERR_MEMORY, "blockade-graph basis has 2^40 codes or more");
    HIP_TRY(hipMalloc(&P->d_hpk, sizeof(uint64_t) * (size_t)(nhi + 1)));
    if (int rc = gr_write(GrHighCount{G, cls_idx, ccnt}, GrHighWrite{G, cls_idx, P->d_hpk}, nhi, bsum))
      return rc;
    k_gr_seal<<<1, 1>>>(P->d_hpk, nhi, bsum + (nhi + GR_TILE - 1) / GR_TILE);
    LAUNCH_CHECK();
    P->dim = dim;
    P->gr_tl = tl;
    P->gr_nc = nc;
    // constant-time ranks (and the segment kernel) when the prefix and row
    // tables are affordable
    std::vector<int64_t> cc((size_t)nc), post_off((size_t)nc);
    HIP_TRY(hipMemcpy(cc.data(), ccnt, sizeof(int64_t) * (size_t)nc, hipMemcpyDeviceToHost));
    int64_t rows_all = 0;
    for (int64_t c = 0; c < nc; ++c) {
      post_off[(size_t)c] = rows_all;
      rows_all += cc[(size_t)c];
    }
    const size_t pre_bytes = sizeof(uint32_t) * (size_t)nc * (size_t)(tl + 1);
    const size_t post_bytes = sizeof(uint32_t) * (size_t)rows_all;
    const size_t rows_bytes = (2 * sizeof(uint32_t) + sizeof(double) +
                               sizeof(uint32_t) * (size_t)L) * (size_t)rows_all;
    if (pre_bytes <= gr_prefix_budget(dim)) {
      HIP_TRY(hipMalloc(&P->d_pre, pre_bytes));
      k_gr_prefix<<<dim3(cg), dim3(256)>>>(P->d_lowtab, tl, P->d_fc, nc, P->d_pre);
      LAUNCH_CHECK();
      P->rank_mode = RANK_GRAPH;
    } else {
      P->rank_mode = RANK_SEARCH;
    }
    if (P->rank_mode == RANK_GRAPH && P->npairs == 0 &&
        pre_bytes + rows_bytes <= gr_prefix_budget(dim)) {
      // tables of the segment kernel (Omega/delta/V models)
      HIP_TRY(hipMalloc(&P->d_post_off, sizeof(int64_t) * (size_t)nc));
      HIP_TRY(hipMemcpy(P->d_post_off, post_off.data(), sizeof(int64_t) * (size_t)nc,
                        hipMemcpyHostToDevice));
      HIP_TRY(hipMalloc(&P->d_post, post_bytes > 0 ? post_bytes : sizeof(uint32_t)));
      k_gr_post<<<dim3(cg), dim3(256)>>>(P->d_lowtab, tl, P->d_fc, nc, P->d_post_off, P->d_post);
      LAUNCH_CHECK();
      HIP_TRY(hipMalloc(&dl_dev, sizeof(double) * (size_t)n));
      HIP_TRY(hipMemcpy(dl_dev, P->dl, sizeof(double) * (size_t)n, hipMemcpyHostToDevice));
      const size_t rows1 = (size_t)(rows_all > 0 ? rows_all : 1);
      HIP_TRY(hipMalloc(&P->d_rcode, sizeof(uint32_t) * rows1));
      HIP_TRY(hipMalloc(&P->d_rdiag, sizeof(double) * rows1));
      // segment positions are < tl: 16-bit row tables halve their traffic
      P->gr_l16 = gr_row16(tl);
      const size_t lsz = P->gr_l16 ? sizeof(uint16_t) : sizeof(uint32_t);
      HIP_TRY(hipMalloc(&P->d_lrow, lsz * rows1 * (size_t)(L > 0 ? L : 1)));
      if (P->gr_l16)
        k_gr_rows<uint16_t><<<dim3(cg), dim3(256)>>>(n, L, P->d_lowtab, tl, P->d_lo_idx, P->d_fc, nc,
                                                     P->d_pre, P->d_post, P->d_post_off, rows_all,
                                                     dl_dev, P->d_vb, P->d_rcode, P->d_rdiag,
                                                     static_cast<uint16_t*>(P->d_lrow));
      else
        k_gr_rows<uint32_t><<<dim3(cg), dim3(256)>>>(n, L, P->d_lowtab, tl, P->d_lo_idx, P->d_fc, nc,
                                                     P->d_pre, P->d_post, P->d_post_off, rows_all,
                                                     dl_dev, P->d_vb, P->d_rcode, P->d_rdiag,
                                                     static_cast<uint32_t*>(P->d_lrow));
      LAUNCH_CHECK();
      P->gr_rows = rows_all;
      int64_t nv = 0;
      if (int rc = gr_total(GrValidFlag{P->d_hpk}, nhi, bsum, nv)) return rc;
      HIP_TRY(hipMalloc(&P->d_hlist, sizeof(uint32_t) * (size_t)nv));
      if (int rc = gr_write(GrValidFlag{P->d_hpk}, GrListWrite{P->d_hlist}, nhi, bsum)) return rc;
      P->gr_nv = nv;
      // top pass: the top K high bits when none of them has a low neighbour
      // and their independent patterns fit a block (large bases only)
      std::vector<uint32_t> tp;
      int K = 0;
      if (dim >= gr_split_min_dim()) {
        for (int k = 3; k <= GR_KMAX && k < H; ++k) {
          bool clear = true;
          for (int q = H - k; q < H; ++q) clear = clear && !(nbb[q + L] & low);
          if (!clear) break;
          std::vector<uint32_t> cand;
          for (uint32_t g = 0; g < (1u << k); ++g) {
            bool ind = true;
            for (int a = 0; a < k && ind; ++a)
              if (((g >> a) & 1u) && (g & (G.nb_hi[H - k + a] >> (H - k)))) ind = false;
            if (ind) cand.push_back(g);
          }
          if ((int)cand.size() > GR_TOPMAX) break;
          K = k;
          tp.swap(cand);
        }
      }
      if (K >= 3 && H - K <= GR_HNEAR) {
        std::vector<int16_t> tpk(tp.size() * (size_t)K, -1);
        for (size_t t = 0; t < tp.size(); ++t)
          for (int k = 0; k < K; ++k) {
            const uint32_t g2 = tp[t] ^ (1u << k);
            auto it = std::lower_bound(tp.begin(), tp.end(), g2);
            if (it != tp.end() && *it == g2) tpk[t * K + k] = (int16_t)(it - tp.begin());
          }
        std::vector<uint32_t> hl((size_t)nv);
        HIP_TRY(hipMemcpy(hl.data(), P->d_hlist, sizeof(uint32_t) * (size_t)nv, hipMemcpyDeviceToHost));
        P->gr_nbody = std::lower_bound(hl.begin(), hl.end(), 1u << (H - K)) - hl.begin();
        HIP_TRY(hipMalloc(&P->d_tp, sizeof(uint32_t) * tp.size()));
        HIP_TRY(hipMemcpy(P->d_tp, tp.data(), sizeof(uint32_t) * tp.size(), hipMemcpyHostToDevice));
        HIP_TRY(hipMalloc(&P->d_tpk, sizeof(int16_t) * tpk.size()));
        HIP_TRY(hipMemcpy(P->d_tpk, tpk.data(), sizeof(int16_t) * tpk.size(), hipMemcpyHostToDevice));
        P->gr_K = K;
        P->gr_nt = (int)tp.size();
      }
    }
    HIP_TRY(hipMalloc(&P->d_codes, sizeof(int64_t) * (size_t)dim));
    const unsigned eg = (unsigned)(nhi < 65536 ? nhi : 65536);
    k_gr_emit<<<dim3(eg), dim3(256)>>>(L, P->d_hpk, nhi, P->d_lowtab, tl, P->d_fc, P->d_codes);
    LAUNCH_CHECK();
    HIP_TRY(hipStreamSynchronize(nullptr));
    return RYD_OK;
  };
  const int rc = run();
  release();
  return rc;
}

}  // namespace

// Three-term recurrence in the complement of a unit vector g (the lock):
// steps j0 .. j0+steps-1 as do_lanczos, except that the orth step leaves
// u_pre = w − (alpha_j / nrm_j) u_j with the sums of <g, u_pre>, a finish
// stores γ_j = <g, u_pre> to gam[2j..2j+1] and nrm_{j+1} = |u_pre − γ_j g|,
// and k_lz_unlock removes γ_j g.  The start vector must be orthogonal to g.
// The replay reads the stored alpha, nrm and gam.
template <class V>
int do_lanczos_locked(const ryd_plan* P, int dtype, void* b0, void* b1, void* b2,
                      const void* lock_v, int64_t j0, int64_t steps, double* alpha, double* nrm,
                      double* gam, void* state, double tiny, void* ritz0, void* ritz1,
                      const double* y0, const double* y1, int replay, void* ws, hipStream_t st) {
  V* B[3] = {static_cast<V*>(b0), static_cast<V*>(b1), static_cast<V*>(b2)};
  const V* g = static_cast<const V*>(lock_v);
  LzState* S = static_cast<LzState*>(state);
  const int64_t D = P->dim;
  const unsigned sg = stream_grid(D);
  if (j0 == 0 && !replay) {
    k_lz_init<<<1, 1, 0, st>>>(S, tiny);
    LAUNCH_CHECK();
    Reduce rd = make_reduce(ws, POST_LZ_NORM, nullptr);
    rd.nrm = nrm;
    rd.st = S;
    rd.j = -1;
    k_dots<V><<<sg, 256, 0, st>>>(nullptr, B[0], D, rd);
    LAUNCH_CHECK();
    k_reduce_finish<<<1, 256, 0, st>>>(rd, sg);
    LAUNCH_CHECK();
  }
  const int* stop = replay ? nullptr : &S->stop;
  for (int64_t j = j0; j < j0 + steps; ++j) {
    V* uj = B[j % 3];
    V* ujm = B[(j + 2) % 3];
    V* w = B[(j + 1) % 3];
    ApplyIO<V> io;
    io.x = uj;
    io.z = j > 0 ? ujm : nullptr;
    io.y = w;
    io.probe = nullptr;
    io.smode = SM_LANCZOS;
    io.s = io.ch = io.cx = io.cz = 0.0;
    io.nrm = nrm;
    io.j = (int)j;
    io.stop = stop;
    io.reduce = !replay;
    Reduce ra = make_reduce(ws, POST_LZ_ALPHA, nullptr);
    ra.alpha = alpha;
    ra.st = S;
    ra.j = (int)j;
    if (int rc = launch_apply<V>(P, dtype, io, ra, st)) return rc;
    Reduce rb = make_reduce(ws, POST_LZ_LOCK, gam + 2 * j);
    rb.nrm = nrm;
    rb.st = S;
    rb.j = (int)j;
    k_lz_orth_lock<V><<<sg, 256, 0, st>>>(uj, w, g, D, alpha, nrm, (int)j, static_cast<V*>(ritz0),
                                          static_cast<V*>(ritz1), y0, y1, stop, !replay, rb);
    LAUNCH_CHECK();
    if (!replay) k_reduce_finish<<<1, 256, 0, st>>>(rb, sg);
    LAUNCH_CHECK();
    k_lz_unlock<V><<<sg, 256, 0, st>>>(w, g, D, gam + 2 * j, stop);
    LAUNCH_CHECK();
  }
  return RYD_OK;
}

#define DISPATCH(dtype, FN, ...)                                 \
  switch (dtype) {                                               \
    case RYD_F32: return FN<float>(__VA_ARGS__);                 \
    case RYD_F64: return FN<double>(__VA_ARGS__);                \
    case RYD_C64: return FN<c32>(__VA_ARGS__);                   \
    case RYD_C128: return FN<c64>(__VA_ARGS__);                  \
    default: return fail(RYD_ERR_INVALID, "unknown dtype");     \
  }

// k_ritz_stored_sliced below RITZ_SLICED_MAX_DIM, else one thread per element
template <class V>
void launch_ritz_stored(const void* basis, int64_t ld, int64_t dim, int count, const double* yd,
                        int two, const double* nrm, int64_t j0, int accumulate, void* out0,
                        void* out1, hipStream_t st) {
  const V* b = static_cast<const V*>(basis);
  V* o0 = static_cast<V*>(out0);
  V* o1 = static_cast<V*>(out1);
  if (dim <= RITZ_SLICED_MAX_DIM) {
    const unsigned g = (unsigned)((dim + RITZ_E - 1) / RITZ_E);
    k_ritz_stored_sliced<V><<<g, 256, 0, st>>>(b, ld, dim, count, yd, two, nrm, j0, accumulate,
                                               o0, o1);
  } else {
    k_ritz_stored<V><<<stream_grid(dim), 256, 0, st>>>(b, ld, dim, count, yd, two, nrm, j0,
                                                       accumulate, o0, o1);
  }
}

extern "C" {

int ryd_abi_version(void) { return RYD_ABI_VERSION; }

int ryd_last_error(char* buf, size_t size) {
  if (!buf || !size) return RYD_ERR_INVALID;
  std::snprintf(buf, size, "%s", g_err.c_str());
  return RYD_OK;
}

int ryd_plan_create(const ryd_model* m, ryd_plan** out) {
  if (!m || !out) return fail(RYD_ERR_INVALID, "null model or output pointer");
  *out = nullptr;
  const int n = m->n;
  if (n < 1 || n > 62) return fail(RYD_ERR_INVALID, "n must be in [1, 62]");
  if (m->basis_kind == RYD_BASIS_CHAIN_FIB && n > 50)
    return fail(RYD_ERR_INVALID, "chain blockade sector needs n <= 50");
  if (!m->om_re || !m->delta) return fail(RYD_ERR_INVALID, "om_re and delta are required");
  if (m->n_pairs < 0 || (m->n_pairs > 0 && (!m->pair_i || !m->pair_j || !m->pair_v)))
    return fail(RYD_ERR_INVALID, "pair arrays are required when n_pairs > 0");
  if (m->basis_kind != RYD_BASIS_FULL && m->basis_kind != RYD_BASIS_CHAIN_FIB &&
      m->basis_kind != RYD_BASIS_CODES && m->basis_kind != RYD_BASIS_GRAPH)
    return fail(RYD_ERR_INVALID, "unknown basis kind");
  if (m->basis_kind == RYD_BASIS_GRAPH) {
    if (n > 2 * GR_MAXB) return fail(RYD_ERR_INVALID, "blockade-graph basis needs n <= 60");
    if (m->n_edges < 0 || (m->n_edges > 0 && (!m->edge_i || !m->edge_j)))
      return fail(RYD_ERR_INVALID, "edge arrays are required when n_edges > 0");
    for (int k = 0; k < m->n_edges; ++k) {
      const int i = m->edge_i[k], j = m->edge_j[k];
      if (i < 0 || j < 0 || i >= n || j >= n || i == j)
        return fail(RYD_ERR_INVALID, "blockade edge out of range or a self loop");
    }
  }
  bool pairflip = false;
  for (int k = 0; k < m->n_pairs; ++k) {
    const double jr = m->pair_jr ? m->pair_jr[k] : 0.0, ji = m->pair_ji ? m->pair_ji[k] : 0.0;
    const double ar = m->pair_ar ? m->pair_ar[k] : 0.0, ai = m->pair_ai ? m->pair_ai[k] : 0.0;
    if (!std::isfinite(jr) || !std::isfinite(ji) || !std::isfinite(ar) || !std::isfinite(ai))
      return fail(RYD_ERR_INVALID, "J/A must be finite");
    if (jr != 0.0 || ji != 0.0 || ar != 0.0 || ai != 0.0) pairflip = true;
  }
  if (pairflip && m->basis_kind == RYD_BASIS_CHAIN_FIB)
    return fail(RYD_ERR_UNSUPPORTED, "the chain sector kernels take Omega/delta/V only; give the "
                                     "sector as codes for J/A models");
  if (m->basis_kind == RYD_BASIS_CODES) {
    if (!m->codes || m->n_codes < 1) return fail(RYD_ERR_INVALID, "codes basis needs codes");
    for (int64_t k = 0; k < m->n_codes; ++k) {
      if (m->codes[k] < 0 || (n < 63 && (m->codes[k] >> n)))
        return fail(RYD_ERR_INVALID, "codes must lie in [0, 2^n)");
      if (k && m->codes[k] <= m->codes[k - 1])
        return fail(RYD_ERR_INVALID, "codes must be ascending and unique");
    }
  }
  if (m->basis_kind == RYD_BASIS_FULL && n > 40)
    return fail(RYD_ERR_INVALID, "full basis needs n <= 40");
  for (int i = 0; i < n; ++i) {
    if (!std::isfinite(m->om_re[i]) || !std::isfinite(m->delta[i]) ||
        (m->om_im && !std::isfinite(m->om_im[i])))
      return fail(RYD_ERR_INVALID, "Om/delta must be finite");
  }
  ryd_plan* P = new ryd_plan();
  P->n = n;
  P->basis = m->basis_kind;
  P->dim = m->basis_kind == RYD_BASIS_CHAIN_FIB ? fib(n + 2)
           : m->basis_kind == RYD_BASIS_CODES   ? m->n_codes
           : m->basis_kind == RYD_BASIS_GRAPH   ? 0
                                                : (1ll << n);
  P->generic = pairflip || m->basis_kind == RYD_BASIS_CODES || m->basis_kind == RYD_BASIS_GRAPH;
  (void)hipGetDevice(&P->device);
  std::vector<double> dl(n), vb((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) {
    const int p = n - 1 - i;
    P->ar[p] = 0.5 * m->om_re[i];
    P->ai[p] = m->om_im ? 0.5 * m->om_im[i] : 0.0;
    if (P->ai[p] != 0.0) P->cplx = true;
    dl[p] = m->delta[i];
    P->dl[p] = m->delta[i];
  }
  for (int k = 0; k < m->n_pairs; ++k) {
    const int i = m->pair_i[k], j = m->pair_j[k];
    if (i < 0 || j < 0 || i >= n || j >= n || i == j || !std::isfinite(m->pair_v[k])) {
      delete P;
      return fail(RYD_ERR_INVALID, "pair indices out of range, self pair or non-finite V");
    }
    const int p = n - 1 - i, q = n - 1 - j;
    vb[(size_t)p * n + q] += m->pair_v[k];
    vb[(size_t)q * n + p] += m->pair_v[k];
  }
  auto cleanup = [&](int code, const std::string& msg) {
    ryd_plan_destroy(P);
    return fail(code, msg);
  };
  hipError_t e = hipMalloc(&P->d_vb, sizeof(double) * (size_t)n * n);
  if (e != hipSuccess) return cleanup(RYD_ERR_DEVICE, std::string("hipMalloc: ") + hipGetErrorString(e));
  e = hipMemcpy(P->d_vb, vb.data(), sizeof(double) * (size_t)n * n, hipMemcpyHostToDevice);
  if (e != hipSuccess) return cleanup(RYD_ERR_DEVICE, std::string("hipMemcpy: ") + hipGetErrorString(e));
  if (P->generic) {
    // pair flip channels by bit position, the sorted codes and the rank rule
    std::vector<PairChan> pc;
    for (int k = 0; k < m->n_pairs; ++k) {
      PairChan c;
      c.pi = n - 1 - m->pair_i[k];
      c.pj = n - 1 - m->pair_j[k];
      c.jr = m->pair_jr ? m->pair_jr[k] : 0.0;
      c.ji = m->pair_ji ? m->pair_ji[k] : 0.0;
      c.ar = m->pair_ar ? m->pair_ar[k] : 0.0;
      c.ai = m->pair_ai ? m->pair_ai[k] : 0.0;
      if (c.jr != 0.0 || c.ji != 0.0 || c.ar != 0.0 || c.ai != 0.0) pc.push_back(c);
      if (c.ji != 0.0 || c.ai != 0.0) P->cplx = true;     // complex H: complex vectors only
    }
    P->npairs = (int)pc.size();
    std::vector<int64_t> binom(64 * BINOM_LD, 0);
    for (int a = 0; a < 64; ++a) {
      binom[a * BINOM_LD] = 1;
      for (int b = 1; b <= a; ++b)
        binom[a * BINOM_LD + b] = binom[(a - 1) * BINOM_LD + b - 1] +
                                  (b <= a - 1 ? binom[(a - 1) * BINOM_LD + b] : 0);
    }
    P->rank_mode = RANK_FULL;
    if (m->basis_kind == RYD_BASIS_CODES) {
      // a complete fixed-popcount sector ranks in closed form
      const int pc0 = __builtin_popcountll((uint64_t)m->codes[0]);
      bool same = true;
      for (int64_t k = 1; k < m->n_codes && same; ++k)
        same = __builtin_popcountll((uint64_t)m->codes[k]) == pc0;
      P->popcount = pc0;
      P->rank_mode = (same && m->n_codes == binom[n * BINOM_LD + pc0]) ? RANK_NUMBER : RANK_SEARCH;
    }
    e = hipMalloc(&P->d_binom, sizeof(int64_t) * binom.size());
    if (e == hipSuccess)
      e = hipMemcpy(P->d_binom, binom.data(), sizeof(int64_t) * binom.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess && P->rank_mode == RANK_NUMBER && n <= RS_MAXN) {
      // r = Σ_i C(p_i, i) over the set bits p_1 < p_2 < ...: the low rs_L bits
      // give the first terms, the high bits the rest with indices shifted by
      // the low popcount
      const int L = n / 2, Hn = n - L;
      std::vector<int64_t> rs(((size_t)1 << L) + ((size_t)1 << Hn) * (size_t)(L + 1), 0);
      for (uint64_t l = 0; l < (1ull << L); ++l) {
        int64_t r = 0;
        int i = 1;
        for (uint64_t c = l; c; c &= c - 1, ++i) r += binom[__builtin_ctzll(c) * BINOM_LD + i];
        rs[l] = r;
      }
      for (uint64_t h = 0; h < (1ull << Hn); ++h)
        for (int kl = 0; kl <= L; ++kl) {
          int64_t r = 0;
          int i = kl + 1;
          for (uint64_t c = h; c; c &= c - 1, ++i) {
            const int p = L + __builtin_ctzll(c);
            r += i <= p ? binom[p * BINOM_LD + i] : 0;
          }
          rs[((size_t)1 << L) + h * (size_t)(L + 1) + (size_t)kl] = r;
        }
      P->rs_L = L;
      e = hipMalloc(&P->d_rsplit, sizeof(int64_t) * rs.size());
      if (e == hipSuccess)
        e = hipMemcpy(P->d_rsplit, rs.data(), sizeof(int64_t) * rs.size(), hipMemcpyHostToDevice);
    }
    if (e == hipSuccess && !pc.empty()) e = hipMalloc(&P->d_pairs, sizeof(PairChan) * pc.size());
    if (e == hipSuccess && !pc.empty())
      e = hipMemcpy(P->d_pairs, pc.data(), sizeof(PairChan) * pc.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess && m->basis_kind == RYD_BASIS_CODES) {
      e = hipMalloc(&P->d_codes, sizeof(int64_t) * (size_t)m->n_codes);
      if (e == hipSuccess)
        e = hipMemcpy(P->d_codes, m->codes, sizeof(int64_t) * (size_t)m->n_codes,
                      hipMemcpyHostToDevice);
    }
    if (e != hipSuccess)
      return cleanup(e == hipErrorOutOfMemory ? RYD_ERR_MEMORY : RYD_ERR_DEVICE,
                     std::string("channel table upload: ") + hipGetErrorString(e));
    if (m->basis_kind == RYD_BASIS_GRAPH) {
      if (int rc = build_graph(P, m)) {
        const std::string msg = g_err;
        ryd_plan_destroy(P);
        return fail(rc, msg);
      }
    }
    size_t blocks = CHAN_MAX_BLOCKS;
    P->ws_bytes = WS_HEADER + blocks * NRED * sizeof(double);
    const size_t occ = (size_t)OCC_BLOCKS * (MAXN_OCC + 1) * sizeof(double);
    if (WS_HEADER + occ > P->ws_bytes) P->ws_bytes = WS_HEADER + occ;
    if (m->basis_kind == RYD_BASIS_CODES || m->basis_kind == RYD_BASIS_GRAPH) {
      *out = P;
      return RYD_OK;
    }
  }
  if (P->basis == RYD_BASIS_CHAIN_FIB) {
    const int L = n < FIB_L ? n : FIB_L;
    P->fib_L = L;
    P->fib_cnt[0] = fib(L + 2);
    P->fib_cnt[1] = fib(L + 1);
    std::vector<uint32_t> codes;
    std::vector<double> dlo;
    for (uint32_t c = 0; c < (1u << L); ++c) {
      if (c & (c >> 1)) continue;       // adjacent excitations are blockaded
      codes.push_back(c);
      double d = 0.0;
      for (int p = 0; p < L; ++p) {
        if (!((c >> p) & 1u)) continue;
        d += dl[p];
        for (int q = p + 1; q < L; ++q)
          if ((c >> q) & 1u) d += vb[(size_t)p * n + q];
      }
      dlo.push_back(d);
    }
    // high parts in ascending order: Zeckendorf unrank of k in the (n-L)-bit
    // system, each with its first rank and the diagonal of its own bits
    const int mh = n - L;
    std::vector<int64_t> F(mh + 3 > n + 3 ? mh + 3 : n + 3);
    F[0] = 0;
    F[1] = 1;
    for (size_t i = 2; i < F.size(); ++i) F[i] = F[i - 1] + F[i - 2];
    const int64_t nh = F[mh + 2];
    std::vector<FibHi> hinfo((size_t)nh);
    int64_t next_rank = 0;
    for (int64_t k = 0; k < nh; ++k) {
      uint64_t h = 0;
      int64_t r = k;
      int prev = 0;
      for (int q = mh - 1; q >= 0; --q) {
        if (!prev && r >= F[q + 2]) { h |= 1ull << (q + L); r -= F[q + 2]; prev = 1; }
        else prev = 0;
      }
      FibHi& H = hinfo[(size_t)k];
      H.code = h;
      H.R = 0;
      H.ehi = 0.0;
      H.pad = 0.0;
      for (uint64_t c = h; c; c &= c - 1) {
        const int p = __builtin_ctzll(c);
        H.R += F[p + 2];
        H.ehi += dl[p];
        for (uint64_t c2 = c & (c - 1); c2; c2 &= c2 - 1) H.ehi += vb[(size_t)p * n + __builtin_ctzll(c2)];
      }
      if (H.R != next_rank) return cleanup(RYD_ERR_DEVICE, "internal: sector high-part ranks out of order");
      next_rank += ((h >> L) & 1ull) ? F[L + 1] : F[L + 2];
    }
    if (next_rank != P->dim) return cleanup(RYD_ERR_DEVICE, "internal: sector dimension mismatch");
    P->fib_nh = nh;
    // split H·v: the top FIB_K bits in a second pass once the far partners
    // of those bits no longer stay in the caches
    P->fib_split = n >= FIB_SPLIT_MIN_N && n - FIB_K >= L;
    std::vector<FibTop> top(FIB_NG);
    if (P->fib_split) {
      const int S = n - FIB_K;
      for (int gi = 0; gi < FIB_NG; ++gi) {
        uint64_t g = 0;
        int64_t r = gi;
        int prev = 0;
        for (int q = FIB_K - 1; q >= 0; --q) {
          if (!prev && r >= F[q + 2]) { g |= 1ull << (q + S); r -= F[q + 2]; prev = 1; }
          else prev = 0;
        }
        top[gi].g = g;
        top[gi].R = 0;
        for (uint64_t c = g; c; c &= c - 1) top[gi].R += F[__builtin_ctzll(c) + 2];
      }
    }
    e = hipMalloc(&P->d_lo_code, sizeof(uint32_t) * codes.size());
    if (e == hipSuccess) e = hipMalloc(&P->d_lo_diag, sizeof(double) * dlo.size());
    if (e == hipSuccess) e = hipMalloc(&P->d_hi, sizeof(FibHi) * hinfo.size());
    if (e == hipSuccess)
      e = hipMemcpy(P->d_lo_code, codes.data(), sizeof(uint32_t) * codes.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess)
      e = hipMemcpy(P->d_lo_diag, dlo.data(), sizeof(double) * dlo.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(P->d_hi, hinfo.data(), sizeof(FibHi) * hinfo.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess && P->fib_split) e = hipMalloc(&P->d_top, sizeof(FibTop) * FIB_NG);
    if (e == hipSuccess && P->fib_split)
      e = hipMemcpy(P->d_top, top.data(), sizeof(FibTop) * FIB_NG, hipMemcpyHostToDevice);
    if (e != hipSuccess)
      return cleanup(e == hipErrorOutOfMemory ? RYD_ERR_MEMORY : RYD_ERR_DEVICE,
                     std::string("sector table upload: ") + hipGetErrorString(e));
    size_t blocks = fib_grid(P);
    if (P->fib_split && fib_top_grid(P, 16) > blocks) blocks = fib_top_grid(P, 16);
    if (blocks < 8192) blocks = 8192;
    P->ws_bytes = WS_HEADER + blocks * NRED * sizeof(double);
    const size_t occ = (size_t)OCC_BLOCKS * (MAXN_OCC + 1) * sizeof(double);
    if (WS_HEADER + occ > P->ws_bytes) P->ws_bytes = WS_HEADER + occ;
    *out = P;
    return RYD_OK;
  }
  bool want[NTBW] = {};
  for (int dt = RYD_F32; dt <= RYD_C128; ++dt) {
    want[tile_bits_of(dt) - 10] = true;
    const size_t eb = elem_bytes(dt);
    const int smin = std::min(split_min_n(eb, false), split_min_n(eb, true));
    if (n >= smin) want[passa_bits_of(dt) - 10] = true;
  }
  for (int k = 0; k < NTBW; ++k) {
    if (!want[k]) continue;
    const int tb = n < 10 + k ? n : 10 + k;
    P->tbits[k] = tb;
    std::vector<double> elo, ehi;
    build_tables(n, tb, dl, vb, elo, ehi);
    e = hipMalloc(&P->d_elo[k], sizeof(double) * elo.size());
    if (e == hipSuccess) e = hipMalloc(&P->d_ehi[k], sizeof(double) * ehi.size());
    if (e == hipSuccess)
      e = hipMemcpy(P->d_elo[k], elo.data(), sizeof(double) * elo.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess)
      e = hipMemcpy(P->d_ehi[k], ehi.data(), sizeof(double) * ehi.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess)
      return cleanup(e == hipErrorOutOfMemory ? RYD_ERR_MEMORY : RYD_ERR_DEVICE,
                     std::string("table upload: ") + hipGetErrorString(e));
  }
  // workspace: header + NRED partials for the largest grid we launch
  size_t blocks = P->generic ? (size_t)CHAN_MAX_BLOCKS : 8192;
  for (int dt = RYD_F32; dt <= RYD_C128; ++dt) {
    size_t g = apply_grid(P, dt);
    if (g > blocks) blocks = g;
  }
  P->ws_bytes = WS_HEADER + blocks * NRED * sizeof(double);
  const size_t occ = (size_t)OCC_BLOCKS * (MAXN_OCC + 1) * sizeof(double);
  if (WS_HEADER + occ > P->ws_bytes) P->ws_bytes = WS_HEADER + occ;
  *out = P;
  return RYD_OK;
}

int ryd_plan_create_csr(int64_t dim, const int64_t* indptr, const int64_t* indices,
                        const double* val_re, const double* val_im, ryd_plan** out) {
  if (!out || !indptr) return fail(RYD_ERR_INVALID, "null argument");
  *out = nullptr;
  if (dim < 1 || dim > (int64_t)INT32_MAX) return fail(RYD_ERR_INVALID, "csr: dim must be 1..2^31-1");
  if (indptr[0] != 0) return fail(RYD_ERR_INVALID, "csr: indptr[0] must be 0");
  for (int64_t k = 0; k < dim; ++k)
    if (indptr[k + 1] < indptr[k]) return fail(RYD_ERR_INVALID, "csr: indptr must not decrease");
  const int64_t nnz = indptr[dim];
  if (nnz > 0 && (!indices || !val_re)) return fail(RYD_ERR_INVALID, "null argument");
  std::vector<int32_t> col((size_t)(nnz > 0 ? nnz : 1), 0);
  bool cplx = false;
  for (int64_t q = 0; q < nnz; ++q) {
    if (indices[q] < 0 || indices[q] >= dim)
      return fail(RYD_ERR_INVALID, "csr: column index out of range");
    if (!std::isfinite(val_re[q]) || (val_im && !std::isfinite(val_im[q])))
      return fail(RYD_ERR_INVALID, "csr: non-finite value");
    col[(size_t)q] = (int32_t)indices[q];
    if (val_im && val_im[q] != 0.0) cplx = true;
  }
  ryd_plan* P = new ryd_plan();
  P->n = 0;
  P->basis = RYD_BASIS_CSR;
  P->dim = dim;
  P->cplx = cplx;
  P->generic = true;
  (void)hipGetDevice(&P->device);
  const size_t nb = (size_t)(nnz > 0 ? nnz : 1);
  hipError_t e = hipMalloc(&P->d_csr_ptr, sizeof(int64_t) * (size_t)(dim + 1));
  if (e == hipSuccess) e = hipMemcpy(P->d_csr_ptr, indptr, sizeof(int64_t) * (size_t)(dim + 1), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMalloc(&P->d_csr_col, sizeof(int32_t) * nb);
  if (e == hipSuccess) e = hipMemcpy(P->d_csr_col, col.data(), sizeof(int32_t) * nb, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMalloc(&P->d_csr_re, sizeof(double) * nb);
  if (e == hipSuccess && nnz > 0)
    e = hipMemcpy(P->d_csr_re, val_re, sizeof(double) * (size_t)nnz, hipMemcpyHostToDevice);
  if (e == hipSuccess && cplx) e = hipMalloc(&P->d_csr_im, sizeof(double) * nb);
  if (e == hipSuccess && cplx)
    e = hipMemcpy(P->d_csr_im, val_im, sizeof(double) * (size_t)nnz, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    ryd_plan_destroy(P);
    return fail(e == hipErrorOutOfMemory ? RYD_ERR_MEMORY : RYD_ERR_DEVICE,
                std::string("csr upload: ") + hipGetErrorString(e));
  }
  P->ws_bytes = WS_HEADER + (size_t)CHAN_MAX_BLOCKS * NRED * sizeof(double);
  *out = P;
  return RYD_OK;
}

int ryd_plan_destroy(ryd_plan* P) {
  if (!P) return RYD_OK;
  if (P->d_csr_ptr) (void)hipFree(P->d_csr_ptr);
  if (P->d_csr_col) (void)hipFree(P->d_csr_col);
  if (P->d_csr_re) (void)hipFree(P->d_csr_re);
  if (P->d_csr_im) (void)hipFree(P->d_csr_im);
  if (P->d_vb) (void)hipFree(P->d_vb);
  if (P->d_lo_code) (void)hipFree(P->d_lo_code);
  if (P->d_lo_diag) (void)hipFree(P->d_lo_diag);
  if (P->d_hi) (void)hipFree(P->d_hi);
  if (P->d_top) (void)hipFree(P->d_top);
  if (P->d_codes) (void)hipFree(P->d_codes);
  if (P->d_binom) (void)hipFree(P->d_binom);
  if (P->d_rsplit) (void)hipFree(P->d_rsplit);
  if (P->d_pairs) (void)hipFree(P->d_pairs);
  if (P->d_hpk) (void)hipFree(P->d_hpk);
  if (P->d_lo_idx) (void)hipFree(P->d_lo_idx);
  if (P->d_lowtab) (void)hipFree(P->d_lowtab);
  if (P->d_fc) (void)hipFree(P->d_fc);
  if (P->d_pre) (void)hipFree(P->d_pre);
  if (P->d_hlist) (void)hipFree(P->d_hlist);
  if (P->d_post) (void)hipFree(P->d_post);
  if (P->d_post_off) (void)hipFree(P->d_post_off);
  if (P->d_rcode) (void)hipFree(P->d_rcode);
  if (P->d_rdiag) (void)hipFree(P->d_rdiag);
  if (P->d_lrow) (void)hipFree(P->d_lrow);
  if (P->d_tp) (void)hipFree(P->d_tp);
  if (P->d_tpk) (void)hipFree(P->d_tpk);
  for (int k = 0; k < NTBW; ++k) {
    if (P->d_elo[k]) (void)hipFree(P->d_elo[k]);
    if (P->d_ehi[k]) (void)hipFree(P->d_ehi[k]);
  }
  delete P;
  return RYD_OK;
}

int64_t ryd_plan_dim(const ryd_plan* P) { return P ? P->dim : 0; }

int ryd_plan_rank_mode(const ryd_plan* P) {
  if (!P) return -1;
  if (P->basis == RYD_BASIS_CHAIN_FIB) return 4;
  return P->generic ? P->rank_mode : RANK_FULL;
}

size_t ryd_workspace_bytes(const ryd_plan* P) { return P ? P->ws_bytes : 0; }

int ryd_hv(const ryd_plan* P, int dtype, const void* x, void* y, void* stream) {
  if (!P || !x || !y) return fail(RYD_ERR_INVALID, "null plan or vector");
  if (int rc = check_dtype(P, dtype)) return rc;
  hipStream_t st = static_cast<hipStream_t>(stream);
  DISPATCH(dtype, do_apply, P, dtype, x, nullptr, y, 1.0, 1.0, 0.0, 0.0, nullptr, nullptr, 0,
           nullptr, st);
}

int ryd_apply(const ryd_plan* P, int dtype, const void* x, const void* z, void* y, double s,
              double ch, double cx, double cz, const uint32_t* probe, double* sums,
              int accumulate, void* ws, void* stream) {
  if (!P || !x || !y) return fail(RYD_ERR_INVALID, "null plan or vector");
  if (sums && !ws) return fail(RYD_ERR_INVALID, "reductions need a workspace");
  if (int rc = check_dtype(P, dtype)) return rc;
  hipStream_t st = static_cast<hipStream_t>(stream);
  DISPATCH(dtype, do_apply, P, dtype, x, z, y, s, ch, cx, cz, probe, sums, accumulate, ws, st);
}

size_t ryd_lanczos_state_bytes(void) { return sizeof(LzState); }

int ryd_lanczos_stored(const ryd_plan* P, int dtype, void* basis, int64_t ld, int64_t jbase,
                       int64_t j0, int64_t steps, double* alpha, double* nrm, void* state,
                       double tiny, void* ws, void* stream) {
  if (!P || !basis || !alpha || !nrm || !state || !ws || ld < P->dim || j0 < jbase ||
      (j0 > 0 && j0 - 1 < jbase) || steps < 0)
    return fail(RYD_ERR_INVALID, "bad argument to ryd_lanczos_stored");
  if (int rc = check_dtype(P, dtype)) return rc;
  hipStream_t st = static_cast<hipStream_t>(stream);
  DISPATCH(dtype, do_lanczos_stored, P, dtype, basis, ld, jbase, j0, steps, alpha, nrm, state,
           tiny, ws, st);
}

int ryd_ritz_stored(int dtype, const void* basis, int64_t ld, int64_t dim, int64_t count,
                    const double* y0, const double* y1, const double* nrm, int64_t j0,
                    int accumulate, void* out0, void* out1, void* ws, void* stream) {
  if (!basis || !y0 || !nrm || !out0 || !ws || dim < 1 || ld < dim || count < 0 ||
      count > RITZ_MAX || (y1 && !out1))
    return fail(RYD_ERR_INVALID, "bad argument to ryd_ritz_stored");
  if (count == 0 && accumulate) return RYD_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  double* yd = reinterpret_cast<double*>(static_cast<char*>(ws) + WS_HEADER);
  HIP_TRY(hipMemcpyAsync(yd, y0, sizeof(double) * count, hipMemcpyHostToDevice, st));
  if (y1) HIP_TRY(hipMemcpyAsync(yd + RITZ_MAX, y1, sizeof(double) * count, hipMemcpyHostToDevice, st));
  const int two = y1 ? 1 : 0;
  switch (dtype) {
    case RYD_F32:
      launch_ritz_stored<float>(basis, ld, dim, (int)count, yd, two, nrm, j0, accumulate, out0, out1, st);
      break;
    case RYD_F64:
      launch_ritz_stored<double>(basis, ld, dim, (int)count, yd, two, nrm, j0, accumulate, out0, out1, st);
      break;
    case RYD_C64:
      launch_ritz_stored<c32>(basis, ld, dim, (int)count, yd, two, nrm, j0, accumulate, out0, out1, st);
      break;
    case RYD_C128:
      launch_ritz_stored<c64>(basis, ld, dim, (int)count, yd, two, nrm, j0, accumulate, out0, out1, st);
      break;
    default:
      return fail(RYD_ERR_INVALID, "unknown dtype");
  }
  LAUNCH_CHECK();
  // the host coefficient copies must finish before the caller's arrays go away
  HIP_TRY(hipStreamSynchronize(st));
  return RYD_OK;
}

int ryd_lanczos(const ryd_plan* P, int dtype, void* b0, void* b1, void* b2, int64_t j0,
                int64_t steps, double* alpha, double* nrm, void* state, double tiny, void* ritz0,
                void* ritz1, const double* y0, const double* y1, int replay, void* ws,
                void* stream) {
  if (!P || !b0 || !b1 || !b2 || !alpha || !nrm || !state || !ws)
    return fail(RYD_ERR_INVALID, "null argument to ryd_lanczos");
  if (j0 < 0 || steps < 0) return fail(RYD_ERR_INVALID, "negative step range");
  if ((ritz0 && !y0) || (ritz1 && !y1)) return fail(RYD_ERR_INVALID, "Ritz vector without coefficients");
  if (int rc = check_dtype(P, dtype)) return rc;
  hipStream_t st = static_cast<hipStream_t>(stream);
  DISPATCH(dtype, do_lanczos, P, dtype, b0, b1, b2, j0, steps, alpha, nrm, state, tiny, ritz0,
           ritz1, y0, y1, replay, ws, st);
}

int ryd_lanczos_locked(const ryd_plan* P, int dtype, void* b0, void* b1, void* b2,
                       const void* lock, int64_t j0, int64_t steps, double* alpha, double* nrm,
                       double* gam, void* state, double tiny, void* ritz0, void* ritz1,
                       const double* y0, const double* y1, int replay, void* ws, void* stream) {
  if (!P || !b0 || !b1 || !b2 || !lock || !alpha || !nrm || !gam || !state || !ws)
    return fail(RYD_ERR_INVALID, "null argument to ryd_lanczos_locked");
  if (j0 < 0 || steps < 0) return fail(RYD_ERR_INVALID, "negative step range");
  if ((ritz0 && !y0) || (ritz1 && !y1)) return fail(RYD_ERR_INVALID, "Ritz vector without coefficients");
  if (int rc = check_dtype(P, dtype)) return rc;
  hipStream_t st = static_cast<hipStream_t>(stream);
  DISPATCH(dtype, do_lanczos_locked, P, dtype, b0, b1, b2, lock, j0, steps, alpha, nrm, gam,
           state, tiny, ritz0, ritz1, y0, y1, replay, ws, st);
}

size_t ryd_reortho_workspace_bytes(int64_t dim, int64_t k) {
  if (dim < 1 || k < 1) return 0;
  const int64_t nchunk = (dim + GS_CHUNK - 1) / GS_CHUNK;
  return sizeof(double) * (size_t)(2 * k * nchunk + 2 * k);
}

int ryd_lanczos_reortho(const ryd_plan* P, int dtype, void* basis, int64_t ld, void* w, int64_t k,
                        double phi_scale, double* alpha, double* nrm, void* state, double tiny,
                        int passes, void* gs_ws, void* ws, void* stream) {
  if (!P || !basis || !w || !alpha || !nrm || !state || !gs_ws || !ws)
    return fail(RYD_ERR_INVALID, "null argument to ryd_lanczos_reortho");
  if (k < 1 || k > P->dim) return fail(RYD_ERR_INVALID, "ryd_lanczos_reortho: need 1 <= k <= dim");
  if (ld < P->dim) return fail(RYD_ERR_INVALID, "ryd_lanczos_reortho: row stride below dim");
  if (passes < 0 || passes > 4) return fail(RYD_ERR_INVALID, "ryd_lanczos_reortho: passes in 0..4");
  if (!(phi_scale > 0.0) || !std::isfinite(phi_scale))
    return fail(RYD_ERR_INVALID, "ryd_lanczos_reortho: phi_scale must be finite and > 0");
  if (int rc = check_dtype(P, dtype)) return rc;
  hipStream_t st = static_cast<hipStream_t>(stream);
  DISPATCH(dtype, do_lanczos_reortho, P, dtype, basis, ld, (int64_t)0, w, k, phi_scale, alpha, nrm,
           state, tiny, passes, static_cast<double*>(gs_ws), ws, st);
}

int ryd_lanczos_reortho_locked(const ryd_plan* P, int dtype, void* basis, int64_t ld, int64_t nlock,
                               void* w, int64_t k, double phi_scale, double* alpha, double* nrm,
                               void* state, double tiny, int passes, void* gs_ws, void* ws,
                               void* stream) {
  if (!P || !basis || !w || !alpha || !nrm || !state || !gs_ws || !ws)
    return fail(RYD_ERR_INVALID, "null argument to ryd_lanczos_reortho_locked");
  if (nlock < 0 || k < 1 || nlock + k > P->dim)
    return fail(RYD_ERR_INVALID, "ryd_lanczos_reortho_locked: need nlock >= 0, k >= 1, nlock + k <= dim");
  if (nlock > 0 && passes < 1)
    return fail(RYD_ERR_INVALID, "ryd_lanczos_reortho_locked: locked rows need passes >= 1");
  if (ld < P->dim) return fail(RYD_ERR_INVALID, "ryd_lanczos_reortho_locked: row stride below dim");
  if (passes < 0 || passes > 4)
    return fail(RYD_ERR_INVALID, "ryd_lanczos_reortho_locked: passes in 0..4");
  if (!(phi_scale > 0.0) || !std::isfinite(phi_scale))
    return fail(RYD_ERR_INVALID, "ryd_lanczos_reortho_locked: phi_scale must be finite and > 0");
  if (int rc = check_dtype(P, dtype)) return rc;
  hipStream_t st = static_cast<hipStream_t>(stream);
  DISPATCH(dtype, do_lanczos_reortho, P, dtype, basis, ld, nlock, w, k, phi_scale, alpha, nrm,
           state, tiny, passes, static_cast<double*>(gs_ws), ws, st);
}

int ryd_lanczos_status(const void* state, int* stopped, int64_t* steps, void* stream) {
  if (!state || !stopped || !steps) return fail(RYD_ERR_INVALID, "null argument");
  LzState h;
  HIP_TRY(hipMemcpyAsync(&h, state, sizeof(LzState), hipMemcpyDeviceToHost,
                         static_cast<hipStream_t>(stream)));
  HIP_TRY(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
  *stopped = h.stop;
  *steps = h.steps;
  return RYD_OK;
}

int ryd_cheb_moments(const ryd_plan* P, int dtype, const uint32_t* probe, void* t0, void* t1,
                     void* t2, int64_t nsteps, double c, double r, double* sums, void* ws,
                     void* stream) {
  if (!P || !probe || !t0 || !t1 || !t2 || !sums || !ws)
    return fail(RYD_ERR_INVALID, "null argument to ryd_cheb_moments");
  if (!(r > 0.0) || !std::isfinite(c)) return fail(RYD_ERR_INVALID, "need finite c and r > 0");
  if (nsteps < 1) return fail(RYD_ERR_INVALID, "nsteps must be >= 1");
  if (int rc = check_dtype(P, dtype)) return rc;
  hipStream_t st = static_cast<hipStream_t>(stream);
  DISPATCH(dtype, do_cheb, P, dtype, probe, t0, t1, t2, nsteps, c, r, sums, ws, st);
}

}  // extern "C"

// Allocation-free variant of ryd_cheb_moments: the reduction partials live in
// this file-scope device array instead of a caller workspace.  Each GPU holds
// its own copy until the process ends; host code uses the per-device address
// from hipGetSymbolAddress (the array's name is only a host-side stand-in).
// Every call on a GPU shares that GPU's array, so calls that overlap on one
// GPU race.  Every block writes its partial before the finishing launch reads
// it, so nothing in the array needs clearing between calls.
constexpr size_t SCRATCH_BYTES = (size_t)8 << 20;
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

namespace {
constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[SCRATCH_MAX_DEVICES] = {};

// the address of `scratch` on the current device, looked up once per device
int scratch_on_current_device(void** out) {
  int dev = -1;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) return fail(RYD_ERR_DEVICE, "device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(RYD_ERR_DEVICE, "hipGetSymbolAddress returned a null scratch address");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return RYD_OK;
}
}  // namespace

extern "C" {

int ryd_cheb_moments_noalloc(const ryd_plan* P, int dtype, const uint32_t* probe, void* t0,
                             void* t1, void* t2, int64_t nsteps, double c, double r, double* sums,
                             void* stream) {
  if (!P || !probe || !t0 || !t1 || !t2 || !sums)
    return fail(RYD_ERR_INVALID, "null argument to ryd_cheb_moments_noalloc");
  if (P->ws_bytes > SCRATCH_BYTES)
    return fail(RYD_ERR_INVALID, "ryd_cheb_moments_noalloc: the plan's reductions need " +
                                     std::to_string(P->ws_bytes) + " bytes of scratch, more than the " +
                                     std::to_string(SCRATCH_BYTES) + " held for allocation-free calls");
  int dev = -1;
  HIP_TRY(hipGetDevice(&dev));
  if (dev != P->device)
    return fail(RYD_ERR_INVALID, "ryd_cheb_moments_noalloc: the plan's device must be current");
  void* ws = nullptr;
  if (int rc = scratch_on_current_device(&ws)) return rc;
  return ryd_cheb_moments(P, dtype, probe, t0, t1, t2, nsteps, c, r, sums, ws, stream);
}

}  // extern "C"

namespace {
template <class V>
int do_dots(const void* a, const void* b, int64_t n, double* sums, void* ws, hipStream_t st) {
  Reduce rd = make_reduce(ws, POST_STORE, sums);
  k_dots<V><<<stream_grid(n), 256, 0, st>>>(static_cast<const V*>(a), static_cast<const V*>(b), n, rd);
  LAUNCH_CHECK();
  k_reduce_finish<<<1, 256, 0, st>>>(rd, stream_grid(n));
  LAUNCH_CHECK();
  return RYD_OK;
}
}  // namespace

extern "C" {

int ryd_dots(int dtype, const void* a, const void* b, int64_t n, double* sums, void* ws,
             void* stream) {
  if (!b || !sums || !ws || n < 0) return fail(RYD_ERR_INVALID, "bad argument to ryd_dots");
  hipStream_t st = static_cast<hipStream_t>(stream);
  switch (dtype) {
    case RYD_F32: return do_dots<float>(a, b, n, sums, ws, st);
    case RYD_F64: return do_dots<double>(a, b, n, sums, ws, st);
    case RYD_C64: return do_dots<c32>(a, b, n, sums, ws, st);
    case RYD_C128: return do_dots<c64>(a, b, n, sums, ws, st);
    default: return fail(RYD_ERR_INVALID, "unknown dtype");
  }
}

int ryd_cast(int din, const void* x, int dout, void* y, int64_t n, void* stream) {
  if (!x || !y || n < 0) return fail(RYD_ERR_INVALID, "bad argument to ryd_cast");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const bool cin = din >= RYD_C64, cout = dout >= RYD_C64;
  if (cin != cout) return fail(RYD_ERR_INVALID, "cast cannot change real/complex kind");
  const int64_t m = cin ? 2 * n : n;
  const bool in32 = (din == RYD_F32 || din == RYD_C64), out32 = (dout == RYD_F32 || dout == RYD_C64);
  const unsigned g = stream_grid(m);
  if (in32 && !out32)
    k_cast<float, double><<<g, 256, 0, st>>>(static_cast<const float*>(x), static_cast<double*>(y), m);
  else if (!in32 && out32)
    k_cast<double, float><<<g, 256, 0, st>>>(static_cast<const double*>(x), static_cast<float*>(y), m);
  else if (in32)
    k_cast<float, float><<<g, 256, 0, st>>>(static_cast<const float*>(x), static_cast<float*>(y), m);
  else
    k_cast<double, double><<<g, 256, 0, st>>>(static_cast<const double*>(x), static_cast<double*>(y), m);
  LAUNCH_CHECK();
  return RYD_OK;
}

}  // extern "C"

namespace {
template <class V>
int do_occ(const ryd_plan* P, const void* x, double* out, void* ws, hipStream_t st) {
  if (P->basis == RYD_BASIS_CODES || P->basis == RYD_BASIS_GRAPH) {
    const int64_t g0 = (P->dim + 4095) / 4096;
    const unsigned g = (unsigned)(g0 < OCC_BLOCKS ? g0 : OCC_BLOCKS);
    double* part = reinterpret_cast<double*>(static_cast<char*>(ws) + WS_HEADER);
    k_chan_occ<V><<<g, 256, 0, st>>>(chan_model(P), static_cast<const V*>(x), part);
    LAUNCH_CHECK();
    k_occ_finish<<<1, 256, 0, st>>>(part, (int)g, P->n, out);
    LAUNCH_CHECK();
    return RYD_OK;
  }
  if (P->basis == RYD_BASIS_CHAIN_FIB) {
    const int64_t nh = P->fib_nh;
    const unsigned g = (unsigned)(nh < OCC_BLOCKS ? nh : OCC_BLOCKS);
    double* part = reinterpret_cast<double*>(static_cast<char*>(ws) + WS_HEADER);
    k_fib_occ<V><<<g, 256, 0, st>>>(fib_model(P), static_cast<const V*>(x), part);
    LAUNCH_CHECK();
    k_occ_finish<<<1, 256, 0, st>>>(part, (int)g, P->n, out);
    LAUNCH_CHECK();
    return RYD_OK;
  }
  const uint64_t ob = P->n < OB ? P->n : OB;
  const uint64_t ntiles = (1ull << P->n) >> ob;
  const unsigned g = (unsigned)(ntiles < OCC_BLOCKS ? ntiles : OCC_BLOCKS);
  // partials after the header
  double* part = reinterpret_cast<double*>(static_cast<char*>(ws) + WS_HEADER);
  k_site_occ<V><<<g, 256, 0, st>>>(static_cast<const V*>(x), P->n, part);
  LAUNCH_CHECK();
  k_occ_finish<<<1, 256, 0, st>>>(part, (int)g, P->n, out);
  LAUNCH_CHECK();
  return RYD_OK;
}
}  // namespace

extern "C" int ryd_site_occupation(const ryd_plan* P, int dtype, const void* x, double* out,
                                   void* ws, void* stream) {
  if (P && P->basis == RYD_BASIS_CSR)
    return fail(RYD_ERR_UNSUPPORTED, "a sparse-matrix plan has no site codes");
  if (!P || !x || !out || !ws) return fail(RYD_ERR_INVALID, "null argument to ryd_site_occupation");
  hipStream_t st = static_cast<hipStream_t>(stream);
  switch (dtype) {
    case RYD_F32: return do_occ<float>(P, x, out, ws, st);
    case RYD_F64: return do_occ<double>(P, x, out, ws, st);
    case RYD_C64: return do_occ<c32>(P, x, out, ws, st);
    case RYD_C128: return do_occ<c64>(P, x, out, ws, st);
    default: return fail(RYD_ERR_INVALID, "unknown dtype");
  }
}

namespace {

ObsBasis obs_basis(const ryd_plan* P) {
  ObsBasis B;
  std::memset(&B, 0, sizeof(B));
  B.n = P->n;
  B.dim = P->dim;
  if (P->basis == RYD_BASIS_CHAIN_FIB) {
    B.kind = 2;
    for (int p = 0; p < MAXN; ++p) B.W[p] = fib(p + 2);
  } else if (P->generic) {
    B.kind = 1;
    B.mode = P->rank_mode;
    B.m = P->popcount;
    B.gl = P->gr_L;
    B.tl = P->gr_tl;
    B.codes = P->d_codes;
    B.binom = P->d_binom;
    B.rsL = P->rs_L;
    B.rsplit = P->d_rsplit;
    B.hpk = P->d_hpk;
    B.lo_idx = P->d_lo_idx;
    B.fc = P->d_fc;
    B.pre = P->d_pre;
  } else {
    B.kind = 0;
  }
  return B;
}

template <class V>
int do_pauli(const ryd_plan* P, const void* x, int count, const uint64_t* flip,
             const uint64_t* sign, const int32_t* ny, double* out, void* ws, hipStream_t st) {
  const ObsBasis B = obs_basis(P);
  int64_t g0 = (P->dim + 4095) / 4096;
  const int64_t cap = (int64_t)((P->ws_bytes - WS_HEADER) / (2 * PAULI_BATCH * sizeof(double))) - 1;
  if (g0 > cap) g0 = cap;
  if (g0 > 2048) g0 = 2048;
  if (g0 < 1) return fail(RYD_ERR_INVALID, "workspace too small");
  const unsigned g = (unsigned)g0;
  double* part = reinterpret_cast<double*>(static_cast<char*>(ws) + WS_HEADER);
  for (int q0 = 0; q0 < count; q0 += PAULI_BATCH) {
    PauliBatch S;
    S.count = count - q0 < PAULI_BATCH ? count - q0 : PAULI_BATCH;
    for (int q = 0; q < PAULI_BATCH; ++q) {
      S.flip[q] = q < S.count ? flip[q0 + q] : 0;
      S.sign[q] = q < S.count ? sign[q0 + q] : 0;
    }
    k_pauli<V><<<g, 256, 0, st>>>(B, S, static_cast<const V*>(x), part);
    LAUNCH_CHECK();
    // (re, im) sums, then the i^{ny} factor on the host side of the copy
    k_sum_columns<<<1, 256, 0, st>>>(part, (int)g, 2 * PAULI_BATCH, part + (size_t)g * 2 * PAULI_BATCH);
    LAUNCH_CHECK();
    double h[2 * PAULI_BATCH];
    HIP_TRY(hipMemcpyAsync(h, part + (size_t)g * 2 * PAULI_BATCH, sizeof(h), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    double o[2 * PAULI_BATCH];
    for (int q = 0; q < S.count; ++q) {
      double re = h[2 * q], im = h[2 * q + 1];
      for (int k = 0; k < (ny[q0 + q] & 3); ++k) {   // multiply by i
        const double t = re;
        re = -im;
        im = t;
      }
      o[2 * q] = re;
      o[2 * q + 1] = im;
    }
    HIP_TRY(hipMemcpyAsync(out + 2 * q0, o, sizeof(double) * 2 * S.count, hipMemcpyHostToDevice, st));
    HIP_TRY(hipStreamSynchronize(st));
  }
  return RYD_OK;
}

template <class V>
int do_corr(const ryd_plan* P, const void* x, double* out, void* scratch, size_t bytes,
            hipStream_t st) {
  const ObsBasis B = obs_basis(P);
  const int n = P->n;
  const int width = n * (n + 1) / 2 + 1;
  int64_t g0 = (P->dim + 4095) / 4096;
  const int64_t cap = (int64_t)(bytes / (sizeof(double) * (size_t)width)) - 1;   // + the sums
  if (g0 > cap) g0 = cap;
  if (g0 > 512) g0 = 512;
  if (g0 < 1) return fail(RYD_ERR_INVALID, "scratch too small for the correlation sums");
  const unsigned g = (unsigned)g0;
  double* part = static_cast<double*>(scratch);
  double* sums = part + (size_t)g * width;
  k_corr<V><<<g, 256, 0, st>>>(B, static_cast<const V*>(x), part);
  LAUNCH_CHECK();
  k_sum_columns<<<(width + 255) / 256, 256, 0, st>>>(part, (int)g, width, sums);
  LAUNCH_CHECK();
  k_corr_square<<<(n * n + 256) / 256, 256, 0, st>>>(sums, n, out);
  LAUNCH_CHECK();
  return RYD_OK;
}

}  // namespace

extern "C" int ryd_pauli_expect(const ryd_plan* P, int dtype, const void* x, int count,
                                const uint64_t* flip, const uint64_t* sign, const int32_t* ny,
                                double* out, void* ws, void* stream) {
  if (P && P->basis == RYD_BASIS_CSR)
    return fail(RYD_ERR_UNSUPPORTED, "a sparse-matrix plan has no site codes");
  if (!P || !x || !out || !ws || count < 0 || (count > 0 && (!flip || !sign || !ny)))
    return fail(RYD_ERR_INVALID, "bad argument to ryd_pauli_expect");
  for (int q = 0; q < count; ++q)
    if ((P->n < 64 && ((flip[q] | sign[q]) >> P->n)) || ny[q] < 0)
      return fail(RYD_ERR_INVALID, "Pauli string outside the plan's sites");
  if (count == 0) return RYD_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  switch (dtype) {
    case RYD_F32: return do_pauli<float>(P, x, count, flip, sign, ny, out, ws, st);
    case RYD_F64: return do_pauli<double>(P, x, count, flip, sign, ny, out, ws, st);
    case RYD_C64: return do_pauli<c32>(P, x, count, flip, sign, ny, out, ws, st);
    case RYD_C128: return do_pauli<c64>(P, x, count, flip, sign, ny, out, ws, st);
    default: return fail(RYD_ERR_INVALID, "unknown dtype");
  }
}

extern "C" int ryd_site_correlations(const ryd_plan* P, int dtype, const void* x, double* out,
                                     void* scratch, size_t scratch_bytes, void* stream) {
  if (P && P->basis == RYD_BASIS_CSR)
    return fail(RYD_ERR_UNSUPPORTED, "a sparse-matrix plan has no site codes");
  if (!P || !x || !out || !scratch) return fail(RYD_ERR_INVALID, "bad argument to ryd_site_correlations");
  hipStream_t st = static_cast<hipStream_t>(stream);
  switch (dtype) {
    case RYD_F32: return do_corr<float>(P, x, out, scratch, scratch_bytes, st);
    case RYD_F64: return do_corr<double>(P, x, out, scratch, scratch_bytes, st);
    case RYD_C64: return do_corr<c32>(P, x, out, scratch, scratch_bytes, st);
    case RYD_C128: return do_corr<c64>(P, x, out, scratch, scratch_bytes, st);
    default: return fail(RYD_ERR_INVALID, "unknown dtype");
  }
}

extern "C" int ryd_sector_codes(const ryd_plan* P, const int64_t* ranks, int64_t first,
                                int64_t count, int64_t* codes, void* stream) {
  if (P && P->basis == RYD_BASIS_CSR)
    return fail(RYD_ERR_UNSUPPORTED, "a sparse-matrix plan has no site codes");
  if (!P || !codes || count < 0) return fail(RYD_ERR_INVALID, "bad argument to ryd_sector_codes");
  if (P->basis == RYD_BASIS_FULL)
    return fail(RYD_ERR_UNSUPPORTED, "codes and ranks are defined for restricted bases");
  if (count == 0) return RYD_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (P->basis == RYD_BASIS_CODES || P->basis == RYD_BASIS_GRAPH)
    k_chan_codes<<<stream_grid(count), 256, 0, st>>>(chan_model(P), first, count, ranks, codes);
  else
    k_fib_unrank<<<stream_grid(count), 256, 0, st>>>(fib_model(P), first, count, ranks, codes);
  LAUNCH_CHECK();
  return RYD_OK;
}

extern "C" int ryd_sector_rank(const ryd_plan* P, const int64_t* codes, int64_t count,
                               int64_t* ranks, void* stream) {
  if (P && P->basis == RYD_BASIS_CSR)
    return fail(RYD_ERR_UNSUPPORTED, "a sparse-matrix plan has no site codes");
  if (!P || !codes || !ranks || count < 0) return fail(RYD_ERR_INVALID, "bad argument to ryd_sector_rank");
  if (P->basis == RYD_BASIS_FULL)
    return fail(RYD_ERR_UNSUPPORTED, "codes and ranks are defined for restricted bases");
  if (count == 0) return RYD_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (P->basis == RYD_BASIS_CODES || P->basis == RYD_BASIS_GRAPH)
    k_chan_rank<<<stream_grid(count), 256, 0, st>>>(chan_model(P), codes, count, ranks);
  else
    k_fib_rank<<<stream_grid(count), 256, 0, st>>>(fib_model(P), codes, count, ranks);
  LAUNCH_CHECK();
  return RYD_OK;
}

namespace {

// Launch one fused evolution term; *grid_out receives the block count (the
// number of reduction partials when io.reduce).  The caller finishes the
// reduction.
template <class V>
int launch_evo_term(const ryd_plan* P, EvoIO<V> io, Reduce rd, hipStream_t st, unsigned* grid_out) {
  constexpr int TB = tile_bits<V>();
  unsigned grid;
  if constexpr (sizeof(V) == 8) {
    if (P->n >= EVO_WIDE_TB && P->n <= EVO_WIDE_MAXN) {
      // complex64 up to 2^21 codes: one 2048-code tile per 1024-thread
      // block, every far partner requested at once
      FullModel M = full_model(P, EVO_WIDE_TB);
      if (M.tb != EVO_WIDE_TB || !M.elo)
        return fail(RYD_ERR_DEVICE, "evolution tables: unexpected tile width");
      grid = (unsigned)(1ull << (P->n - EVO_WIDE_TB));
      k_evo_wide<V><<<dim3(grid), dim3(1024), 0, st>>>(M, io, rd);
      LAUNCH_CHECK();
      *grid_out = grid;
      return RYD_OK;
    }
  }
  if (P->n >= TB) {
    FullModel M = full_model(P, TB);
    // below ~1000 tiles one tile per block leaves most CUs idle: spread each
    // tile over tile_r blocks
    constexpr int R = tile_r<V>();
    grid = (unsigned)(1ull << (P->n - TB));
    if (grid < 1024) {
      grid *= R;
      k_evo_tiled<V, 1><<<dim3(grid), dim3(256), 0, st>>>(M, io, rd);
    } else {
      k_evo_tiled<V, R><<<dim3(grid), dim3(256), 0, st>>>(M, io, rd);
    }
  } else {
    FullModel M = full_model(P, TB);
    grid = (unsigned)(((1ull << P->n) + 255) / 256);
    k_evo_small<V><<<dim3(grid), dim3(256), 0, st>>>(M, io, rd);
  }
  LAUNCH_CHECK();
  *grid_out = grid;
  return RYD_OK;
}

template <class V>
int launch_evo(const ryd_plan* P, int dtype, EvoIO<V> io, Reduce rd, hipStream_t st) {
  (void)dtype;
  unsigned grid = 0;
  if (int rc = launch_evo_term<V>(P, io, rd, st, &grid)) return rc;
  if (io.reduce) k_reduce_finish<<<1, 256, 0, st>>>(rd, grid);
  LAUNCH_CHECK();
  return RYD_OK;
}

// Blocks of one evolution term (the reduction partials it writes).
template <class V>
unsigned evo_term_blocks(const ryd_plan* P) {
  constexpr int TB = tile_bits<V>();
  if (sizeof(V) == 8 && P->n >= EVO_WIDE_TB && P->n <= EVO_WIDE_MAXN)
    return (unsigned)(1ull << (P->n - EVO_WIDE_TB));
  if (P->n >= TB) {
    const unsigned grid = (unsigned)(1ull << (P->n - TB));
    return grid < 1024 ? grid * tile_r<V>() : grid;
  }
  return (unsigned)(((1ull << P->n) + 255) / 256);
}

template <class V>
EvoIO<V> evo_io(double om, double dadd) {
  EvoIO<V> io;
  io.x = io.z = io.w = io.q = nullptr;
  io.y = io.acc = nullptr;
  io.om = om;
  io.dadd = dadd;
  io.s = 1.0;
  io.ch = 1.0;
  io.cx = io.cz = 0.0;
  io.cw_re = io.cw_im = io.ca_re = io.ca_im = 0.0;
  io.c0_re = io.c0_im = 0.0;
  io.acc_init = 0;
  io.reduce = 0;
  return io;
}

// Forward slices with the fused kernel.  Slice k: t_0 = ψ_k, t_{j+1} =
// (1|2)·H~ t_j − t_{j−1}, ψ_{k+1} = Σ_j a_j t_j accumulated by the term
// launches themselves (the first term initialises it), so a slice is K
// launches and ψ alternates between the caller's buffer and one work vector
// (one copy at the end when the result sits in the work vector).
// tstore != null keeps t_0..t_{K−1} of slice k at tstore + k·K·D (the
// adjoint's operands) and ψ_k itself in that slot's t_0; work then needs one
// vector (t_K), else four (three t's and the ψ ping-pong).
template <class V>
int do_evolve_forward(const ryd_plan* P, int dtype, void* psi_v, int64_t nslices,
                      const double* om, const double* dadd, const double* c, const double* r,
                      int K, const double* coef, void* work_v, void* tstore_v, hipStream_t st) {
  V* psi = static_cast<V*>(psi_v);
  const int64_t D = P->dim;
  V* W = static_cast<V*>(work_v);
  V* TS = static_cast<V*>(tstore_v);
  Reduce none = make_reduce(nullptr, POST_NONE, nullptr);
  if (nslices <= 0) return RYD_OK;
  V* cur = psi;
  if (TS) {
    HIP_TRY(hipMemcpyAsync(TS, psi, sizeof(V) * D, hipMemcpyDeviceToDevice, st));
    cur = TS;
  }
  V* pong = W + 3 * D;
  for (int64_t k = 0; k < nslices; ++k) {
    const double* a = coef + (size_t)k * (K + 1) * 2;
    const double inv = 1.0 / r[k];
    V* slot = TS ? TS + (size_t)k * K * D : nullptr;
    V* nxt;
    if (TS)
      nxt = k + 1 < nslices ? TS + (size_t)(k + 1) * K * D : psi;
    else
      nxt = cur == psi ? pong : psi;
    auto tvec = [&](int j) -> V* {       // t_j, j >= 1
      if (TS) return j < K ? slot + (size_t)j * D : W;
      return W + (size_t)((j - 1) % 3) * D;
    };
    for (int j = 0; j < K; ++j) {   // produce t_{j+1}
      EvoIO<V> io = evo_io<V>(om[k], dadd[k]);
      io.x = j == 0 ? cur : tvec(j);
      io.z = j == 0 ? nullptr : (j == 1 ? cur : tvec(j - 1));
      io.y = tvec(j + 1);
      io.acc = nxt;
      io.acc_init = j == 0;
      io.c0_re = a[0];
      io.c0_im = a[1];
      io.ch = (j == 0 ? 1.0 : 2.0) * inv;
      io.cx = -(j == 0 ? 1.0 : 2.0) * c[k] * inv;
      io.cz = j == 0 ? 0.0 : -1.0;
      io.ca_re = a[2 * (j + 1)];
      io.ca_im = a[2 * (j + 1) + 1];
      if (int rc = launch_evo<V>(P, dtype, io, none, st)) return rc;
    }
    cur = nxt;
  }
  if (cur != psi) HIP_TRY(hipMemcpyAsync(psi, cur, sizeof(V) * D, hipMemcpyDeviceToDevice, st));
  return RYD_OK;
}

// Adjoint of nslices stored slices (t_0..t_{K−1} of slice k at tstore +
// k·K·D, from do_evolve_forward), last slice first: Clenshaw b_K = conj(a_K)
// g, b_j = conj(a_j) g + (2|1)·H~ b_{j+1} − b_{j+2}, b_0 → g (the cotangent
// of ψ_k), with the gradient dots of every term written as block partials
// and summed per slice by one launch into sums + k·K·4 (device memory).
// work: 3 vectors; partials: K·evo_term_blocks doubles × 4.
template <class V>
int do_evolve_adjoint_stored(const ryd_plan* P, int dtype, void* g_v, int64_t nslices,
                             const double* om, const double* dadd, const double* c,
                             const double* r, int K, const double* coef, const void* tstore_v,
                             void* work_v, double* sums, double* partials, hipStream_t st) {
  V* g = static_cast<V*>(g_v);
  const int64_t D = P->dim;
  const V* TS = static_cast<const V*>(tstore_v);
  V* W = static_cast<V*>(work_v);
  V* B[3] = {W, W + D, W + 2 * D};
  const unsigned nb = evo_term_blocks<V>(P);
  for (int64_t k = nslices - 1; k >= 0; --k) {
    const double* a = coef + (size_t)k * (K + 1) * 2;
    const double inv = 1.0 / r[k];
    const V* slot = TS + (size_t)k * K * D;
    k_cscale<V><<<stream_grid(D), 256, 0, st>>>(g, B[K % 3], D, a[2 * K], -a[2 * K + 1]);
    LAUNCH_CHECK();
    for (int j = K - 1; j >= 0; --j) {
      EvoIO<V> io = evo_io<V>(om[k], dadd[k]);
      io.x = B[(j + 1) % 3];
      io.z = j + 2 <= K ? B[(j + 2) % 3] : nullptr;
      io.w = g;
      io.y = j == 0 ? g : B[j % 3];
      io.q = slot + (size_t)j * D;
      io.ch = (j == 0 ? 1.0 : 2.0) * inv;
      io.cx = -(j == 0 ? 1.0 : 2.0) * c[k] * inv;
      io.cz = io.z ? -1.0 : 0.0;
      io.cw_re = a[2 * j];
      io.cw_im = -a[2 * j + 1];
      io.reduce = 1;
      Reduce rd = make_reduce(nullptr, POST_NONE, nullptr);
      rd.partials = partials + (size_t)j * nb * NRED;
      unsigned grid = 0;
      if (int rc = launch_evo_term<V>(P, io, rd, st, &grid)) return rc;
    }
    k_reduce_terms<<<K, 256, 0, st>>>(partials, nb, sums + (size_t)k * K * NRED);
    LAUNCH_CHECK();
  }
  (void)dtype;
  return RYD_OK;
}

// Forward slices on any plan with the plan's own H (om = 1, dadd = 0): the
// Chebyshev recurrence through the plan's H·v launches (restricted bases,
// J/A models, full bases beyond the fused kernel), ψ accumulated by a
// streaming complex axpy per term.
template <class V>
int do_evolve_generic(const ryd_plan* P, int dtype, void* psi_v, int64_t nslices,
                      const double* c, const double* r, int K, const double* coef,
                      void* work_v, hipStream_t st) {
  V* psi = static_cast<V*>(psi_v);
  const int64_t D = P->dim;
  V* W = static_cast<V*>(work_v);
  V* T[3] = {W, W + D, W + 2 * D};
  V* acc = W + 3 * D;
  Reduce none = make_reduce(nullptr, POST_NONE, nullptr);
  const unsigned sg = stream_grid(D);
  for (int64_t k = 0; k < nslices; ++k) {
    const double* a = coef + (size_t)k * (K + 1) * 2;
    const double inv = 1.0 / r[k];
    k_cscale<V><<<sg, 256, 0, st>>>(psi, acc, D, a[0], a[1]);
    LAUNCH_CHECK();
    for (int j = 0; j < K; ++j) {   // t_{j+1} = (1 or 2)·H~ t_j − t_{j−1}
      ApplyIO<V> io;
      io.x = j == 0 ? psi : T[(j - 1) % 3];
      io.z = j == 0 ? nullptr : (j == 1 ? psi : T[(j - 2) % 3]);
      io.y = T[j % 3];
      io.probe = nullptr;
      io.smode = SM_CONST;
      io.s = 1.0;
      io.ch = (j == 0 ? 1.0 : 2.0) * inv;
      io.cx = -(j == 0 ? 1.0 : 2.0) * c[k] * inv;
      io.cz = j == 0 ? 0.0 : -1.0;
      io.nrm = nullptr;
      io.j = 0;
      io.stop = nullptr;
      io.reduce = 0;
      if (int rc = launch_apply<V>(P, dtype, io, none, st)) return rc;
      k_caxpy<V><<<sg, 256, 0, st>>>(T[j % 3], acc, D, a[2 * (j + 1)], a[2 * (j + 1) + 1]);
      LAUNCH_CHECK();
    }
    HIP_TRY(hipMemcpyAsync(psi, acc, sizeof(V) * D, hipMemcpyDeviceToDevice, st));
  }
  return RYD_OK;
}

// psi <- exp(-tau·(i·H + G/2)) psi by nsub Taylor steps of degree m (tau =
// CONV·dt of the whole interval); work: 2 vectors
template <class V>
int do_taylor(const ryd_plan* P, int dtype, void* psi_v, double tau, int nsub, int m,
              const double* gamma, void* work_v, hipStream_t st) {
  V* psi = static_cast<V*>(psi_v);
  const int64_t D = P->dim;
  V* v = static_cast<V*>(work_v);
  V* w = v + D;
  TaylorDiag T;
  for (int p = 0; p < MAXN; ++p) T.gb[p] = 0.0;
  for (int i = 0; i < P->n; ++i) T.gb[P->n - 1 - i] = 0.5 * gamma[i];
  const ObsBasis B = obs_basis(P);
  Reduce none = make_reduce(nullptr, POST_NONE, nullptr);
  const unsigned sg = stream_grid(D);
  const double h = tau / nsub;
  for (int sub = 0; sub < nsub; ++sub) {
    HIP_TRY(hipMemcpyAsync(v, psi, sizeof(V) * D, hipMemcpyDeviceToDevice, st));
    for (int k = 1; k <= m; ++k) {
      ApplyIO<V> io;
      io.x = v;
      io.z = nullptr;
      io.y = w;
      io.probe = nullptr;
      io.smode = SM_CONST;
      io.s = 1.0;
      io.ch = 1.0;
      io.cx = 0.0;
      io.cz = 0.0;
      io.nrm = nullptr;
      io.j = 0;
      io.stop = nullptr;
      io.reduce = 0;
      if (int rc = launch_apply<V>(P, dtype, io, none, st)) return rc;
      k_taylor_term<V><<<sg, 256, 0, st>>>(B, T, w, v, psi, h / k, D);
      LAUNCH_CHECK();
    }
  }
  return RYD_OK;
}

// the fused slice kernels cover the full basis with Omega/delta/V up to n = 22
bool evo_fused(const ryd_plan* P) {
  return P->basis == RYD_BASIS_FULL && !P->generic && P->n < SPLIT_MIN_N;
}

template <class V>
int do_evolve_vjp(const ryd_plan* P, int dtype, const void* psi_v, void* g_v, double om,
                  double dadd, double c, double r, int K, const double* a, void* tstore_v,
                  void* work_v, double* sums, void* ws, hipStream_t st) {
  const V* psi = static_cast<const V*>(psi_v);
  V* g = static_cast<V*>(g_v);
  const int64_t D = P->dim;
  V* TS = static_cast<V*>(tstore_v);          // t_0 .. t_{K-1}
  V* W = static_cast<V*>(work_v);
  V* B[3] = {W, W + D, W + 2 * D};
  const double inv = 1.0 / r;
  Reduce none = make_reduce(nullptr, POST_NONE, nullptr);
  // recompute the forward recurrence of this slice, keeping t_0 .. t_{K-1}
  HIP_TRY(hipMemcpyAsync(TS, psi, sizeof(V) * D, hipMemcpyDeviceToDevice, st));
  for (int j = 0; j + 1 < K; ++j) {
    EvoIO<V> io = evo_io<V>(om, dadd);
    io.x = TS + (size_t)j * D;
    io.z = j == 0 ? nullptr : TS + (size_t)(j - 1) * D;
    io.y = TS + (size_t)(j + 1) * D;
    io.ch = (j == 0 ? 1.0 : 2.0) * inv;
    io.cx = -(j == 0 ? 1.0 : 2.0) * c * inv;
    io.cz = j == 0 ? 0.0 : -1.0;
    if (int rc = launch_evo<V>(P, dtype, io, none, st)) return rc;
  }
  // Clenshaw: b_K = conj(a_K) g ; b_j = conj(a_j) g + 2H~ b_{j+1} - b_{j+2} ; b_0 with H~
  k_cscale<V><<<stream_grid(D), 256, 0, st>>>(g, B[K % 3], D, a[2 * K], -a[2 * K + 1]);
  LAUNCH_CHECK();
  for (int j = K - 1; j >= 0; --j) {
    EvoIO<V> io = evo_io<V>(om, dadd);
    io.x = B[(j + 1) % 3];
    io.z = j + 2 <= K ? B[(j + 2) % 3] : nullptr;
    io.w = g;
    io.y = j == 0 ? g : B[j % 3];
    io.q = TS + (size_t)j * D;
    io.ch = (j == 0 ? 1.0 : 2.0) * inv;
    io.cx = -(j == 0 ? 1.0 : 2.0) * c * inv;
    io.cz = io.z ? -1.0 : 0.0;
    io.cw_re = a[2 * j];
    io.cw_im = -a[2 * j + 1];
    io.reduce = 1;
    Reduce rd = make_reduce(ws, POST_STORE, sums + 4 * j);
    if (int rc = launch_evo<V>(P, dtype, io, rd, st)) return rc;
  }
  return RYD_OK;
}

int check_evo(const ryd_plan* P, int dtype, int K) {
  if (P->basis != RYD_BASIS_FULL || P->generic)
    return fail(RYD_ERR_UNSUPPORTED, "evolution kernels cover the full basis with Omega/delta/V");
  if (P->n >= SPLIT_MIN_N) return fail(RYD_ERR_UNSUPPORTED, "evolution kernels need n <= 22");
  if (dtype != RYD_C64 && dtype != RYD_C128)
    return fail(RYD_ERR_INVALID, "evolution needs a complex dtype");
  if (K < 1 || K > 4096) return fail(RYD_ERR_INVALID, "K must be in [1, 4096]");
  return RYD_OK;
}

}  // namespace

extern "C" int ryd_evolve_forward(const ryd_plan* P, int dtype, void* psi, int64_t nslices,
                                  const double* om, const double* dadd, const double* c,
                                  const double* r, int K, const double* coef, void* work,
                                  void* stream) {
  if (P && P->basis == RYD_BASIS_CSR)
    return fail(RYD_ERR_UNSUPPORTED, "a sparse-matrix plan has no site codes");
  if (!P || !psi || !om || !dadd || !c || !r || !coef || !work || nslices < 0)
    return fail(RYD_ERR_INVALID, "null argument to ryd_evolve_forward");
  for (int64_t k = 0; k < nslices; ++k)
    if (!(r[k] > 0.0)) return fail(RYD_ERR_INVALID, "spectral radius r must be > 0");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (!evo_fused(P)) {
    // any other plan evolves under its own H
    if (dtype != RYD_C64 && dtype != RYD_C128)
      return fail(RYD_ERR_INVALID, "evolution needs a complex dtype");
    if (K < 1 || K > 4096) return fail(RYD_ERR_INVALID, "K must be in [1, 4096]");
    for (int64_t k = 0; k < nslices; ++k)
      if (om[k] != 1.0 || dadd[k] != 0.0)
        return fail(RYD_ERR_UNSUPPORTED, "restricted bases, pair flips and n > 22 evolve under "
                                         "the plan's own H (om = 1, dadd = 0 per slice)");
    if (dtype == RYD_C64)
      return do_evolve_generic<c32>(P, dtype, psi, nslices, c, r, K, coef, work, st);
    return do_evolve_generic<c64>(P, dtype, psi, nslices, c, r, K, coef, work, st);
  }
  if (int rc = check_evo(P, dtype, K)) return rc;
  if (dtype == RYD_C64)
    return do_evolve_forward<c32>(P, dtype, psi, nslices, om, dadd, c, r, K, coef, work, nullptr, st);
  return do_evolve_forward<c64>(P, dtype, psi, nslices, om, dadd, c, r, K, coef, work, nullptr, st);
}

extern "C" int ryd_evolve_forward_store(const ryd_plan* P, int dtype, void* psi, int64_t nslices,
                                        const double* om, const double* dadd, const double* c,
                                        const double* r, int K, const double* coef, void* tstore,
                                        void* work, void* stream) {
  if (P && P->basis == RYD_BASIS_CSR)
    return fail(RYD_ERR_UNSUPPORTED, "a sparse-matrix plan has no site codes");
  if (!P || !psi || !om || !dadd || !c || !r || !coef || !tstore || !work || nslices < 0)
    return fail(RYD_ERR_INVALID, "null argument to ryd_evolve_forward_store");
  if (int rc = check_evo(P, dtype, K)) return rc;
  for (int64_t k = 0; k < nslices; ++k)
    if (!(r[k] > 0.0)) return fail(RYD_ERR_INVALID, "spectral radius r must be > 0");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == RYD_C64)
    return do_evolve_forward<c32>(P, dtype, psi, nslices, om, dadd, c, r, K, coef, work, tstore, st);
  return do_evolve_forward<c64>(P, dtype, psi, nslices, om, dadd, c, r, K, coef, work, tstore, st);
}

extern "C" int64_t ryd_evolve_partials_bytes(const ryd_plan* P, int dtype, int K) {
  if (!P || P->basis != RYD_BASIS_FULL || P->generic || P->n >= SPLIT_MIN_N || K < 1 || K > 4096)
    return -1;
  const unsigned nb = dtype == RYD_C64 ? evo_term_blocks<c32>(P)
                                       : (dtype == RYD_C128 ? evo_term_blocks<c64>(P) : 0u);
  if (!nb) return -1;
  return (int64_t)K * nb * NRED * (int64_t)sizeof(double);
}

extern "C" int ryd_evolve_adjoint_stored(const ryd_plan* P, int dtype, void* g, int64_t nslices,
                                         const double* om, const double* dadd, const double* c,
                                         const double* r, int K, const double* coef,
                                         const void* tstore, void* work, double* sums,
                                         double* partials, void* stream) {
  if (P && P->basis == RYD_BASIS_CSR)
    return fail(RYD_ERR_UNSUPPORTED, "a sparse-matrix plan has no site codes");
  if (!P || !g || !om || !dadd || !c || !r || !coef || !tstore || !work || !sums || !partials ||
      nslices < 0)
    return fail(RYD_ERR_INVALID, "null argument to ryd_evolve_adjoint_stored");
  if (int rc = check_evo(P, dtype, K)) return rc;
  for (int64_t k = 0; k < nslices; ++k)
    if (!(r[k] > 0.0)) return fail(RYD_ERR_INVALID, "spectral radius r must be > 0");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == RYD_C64)
    return do_evolve_adjoint_stored<c32>(P, dtype, g, nslices, om, dadd, c, r, K, coef, tstore,
                                         work, sums, partials, st);
  return do_evolve_adjoint_stored<c64>(P, dtype, g, nslices, om, dadd, c, r, K, coef, tstore,
                                       work, sums, partials, st);
}

extern "C" int ryd_evolve_taylor(const ryd_plan* P, int dtype, void* psi, double tau, int nsub,
                                 int m, const double* gamma, void* work, void* stream) {
  if (P && P->basis == RYD_BASIS_CSR)
    return fail(RYD_ERR_UNSUPPORTED, "a sparse-matrix plan has no site codes");
  if (!P || !psi || !gamma || !work) return fail(RYD_ERR_INVALID, "null argument to ryd_evolve_taylor");
  if (dtype != RYD_C64 && dtype != RYD_C128)
    return fail(RYD_ERR_INVALID, "non-Hermitian evolution needs a complex dtype");
  if (!(tau >= 0.0) || !std::isfinite(tau)) return fail(RYD_ERR_INVALID, "tau must be finite and >= 0");
  if (nsub < 1 || nsub > (1 << 24) || m < 1 || m > 64)
    return fail(RYD_ERR_INVALID, "nsub must be in [1, 2^24] and m in [1, 64]");
  for (int i = 0; i < P->n; ++i)
    if (!std::isfinite(gamma[i])) return fail(RYD_ERR_INVALID, "gamma must be finite");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == RYD_C64) return do_taylor<c32>(P, dtype, psi, tau, nsub, m, gamma, work, st);
  return do_taylor<c64>(P, dtype, psi, tau, nsub, m, gamma, work, st);
}

extern "C" int ryd_loss_axpy(const ryd_plan* P, int dtype, const void* x, void* y,
                             const double* gamma, void* stream) {
  if (P && P->basis == RYD_BASIS_CSR)
    return fail(RYD_ERR_UNSUPPORTED, "a sparse-matrix plan has no site codes");
  if (!P || !x || !y || !gamma) return fail(RYD_ERR_INVALID, "null argument to ryd_loss_axpy");
  if (dtype != RYD_C64 && dtype != RYD_C128)
    return fail(RYD_ERR_INVALID, "the loss diagonal needs a complex dtype");
  for (int i = 0; i < P->n; ++i)
    if (!std::isfinite(gamma[i])) return fail(RYD_ERR_INVALID, "gamma must be finite");
  TaylorDiag T;
  for (int p = 0; p < MAXN; ++p) T.gb[p] = 0.0;
  for (int i = 0; i < P->n; ++i) T.gb[P->n - 1 - i] = 0.5 * gamma[i];
  const ObsBasis B = obs_basis(P);
  hipStream_t st = static_cast<hipStream_t>(stream);
  const unsigned sg = stream_grid(P->dim);
  if (dtype == RYD_C64)
    k_loss_axpy<c32><<<sg, 256, 0, st>>>(B, T, static_cast<const c32*>(x), static_cast<c32*>(y), P->dim);
  else
    k_loss_axpy<c64><<<sg, 256, 0, st>>>(B, T, static_cast<const c64*>(x), static_cast<c64*>(y), P->dim);
  LAUNCH_CHECK();
  return RYD_OK;
}

extern "C" int ryd_evolve_adjoint(const ryd_plan* P, int dtype, const void* psi, void* g,
                                  double om, double dadd, double c, double r, int K,
                                  const double* coef, void* tstore, void* work, double* sums,
                                  void* ws, void* stream) {
  if (P && P->basis == RYD_BASIS_CSR)
    return fail(RYD_ERR_UNSUPPORTED, "a sparse-matrix plan has no site codes");
  if (!P || !psi || !g || !coef || !tstore || !work || !sums || !ws)
    return fail(RYD_ERR_INVALID, "null argument to ryd_evolve_adjoint");
  if (int rc = check_evo(P, dtype, K)) return rc;
  if (!(r > 0.0)) return fail(RYD_ERR_INVALID, "spectral radius r must be > 0");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == RYD_C64)
    return do_evolve_vjp<c32>(P, dtype, psi, g, om, dadd, c, r, K, coef, tstore, work, sums, ws, st);
  return do_evolve_vjp<c64>(P, dtype, psi, g, om, dadd, c, r, K, coef, tstore, work, sums, ws, st);
}

// ---------------------------------------------------------------------------
// batched Lanczos
// ---------------------------------------------------------------------------
struct ryd_batch {
  int count = 0;
  int dtype = RYD_F32;
  int steps = 0;
  int device = 0;
  int64_t total = 0;                 // Σ 2^n
  std::vector<int> n;
  std::vector<int64_t> off;
  // per tile-bit candidate k (tb = 10 + k): instance table and tile map
  BatchInst* d_inst[3] = {nullptr, nullptr, nullptr};
  int* d_tile_inst[3] = {nullptr, nullptr, nullptr};
  int ntiles[3] = {0, 0, 0};
  double* d_tables = nullptr;        // all elo/ehi/vb tables
  int choice = 2;                    // chosen candidate (tb = 12 by default)
  // owned Krylov workspace (fixed addresses for graph capture): three
  // rotating rows of total elements
  void* d_vec = nullptr;
  // caller-lent memory for the whole basis (steps + 1 rows): the recurrence
  // then writes every row there, and Ritz vectors combine the rows of the
  // last run instead of replaying it (basis_valid: that run has happened)
  void* basis = nullptr;
  bool basis_valid = false;
  double* d_alpha = nullptr;         // [count][steps]
  double* d_nrm = nullptr;           // [count][steps+1]
  int* d_stop = nullptr;             // [count]
  double* d_partial = nullptr;       // [max tiles]
  size_t maxtiles = 0;
  double* d_ycoef = nullptr;         // [count][steps] Ritz coefficients (replay)
  int next_step = -1;                // step where the last ryd_batch_run_steps ended
  double* d_occp = nullptr;          // [max tiles][BMAXN+1] occupation partials
  hipStream_t stream = nullptr;      // capture / replay stream (default priority)
  hipGraphExec_t exec = nullptr;
  int exec_choice = -1;
  hipEvent_t ev = nullptr;
  // CSR instances (ryd_batch_create_csr): tiles of csr_tr rows, candidate 0
  bool csr = false;
  bool csr_wave = false;             // one row per wavefront (long rows)
  int csr_tr = 0;
  int64_t csr_maxdim = 0;            // rows of the largest instance
  int64_t* d_csr_ptr = nullptr;      // [total + 1]
  int32_t* d_csr_col = nullptr;      // [nnz], column inside the instance
  double* d_csr_re = nullptr;        // [nnz]
  double* d_csr_im = nullptr;        // [nnz] or null
  double* d_gam = nullptr;           // [count][4]: γ, <g, u_j>
  const void* lock = nullptr;        // caller's lock vectors (concatenated) or null
};

namespace {

size_t batch_es(const ryd_batch* B) {
  switch (B->dtype) {
    case RYD_F32: return 4;
    case RYD_F64: return 8;
    case RYD_C64: return 8;
    default: return 16;
  }
}

// breakdown threshold of |u_{j+1}|: single or double precision vectors
double batch_tiny(const ryd_batch* B) {
  return (B->dtype == RYD_F32 || B->dtype == RYD_C64) ? 1e-6 : 1e-13;
}

BcLz bc_lz(const ryd_batch* B, int j) {
  BcLz L;
  L.inst = B->d_inst[0];
  L.tile_inst = B->d_tile_inst[0];
  L.alpha = B->d_alpha;
  L.nrm = B->d_nrm;
  L.stop = B->d_stop;
  L.partial = B->d_partial;
  L.gam = B->d_gam;
  L.ldk = B->steps;
  L.j = j;
  L.tr = B->csr_tr;
  L.tiny = batch_tiny(B);
  return L;
}

BatchLz batch_lz(const ryd_batch* B, int k, int j) {
  BatchLz L;
  L.inst = B->d_inst[k];
  L.tile_inst = B->d_tile_inst[k];
  L.alpha = B->d_alpha;
  L.nrm = B->d_nrm;
  L.stop = B->d_stop;
  L.partial = B->d_partial;
  L.ldk = B->steps;
  L.j = j;
  L.tiny = batch_tiny(B);
  return L;
}

__global__ void k_batch_reset(int* stop, int count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) stop[i] = 0;
}

// Row 0 of the recurrence: where the start vectors go.
void* batch_rows(const ryd_batch* B) { return B->basis ? B->basis : B->d_vec; }

template <class V>
int batch_sequence(ryd_batch* B, int k, hipStream_t st, V* ritz = nullptr,
                   const double* ycoef = nullptr, int j0 = 0, int j1 = -1) {
  if (j1 < 0) j1 = B->steps;
  V* base = static_cast<V*>(batch_rows(B));
  // u_j: row j of a lent basis, else one of three rotating rows (u_{-1}
  // is never read: its coefficient is zero at j = 0)
  auto U = [&](int j) -> V* {
    const int r = B->basis ? (j < 0 ? 0 : j) : ((j % 3) + 3) % 3;
    return base + (size_t)r * B->total;
  };
  const unsigned g = (unsigned)B->ntiles[k];
  if (j0 == 0) {
    k_batch_reset<<<(B->count + 255) / 256, 256, 0, st>>>(B->d_stop, B->count);
    LAUNCH_CHECK();
    k_batch_norm0<V><<<g, 256, 0, st>>>(batch_lz(B, k, 0), U(0));
    LAUNCH_CHECK();
    k_batch_finish<<<B->count, 256, 0, st>>>(batch_lz(B, k, -1), 1);
    LAUNCH_CHECK();
  }
  for (int j = j0; j < j1; ++j) {
    BatchLz L = batch_lz(B, k, j);
    k_batch_apply<V><<<g, 256, 0, st>>>(L, U(j), U(j - 1), U(j + 1));
    LAUNCH_CHECK();
    k_batch_finish<<<B->count, 256, 0, st>>>(L, 0);
    LAUNCH_CHECK();
    k_batch_orth<V><<<g, 256, 0, st>>>(L, U(j), U(j + 1), ritz, ycoef);
    LAUNCH_CHECK();
    k_batch_finish<<<B->count, 256, 0, st>>>(L, 1);
    LAUNCH_CHECK();
  }
  return RYD_OK;
}

// The recurrence over CSR instances (one tile shape), optionally locked
// against the caller's vectors and accumulating Ritz vectors (replay).
// One launch per run (k_bc_cycle) when every instance fits one workgroup's
// LDS and rows are short; measured on the dense-stack workload (64 × 1024).
#ifndef RYD_BC_FUSED
#define RYD_BC_FUSED 1
#endif
template <class V>
bool bc_fused(const ryd_batch* B) {
  return RYD_BC_FUSED && B->csr && !B->csr_wave && B->csr_maxdim >= 1 &&
         B->csr_maxdim <= (int64_t)LZS_THREADS * bcc_maxe<V>();
}

template <class V>
int bc_sequence(ryd_batch* B, hipStream_t st, V* ritz = nullptr, const double* ycoef = nullptr) {
  V* base = static_cast<V*>(batch_rows(B));
  auto U = [&](int j) -> V* {
    const int r = B->basis ? (j < 0 ? 0 : j) : ((j % 3) + 3) % 3;
    return base + (size_t)r * B->total;
  };
  const V* lock = static_cast<const V*>(B->lock);
  const BcMat M{B->d_csr_ptr, B->d_csr_col, B->d_csr_re, B->d_csr_im};
  if (bc_fused<V>(B)) {
    k_bc_cycle<V><<<B->count, LZS_THREADS, 0, st>>>(bc_lz(B, 0), M, base, B->total,
                                                    B->basis ? 1 : 0, lock, ritz, ycoef,
                                                    B->steps);
    LAUNCH_CHECK();
    return RYD_OK;
  }
  const unsigned g = (unsigned)B->ntiles[0];
  k_batch_reset<<<(B->count + 255) / 256, 256, 0, st>>>(B->d_stop, B->count);
  LAUNCH_CHECK();
  k_bc_norm0<V><<<g, 256, 0, st>>>(bc_lz(B, 0), U(0), lock);
  LAUNCH_CHECK();
  k_bc_finish<<<B->count, 256, 0, st>>>(bc_lz(B, -1), 1, lock != nullptr);
  LAUNCH_CHECK();
  for (int j = 0; j < B->steps; ++j) {
    const BcLz L = bc_lz(B, j);
    if (B->csr_wave)
      k_bc_apply<V, true><<<g, 256, 0, st>>>(L, M, U(j), U(j - 1), U(j + 1), lock);
    else
      k_bc_apply<V, false><<<g, 256, 0, st>>>(L, M, U(j), U(j - 1), U(j + 1), lock);
    LAUNCH_CHECK();
    k_bc_finish<<<B->count, 256, 0, st>>>(L, 0, lock != nullptr);
    LAUNCH_CHECK();
    k_bc_orth<V><<<g, 256, 0, st>>>(L, U(j), U(j + 1), ritz, ycoef, lock);
    LAUNCH_CHECK();
    k_bc_finish<<<B->count, 256, 0, st>>>(L, 1, lock != nullptr);
    LAUNCH_CHECK();
  }
  return RYD_OK;
}

// one recurrence of the batch in its element type
int batch_sequence_any(ryd_batch* B, int k, hipStream_t st) {
  if (B->csr) {
    switch (B->dtype) {
      case RYD_F32: return bc_sequence<float>(B, st);
      case RYD_F64: return bc_sequence<double>(B, st);
      case RYD_C64: return bc_sequence<c32>(B, st);
      default: return bc_sequence<c64>(B, st);
    }
  }
  return B->dtype == RYD_F32 ? batch_sequence<float>(B, k, st) : batch_sequence<double>(B, k, st);
}

int batch_launch(ryd_batch* B, int k, int use_graph) {
  hipStream_t st = B->stream;
  if (!use_graph) return batch_sequence_any(B, k, st);
  if (B->exec == nullptr || B->exec_choice != k) {
    if (B->exec) { (void)hipGraphExecDestroy(B->exec); B->exec = nullptr; }
    hipGraph_t graph;
    HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    int rc = batch_sequence_any(B, k, st);
    hipError_t e = hipStreamEndCapture(st, &graph);
    if (rc) return rc;
    if (e != hipSuccess) return fail(RYD_ERR_DEVICE, std::string("graph capture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&B->exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(RYD_ERR_DEVICE, std::string("graph instantiate: ") + hipGetErrorString(e));
    B->exec_choice = k;
  }
  HIP_TRY(hipGraphLaunch(B->exec, st));
  return RYD_OK;
}

}  // namespace

extern "C" int ryd_batch_create(const ryd_model* models, int32_t count, int dtype, int32_t steps,
                                ryd_batch** out) {
  if (!models || !out || count < 1 || steps < 1)
    return fail(RYD_ERR_INVALID, "bad argument to ryd_batch_create");
  if (dtype != RYD_F32 && dtype != RYD_F64)
    return fail(RYD_ERR_INVALID, "batched Lanczos runs real FP32 or FP64 vectors");
  *out = nullptr;
  ryd_batch* B = new ryd_batch();
  B->count = count;
  B->dtype = dtype;
  B->steps = steps;
  (void)hipGetDevice(&B->device);
  auto bad = [&](int code, const std::string& msg) {
    ryd_batch_destroy(B);
    return fail(code, msg);
  };
  // host tables for every candidate tile size
  std::vector<double> tab;
  struct Loc { size_t elo, ehi, vb; };
  std::vector<Loc> loc[3];
  std::vector<BatchInst> hinst[3];
  std::vector<int> tmap[3];
  for (int b = 0; b < count; ++b) {
    const ryd_model& m = models[b];
    const int n = m.n;
    if (n < 1 || n > BMAXN || m.basis_kind != RYD_BASIS_FULL || !m.om_re || !m.delta)
      return bad(RYD_ERR_INVALID, "batch instances need the full basis with 1 <= n <= 24");
    for (int q = 0; q < m.n_pairs; ++q)
      if ((m.pair_jr && m.pair_jr[q] != 0.0) || (m.pair_ji && m.pair_ji[q] != 0.0) ||
          (m.pair_ar && m.pair_ar[q] != 0.0) || (m.pair_ai && m.pair_ai[q] != 0.0))
        return bad(RYD_ERR_UNSUPPORTED, "batch instances take Omega/delta/V only (no J/A)");
    for (int i = 0; i < n; ++i)
      if (m.om_im && m.om_im[i] != 0.0) return bad(RYD_ERR_INVALID, "batch instances need real drives");
    B->n.push_back(n);
    B->off.push_back(B->total);
    std::vector<double> dl(n), vb((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) dl[n - 1 - i] = m.delta[i];
    for (int q = 0; q < m.n_pairs; ++q) {
      const int i = m.pair_i[q], j = m.pair_j[q];
      if (i < 0 || j < 0 || i >= n || j >= n || i == j)
        return bad(RYD_ERR_INVALID, "pair indices out of range");
      vb[(size_t)(n - 1 - i) * n + (n - 1 - j)] += m.pair_v[q];
      vb[(size_t)(n - 1 - j) * n + (n - 1 - i)] += m.pair_v[q];
    }
    const size_t vbo = tab.size();
    tab.insert(tab.end(), vb.begin(), vb.end());
    for (int k = 0; k < 3; ++k) {
      const int tb = n < 10 + k ? n : 10 + k;
      std::vector<double> elo, ehi;
      build_tables(n, tb, dl, vb, elo, ehi);
      Loc L;
      L.vb = vbo;
      L.elo = tab.size();
      tab.insert(tab.end(), elo.begin(), elo.end());
      L.ehi = tab.size();
      tab.insert(tab.end(), ehi.begin(), ehi.end());
      loc[k].push_back(L);
      BatchInst I;
      I.n = n;
      I.tb = tb;
      I.ntiles = 1 << (n - tb);
      I.pad = 0;
      I.off = B->total;
      I.tile0 = (int64_t)tmap[k].size();
      for (int p = 0; p < BMAXN; ++p) I.ar[p] = 0.0;
      for (int i = 0; i < n; ++i) I.ar[n - 1 - i] = 0.5 * m.om_re[i];
      hinst[k].push_back(I);
      for (int t = 0; t < I.ntiles; ++t) tmap[k].push_back(b);
    }
    B->total += 1ll << n;
  }
  hipError_t e = hipMalloc(&B->d_tables, sizeof(double) * tab.size());
  if (e == hipSuccess) e = hipMemcpy(B->d_tables, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice);
  size_t maxtiles = 0;
  for (int k = 0; k < 3 && e == hipSuccess; ++k) {
    for (int b = 0; b < count; ++b) {
      hinst[k][b].elo = B->d_tables + loc[k][b].elo;
      hinst[k][b].ehi = B->d_tables + loc[k][b].ehi;
      hinst[k][b].vb = B->d_tables + loc[k][b].vb;
    }
    B->ntiles[k] = (int)tmap[k].size();
    if (tmap[k].size() > maxtiles) maxtiles = tmap[k].size();
    e = hipMalloc(&B->d_inst[k], sizeof(BatchInst) * count);
    if (e == hipSuccess) e = hipMemcpy(B->d_inst[k], hinst[k].data(), sizeof(BatchInst) * count, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc(&B->d_tile_inst[k], sizeof(int) * tmap[k].size());
    if (e == hipSuccess) e = hipMemcpy(B->d_tile_inst[k], tmap[k].data(), sizeof(int) * tmap[k].size(), hipMemcpyHostToDevice);
  }
  const size_t es = dtype == RYD_F32 ? 4 : 8;
  if (e == hipSuccess) e = hipMalloc(&B->d_vec, es * 3 * (size_t)B->total);
  if (e == hipSuccess) e = hipMalloc(&B->d_alpha, sizeof(double) * (size_t)count * steps);
  if (e == hipSuccess) e = hipMalloc(&B->d_nrm, sizeof(double) * (size_t)count * (steps + 1));
  if (e == hipSuccess) e = hipMalloc(&B->d_stop, sizeof(int) * count);
  if (e == hipSuccess) e = hipMalloc(&B->d_partial, sizeof(double) * maxtiles);
  B->maxtiles = maxtiles;
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&B->ev, hipEventDisableTiming);
  if (e != hipSuccess)
    return bad(e == hipErrorOutOfMemory ? RYD_ERR_MEMORY : RYD_ERR_DEVICE,
               std::string("batch allocation: ") + hipGetErrorString(e));
  *out = B;
  return RYD_OK;
}

extern "C" int ryd_batch_create_csr(int32_t count, const int64_t* dims, const int64_t* indptr,
                                    const int64_t* indices, const double* val_re,
                                    const double* val_im, int dtype, int32_t steps,
                                    ryd_batch** out) {
  if (!dims || !indptr || !out || count < 1 || steps < 1)
    return fail(RYD_ERR_INVALID, "bad argument to ryd_batch_create_csr");
  if (dtype < RYD_F32 || dtype > RYD_C128)
    return fail(RYD_ERR_INVALID, "ryd_batch_create_csr: unknown dtype");
  *out = nullptr;
  // instances, tiles and the row pointers, all checked on the host
  int64_t total = 0;
  for (int32_t b = 0; b < count; ++b) {
    if (dims[b] < 1 || dims[b] > (int64_t)INT32_MAX)
      return fail(RYD_ERR_INVALID, "csr batch: every dim must be 1..2^31-1");
    total += dims[b];
  }
  if (indptr[0] != 0) return fail(RYD_ERR_INVALID, "csr batch: indptr[0] must be 0");
  for (int64_t r = 0; r < total; ++r)
    if (indptr[r + 1] < indptr[r]) return fail(RYD_ERR_INVALID, "csr batch: indptr must not decrease");
  const int64_t nnz = indptr[total];
  if (nnz > 0 && (!indices || !val_re)) return fail(RYD_ERR_INVALID, "null argument");
  const bool cvec = dtype == RYD_C64 || dtype == RYD_C128;
  std::vector<int32_t> col((size_t)(nnz > 0 ? nnz : 1), 0);
  bool cvals = false;
  {
    int64_t r = 0;
    for (int32_t b = 0; b < count; ++b) {
      for (int64_t k = 0; k < dims[b]; ++k, ++r) {
        for (int64_t q = indptr[r]; q < indptr[r + 1]; ++q) {
          if (indices[q] < 0 || indices[q] >= dims[b])
            return fail(RYD_ERR_INVALID, "csr batch: column index outside its instance");
          if (!std::isfinite(val_re[q]) || (val_im && !std::isfinite(val_im[q])))
            return fail(RYD_ERR_INVALID, "csr batch: non-finite value");
          col[(size_t)q] = (int32_t)indices[q];
          if (val_im && val_im[q] != 0.0) cvals = true;
        }
      }
    }
  }
  if (cvals && !cvec)
    return fail(RYD_ERR_INVALID, "csr batch: complex values need a complex dtype");
  ryd_batch* B = new ryd_batch();
  B->csr = true;
  B->count = count;
  B->dtype = dtype;
  B->steps = steps;
  B->choice = 0;
  B->total = total;
  (void)hipGetDevice(&B->device);
  auto bad = [&](int code, const std::string& msg) {
    ryd_batch_destroy(B);
    return fail(code, msg);
  };
  // long rows (a dense or nearly dense matrix) run one row per wavefront
  B->csr_wave = total > 0 && nnz >= 16 * total;
  B->csr_tr = B->csr_wave ? 64 : 1024;
  std::vector<BatchInst> hinst((size_t)count);
  std::vector<int> tmap;
  int64_t off = 0;
  for (int32_t b = 0; b < count; ++b) {
    BatchInst I{};
    const int64_t d = dims[b];
    int bits = 0;
    while (bits < 62 && (1ll << bits) < d) ++bits;
    I.n = (1ll << bits) == d && bits <= BMAXN ? bits : 0;
    I.tb = 0;
    I.ntiles = (int)((d + B->csr_tr - 1) / B->csr_tr);
    I.off = off;
    I.tile0 = (int64_t)tmap.size();
    I.dim = d;
    hinst[(size_t)b] = I;
    if (d > B->csr_maxdim) B->csr_maxdim = d;
    for (int t = 0; t < I.ntiles; ++t) tmap.push_back(b);
    B->n.push_back(I.n);
    B->off.push_back(off);
    off += d;
  }
  if (tmap.size() > (size_t)INT32_MAX) return bad(RYD_ERR_INVALID, "csr batch: too many tiles");
  B->ntiles[0] = (int)tmap.size();
  B->maxtiles = tmap.size();
  hipError_t e = hipMalloc(&B->d_inst[0], sizeof(BatchInst) * count);
  if (e == hipSuccess) e = hipMemcpy(B->d_inst[0], hinst.data(), sizeof(BatchInst) * count, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMalloc(&B->d_tile_inst[0], sizeof(int) * tmap.size());
  if (e == hipSuccess) e = hipMemcpy(B->d_tile_inst[0], tmap.data(), sizeof(int) * tmap.size(), hipMemcpyHostToDevice);
  const size_t nb = (size_t)(nnz > 0 ? nnz : 1);
  if (e == hipSuccess) e = hipMalloc(&B->d_csr_ptr, sizeof(int64_t) * (size_t)(total + 1));
  if (e == hipSuccess) e = hipMemcpy(B->d_csr_ptr, indptr, sizeof(int64_t) * (size_t)(total + 1), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMalloc(&B->d_csr_col, sizeof(int32_t) * nb);
  if (e == hipSuccess) e = hipMemcpy(B->d_csr_col, col.data(), sizeof(int32_t) * nb, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMalloc(&B->d_csr_re, sizeof(double) * nb);
  if (e == hipSuccess && nnz > 0)
    e = hipMemcpy(B->d_csr_re, val_re, sizeof(double) * (size_t)nnz, hipMemcpyHostToDevice);
  if (e == hipSuccess && cvals) e = hipMalloc(&B->d_csr_im, sizeof(double) * nb);
  if (e == hipSuccess && cvals)
    e = hipMemcpy(B->d_csr_im, val_im, sizeof(double) * (size_t)nnz, hipMemcpyHostToDevice);
  const size_t es = batch_es(B);
  if (e == hipSuccess) e = hipMalloc(&B->d_vec, es * 3 * (size_t)total);
  if (e == hipSuccess) e = hipMalloc(&B->d_alpha, sizeof(double) * (size_t)count * steps);
  if (e == hipSuccess) e = hipMalloc(&B->d_nrm, sizeof(double) * (size_t)count * (steps + 1));
  if (e == hipSuccess) e = hipMalloc(&B->d_stop, sizeof(int) * count);
  if (e == hipSuccess) e = hipMalloc(&B->d_partial, sizeof(double) * 3 * B->maxtiles);
  if (e == hipSuccess) e = hipMalloc(&B->d_gam, sizeof(double) * 4 * (size_t)count);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&B->ev, hipEventDisableTiming);
  if (e != hipSuccess)
    return bad(e == hipErrorOutOfMemory ? RYD_ERR_MEMORY : RYD_ERR_DEVICE,
               std::string("csr batch allocation: ") + hipGetErrorString(e));
  *out = B;
  return RYD_OK;
}

extern "C" int ryd_batch_set_lock(ryd_batch* B, const void* lock) {
  if (!B) return fail(RYD_ERR_INVALID, "null batch");
  if (!B->csr) return fail(RYD_ERR_UNSUPPORTED, "lock vectors need a CSR batch");
  // the captured graph holds the lock pointer
  if (B->exec) {
    HIP_TRY(hipStreamSynchronize(B->stream));
    (void)hipGraphExecDestroy(B->exec);
    B->exec = nullptr;
    B->exec_choice = -1;
  }
  B->lock = lock;
  return RYD_OK;
}

extern "C" int ryd_batch_destroy(ryd_batch* B) {
  if (!B) return RYD_OK;
  if (B->exec) (void)hipGraphExecDestroy(B->exec);
  if (B->d_ycoef) (void)hipFree(B->d_ycoef);
  if (B->d_occp) (void)hipFree(B->d_occp);
  for (int k = 0; k < 3; ++k) {
    if (B->d_inst[k]) (void)hipFree(B->d_inst[k]);
    if (B->d_tile_inst[k]) (void)hipFree(B->d_tile_inst[k]);
  }
  void* ptrs[] = {B->d_tables, B->d_vec, B->d_alpha, B->d_nrm, B->d_stop, B->d_partial,
                  B->d_csr_ptr, B->d_csr_col, B->d_csr_re, B->d_csr_im, B->d_gam};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (B->ev) (void)hipEventDestroy(B->ev);
  if (B->stream) (void)hipStreamDestroy(B->stream);
  delete B;
  return RYD_OK;
}

extern "C" int64_t ryd_batch_offset(const ryd_batch* B, int32_t index) {
  if (!B || index < 0 || index > B->count) return -1;
  return index == B->count ? B->total : B->off[index];
}

extern "C" int ryd_batch_run(ryd_batch* B, const void* v0, int use_graph, void* stream) {
  if (!B || !v0) return fail(RYD_ERR_INVALID, "bad argument to ryd_batch_run");
  hipStream_t caller = static_cast<hipStream_t>(stream);
  const size_t es = batch_es(B);
  // order after the caller's stream, copy the start vectors, run, order back
  HIP_TRY(hipEventRecord(B->ev, caller));
  HIP_TRY(hipStreamWaitEvent(B->stream, B->ev, 0));
  HIP_TRY(hipMemcpyAsync(batch_rows(B), v0, es * (size_t)B->total, hipMemcpyDeviceToDevice,
                         B->stream));
  B->basis_valid = false;
  B->next_step = -1;
  if (int rc = batch_launch(B, B->choice, use_graph)) return rc;
  B->basis_valid = B->basis != nullptr;
  HIP_TRY(hipEventRecord(B->ev, B->stream));
  HIP_TRY(hipStreamWaitEvent(caller, B->ev, 0));
  return RYD_OK;
}

extern "C" int ryd_batch_run_steps(ryd_batch* B, const void* v0, int32_t j0, int32_t j1,
                                   void* stream) {
  if (!B || (j0 == 0 && !v0) || j0 < 0 || j1 <= j0 || j1 > B->steps)
    return fail(RYD_ERR_INVALID, "bad argument to ryd_batch_run_steps");
  if (B->csr) return fail(RYD_ERR_UNSUPPORTED, "steps in parts run on full-basis batches");
  if (j0 > 0 && B->next_step != j0)
    return fail(RYD_ERR_INVALID, "ryd_batch_run_steps continues at the step where the last call ended");
  hipStream_t caller = static_cast<hipStream_t>(stream);
  HIP_TRY(hipEventRecord(B->ev, caller));
  HIP_TRY(hipStreamWaitEvent(B->stream, B->ev, 0));
  if (j0 == 0) {
    HIP_TRY(hipMemcpyAsync(batch_rows(B), v0, batch_es(B) * (size_t)B->total,
                           hipMemcpyDeviceToDevice, B->stream));
    B->basis_valid = false;
  }
  B->next_step = -1;
  const int rc = B->dtype == RYD_F32
                     ? batch_sequence<float>(B, B->choice, B->stream, nullptr, nullptr, j0, j1)
                     : batch_sequence<double>(B, B->choice, B->stream, nullptr, nullptr, j0, j1);
  if (rc) return rc;
  B->next_step = j1;
  B->basis_valid = j1 == B->steps && B->basis != nullptr;
  HIP_TRY(hipEventRecord(B->ev, B->stream));
  HIP_TRY(hipStreamWaitEvent(caller, B->ev, 0));
  return RYD_OK;
}

extern "C" int ryd_batch_stop(ryd_batch* B, const int32_t* flags) {
  if (!B || !flags) return fail(RYD_ERR_INVALID, "bad argument to ryd_batch_stop");
  std::vector<int> st((size_t)B->count);
  HIP_TRY(hipMemcpyAsync(st.data(), B->d_stop, sizeof(int) * st.size(), hipMemcpyDeviceToHost,
                         B->stream));
  HIP_TRY(hipStreamSynchronize(B->stream));
  bool all = true;
  for (int b = 0; b < B->count; ++b) {
    if (flags[b]) st[b] = 1;
    all = all && st[b] != 0;
  }
  HIP_TRY(hipMemcpyAsync(B->d_stop, st.data(), sizeof(int) * st.size(), hipMemcpyHostToDevice,
                         B->stream));
  HIP_TRY(hipStreamSynchronize(B->stream));
  // every instance stopped: the rows each one ran are final (Ritz vectors
  // combine them; coefficients past an instance's stop are zero, unread)
  if (all && B->next_step > 0) B->basis_valid = B->basis != nullptr;
  return RYD_OK;
}

extern "C" int ryd_batch_autotune(ryd_batch* B, const void* v0, int32_t* chosen_tile_bits,
                                  double* ms_out /* [3] or NULL */) {
  if (!B || !v0) return fail(RYD_ERR_INVALID, "bad argument to ryd_batch_autotune");
  if (B->csr) return fail(RYD_ERR_UNSUPPORTED, "a CSR batch has one tile shape");
  const size_t es = batch_es(B);
  hipEvent_t a, b;
  HIP_TRY(hipEventCreate(&a));
  HIP_TRY(hipEventCreate(&b));
  double best = 1e300;
  int bestk = 2;
  for (int k = 0; k < 3; ++k) {
    float ms = 0.f;
    for (int rep = 0; rep < 2; ++rep) {   // first pass warms the kernels
      HIP_TRY(hipMemcpyAsync(batch_rows(B), v0, es * (size_t)B->total, hipMemcpyDeviceToDevice,
                             B->stream));
      HIP_TRY(hipEventRecord(a, B->stream));
      if (int rc = batch_launch(B, k, 0)) return rc;
      HIP_TRY(hipEventRecord(b, B->stream));
      HIP_TRY(hipEventSynchronize(b));
      HIP_TRY(hipEventElapsedTime(&ms, a, b));
    }
    if (ms_out) ms_out[k] = ms;
    if (ms < best) { best = ms; bestk = k; }
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  B->choice = bestk;
  if (chosen_tile_bits) *chosen_tile_bits = 10 + bestk;
  return RYD_OK;
}

extern "C" int ryd_ritz_lowest(int32_t count, const int32_t* steps, const double* alpha, int64_t lda,
                               const double* offd, int64_t ldo, const double* beta_last,
                               double conv_tol, double merge_rel, double* E, double* bound, double* y,
                               int64_t ldy) {
  if (count < 0 || (count > 0 && (!steps || !alpha || !offd || !beta_last || !E || !bound)))
    return fail(RYD_ERR_INVALID, "ryd_ritz_lowest: null argument");
  if (!(merge_rel >= 0.0)) return fail(RYD_ERR_INVALID, "ryd_ritz_lowest: merge_rel must be >= 0");
  for (int32_t b = 0; b < count; ++b) {
    const int32_t k = steps[b];
    if (k < 0 || k > lda || (k > 1 && k - 1 > ldo) || (y && k > ldy))
      return fail(RYD_ERR_INVALID, "ryd_ritz_lowest: steps exceed the leading dimensions");
  }
  // instances are independent: split them over host threads
  auto run = [&](int32_t b0, int32_t b1) {
    tri::Work w;
    for (int32_t b = b0; b < b1; ++b) {
      double* y0 = y ? y + (size_t)2 * b * ldy : nullptr;
      double* y1 = y ? y0 + ldy : nullptr;
      tri::lowest_pairs(alpha + (size_t)b * lda, offd + (size_t)b * ldo, steps[b], beta_last[b],
                        conv_tol, merge_rel, E + 2 * (size_t)b, bound + 2 * (size_t)b, y0, y1, w);
    }
  };
  const int hw = (int)std::thread::hardware_concurrency();
  const int nt = std::max(1, std::min({hw > 0 ? hw : 1, 16, (int)(count / 8)}));
  if (nt <= 1) {
    run(0, count);
    return RYD_OK;
  }
  std::vector<std::thread> pool;
  for (int t = 0; t < nt; ++t) {
    const int32_t b0 = (int32_t)((int64_t)count * t / nt), b1 = (int32_t)((int64_t)count * (t + 1) / nt);
    pool.emplace_back(run, b0, b1);
  }
  for (auto& th : pool) th.join();
  return RYD_OK;
}

extern "C" int64_t ryd_batch_basis_bytes(const ryd_batch* B) {
  if (!B) return -1;
  return (int64_t)batch_es(B) * (int64_t)(B->steps + 1) * B->total;
}

extern "C" int ryd_batch_attach_basis(ryd_batch* B, void* basis, int64_t bytes) {
  if (!B) return fail(RYD_ERR_INVALID, "null batch");
  if (basis && bytes < ryd_batch_basis_bytes(B))
    return fail(RYD_ERR_INVALID, "ryd_batch_attach_basis: the basis needs ryd_batch_basis_bytes");
  if (basis && (reinterpret_cast<uintptr_t>(basis) & 15))
    return fail(RYD_ERR_INVALID, "ryd_batch_attach_basis: the basis must be 16-byte aligned");
  // the captured graph holds the old row addresses
  if (B->exec) {
    HIP_TRY(hipStreamSynchronize(B->stream));
    (void)hipGraphExecDestroy(B->exec);
    B->exec = nullptr;
    B->exec_choice = -1;
  }
  B->basis = basis;
  B->basis_valid = false;
  return RYD_OK;
}

extern "C" int ryd_batch_set_tile_bits(ryd_batch* B, int32_t tile_bits) {
  if (!B || tile_bits < 10 || tile_bits > 12)
    return fail(RYD_ERR_INVALID, "ryd_batch_set_tile_bits: tile bits must be 10, 11 or 12");
  if (B->csr) return fail(RYD_ERR_UNSUPPORTED, "a CSR batch has one tile shape");
  B->choice = tile_bits - 10;
  return RYD_OK;
}

// Ritz vectors by combining the lent basis rows of the last run (the
// coefficient table of k_batch_ritz_stored holds 2048 rows) or by replay.
bool batch_combines(const ryd_batch* B) {
  return B->basis && B->basis_valid && B->steps <= RYD_BATCH_COMBINE_MAX;
}

template <class V>
int bc_ritz(ryd_batch* B, V* ritz, double* occ) {
  hipStream_t st = B->stream;
  const BcLz L = bc_lz(B, 0);
  const unsigned g = (unsigned)B->ntiles[0];
  if (batch_combines(B)) {
    k_bc_ritz_stored<V><<<g, 256, 0, st>>>(L, static_cast<const V*>(B->basis), B->total, B->steps,
                                           B->d_ycoef, ritz);
    LAUNCH_CHECK();
  } else {
    k_batch_zero<V><<<stream_grid(B->total), 256, 0, st>>>(ritz, B->total);
    LAUNCH_CHECK();
    if (int rc = bc_sequence<V>(B, st, ritz, B->d_ycoef)) return rc;
  }
  k_bc_occ<V><<<g, 256, 0, st>>>(L, ritz, B->d_occp);
  LAUNCH_CHECK();
  k_batch_occ_finish<<<B->count, 64, 0, st>>>(batch_lz(B, 0, 0), B->d_occp, occ);
  LAUNCH_CHECK();
  return RYD_OK;
}

template <class V>
int batch_ritz(ryd_batch* B, V* ritz, double* occ) {
  if (B->csr) return bc_ritz<V>(B, ritz, occ);
  hipStream_t st = B->stream;
  const int k = B->choice;
  if (batch_combines(B)) {
    // the rows of the last run are still there: one streaming combination
    const BatchLz L = batch_lz(B, k, 0);
    k_batch_ritz_stored<V><<<(unsigned)B->ntiles[k], 256, 0, st>>>(
        L, static_cast<const V*>(B->basis), B->total, B->steps, B->d_ycoef, ritz);
    LAUNCH_CHECK();
  } else {
    k_batch_zero<V><<<stream_grid(B->total), 256, 0, st>>>(ritz, B->total);
    LAUNCH_CHECK();
    if (int rc = batch_sequence<V>(B, k, st, ritz, B->d_ycoef)) return rc;
  }
  const BatchLz L = batch_lz(B, k, 0);
  k_batch_occ<V><<<(unsigned)B->ntiles[k], 256, 0, st>>>(L, ritz, B->d_occp);
  LAUNCH_CHECK();
  k_batch_occ_finish<<<B->count, 64, 0, st>>>(L, B->d_occp, occ);
  LAUNCH_CHECK();
  return RYD_OK;
}

extern "C" int ryd_batch_ritz(ryd_batch* B, const void* v0, const double* ycoef, void* ritz,
                              double* occ, void* stream) {
  if (!B || !v0 || !ycoef || !ritz || !occ) return fail(RYD_ERR_INVALID, "bad argument to ryd_batch_ritz");
  hipStream_t caller = static_cast<hipStream_t>(stream);
  const size_t es = batch_es(B);
  if (!B->d_ycoef) HIP_TRY(hipMalloc(&B->d_ycoef, sizeof(double) * (size_t)B->count * B->steps));
  if (!B->d_occp) HIP_TRY(hipMalloc(&B->d_occp, sizeof(double) * B->maxtiles * (BMAXN + 1)));
  HIP_TRY(hipEventRecord(B->ev, caller));
  HIP_TRY(hipStreamWaitEvent(B->stream, B->ev, 0));
  HIP_TRY(hipMemcpyAsync(B->d_ycoef, ycoef, sizeof(double) * (size_t)B->count * B->steps,
                         hipMemcpyHostToDevice, B->stream));
  if (!batch_combines(B))   // a replay starts again from v0
    HIP_TRY(hipMemcpyAsync(batch_rows(B), v0, es * (size_t)B->total, hipMemcpyDeviceToDevice,
                           B->stream));
  int rc;
  switch (B->dtype) {
    case RYD_F32: rc = batch_ritz<float>(B, static_cast<float*>(ritz), occ); break;
    case RYD_F64: rc = batch_ritz<double>(B, static_cast<double*>(ritz), occ); break;
    case RYD_C64: rc = bc_ritz<c32>(B, static_cast<c32*>(ritz), occ); break;
    default: rc = bc_ritz<c64>(B, static_cast<c64*>(ritz), occ); break;
  }
  if (rc) return rc;
  HIP_TRY(hipEventRecord(B->ev, B->stream));
  HIP_TRY(hipStreamWaitEvent(caller, B->ev, 0));
  // the host coefficients must stay valid until the copy has run
  HIP_TRY(hipStreamSynchronize(B->stream));
  return RYD_OK;
}

extern "C" int ryd_batch_results(const ryd_batch* B, double* alpha, double* nrm, int32_t* stop) {
  if (!B) return fail(RYD_ERR_INVALID, "null batch");
  HIP_TRY(hipStreamSynchronize(B->stream));
  if (alpha) HIP_TRY(hipMemcpy(alpha, B->d_alpha, sizeof(double) * (size_t)B->count * B->steps, hipMemcpyDeviceToHost));
  if (nrm) HIP_TRY(hipMemcpy(nrm, B->d_nrm, sizeof(double) * (size_t)B->count * (B->steps + 1), hipMemcpyDeviceToHost));
  if (stop) HIP_TRY(hipMemcpy(stop, B->d_stop, sizeof(int) * B->count, hipMemcpyDeviceToHost));
  return RYD_OK;
}
