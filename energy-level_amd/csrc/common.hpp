// Shared device helpers for the Rydberg H·v library (gfx950 / CDNA4).
//
// Element types of state vectors: float, double, c32 (complex64), c64
// (complex128, interleaved re/im).  Reductions are always accumulated in
// FP64.  Each block stores its partial sums in the caller's workspace and a
// one-block finishing launch adds them in block order (no atomics, no
// device-scope fences), so every reduction is deterministic for a fixed
// launch shape.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ryd {

struct c32 { float re, im; };
struct c64 { double re, im; };

template <class V> struct VT;
template <> struct VT<float>  { using R = float;  static constexpr bool cplx = false; };
template <> struct VT<double> { using R = double; static constexpr bool cplx = false; };
template <> struct VT<c32>    { using R = float;  static constexpr bool cplx = true; };
template <> struct VT<c64>    { using R = double; static constexpr bool cplx = true; };

__device__ __forceinline__ float  vzero(float)  { return 0.f; }
__device__ __forceinline__ double vzero(double) { return 0.0; }
__device__ __forceinline__ c32    vzero(c32)    { return {0.f, 0.f}; }
__device__ __forceinline__ c64    vzero(c64)    { return {0.0, 0.0}; }

// x * s (real s)
__device__ __forceinline__ float  vscale(float x, float s)   { return x * s; }
__device__ __forceinline__ double vscale(double x, double s) { return x * s; }
__device__ __forceinline__ c32    vscale(c32 x, float s)     { return {x.re * s, x.im * s}; }
__device__ __forceinline__ c64    vscale(c64 x, double s)    { return {x.re * s, x.im * s}; }

// acc += a * x (real a)
__device__ __forceinline__ void vfma(float& acc, float a, float x)    { acc = fmaf(a, x, acc); }
__device__ __forceinline__ void vfma(double& acc, double a, double x) { acc = fma(a, x, acc); }
__device__ __forceinline__ void vfma(c32& acc, float a, c32 x)  { acc.re = fmaf(a, x.re, acc.re); acc.im = fmaf(a, x.im, acc.im); }
__device__ __forceinline__ void vfma(c64& acc, double a, c64 x) { acc.re = fma(a, x.re, acc.re); acc.im = fma(a, x.im, acc.im); }

// acc += (ar + i ai) * x ; real element types ignore ai (the host refuses
// complex amplitudes on real vectors)
__device__ __forceinline__ void vcfma(float& acc, float ar, float, float x)     { acc = fmaf(ar, x, acc); }
__device__ __forceinline__ void vcfma(double& acc, double ar, double, double x) { acc = fma(ar, x, acc); }
__device__ __forceinline__ void vcfma(c32& acc, float ar, float ai, c32 x) {
  acc.re = fmaf(ar, x.re, fmaf(-ai, x.im, acc.re));
  acc.im = fmaf(ar, x.im, fmaf(ai, x.re, acc.im));
}
__device__ __forceinline__ void vcfma(c64& acc, double ar, double ai, c64 x) {
  acc.re = fma(ar, x.re, fma(-ai, x.im, acc.re));
  acc.im = fma(ar, x.im, fma(ai, x.re, acc.im));
}

// Re(conj(a) b) and |a|^2 in FP64
__device__ __forceinline__ double rdot(float a, float b)   { return (double)a * (double)b; }
__device__ __forceinline__ double rdot(double a, double b) { return a * b; }
__device__ __forceinline__ double rdot(c32 a, c32 b) { return (double)a.re * b.re + (double)a.im * b.im; }
__device__ __forceinline__ double rdot(c64 a, c64 b) { return a.re * b.re + a.im * b.im; }
template <class V> __device__ __forceinline__ double nrm2(V a) { return rdot(a, a); }

// probe weight (±1) times element, real part only: Re(z * a)
__device__ __forceinline__ double zdot(double z, float a)  { return z * (double)a; }
__device__ __forceinline__ double zdot(double z, double a) { return z * a; }
__device__ __forceinline__ double zdot(double z, c32 a)    { return z * (double)a.re; }
__device__ __forceinline__ double zdot(double z, c64 a)    { return z * a.re; }

template <class V> __device__ __forceinline__ V vfrom_real(double v);
template <> __device__ __forceinline__ float  vfrom_real<float>(double v)  { return (float)v; }
template <> __device__ __forceinline__ double vfrom_real<double>(double v) { return v; }
template <> __device__ __forceinline__ c32    vfrom_real<c32>(double v)    { return {(float)v, 0.f}; }
template <> __device__ __forceinline__ c64    vfrom_real<c64>(double v)    { return {v, 0.0}; }

template <class V> struct alignas(4 * sizeof(V)) V4 { V v[4]; };

// Streaming (non-temporal) stores for outputs that are not read again soon:
// they keep the partner data of a tiled H·v in L2.
__device__ __forceinline__ void nt_store(float* p, float v)   { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ void nt_store(double* p, double v) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ void nt_store(c32* p, c32 v) {
  __builtin_nontemporal_store(v.re, &p->re);
  __builtin_nontemporal_store(v.im, &p->im);
}
__device__ __forceinline__ void nt_store(c64* p, c64 v) {
  __builtin_nontemporal_store(v.re, &p->re);
  __builtin_nontemporal_store(v.im, &p->im);
}

__device__ __forceinline__ float  nt_load(const float* p)  { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ double nt_load(const double* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ c32 nt_load(const c32* p) {
  return {__builtin_nontemporal_load(&p->re), __builtin_nontemporal_load(&p->im)};
}
__device__ __forceinline__ c64 nt_load(const c64* p) {
  return {__builtin_nontemporal_load(&p->re), __builtin_nontemporal_load(&p->im)};
}

// 16-byte vector types for non-temporal stores
typedef float  nt_f4 __attribute__((ext_vector_type(4)));
typedef double nt_d2 __attribute__((ext_vector_type(2)));

// One 16-byte non-temporal store of a 16-byte-aligned group of elements.
template <class G>
__device__ __forceinline__ void nt_store16(void* p, const G& g) {
  static_assert(sizeof(G) == 16, "16-byte groups");
  const nt_f4 v = __builtin_bit_cast(nt_f4, g);
  __builtin_nontemporal_store(v, reinterpret_cast<nt_f4*>(p));
}

// Raw buffer loads over [p, p + bytes): the hardware range check returns zero
// for an offset at or past the end, so a partner range shorter than the
// block's own needs neither a clamp nor a branch.  (Buffer descriptor word 3
// of the gfx9 family: 32-bit data format, no swizzle.)
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}
template <class V> __device__ __forceinline__ V buf_load(__amdgpu_buffer_rsrc_t r, uint32_t off);
template <> __device__ __forceinline__ float buf_load<float>(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
template <> __device__ __forceinline__ double buf_load<double>(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
template <> __device__ __forceinline__ c32 buf_load<c32>(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  return __builtin_bit_cast(c32, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
template <> __device__ __forceinline__ c64 buf_load<c64>(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  return __builtin_bit_cast(c64, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
}

// the same with a wave-uniform byte offset (soffset) added to the lane's
template <class V> __device__ __forceinline__ V buf_load_so(__amdgpu_buffer_rsrc_t r, uint32_t off,
                                                            uint32_t soff);
template <> __device__ __forceinline__ float buf_load_so<float>(__amdgpu_buffer_rsrc_t r, uint32_t off,
                                                                uint32_t soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, soff, 0));
}
template <> __device__ __forceinline__ double buf_load_so<double>(__amdgpu_buffer_rsrc_t r, uint32_t off,
                                                                  uint32_t soff) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, soff, 0));
}
template <> __device__ __forceinline__ c32 buf_load_so<c32>(__amdgpu_buffer_rsrc_t r, uint32_t off,
                                                            uint32_t soff) {
  return __builtin_bit_cast(c32, __builtin_amdgcn_raw_buffer_load_b64(r, off, soff, 0));
}
template <> __device__ __forceinline__ c64 buf_load_so<c64>(__amdgpu_buffer_rsrc_t r, uint32_t off,
                                                            uint32_t soff) {
  return __builtin_bit_cast(c64, __builtin_amdgcn_raw_buffer_load_b128(r, off, soff, 0));
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Coherent (L1-bypassing) loads for data written by other blocks.
__device__ __forceinline__ double load_coherent(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int load_coherent(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Reduction bookkeeping in the caller's workspace:
//   ws: [header : 256 B][partials : nblocks * NRED doubles]
constexpr int NRED = 4;
constexpr size_t WS_HEADER = 256;

// Post-operations applied to the final sums by the finishing launch.
enum PostOp : int {
  POST_NONE = 0,      // sums discarded
  POST_STORE = 1,     // out[k] = sum_k
  POST_ACCUM = 2,     // out[k] += sum_k
  POST_LZ_ALPHA = 3,  // Lanczos: alpha[j] = sum_0
  POST_LZ_NORM = 4,   // Lanczos: nrm[j+1] = sqrt(sum_1); breakdown check
  POST_LZ_LOCK = 5,   // locked Lanczos: γ = (sum_0, sum_2) to out[0..1],
                      // nrm[j+1] = sqrt(sum_1 − |γ|^2); breakdown check
};

// Device-side Lanczos state (caller-owned device memory).
struct LzState {
  int stop;          // set when the recurrence broke down
  int steps;         // number of valid steps when stopped
  int pad[2];
  double tiny;       // absolute breakdown threshold on |u_{j+1}|
};

struct Reduce {
  double* partials;      // per-block partial sums (NRED each)
  int post;              // PostOp
  double* out;           // POST_STORE / POST_ACCUM target (NRED doubles)
  double* alpha;         // Lanczos alpha array
  double* nrm;           // Lanczos norms array (nrm[j] = |u_j|)
  LzState* st;
  int j;                 // Lanczos step
};

// Block-level FP64 reduction of NRED values into partials[blockIdx.x][NRED]
// with plain stores.  Every thread of the block calls it (blockDim.x a
// multiple of 64, at most 1024).  The sums over blocks are formed by
// k_reduce_finish in the next launch on the same stream: the kernel boundary
// makes the partials visible on every XCD, so no device-scope fence (which
// writes back L2 on this multi-XCD part) is needed, and the fixed-order sum
// keeps every reduction deterministic.
__device__ inline void block_reduce_store(double v0, double v1, double v2, double v3,
                                          double* partials) {
  __shared__ double red_lds[16 * NRED];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  v0 = wave_sum(v0); v1 = wave_sum(v1); v2 = wave_sum(v2); v3 = wave_sum(v3);
  if (lane == 0) {
    red_lds[wid * NRED + 0] = v0; red_lds[wid * NRED + 1] = v1;
    red_lds[wid * NRED + 2] = v2; red_lds[wid * NRED + 3] = v3;
  }
  __syncthreads();
  if (threadIdx.x < NRED) {
    const int k = threadIdx.x;
    const int nw = blockDim.x >> 6;
    double s = 0.0;
    for (int w = 0; w < nw; ++w) s += red_lds[w * NRED + k];
    partials[(size_t)blockIdx.x * NRED + k] = s;
  }
}

// One block: fixed-order sum of nblocks partials and the post-operation.
__global__ __launch_bounds__(256) void k_reduce_finish(Reduce rd, unsigned nblocks) {
  __shared__ double red_lds[4 * NRED];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (rd.st && rd.post != POST_LZ_NORM && rd.post != POST_LZ_LOCK && rd.st->stop) return;
  double s[NRED] = {0.0, 0.0, 0.0, 0.0};
  for (unsigned b = threadIdx.x; b < nblocks; b += blockDim.x) {
#pragma unroll
    for (int k = 0; k < NRED; ++k) s[k] += rd.partials[(size_t)b * NRED + k];
  }
#pragma unroll
  for (int k = 0; k < NRED; ++k) s[k] = wave_sum(s[k]);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NRED; ++k) red_lds[wid * NRED + k] = s[k];
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double t[NRED];
  for (int k = 0; k < NRED; ++k) t[k] = red_lds[k] + red_lds[NRED + k] + red_lds[2 * NRED + k] +
                                        red_lds[3 * NRED + k];
  switch (rd.post) {
    case POST_STORE:
      for (int k = 0; k < NRED; ++k) rd.out[k] = t[k];
      break;
    case POST_ACCUM:
      for (int k = 0; k < NRED; ++k) rd.out[k] += t[k];
      break;
    case POST_LZ_ALPHA:
      rd.alpha[rd.j] = t[0];
      break;
    case POST_LZ_NORM: {
      if (rd.st->stop) break;
      const double nv = sqrt(t[1] > 0.0 ? t[1] : 0.0);
      rd.nrm[rd.j + 1] = nv;
      if (!(nv > rd.st->tiny)) {   // also catches NaN
        rd.st->stop = 1;
        rd.st->steps = rd.j + 1;
      }
      break;
    }
    case POST_LZ_LOCK: {
      if (rd.st->stop) break;
      rd.out[0] = t[0];
      rd.out[1] = t[2];
      const double n2 = t[1] - (t[0] * t[0] + t[2] * t[2]);
      const double nv = sqrt(n2 > 0.0 ? n2 : 0.0);
      rd.nrm[rd.j + 1] = nv;
      if (!(nv > rd.st->tiny)) {
        rd.st->stop = 1;
        rd.st->steps = rd.j + 1;
      }
      break;
    }
    default:
      break;
  }
}

}  // namespace ryd
