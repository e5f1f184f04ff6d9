// Register-resident Stockham passes with LDS transposes.
//
// A transform of M points is spread over T = M / R threads, R points per
// thread.  A pass of radix r over sub-transforms of current size p gives each
// thread R / r butterflies; butterfly b (= t + T*u) reads points
// b + j*(M/r), multiplies point j by exp(SIGN 2 pi i j k / (p r)) with
// k = b mod p, runs an r-point DFT in registers and writes point q to
// (b - k) r + k + q p.  After the last pass (p r = M) the result is in
// natural order, so no bit reversal is ever needed.
//
// Twiddle tables (built on the host in double, stored as f32) hold, for a
// pass (r, p), the entries exp(SIGN 2 pi i j k / (p r)) at [(j - 1) p + k],
// j = 1..r-1, k = 0..p-1, so that lanes with consecutive k read consecutive
// addresses.
#pragma once
#include "fft_device.h"

namespace spk {

// LDS addresses are padded (one slot per 32 points, lds_pad).  Every access
// pattern below is split into a per-butterfly base and a compile-time
// offset, padded separately: lds_pad(base + off) == lds_pad(base) +
// pad_off(off) holds whenever (base mod 32) + (off mod 32) < 32, which is
// true for all radix plans in use (checked on the host by the plan tests).
// The split lets the compiler use immediate DS offsets instead of keeping one
// address register per point.
__host__ __device__ constexpr int pad_off(int o) { return o + (o >> 5); }

// Logical index of input point j of butterfly b = t + T u in pass (r, .).
template <int M, int R, int T, int r>
__device__ __forceinline__ int st_in_index(int t, int e) {
  const int u = e / r, j = e % r;
  return t + T * u + j * (M / r);
}

// Padded address of the same point.
template <int M, int R, int T, int r>
__device__ __forceinline__ int st_in_addr(int t, int e) {
  const int u = e / r, j = e % r;
  const int b = t + T * u;
  return lds_pad(b) + pad_off(j * (M / r));
}

// Padded address of output point q of butterfly b in pass (r, p).
template <int M, int R, int T, int r, int p>
__device__ __forceinline__ int st_out_addr(int t, int e) {
  const int u = e / r, q = e % r;
  const int b = t + T * u;
  const int k = b & (p - 1);
  return lds_pad((b - k) * r + k) + pad_off(q * p);
}

// Multiply v[j] (j >= 1) by the pass-table twiddle of (j, k).  For r >= 8
// only the rows j = 1, 2, 3 and j = 4a are loaded and the rest are formed as
// products (j = 4a + b): 10 instead of 31 loads in flight for r = 32, which
// keeps the 1024-thread kernels inside 128 registers.  Each twiddle then
// carries at most two f32 roundings.
template <int r, int p>
__device__ __forceinline__ void apply_pass_twiddles(cpx (&v)[r], const cpx* __restrict__ tw, int k) {
  if constexpr (r < 8) {
#pragma unroll
    for (int j = 1; j < r; ++j) v[j] = cmul(v[j], tw[(j - 1) * p + k]);
  } else {
    cpx lo[4], hi[r / 4];
#pragma unroll
    for (int b = 1; b < 4; ++b) lo[b] = tw[(b - 1) * p + k];
#pragma unroll
    for (int a = 1; a < r / 4; ++a) hi[a] = tw[(4 * a - 1) * p + k];
#pragma unroll
    for (int j = 1; j < r; ++j) {
      const int a = j >> 2, b = j & 3;
      const cpx w = (b == 0) ? hi[a] : (a == 0 ? lo[b] : cmul(lo[b], hi[a]));
      v[j] = cmul(v[j], w);
    }
  }
}

// The rows of a pass table that apply_pass_twiddles reads (all r - 1 for
// r < 8; j = 1, 2, 3 and 4a otherwise), stored one after the other: a
// compact copy, for tables kept in LDS.  compact_row(s) is the table row
// (j - 1) of compact slot s.
template <int r>
__host__ __device__ constexpr int compact_rows() {
  return r < 8 ? r - 1 : r / 4 + 2;
}
template <int r>
__host__ __device__ constexpr int compact_row(int s) {
  return (r < 8 || s < 3) ? s : 4 * (s - 2) - 1;
}

// apply_pass_twiddles over a compact copy of the table (same products).
template <int r, int p>
__device__ __forceinline__ void apply_compact_twiddles(cpx (&v)[r], const cpx* tw, int k) {
  if constexpr (r < 8) {
#pragma unroll
    for (int j = 1; j < r; ++j) v[j] = cmul(v[j], tw[(j - 1) * p + k]);
  } else {
    cpx lo[4], hi[r / 4];
#pragma unroll
    for (int b = 1; b < 4; ++b) lo[b] = tw[(b - 1) * p + k];
#pragma unroll
    for (int a = 1; a < r / 4; ++a) hi[a] = tw[(2 + a) * p + k];
#pragma unroll
    for (int j = 1; j < r; ++j) {
      const int a = j >> 2, b = j & 3;
      const cpx w = (b == 0) ? hi[a] : (a == 0 ? lo[b] : cmul(lo[b], hi[a]));
      v[j] = cmul(v[j], w);
    }
  }
}

// The table twiddles of one butterfly of pass (r, p) (r >= 8), loaded ahead
// of their use: rows j = 1, 2, 3 and 4a as in apply_pass_twiddles.
template <int r>
struct PassTwiddles {
  cpx lo[4], hi[r / 4];
};
template <int r, int p>
__device__ __forceinline__ PassTwiddles<r> load_pass_twiddles(const cpx* __restrict__ tw, int k) {
  static_assert(r >= 8, "table rows 1, 2, 3 and 4a");
  PassTwiddles<r> w;
#pragma unroll
  for (int b = 1; b < 4; ++b) w.lo[b] = tw[(b - 1) * p + k];
#pragma unroll
  for (int a = 1; a < r / 4; ++a) w.hi[a] = tw[(4 * a - 1) * p + k];
  return w;
}
template <int r>
__device__ __forceinline__ void apply_loaded_twiddles(cpx (&v)[r], const PassTwiddles<r>& w) {
#pragma unroll
  for (int j = 1; j < r; ++j) {
    const int a = j >> 2, b = j & 3;
    const cpx t = (b == 0) ? w.hi[a] : (a == 0 ? w.lo[b] : cmul(w.lo[b], w.hi[a]));
    v[j] = cmul(v[j], t);
  }
}

// Twiddle then DFT for every butterfly a thread owns in pass (r, p).
template <int M, int R, int T, int r, int p, int SIGN, bool CONJ_HI = false>
__device__ __forceinline__ void st_butterflies(cpx (&x)[R], int t, const cpx* __restrict__ tw) {
#pragma unroll
  for (int u = 0; u < R / r; ++u) {
    const int b = t + T * u;
    const int k = b & (p - 1);
    cpx v[r];
#pragma unroll
    for (int j = 0; j < r; ++j) v[j] = x[u * r + j];
    if constexpr (p > 1) apply_pass_twiddles<r, p>(v, tw, k);
    dft<r, SIGN, CONJ_HI>(v);
#pragma unroll
    for (int q = 0; q < r; ++q) x[u * r + q] = v[q];
  }
}

// Move the R points of every thread through LDS: point e is written at the
// padded address waddr(e) and point e is then read back from raddr(e).
// SPLIT moves the real and imaginary planes one after the other so that
// transforms twice the LDS size fit.  Starts with a barrier so earlier LDS
// readers are done.
// (mid(): called once the real plane is in LDS and the real parts are dead
// in registers: a place to issue loads for the next pass.)
struct NoMid {
  __device__ void operator()() const {}
};
template <int R, bool SPLIT, typename WA, typename RA, typename MID = NoMid>
__device__ __forceinline__ void lds_exchange(cpx (&x)[R], float* lds, WA waddr, RA raddr, MID mid = MID{}) {
  if constexpr (SPLIT) {
    __syncthreads();
#pragma unroll
    for (int e = 0; e < R; ++e) lds[waddr(e)] = x[e].x;
    mid();
    __syncthreads();
#pragma unroll
    for (int e = 0; e < R; ++e) x[e].x = lds[raddr(e)];
    __syncthreads();
#pragma unroll
    for (int e = 0; e < R; ++e) lds[waddr(e)] = x[e].y;
    __syncthreads();
#pragma unroll
    for (int e = 0; e < R; ++e) x[e].y = lds[raddr(e)];
  } else {
    cpx* l = reinterpret_cast<cpx*>(lds);
    __syncthreads();
#pragma unroll
    for (int e = 0; e < R; ++e) l[waddr(e)] = x[e];
    mid();
    __syncthreads();
#pragma unroll
    for (int e = 0; e < R; ++e) x[e] = l[raddr(e)];
  }
}

// Wave-level ordering point for LDS traffic that stays inside one wave: a
// wave's LDS instructions are executed in order, so only the compiler has
// to be kept from moving accesses across it (no s_barrier is issued).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// lds_exchange for a transform whose T threads lie in one wave (T divides
// 64) and whose LDS region no other wave touches between workgroup
// barriers: the exchange then needs no workgroup barrier.
template <int R, typename WA, typename RA>
__device__ __forceinline__ void lds_exchange_wave(cpx (&x)[R], float* lds, WA waddr, RA raddr) {
  cpx* l = reinterpret_cast<cpx*>(lds);
  wave_sync();
#pragma unroll
  for (int e = 0; e < R; ++e) l[waddr(e)] = x[e];
  wave_sync();
#pragma unroll
  for (int e = 0; e < R; ++e) x[e] = l[raddr(e)];
}

// Workgroup-wide exchange, or the wave-local form when every tile of T
// threads lies inside one wave.
template <int T, int R, typename WA, typename RA>
__device__ __forceinline__ void tile_exchange(cpx (&x)[R], float* lds, WA waddr, RA raddr) {
  if constexpr (T <= 64) lds_exchange_wave<R>(x, lds, waddr, raddr);
  else lds_exchange<R, false>(x, lds, waddr, raddr);
}

}  // namespace spk
