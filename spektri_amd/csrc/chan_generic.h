// Channel stage for IFFT sizes that are multiples of 4 without a power-of-two
// plan (the reference accepts any multiple of 4, fcfb.rs:225-248, and plans
// rustfft for it, fcfb.rs:144-150).  Same arithmetic contract as the
// power-of-two kernels (fcfb.rs:154-188):
//   B[i] = W[j] get_bin(first + j), j = (i + L/2) mod L,
//   y = unnormalised inverse DFT_L(B), output y[L/8 .. (L/8)*7)
// with L/8 rounded down as in the reference's integer slice, so a channel
// gives 6 (L/8) samples per block.
//
// Mixed-radix Stockham transform over the prime factors of L (2s paired into
// 4s), one stage per factor, ping-ponging between two LDS buffers of the
// tile.  Stage (p, m) computes every output o = m (c p + q) + k (k < m,
// q < p) directly as
//   out[o] = sum_j in[c m + k + j L/p] * w_L^(j (k + q m) L/(m p)),
// which merges the inter-stage twiddle and the p-point DFT kernel into one
// table entry per term (w_L = exp(2 pi i / L), a table of L f32 values
// made in double on the host).  A stage costs p complex multiply-adds per
// output; the sum over stages is the sum of L's prime factors.
#pragma once
#include "kernels.h"

namespace spk {

struct GenChanParams {
  const cpx* spec;
  long long spec_stride;
  int N, K, real_input;
  const float* weights;
  const ChanDesc* chans;
  int nch, nblocks;
  cpx* out;            // channel c's block g at out + out_off + g * 6 (L/8)
  const cpx* tw;       // w_L^e, e = 0..L-1 (inverse sign)
  int L;
  int nstages;
  int radix[kGenMaxStages];
  int tile_threads;    // threads per (channel, block) tile; divides the workgroup
};

// Dynamic LDS: tiles_per_workgroup * 2 * L cpx.
__global__ __launch_bounds__(kGenThreads) void chan_generic_kernel(GenChanParams P) {
  extern __shared__ __align__(16) float gen_lds[];
  const int T = P.tile_threads;
  const int L = P.L;
  const int tl = threadIdx.x / T, t = threadIdx.x % T;
  const int tpw = kGenThreads / T;
  cpx* a = reinterpret_cast<cpx*>(gen_lds) + (long long)tl * 2 * L;

  const long long ntiles = (long long)P.nch * P.nblocks;
  long long tile = (long long)blockIdx.x * tpw + tl;
  const bool active = tile < ntiles;
  if (!active) tile = 0;  // computes a valid tile, stores nothing (barriers stay uniform)
  const int c = (int)(tile % P.nch);
  const int g = (int)(tile / P.nch);
  const ChanDesc d = P.chans[c];
  gen_chan_tile(P.spec, P.spec_stride, P.N, P.K, P.real_input, P.weights, d, g, active, a, a + L, t, T, L,
                P.nstages, P.radix, P.tw, P.out);
}

// Wide channels (kGenMaxL < L <= kGenBigMaxL, any multiple of 4, and the
// power-of-two 16384): one (channel, block) tile per workgroup of
// kGenBigThreads threads over one static LDS buffer of L values.  Each
// stage computes every thread's outputs (at most kGenBigMaxL /
// kGenBigThreads) into registers from the buffer, waits, and writes them
// back in place; the arithmetic is chan_generic_kernel's.
constexpr int kGenBigMaxL = 20480;  // 160 KiB of LDS
constexpr int kGenBigThreads = 1024;

__global__ __launch_bounds__(kGenBigThreads) void chan_generic_big_kernel(GenChanParams P) {
  __shared__ cpx buf[kGenBigMaxL];
  constexpr int U = kGenBigMaxL / kGenBigThreads;
  const int L = P.L;
  const int t = (int)threadIdx.x;
  const long long tile = blockIdx.x;  // grid = nch * nblocks exactly
  const int c = (int)(tile % P.nch);
  const int g = (int)(tile / P.nch);
  const ChanDesc d = P.chans[c];
  const cpx* X = P.spec + (long long)g * P.spec_stride;
  const float* W = P.weights + d.w_off;
  for (int i = t; i < L; i += kGenBigThreads) {
    int jj = i + L / 2;
    if (jj >= L) jj -= L;
    int m = d.f_mod + jj;
    if (m >= P.N) m -= P.N;
    const bool mirror = P.real_input && m >= P.K;
    cpx v = X[mirror ? P.N - m : m];
    if (mirror) v.y = -v.y;
    buf[i] = cscale(v, W[jj]);
  }
  int m = 1;
  for (int s = 0; s < P.nstages; ++s) {
    __syncthreads();
    const int p = P.radix[s];
    const int Lp = L / p;
    const int mp = m * p;
    const int scale = L / mp;
    cpx val[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int o = t + kGenBigThreads * u;
      if (o >= L) break;
      const int k = o % m;
      const int r = o / m;
      const int q = r % p;
      const int base = (r / p) * m + k;
      const int step = (k + q * m) * scale;
      cpx acc = buf[base];
      int e = step;
      for (int j = 1; j < p; ++j) {
        const cpx v = buf[base + j * Lp];
        const cpx w = P.tw[e];
        acc.x += v.x * w.x - v.y * w.y;
        acc.y += v.x * w.y + v.y * w.x;
        e += step;
        if (e >= L) e -= L;
      }
      val[u] = acc;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int o = t + kGenBigThreads * u;
      if (o >= L) break;
      buf[o] = val[u];
    }
    m = mp;
  }
  __syncthreads();
  const int lo = L / 8, n = 6 * (L / 8);
  cpx* out = P.out + d.out_off + (long long)g * n;
  for (int i = t; i < n; i += kGenBigThreads) out[i] = buf[lo + i];
}

// Adjoint of the generic channel stage (the same contract as
// chan_adjoint_kernel): for output gradient ybar (packed layout), x = DFT_L
// (forward sign) of ybar zero-padded to L at positions L/8 .. L/8 + 6 (L/8),
// then per bin i (j = (i + L/2) mod L): wbar[j] += Re(conj(G[j]) x[i]) over
// blocks and xbar[j] = W[j] x[i].  Tile (channel c, split s) takes blocks
// s, s + S, ...; thread t keeps the weight sums of bins t + T u in
// registers (at most kGenMaxL / kGenThreads of them: T = 256 above L = 256)
// and writes them to wpart[c][s][j] (the splits are added in order by
// adjoint_reduce_kernel).  Dynamic LDS: tiles_per_workgroup * 2 L cpx.
struct GenAdjParams {
  const cpx* spec;
  long long spec_stride;
  int N, K, real_input;
  const float* weights;
  const ChanDesc* chans;
  const long long* pk_off;
  int nch, nblocks, S;
  const cpx* ybar;
  cpx* xbar;           // may be null
  float* wpart;        // [nch][S][L]
  const cpx* tw;       // w_L^e, e = 0..L-1 (inverse sign; conjugated here)
  int L;
  int nstages;
  int radix[kGenMaxStages];
  int tile_threads;
};

__global__ __launch_bounds__(kGenThreads) void chan_generic_adjoint_kernel(GenAdjParams P) {
  extern __shared__ __align__(16) float gen_lds[];
  const int T = P.tile_threads;
  const int L = P.L;
  const int tl = threadIdx.x / T, t = threadIdx.x % T;
  const int tpw = kGenThreads / T;
  cpx* a0 = reinterpret_cast<cpx*>(gen_lds) + (long long)tl * 2 * L;
  cpx* b0 = a0 + L;
  constexpr int U = kGenMaxL / kGenThreads;
  float wacc[U];

  long long tile = (long long)blockIdx.x * tpw + tl;
  const bool active = tile < (long long)P.nch * P.S;
  if (!active) tile = 0;  // computes a valid tile, stores nothing (barriers stay uniform)
  const int c = (int)(tile / P.S), split = (int)(tile % P.S);
  const ChanDesc d = P.chans[c];
  const float* W = P.weights + d.w_off;
  const int lo = L / 8, n_out = 6 * (L / 8);
#pragma unroll
  for (int u = 0; u < U; ++u) wacc[u] = 0.f;

  const int iters = (P.nblocks + P.S - 1) / P.S;
  for (int it = 0; it < iters; ++it) {  // uniform trip count: the stages hold barriers
    const int g = it * P.S + split;
    const bool act = active && g < P.nblocks;
    const int gg = g < P.nblocks ? g : 0;
    const cpx* yb = P.ybar + (long long)P.nblocks * P.pk_off[c] + (long long)gg * n_out;
    cpx* a = a0;
    cpx* b = b0;
    __syncthreads();
    for (int i = t; i < L; i += T) a[i] = (i >= lo && i < lo + n_out) ? yb[i - lo] : cpx{0.f, 0.f};
    int m = 1;
    for (int s = 0; s < P.nstages; ++s) {
      __syncthreads();
      const int p = P.radix[s];
      const int Lp = L / p;
      const int mp = m * p;
      const int scale = L / mp;
      for (int o = t; o < L; o += T) {
        const int k = o % m;
        const int r = o / m;
        const int q = r % p;
        const int base = (r / p) * m + k;
        const int step = (k + q * m) * scale;
        cpx acc = a[base];
        int e = step;
        for (int j = 1; j < p; ++j) {
          const cpx v = a[base + j * Lp];
          const cpx w = P.tw[e];  // conj(w) below: forward sign
          acc.x += v.x * w.x + v.y * w.y;
          acc.y += v.y * w.x - v.x * w.y;
          e += step;
          if (e >= L) e -= L;
        }
        b[o] = acc;
      }
      cpx* tmp = a;
      a = b;
      b = tmp;
      m = mp;
    }
    __syncthreads();
    const cpx* X = P.spec + (long long)gg * P.spec_stride;
    cpx* xb = P.xbar ? P.xbar + (long long)P.nblocks * d.w_off + (long long)gg * L : nullptr;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = t + T * u;
      if (i >= L) break;
      int j = i + L / 2;
      if (j >= L) j -= L;
      int mm = d.f_mod + j;
      if (mm >= P.N) mm -= P.N;
      const bool mirror = P.real_input && mm >= P.K;
      cpx G = X[mirror ? P.N - mm : mm];
      if (mirror) G.y = -G.y;
      if (act) {
        wacc[u] += G.x * a[i].x + G.y * a[i].y;
        if (xb) xb[j] = cscale(a[i], W[j]);
      }
    }
  }
  if (active) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = t + T * u;
      if (i >= L) break;
      int j = i + L / 2;
      if (j >= L) j -= L;
      P.wpart[((long long)c * P.S + split) * L + j] = wacc[u];
    }
  }
}

// Adjoint of the wide channels (chan_generic_big_kernel's sizes): one
// (channel, block split) tile per workgroup of kGenBigThreads threads over
// one static LDS buffer; each stage computes a thread's outputs into
// registers, waits and writes them back (forward sign: conjugated
// twiddles).  Thread t keeps the weight sums of bins t + kGenBigThreads u.
__global__ __launch_bounds__(kGenBigThreads) void chan_generic_big_adjoint_kernel(GenAdjParams P) {
  __shared__ cpx buf[kGenBigMaxL];
  constexpr int U = kGenBigMaxL / kGenBigThreads;
  const int L = P.L;
  const int t = (int)threadIdx.x;
  const int c = (int)blockIdx.x / P.S, split = (int)blockIdx.x % P.S;  // grid = nch * S exactly
  const ChanDesc d = P.chans[c];
  const float* W = P.weights + d.w_off;
  const int lo = L / 8, n_out = 6 * (L / 8);
  float wacc[U];
#pragma unroll
  for (int u = 0; u < U; ++u) wacc[u] = 0.f;
  for (int g = split; g < P.nblocks; g += P.S) {  // the same trip count for the whole workgroup
    const cpx* yb = P.ybar + (long long)P.nblocks * P.pk_off[c] + (long long)g * n_out;
    __syncthreads();
    for (int i = t; i < L; i += kGenBigThreads) buf[i] = (i >= lo && i < lo + n_out) ? yb[i - lo] : cpx{0.f, 0.f};
    int m = 1;
    for (int s = 0; s < P.nstages; ++s) {
      __syncthreads();
      const int p = P.radix[s];
      const int Lp = L / p;
      const int mp = m * p;
      const int scale = L / mp;
      cpx val[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int o = t + kGenBigThreads * u;
        if (o >= L) break;
        const int k = o % m;
        const int r = o / m;
        const int q = r % p;
        const int base = (r / p) * m + k;
        const int step = (k + q * m) * scale;
        cpx acc = buf[base];
        int e = step;
        for (int j = 1; j < p; ++j) {
          const cpx v = buf[base + j * Lp];
          const cpx w = P.tw[e];  // conj(w) below: forward sign
          acc.x += v.x * w.x + v.y * w.y;
          acc.y += v.y * w.x - v.x * w.y;
          e += step;
          if (e >= L) e -= L;
        }
        val[u] = acc;
      }
      __syncthreads();
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int o = t + kGenBigThreads * u;
        if (o >= L) break;
        buf[o] = val[u];
      }
      m = mp;
    }
    __syncthreads();
    const cpx* X = P.spec + (long long)g * P.spec_stride;
    cpx* xb = P.xbar ? P.xbar + (long long)P.nblocks * d.w_off + (long long)g * L : nullptr;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = t + kGenBigThreads * u;
      if (i >= L) break;
      int j = i + L / 2;
      if (j >= L) j -= L;
      int mm = d.f_mod + j;
      if (mm >= P.N) mm -= P.N;
      const bool mirror = P.real_input && mm >= P.K;
      cpx G = X[mirror ? P.N - mm : mm];
      if (mirror) G.y = -G.y;
      wacc[u] += G.x * buf[i].x + G.y * buf[i].y;
      if (xb) xb[j] = cscale(buf[i], W[j]);
    }
  }
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int i = t + kGenBigThreads * u;
    if (i >= L) break;
    int j = i + L / 2;
    if (j >= L) j -= L;
    P.wpart[((long long)c * P.S + split) * L + j] = wacc[u];
  }
}

}  // namespace spk
