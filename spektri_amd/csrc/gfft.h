// Forward transform for FFT sizes that are not powers of two (rustfft plans
// any N: multifft.rs:26-28,45).  Mixed-radix Stockham passes through global
// memory over the Nc-point complex transform of every block (Nc = N for
// complex input, N/2 packed sample pairs for real input, which the untangle
// kernel of cplx.h then turns into the real spectrum).  The first pass reads
// the framed, windowed input (history of the last O samples, then this
// call's samples) straight from the caller's buffer; the last pass writes
// the spectrum rows in natural order.
//
// Pass with radix R over Ns = product of the radices before it: butterfly j
// (0 <= j < S = Nc / R) with k = j mod Ns reads v_q = in[j + q S], applies
// the twiddle W_{Ns R}^{q k}, takes the R-point DFT and writes
// out[(j / Ns) Ns R + k + p Ns] = sum_q v_q W_R^{p q}.  All twiddles come
// from one table of W_Nc^m, m = 0..Nc-1 (W_{Ns R}^{q k} = W_Nc^{q k Nc / (Ns R)},
// W_R^m = W_Nc^{m Nc / R}), computed in double on the host.
#pragma once
#include <hip/hip_runtime.h>

#include "stockham.h"

namespace spk {

constexpr int kGfftThreads = 256;

struct GfftPassParams {
  const cpx* in;           // previous pass's rows (first pass: unused)
  long long in_stride;
  const void* src;         // first pass: this call's samples (pairs)
  const void* hist;        // first pass: the previous call's last Oc pairs
  int Hc, Oc;              // hop and overlap in pairs
  float wscale;            // window s / N (dsp.rs:70-73)
  cpx* out;
  long long out_stride;
  const cpx* tw;           // W_Nc^m
  int Nc, Ns;
  int nblocks;
};

// LOAD: 0 = rows of the previous pass, 1 = s16 pairs, 2 = f32 pairs.
template <int LOAD>
__device__ __forceinline__ cpx gfft_load(const GfftPassParams& P, int b, int n) {
  if constexpr (LOAD == 0) {
    return P.in[(long long)b * P.in_stride + n];
  } else {
    const long long g = (long long)b * P.Hc - P.Oc + n;  // main.rs:173-181 framing, in pairs
    const bool from_hist = g < 0;
    const long long idx = from_hist ? g + P.Oc : g;
    if constexpr (LOAD == 1) {
      const int w = (from_hist ? static_cast<const int*>(P.hist) : static_cast<const int*>(P.src))[idx];
      return cpx{(float)(short)(w & 0xffff) * P.wscale, (float)(short)(w >> 16) * P.wscale};
    } else {
      const cpx v = (from_hist ? static_cast<const cpx*>(P.hist) : static_cast<const cpx*>(P.src))[idx];
      return cpx{v.x * P.wscale, v.y * P.wscale};
    }
  }
}

template <int R, int LOAD>
__global__ __launch_bounds__(kGfftThreads) void gfft_pass_kernel(GfftPassParams P) {
  const int S = P.Nc / R;
  const int j = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int b = (int)blockIdx.y;
  if (j >= S || b >= P.nblocks) return;
  const int k = j % P.Ns;
  cpx v[R];
#pragma unroll
  for (int q = 0; q < R; ++q) v[q] = gfft_load<LOAD>(P, b, j + q * S);
  if (P.Ns > 1) {
    // q k Nc / (Ns R) < Nc for q < R, k < Ns
    const int step = k * (P.Nc / (P.Ns * R));
#pragma unroll
    for (int q = 1; q < R; ++q) v[q] = cmul(v[q], P.tw[q * step]);
  }
  cpx w[R];
#pragma unroll
  for (int m = 1; m < R; ++m) w[m] = P.tw[m * S];
  cpx* out = P.out + (long long)b * P.out_stride + (long long)(j / P.Ns) * P.Ns * R + k;
#pragma unroll
  for (int p = 0; p < R; ++p) {
    cpx y = v[0];
#pragma unroll
    for (int q = 1; q < R; ++q) {
      const int m = (p * q) % R;
      if (m == 0) {
        y.x += v[q].x;
        y.y += v[q].y;
      } else {
        y.x += v[q].x * w[m].x - v[q].y * w[m].y;
        y.y += v[q].x * w[m].y + v[q].y * w[m].x;
      }
    }
    out[(long long)p * P.Ns] = y;
  }
}

}  // namespace spk
