// Forward transform for FFT sizes outside the power-of-two plans (rustfft
// plans any N: multifft.rs:26-28,45), in two forms: the four-step form
// below (gfft4_kernel) when Nc splits into two lines of at most 480
// points, otherwise mixed-radix Stockham passes through global memory,
// one per radix, over the Nc-point complex transform of every block (Nc = N for
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
  int Hc, Oc;              // hop and overlap in pairs (in samples when samples != 0)
  int samples;             // real input framed in samples: 1 = a pair is two samples
                           // read one by one (hop or overlap odd), 2 = one sample per
                           // point, imaginary part 0 (odd N)
  float wscale;            // window s / N (dsp.rs:70-73)
  cpx* out;
  long long out_stride;
  const cpx* tw;           // W_Nc^m
  int Nc, Ns;
  int nblocks;
};

// LOAD: 0 = rows of the previous pass, 1 = s16 pairs, 2 = f32 pairs.
template <int LOAD>
__device__ __forceinline__ float gfft_sample(const GfftPassParams& P, long long s) {
  const bool from_hist = s < 0;
  const long long idx = from_hist ? s + P.Oc : s;
  if constexpr (LOAD == 1)
    return (float)(from_hist ? static_cast<const short*>(P.hist) : static_cast<const short*>(P.src))[idx];
  else
    return (from_hist ? static_cast<const float*>(P.hist) : static_cast<const float*>(P.src))[idx];
}
template <int LOAD>
__device__ __forceinline__ cpx gfft_load(const GfftPassParams& P, int b, int n) {
  if constexpr (LOAD == 0) {
    return P.in[(long long)b * P.in_stride + n];
  } else {
    if (P.samples) {  // main.rs:173-181 framing in samples; a pair may straddle the history
      const long long s = (long long)b * P.Hc - P.Oc + (P.samples == 2 ? n : 2LL * n);
      if (P.samples == 2) return cpx{gfft_sample<LOAD>(P, s) * P.wscale, 0.f};
      return cpx{gfft_sample<LOAD>(P, s) * P.wscale, gfft_sample<LOAD>(P, s + 1) * P.wscale};
    }
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

// Four-step form, Nc = N1 N2 with N1, N2 <= kG4MaxLine: two launches
// instead of one per radix.  Input pair n = N2 n1 + n2, output bin
// k = k1 + N1 k2:
//   A (columns): Y[k1 N2 + n2] = W_Nc^(n2 k1) sum_n1 x[N2 n1 + n2] W_N1^(n1 k1)
//   B (rows):    X[k1 + N1 k2] = sum_n2 Y[k1 N2 + n2] W_N2^(n2 k2)
// A workgroup takes LINES adjacent lines (columns n2 in A, rows k1 in B),
// so every global read and write is a run of LINES consecutive values or
// a whole row: 8 lines of up to 480 points, or 16 of up to 247 (A, where
// the runs of s16 pairs are short).  Each line's DFT runs in LDS over the
// radices of its length (forward sign), ping-ponging between two LDS
// buffers: radices 2, 3, 4, 5, 7 and 8 as butterflies with their R-point
// DFT unrolled (g4_stage), larger primes as one table sum per output
// (g4_stage_any).
constexpr int kG4Lines = 8;
constexpr int kG4MaxLine = 480;  // two buffers of 8 padded lines and the twiddles within 64 KiB of dynamic LDS
constexpr int kG4LinesWide = 16;
constexpr int kG4MaxLineWide = 247;  // the same for 16 lines
constexpr int g4_max_line(int lines) { return lines == kG4LinesWide ? kG4MaxLineWide : kG4MaxLine; }
constexpr int kG4Threads = 256;
constexpr int kG4MaxStages = 12;

struct Gfft4Params {
  const void* src;         // A: this call's samples (pairs)
  const void* hist;        // A: the previous call's last Oc pairs
  int Hc, Oc;
  int samples;             // as GfftPassParams::samples
  float wscale;
  const cpx* Y;            // B: column results
  cpx* out;                // A: Y rows (stride Nc); B: spectrum rows
  long long out_stride;
  const cpx* tw_line;      // w_n^e, e = 0..n-1, forward sign (n = line length)
  const cpx* tw_full;      // A: W_Nc^m, m = 0..Nc-1
  int Nc, N1, N2;
  int n;                   // line length: N1 in A, N2 in B
  int nlines;              // N2 in A, N1 in B
  int nstages;
  int radix[kG4MaxStages];
  int nblocks;
};

// a / d for 0 <= a < 2^12, 1 <= d <= 2^8 through the float reciprocal
// (butterfly numbers stay below LINES g4_max_line(LINES) / 2 <= 1976,
// line positions below 240, sub-transform sizes m at most 240):
// (a + 1/2) / d is at least 1/(2 d) >= 2^-9 from the nearest integer and
// the two roundings move it by at most 2^-10 (relative 2^-23 times
// a / d < 2^12), so truncation gives the integer quotient.
__device__ __forceinline__ int g4_div(int a, float inv_d) { return (int)(((float)a + 0.5f) * inv_d); }

// One stage of radix R (compile time) over the LINES lines, from LDS
// buffer `in` to `out`: butterfly (l, c, k), k < m, reads
// v_j = in_l[c m + k + j n/R], applies w_n^(j k n/(m R)) and the R-point
// DFT (roots w_R^(jq) = w_n^(jq n/R)), and writes out_l[m (c R + q) + k].
template <int R, int LINES>
__device__ __forceinline__ void g4_stage(const cpx* in, cpx* out, int LS, int n, int m, const cpx* tw, int t) {
  const int nbf = n / R;
  const int tot = LINES * nbf;
  const int s = n / (m * R);
  const float inv_nbf = 1.0f / (float)nbf, inv_m = 1.0f / (float)m;
  cpx w[R];
#pragma unroll
  for (int q = 1; q < R; ++q) w[q] = tw[q * nbf];
  for (int f = t; f < tot; f += kG4Threads) {
    const int l = g4_div(f, inv_nbf), bidx = f - l * nbf;
    const int c = g4_div(bidx, inv_m), k = bidx - c * m;
    const cpx* a = in + l * LS + c * m + k;
    cpx v[R];
#pragma unroll
    for (int j = 0; j < R; ++j) v[j] = a[j * nbf];
    if (m > 1) {
#pragma unroll
      for (int j = 1; j < R; ++j) v[j] = cmul(v[j], tw[j * k * s]);  // j k s < n
    }
    cpx* o = out + l * LS + m * (c * R) + k;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      cpx acc = v[0];
#pragma unroll
      for (int j = 1; j < R; ++j) {
        const int e = (j * q) % R;
        if (e == 0) {
          acc.x += v[j].x;
          acc.y += v[j].y;
        } else {
          acc.x += v[j].x * w[e].x - v[j].y * w[e].y;
          acc.y += v[j].x * w[e].y + v[j].y * w[e].x;
        }
      }
      o[q * m] = acc;
    }
  }
}

// Radices without a butterfly of their own: every output as one sum.
template <int LINES>
__device__ __forceinline__ void g4_stage_any(const cpx* in, cpx* out, int LS, int n, int m, int p, const cpx* tw,
                                             int t) {
  const int total = LINES * n;
  const int np = n / p;
  const int mp = m * p;
  const int scale = n / mp;
  for (int f = t; f < total; f += kG4Threads) {
    const int l = f / n, o = f - l * n;
    const cpx* a = in + l * LS;
    const int k = o % m;
    const int r = o / m;
    const int q = r % p;
    const int base = (r / p) * m + k;
    const int step = (k + q * m) * scale;
    cpx acc = a[base];
    int e = step;
    for (int j = 1; j < p; ++j) {
      const cpx x = a[base + j * np];
      const cpx w = tw[e];
      acc.x += x.x * w.x - x.y * w.y;
      acc.y += x.x * w.y + x.y * w.x;
      e += step;
      if (e >= n) e -= n;
    }
    out[l * LS + o] = acc;
  }
}

template <int LOAD, bool STEP_A, int LINES>
__global__ __launch_bounds__(kG4Threads) void gfft4_kernel(Gfft4Params P) {
  extern __shared__ __align__(16) float g4_lds[];  // 2 LINES (n + 1) + n cpx
  const int n = P.n;
  const int LS = n + 1;  // padded line stride
  cpx* bufs[2] = {reinterpret_cast<cpx*>(g4_lds), reinterpret_cast<cpx*>(g4_lds) + LINES * LS};
  cpx* buf = bufs[0];
  cpx* tw = bufs[1] + LINES * LS;  // the line's w_n^e, read by every butterfly
  for (int e = (int)threadIdx.x; e < n; e += kG4Threads) tw[e] = P.tw_line[e];
  const int t = (int)threadIdx.x;
  const int c0 = (int)blockIdx.x * LINES;
  const int b = (int)blockIdx.y;
  const int total = LINES * n;

  // load: A reads x[N2 n1 + n2] with n2 = c0 + l fastest; B reads row k1 = c0 + l.
  // Every load of the thread is issued before the first LDS store.
  constexpr int UL = (LINES * g4_max_line(LINES) + kG4Threads - 1) / kG4Threads;
  GfftPassParams Q{};
  Q.src = P.src;
  Q.hist = P.hist;
  Q.Hc = P.Hc;
  Q.Oc = P.Oc;
  Q.samples = P.samples;
  Q.wscale = P.wscale;
  cpx vals[UL];
#pragma unroll
  for (int u = 0; u < UL; ++u) {
    const int f = t + kG4Threads * u;
    vals[u] = cpx{0.f, 0.f};
    if (f < total) {
      const int l = STEP_A ? f % LINES : f / n;
      const int i = STEP_A ? f / LINES : f % n;
      const int line = c0 + l;
      if (line < P.nlines) {
        if constexpr (STEP_A)
          vals[u] = gfft_load<LOAD>(Q, b, P.N2 * i + line);
        else
          vals[u] = P.Y[(long long)b * P.Nc + (long long)line * P.N2 + i];
      }
    }
  }
#pragma unroll
  for (int u = 0; u < UL; ++u) {
    const int f = t + kG4Threads * u;
    if (f < total) {
      const int l = STEP_A ? f % LINES : f / n;
      const int i = STEP_A ? f / LINES : f % n;
      buf[l * LS + i] = vals[u];
    }
  }

  int m = 1;
  for (int s = 0; s < P.nstages; ++s) {
    __syncthreads();
    const int p = P.radix[s];
    cpx* nxt = bufs[(s + 1) & 1];
    switch (p) {
      case 2: g4_stage<2, LINES>(buf, nxt, LS, n, m, tw, t); break;
      case 3: g4_stage<3, LINES>(buf, nxt, LS, n, m, tw, t); break;
      case 4: g4_stage<4, LINES>(buf, nxt, LS, n, m, tw, t); break;
      case 5: g4_stage<5, LINES>(buf, nxt, LS, n, m, tw, t); break;
      case 7: g4_stage<7, LINES>(buf, nxt, LS, n, m, tw, t); break;
      case 8: g4_stage<8, LINES>(buf, nxt, LS, n, m, tw, t); break;
      default: g4_stage_any<LINES>(buf, nxt, LS, n, m, p, tw, t); break;
    }
    buf = nxt;
    m *= p;
  }
  __syncthreads();

  // store: A writes Y[k1 N2 + n2] W_Nc^(n2 k1); B writes X[k1 + N1 k2] (line = k1)
  cpx* out = P.out + (long long)b * P.out_stride;
  for (int f = t; f < total; f += kG4Threads) {
    const int l = f % LINES, kk = f / LINES;
    const int line = c0 + l;
    if (line >= P.nlines) continue;
    const cpx v = buf[l * LS + kk];
    if constexpr (STEP_A) {
      out[(long long)kk * P.N2 + line] = cmul(v, P.tw_full[line * kk]);  // n2 k1 < Nc
    } else {
      out[line + (long long)P.N1 * kk] = v;
    }
  }
}

}  // namespace spk
