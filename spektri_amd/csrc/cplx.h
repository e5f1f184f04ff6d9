// Complex-input path (cf32le): windowed complex forward FFT of N points as
// two passes through HBM, Welch accumulated directly from the spectra.
// spektri/src/dsp.rs:114-136 (process_complex), spektri/src/dsp/multifft.rs:11-47
// (apply_window + FFT), spektri/src/dsp/spectrum.rs:82-116 (Welch).
//
// N = 64 N1, n = 64 n1 + n2, k = k1 + N1 k2:
//   X[k1 + N1 k2] = sum_{n2} W64^{n2 k2} W_N^{n2 k1} [ sum_{n1} x[64 n1 + n2] W_{N1}^{n1 k1} ]
// cplx_row_kernel: for every (block, n2) the N1-point FFT over n1 (strided
// input, read through the L2 of one XCD per block), times W_N^{n2 k1}, stored
// as rows Y[n2][k1].  cplx_col_kernel: for every k1 the 64-point DFT over
// n2, stored in natural bin order.
#pragma once
#include "kernels.h"

namespace spk {

struct CplxRowParams {
  const void* src;   // stream samples (cf32, or s16 pairs for the S16 loader); sample s >= 0 at src[s]
  const void* hist;  // the O samples before src[0], same form
  int H, O;
  float wscale;
  int nblocks;
  cpx* Y;  // [block][64][N1]
  const cpx* tw2;   // forward-sign pass tables of the N1 plan
  const cpx* tw3;
  const cpx* ta;    // W_N^{n2 a}, [64][64]
  const cpx* tb;    // W_N^{64 n2 b}, [64][N1/64]
};

// Workgroup w serves XCD w mod 8; block g goes to XCD g mod 8, so the 64
// row transforms of one block (which read interleaved samples of the same
// cache lines) share one L2.  A workgroup of WGT threads holds TPW = WGT / T
// consecutive rows n2 and loads them together: element f of its share is
// sample (n1 = f / TPW, row f mod TPW), so a wave reads runs of TPW
// adjacent samples (a 128-byte line holds 16 consecutive n2 of one n1)
// rather than 64 samples 512 bytes apart, and the rows are staged in LDS
// before the first pass.  Grid: 8 * ceil(nblocks / 8) * (64 / TPW).
// S16: a complex sample is a 32-bit pair of s16 (cs16le input, and real
// s16le input read as x[2n] + i x[2n+1]), converted in the load.
template <int N1, int R, int R1, int R2, int R3, bool S16, int WGT = kChanThreads>
__global__ __launch_bounds__(WGT) void cplx_row_kernel(CplxRowParams P) {
  constexpr int T = N1 / R;
  static_assert(R1 * R2 * R3 == N1, "radix plan must cover N1");
  static_assert(T <= WGT && WGT % T == 0, "tile must divide the workgroup");
  constexpr int TPW = WGT / T;
  static_assert(64 % TPW == 0, "row tiles divide the 64 rows");
  constexpr int TILES = 64 / TPW;  // workgroups per block
  constexpr int LP = lds_padded_size(N1);
  __shared__ float lds_all[TPW * LP * 2];

  const int x = (int)blockIdx.x % 8, r = (int)blockIdx.x / 8;
  const int g = 8 * (r / TILES) + x;
  if (g >= P.nblocks) return;  // whole workgroup
  const int tl = threadIdx.x / T, t = threadIdx.x % T;
  const int n20 = (r % TILES) * TPW;  // the workgroup's first row
  const int n2 = n20 + tl;
  float* lds = lds_all + tl * LP * 2;

  const long long s0 = (long long)g * P.H - P.O;  // stream index of sample 0 of the block
  cpx* rows = reinterpret_cast<cpx*>(lds_all);
#pragma unroll
  for (int k = 0; k < R; ++k) {
    const int f = k * WGT + (int)threadIdx.x;
    const int n1 = f / TPW, j = f % TPW;
    const long long si = s0 + 64LL * n1 + n20 + j;
    cpx a;
    if constexpr (S16) {
      const int w = si < 0 ? static_cast<const int*>(P.hist)[P.O + si] : static_cast<const int*>(P.src)[si];
      a = {(float)(int16_t)w, (float)(w >> 16)};
    } else {
      a = si < 0 ? static_cast<const cpx*>(P.hist)[P.O + si] : static_cast<const cpx*>(P.src)[si];
    }
    rows[j * LP + lds_pad(n1)] = cscale(a, P.wscale);
  }
  __syncthreads();
  cpx v[R];
#pragma unroll
  for (int e = 0; e < R; ++e) v[e] = reinterpret_cast<const cpx*>(lds)[lds_pad(st_in_index<N1, R, T, R1>(t, e))];
  st_butterflies<N1, R, T, R1, 1, -1>(v, t, nullptr);
  tile_exchange<T>(  // wave-local when the row's T threads lie in one wave
      v, lds, [&](int e) { return st_out_addr<N1, R, T, R1, 1>(t, e); },
      [&](int e) { return st_in_addr<N1, R, T, R2>(t, e); });
  st_butterflies<N1, R, T, R2, R1, -1>(v, t, P.tw2);
  constexpr int RL = (R3 > 1) ? R3 : R2;
  if constexpr (R3 > 1) {
    tile_exchange<T>(
        v, lds, [&](int e) { return st_out_addr<N1, R, T, R2, R1>(t, e); },
        [&](int e) { return st_in_addr<N1, R, T, R3>(t, e); });
    st_butterflies<N1, R, T, R3, R1 * R2, -1>(v, t, P.tw3);
  }
  constexpr int PL = N1 / RL;
  cpx* y = P.Y + ((long long)g * 64 + n2) * N1;
  const cpx* ta = P.ta + n2 * 64;
  const cpx* tb = P.tb + n2 * (N1 / 64 > 0 ? N1 / 64 : 1);
#pragma unroll
  for (int e = 0; e < R; ++e) {
    const int u = e / RL, q = e % RL;
    const int k1 = t + T * u + q * PL;
    y[k1] = cmul(v[e], cmul(ta[k1 & 63], tb[k1 >> 6]));
  }
}

// Row stride (points) of the persistent form's staged rows: the padded row,
// rounded up to 32 / TPW mod 32, so that the staging stores of a wave (TPW
// rows, 64 / TPW consecutive n1 each) fall in distinct 8-byte bank pairs.
template <int N1, int TPW>
constexpr int cplx_row_stride() {
  constexpr int base = lds_padded_size(N1), want = 32 / TPW;
  return base + (((want - base) % 32) + 32) % 32;
}

// Persistent form of cplx_row_kernel for long rows (N1 = 2048 and 4096,
// whose staged rows leave room for one workgroup per CU, so a workgroup's
// loads cannot overlap another's transform).  Workgroup w serves XCD
// x = w mod 8 and keeps one tile of TPW rows for all its blocks
// g = x + 8 i, i = par, par + npar, ... (the TILES * npar workgroups of an
// XCD take its blocks npar at a time, every tile of a block together).
// With the rows fixed, every twiddle is fixed: the pass-2 and pass-3 table
// entries are loaded into registers once, and the output factor
// W_N^{n2 k1} = ta[n2][k1 mod 64] tb[n2][k1 / 64] is one per-thread ta
// entry (k1 mod 64 = t mod 64 for every point of a thread) times tb entries
// that are uniform over the wave.  The only vector loads inside the loop are
// then the next block's input, issued as soon as the current block's input
// is staged in LDS, so they are in flight for the whole transform.  Same
// arithmetic as cplx_row_kernel, so the rows are bit-identical.
// Grid: 8 * TILES * npar.
template <int N1, int R, int R1, int R2, int R3, bool S16, int WGT>
__global__ __launch_bounds__(WGT) void cplx_row_pf_kernel(CplxRowParams P) {
  constexpr int T = N1 / R;
  static_assert(R1 == R && R2 == R && R1 * R2 * R3 == N1, "one butterfly per thread in passes 1 and 2");
  static_assert(R3 == R || R3 < 8, "pass 3: one radix-R butterfly, or R / R3 small ones");
  constexpr int TPW = WGT / T, TILES = 64 / TPW, P3 = R1 * R2, PL = N1 / R3;
  static_assert(T % 64 == 0 && PL % 64 == 0 && TPW >= 2 && WGT % T == 0, "lane layout");
  constexpr int LP = cplx_row_stride<N1, TPW>();
  __shared__ float lds_all[TPW * LP * 2];
  using In = std::conditional_t<S16, int, cpx>;

  const int x = (int)blockIdx.x % 8, wj = (int)blockIdx.x / 8;
  const int tile = wj % TILES, par = wj / TILES, npar = (int)gridDim.x / (8 * TILES);
  const int tl = threadIdx.x / T, t = threadIdx.x % T;
  const int n20 = tile * TPW;
  const int n2 = __builtin_amdgcn_readfirstlane(n20 + tl);  // T is a multiple of the wave
  float* lds = lds_all + tl * LP * 2;
  cpx* rows = reinterpret_cast<cpx*>(lds_all);

  // Twiddles in LDS, copied once: the pass-table rows a butterfly reads
  // (compact_rows) and the tile's tb rows.
  constexpr int C2 = compact_rows<R2>() * R1, C3 = compact_rows<R3>() * P3, CB = N1 / 64;
  __shared__ cpx ltw[C2 + C3 + TPW * CB];
  cpx* ltw2 = ltw;
  cpx* ltw3 = ltw + C2;
  cpx* ltb = ltw + C2 + C3;
  for (int i = threadIdx.x; i < C2 + C3 + TPW * CB; i += WGT) {
    if (i < C2)
      ltw[i] = P.tw2[compact_row<R2>(i / R1) * R1 + i % R1];
    else if (i < C2 + C3)
      ltw[i] = P.tw3[compact_row<R3>((i - C2) / P3) * P3 + (i - C2) % P3];
    else
      ltw[i] = P.tb[(n20 + (i - C2 - C3) / CB) * CB + (i - C2 - C3) % CB];
  }
  const cpx ta = P.ta[n2 * 64 + (t & 63)];
  const cpx* tbl = ltb + tl * CB + (t >> 6);

  // Input of block g: blocks with g H >= O read only the stream, through a
  // buffer resource at the tile's first sample (one lane offset register,
  // the point offsets as constants); the others (before the loop only: the
  // host takes this form when 8 H >= O, so every prefetched block qualifies)
  // select the history or the stream per sample.
  In pf[R];
  const unsigned voff = (unsigned)(64 * (threadIdx.x / TPW) + threadIdx.x % TPW) * (unsigned)sizeof(In);
  // (no block: an empty resource at the stream start, whose loads read
  // nothing and return zero)
  auto load_stream = [&](int g, bool any) {
    const char* blk = static_cast<const char*>(P.src) +
                      (any ? ((long long)g * P.H - P.O + n20) * (long long)sizeof(In) : 0LL);
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<char*>(blk), (short)0, any ? 64 * N1 * (int)sizeof(In) : 0, kBufferRsrcWord3);
#pragma unroll
    for (int k = 0; k < R; ++k) {
      const int so = k * 64 * T * (int)sizeof(In);
      if constexpr (S16)
        pf[k] = (int)__builtin_amdgcn_raw_buffer_load_b32(rs, voff, so, 0);
      else
        pf[k] = __builtin_bit_cast(cpx, __builtin_amdgcn_raw_buffer_load_b64(rs, voff, so, 0));
    }
  };

  int g = x + 8 * par;
  if (g >= P.nblocks) return;  // whole workgroup
  if ((long long)g * P.H >= P.O) {
    load_stream(g, true);
  } else {
    const In* src = static_cast<const In*>(P.src);
    const In* hist = static_cast<const In*>(P.hist);
    const long long s0 = (long long)g * P.H - P.O;
#pragma unroll
    for (int k = 0; k < R; ++k) {
      const int f = k * WGT + (int)threadIdx.x;
      const long long si = s0 + 64LL * (f / TPW) + n20 + f % TPW;
      pf[k] = si < 0 ? hist[P.O + si] : src[si];
    }
  }
  // Staging address of point k: the lane's base plus a constant (k T is a
  // multiple of 32, so lds_pad(k T + a) = lds_pad(a) + k (T + T / 32)).
  // The staging closes the loop body, after the output stores, so that the
  // waits for the prefetched words count the same stores on every path (a
  // staging at the top would be reached with and without them and wait for
  // the stores as well).
  const int stage0 = (int)(threadIdx.x % TPW) * LP + lds_pad((int)threadIdx.x / TPW);
  auto stage = [&]() {
#pragma unroll
    for (int k = 0; k < R; ++k) {
      cpx a;
      if constexpr (S16) {
        a = {(float)(int16_t)pf[k], (float)(pf[k] >> 16)};
      } else {
        a = pf[k];
      }
      rows[stage0 + k * (T + T / 32)] = cscale(a, P.wscale);
    }
  };
  __syncthreads();  // the twiddle copies are done
  stage();
  for (;;) {
    const int gn = g + 8 * npar;
    load_stream(gn, gn < P.nblocks);
    __syncthreads();
    cpx v[R];
#pragma unroll
    for (int e = 0; e < R; ++e) v[e] = reinterpret_cast<const cpx*>(lds)[lds_pad(st_in_index<N1, R, T, R1>(t, e))];
    st_butterflies<N1, R, T, R1, 1, -1>(v, t, nullptr);
    tile_exchange<T>(
        v, lds, [&](int e) { return st_out_addr<N1, R, T, R1, 1>(t, e); },
        [&](int e) { return st_in_addr<N1, R, T, R2>(t, e); });
    apply_compact_twiddles<R, R1>(v, ltw2, t & (R1 - 1));
    dft<R, -1>(v);
    tile_exchange<T>(
        v, lds, [&](int e) { return st_out_addr<N1, R, T, R2, R1>(t, e); },
        [&](int e) { return st_in_addr<N1, R, T, R3>(t, e); });
#pragma unroll
    for (int u = 0; u < R / R3; ++u) {
      cpx b[R3];
#pragma unroll
      for (int j = 0; j < R3; ++j) b[j] = v[u * R3 + j];
      apply_compact_twiddles<R3, P3>(b, ltw3, (t + T * u) & (P3 - 1));
      dft<R3, -1>(b);
#pragma unroll
      for (int q = 0; q < R3; ++q) v[u * R3 + q] = b[q];
    }
    cpx* y = P.Y + ((long long)g * 64 + n2) * N1;
#pragma unroll
    for (int e = 0; e < R; ++e)
      y[t + T * (e / R3) + (e % R3) * PL] = cmul(v[e], cmul(ta, tbl[(T / 64) * (e / R3) + (PL / 64) * (e % R3)]));
    if (gn >= P.nblocks) break;
    __syncthreads();  // the transform has read its rows
    stage();
    g = gn;
  }
}

struct CplxColParams {
  const cpx* Y;  // [block][64][N1]
  int N1;
  int nblocks;
  cpx* spec;
  long long spec_stride;
  const cpx* tw64;  // W64^m, m < 64 (forward sign)
};

// 64 columns per 256-thread workgroup, 4 threads per column: column
// gc = 64 x + c is column k1 = gc mod N1 of block g = gc / N1 (so for
// N1 < 64 a workgroup spans several blocks).  Thread (c, s) takes rows
// n2 = s + 4 j, j < 16, runs the 16-point DFT over j, twiddles by
// W64^{s q}, and after an LDS exchange each thread finishes 4 of the 16 q
// values with 4-point DFTs over s: X[k1 + N1 (q + 16 p)].
// Grid: ceil(nblocks * N1 / 64).
__global__ __launch_bounds__(kChanThreads) void cplx_col_kernel(CplxColParams P) {
  __shared__ cpx lds[64 * 16 * 4 + 64];  // [c][q][s], padded per column
  const int c = threadIdx.x & 63, s = threadIdx.x >> 6;
  const long long gc = (long long)blockIdx.x * 64 + c;
  const bool active = gc < (long long)P.nblocks * P.N1;  // columns past the end compute column 0, store nothing
  const long long gcc = active ? gc : 0;
  const int g = (int)(gcc / P.N1), k1 = (int)(gcc % P.N1);
  const cpx* y = P.Y + (long long)g * 64 * P.N1 + k1;
  cpx z[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) z[j] = y[(long long)(s + 4 * j) * P.N1];
  dft<16, -1>(z);
#pragma unroll
  for (int q = 1; q < 16; ++q) z[q] = cmul(z[q], P.tw64[(s * q) & 63]);
  auto la = [](int cc, int q, int ss) { return cc * 65 + q * 4 + ss; };
#pragma unroll
  for (int q = 0; q < 16; ++q) lds[la(c, q, s)] = z[q];
  __syncthreads();
  cpx* out = P.spec + (long long)g * P.spec_stride + k1;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int q = 4 * s + i;
    cpx b[4];
#pragma unroll
    for (int ss = 0; ss < 4; ++ss) b[ss] = lds[la(c, q, ss)];
    dft<4, -1>(b);
    if (active) {
#pragma unroll
      for (int p = 0; p < 4; ++p) out[(long long)(q + 16 * p) * P.N1] = b[p];
    }
  }
}

// Welch records straight from the spectra (complex input), spectrum.rs:95-116:
// for each group of A blocks, sum over its blocks of
// |X[k] - (X[k-1] + X[k+1]) / 2|^2 with get_bin neighbours.  Grid: x over
// bins, y over the groups the launch touches (rows past the count exit).
// Row 0 also saves the next call's input history and resets the batched
// channel queue counters.
struct WelchDirectParams {
  const cpx* spec;
  long long spec_stride;
  int N, K, real_input;
  int nblocks, A;
  const long long* pos;
  long long* pos_next;
  // open-group sums by bin: the previous call's (read) and this call's
  // (written), two buffers alternated per call
  const float* carry_in;
  float* carry_out;
  float* out;
  // input history for the next call: the last hist_words 32-bit words of
  // this call's input, or (hist_shorts != 0: s16 samples with an odd
  // overlap or offset) its last hist_shorts 16-bit samples
  const void* hist_src;
  void* hist_dst;
  int hist_words;
  int hist_shorts;
  int* chan_counters;
  int ncounters;
  unsigned char* qout;
  int qfmt;
};

__device__ __forceinline__ void copy_history(const WelchDirectParams& F, int i0, int stride) {
  if (F.hist_shorts) {
    const unsigned short* src = static_cast<const unsigned short*>(F.hist_src);
    unsigned short* dst = static_cast<unsigned short*>(F.hist_dst);
    for (int i = i0; i < F.hist_shorts; i += stride) dst[i] = src[i];
  } else {
    const int* src = static_cast<const int*>(F.hist_src);
    int* dst = static_cast<int*>(F.hist_dst);
    for (int i = i0; i < F.hist_words; i += stride) dst[i] = src[i];
  }
}

// Real input outside the forward kernel's plans goes through the
// complex pipeline: the complex FFT Z of the M = N/2 packed pairs
// z[n] = x[2n] + i x[2n+1] sits in the spectrum rows, and this kernel
// untangles it in place into the real spectrum X[0..M]:
//   X[k] = E + W^k O,  X[M-k] = conj(E - W^k O),
//   E = (Z[k] + conj Z[M-k]) / 2,  O = -i (Z[k] - conj Z[M-k]) / 2,
// W = exp(-2 pi i / N) (multifft.rs:57-95 untangles a different packing,
// two blocks per transform, into the same X).  Thread k < M/2 + 1 of block
// y writes bins k and M - k.  Grid: x over k, y over blocks.
struct UntangleParams {
  cpx* spec;
  long long spec_stride;
  int M;
  const cpx* tw;  // W^k, k = 0..M/2
};

__global__ void untangle_kernel(UntangleParams P) {
  const int k = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (k > P.M / 2) return;
  cpx* Z = P.spec + (long long)blockIdx.y * P.spec_stride;
  const int km = k == 0 ? 0 : P.M - k;  // k = 0: its own mirror
  const cpx A = Z[k], B = Z[km];
  const cpx E = {0.5f * (A.x + B.x), 0.5f * (A.y - B.y)};
  const cpx O = {0.5f * (A.y + B.y), -0.5f * (A.x - B.x)};
  const cpx WO = cmul(P.tw[k], O);
  Z[k] = {E.x + WO.x, E.y + WO.y};
  const cpx Xm = {E.x - WO.x, WO.y - E.y};
  if (k == 0) Z[P.M] = Xm;           // Nyquist bin
  else if (2 * k != P.M) Z[km] = Xm;  // (k = M/2 is its own mirror)
}

// Real input through the row and column kernels (N = 2^17..2^19): the
// column transforms and the untangle in one launch.  Workgroup w of a block
// takes columns k1 = 32 w + i (slots i < 32) and their mirrors N1 - k1
// (slots 32 + i; slot 32 of workgroup 0 takes column N1/2, whose mirror is
// itself, as column 0's is), so both Z[k] and Z[M - k] of every bin it owns
// are in its LDS: bin k = k1 + N1 m has its mirror in column N1 - k1 at
// m' = 63 - m (column 0: m' = -m mod 64).  Each column's 64-point DFT runs
// as in cplx_col_kernel; then every thread forms X[k] = E + W_N^k O for 16
// bins (untangle_kernel's formulas; W_N^k = W_N^k1 W_128^m from the table
// of W_N^j, j <= N/4) and writes the spectrum row, X[M] included.
// Grid: nblocks * N1 / 64 workgroups.
__global__ __launch_bounds__(kChanThreads) void cplx_col_untangle_kernel(CplxColParams P, const cpx* __restrict__ tu) {
  __shared__ cpx lds[64 * 16 * 4 + 64];  // [c][q][s], then [c][m] (stride 65)
  const int c = threadIdx.x & 63, s = threadIdx.x >> 6;
  const int wpb = P.N1 / 64;  // workgroups per block
  const int g = (int)blockIdx.x / wpb, w = (int)blockIdx.x % wpb;
  auto col_of = [&](int slot) {
    if (slot < 32) return 32 * w + slot;
    return (w == 0 && slot == 32) ? P.N1 / 2 : P.N1 - 32 * w - (slot - 32);
  };
  const int k1 = col_of(c);
  const cpx* y = P.Y + (long long)g * 64 * P.N1 + k1;
  cpx z[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) z[j] = y[(long long)(s + 4 * j) * P.N1];
  dft<16, -1>(z);
#pragma unroll
  for (int q = 1; q < 16; ++q) z[q] = cmul(z[q], P.tw64[(s * q) & 63]);
  auto la = [](int cc, int q, int ss) { return cc * 65 + q * 4 + ss; };
#pragma unroll
  for (int q = 0; q < 16; ++q) lds[la(c, q, s)] = z[q];
  __syncthreads();
  cpx zb[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int q = 4 * s + i;
#pragma unroll
    for (int ss = 0; ss < 4; ++ss) zb[i][ss] = lds[la(c, q, ss)];
    dft<4, -1>(zb[i]);
  }
  __syncthreads();  // every exchange value has been read
  // Z[k1 + N1 m] of slot c at lds[c * 65 + m]
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int p = 0; p < 4; ++p) lds[c * 65 + 4 * s + i + 16 * p] = zb[i][p];
  __syncthreads();
  // untangle: this thread's slot c, bins m = s + 4 j
  const int M = 64 * P.N1;
  const cpx wk1 = tu[k1];  // W_N^k1 (k1 < N1 <= N/4)
  const bool self0 = k1 == 0, selfh = k1 == P.N1 / 2;
  const int c2 = (self0 || selfh) ? c : (c ^ 32);
  cpx* out = P.spec + (long long)g * P.spec_stride + k1;
#pragma unroll 4
  for (int j = 0; j < 16; ++j) {
    const int m = s + 4 * j;
    const int m2 = self0 ? ((64 - m) & 63) : 63 - m;
    const cpx A = lds[c * 65 + m], B = lds[c2 * 65 + m2];
    // W_128^m = W_N^(N1 m): the table holds N1 m' for m' <= 32; beyond,
    // W_128^m = -conj(W_128^(64 - m))
    cpx w128 = tu[(long long)P.N1 * (m <= 32 ? m : 64 - m)];
    if (m > 32) w128 = cpx{-w128.x, w128.y};
    const cpx W = cmul(wk1, w128);
    const cpx E = {0.5f * (A.x + B.x), 0.5f * (A.y - B.y)};
    const cpx O = {0.5f * (A.y + B.y), -0.5f * (A.x - B.x)};
    const cpx WO = cmul(W, O);
    out[(long long)P.N1 * m] = cpx{E.x + WO.x, E.y + WO.y};
    if (self0 && m == 0) P.spec[(long long)g * P.spec_stride + M] = cpx{E.x - WO.x, WO.y - E.y};  // Nyquist bin
  }
}

// Welch records from the kept spectra (spectrum.rs:95-116) for real input,
// K = N/2 + 1 bins, and for complex input of any N (K = N: get_bin wraps
// without a mirror): one thread per four bins k0..k0+3 (two 16-byte loads
// per block; the row stride leaves room past the Nyquist bin), the outer
// neighbours X[k0 - 1] and X[k0 + 4] from the adjacent lanes by DPP; lanes 0
// and 63 of a wave, and the quads at either end of the spectrum, load them
// with get_bin's mirror rule (fftutil.rs:8-22: beyond the stored half,
// conj X[N - m]).  Grid: x over quads (whole waves), y over groups.
__global__ void welch_direct_real_kernel(WelchDirectParams F) {
  const int gi = (int)blockIdx.y;
  const int tid = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int k0 = 4 * tid;
  const long long pos = *F.pos;
  if (gi == 0) {
    copy_history(F, tid, (int)(gridDim.x * blockDim.x));
    if (tid < F.ncounters) F.chan_counters[tid] = 0;
    if (tid == 0) *F.pos_next = pos + F.nblocks;
  }
  const int blk0_mod = (int)(pos % F.A);
  if (gi > (blk0_mod + F.nblocks - 1) / F.A) return;  // whole workgroup
  const int g_lo = gi * F.A - blk0_mod, g_hi = g_lo + F.A;
  const int b_lo = g_lo > 0 ? g_lo : 0;
  const int b_hi = g_hi < F.nblocks ? g_hi : F.nblocks;
  const int lane = (int)(threadIdx.x & 63);
  const bool own = k0 < F.K;  // quads past the spectrum compute (DPP partners) but store nothing
  const int kq = own ? k0 : 0;
  auto bin = [&](const cpx* X, int i) {
    const int m = i < 0 ? i + F.N : (i >= F.N ? i - F.N : i);  // -1 <= i < N + 4
    if (m >= F.K) {
      const cpx v = X[F.N - m];
      return cpx{v.x, -v.y};
    }
    return X[m];
  };
  // neighbours that the lanes next door do not hold
  const bool load_l = lane == 0 || kq == 0;
  const bool load_r = lane == 63 || kq + 4 >= F.K;
  float4 s = {0.f, 0.f, 0.f, 0.f};
  if (gi == 0 && blk0_mod != 0 && own) s = *reinterpret_cast<const float4*>(F.carry_in + kq);
  for (int b = b_lo; b < b_hi; ++b) {
    const cpx* X = F.spec + (long long)b * F.spec_stride;
    const float4 u = *reinterpret_cast<const float4*>(X + kq);      // X[k0], X[k0+1]
    const float4 v = *reinterpret_cast<const float4*>(X + kq + 2);  // X[k0+2], X[k0+3]
    cpx L = {wave_shr1(v.z), wave_shr1(v.w)};
    cpx Rn = {wave_shl1(u.x), wave_shl1(u.y)};
    if (load_l) L = bin(X, kq - 1);
    if (load_r) Rn = bin(X, kq + 4);
    cpx x0 = {u.x, u.y}, x1 = {u.z, u.w}, x2 = {v.x, v.y}, x3 = {v.z, v.w};
    // bins past K - 1 (the last quad) are neighbours by get_bin, not stored values
    if (kq + 1 >= F.K) x1 = bin(X, kq + 1);
    if (kq + 2 >= F.K) x2 = bin(X, kq + 2);
    if (kq + 3 >= F.K) x3 = bin(X, kq + 3);
    s.x += welch_term(x0, L, x1);
    s.y += welch_term(x1, x0, x2);
    s.z += welch_term(x2, x1, x3);
    s.w += welch_term(x3, x2, Rn);
  }
  if (!own) return;
  const float sv[4] = {s.x, s.y, s.z, s.w};
  if (g_hi <= F.nblocks) {
    const long long o = (long long)gi * F.K + kq;
    for (int j = 0; j < 4 && kq + j < F.K; ++j) {
      F.out[o + j] = sv[j] / (float)F.A;
      if (F.qfmt) quantize_bin(F.qout, F.qfmt, o + j, sv[j], F.A);
    }
  } else {
    for (int j = 0; j < 4 && kq + j < F.K; ++j) F.carry_out[kq + j] = sv[j];
  }
}

// Complex input (K = N, a multiple of 1024): one thread per 4 consecutive
// bins k0..k0+3, read as two 16-byte loads per block; the outer neighbours
// X[k0 - 1] and X[k0 + 4] come from the adjacent lanes by DPP, and lanes 0
// and 63 of each wave load theirs (modulo N: get_bin's wrap, fftutil.rs:8-22).
// Grid: x = K / 1024 workgroups of 256 threads, y = Welch groups.
__global__ void welch_direct_kernel(WelchDirectParams F) {
  const int gi = (int)blockIdx.y;
  const int tid = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int k0 = 4 * tid;
  const long long pos = *F.pos;
  if (gi == 0) {
    copy_history(F, tid, (int)(gridDim.x * blockDim.x));
    if (tid < F.ncounters) F.chan_counters[tid] = 0;
    if (tid == 0) *F.pos_next = pos + F.nblocks;
  }
  const int blk0_mod = (int)(pos % F.A);
  // whole workgroups leave together (the DPP below needs every lane)
  if (gi > (blk0_mod + F.nblocks - 1) / F.A) return;
  const int g_lo = gi * F.A - blk0_mod, g_hi = g_lo + F.A;
  const int b_lo = g_lo > 0 ? g_lo : 0;
  const int b_hi = g_hi < F.nblocks ? g_hi : F.nblocks;
  const int lane = (int)(threadIdx.x & 63);
  const int kl = (k0 - 1) & (F.N - 1), kr = (k0 + 4) & (F.N - 1);
  float4 s = {0.f, 0.f, 0.f, 0.f};
  if (gi == 0 && blk0_mod != 0) s = *reinterpret_cast<const float4*>(F.carry_in + k0);
  for (int b = b_lo; b < b_hi; ++b) {
    const cpx* X = F.spec + (long long)b * F.spec_stride;
    const float4 u = *reinterpret_cast<const float4*>(X + k0);      // X[k0], X[k0+1]
    const float4 v = *reinterpret_cast<const float4*>(X + k0 + 2);  // X[k0+2], X[k0+3]
    cpx L = {wave_shr1(v.z), wave_shr1(v.w)};  // lane - 1's X[k0 - 1]
    cpx Rn = {wave_shl1(u.x), wave_shl1(u.y)}; // lane + 1's X[k0 + 4]
    if (lane == 0) L = X[kl];
    if (lane == 63) Rn = X[kr];
    const cpx x0 = {u.x, u.y}, x1 = {u.z, u.w}, x2 = {v.x, v.y}, x3 = {v.z, v.w};
    s.x += welch_term(x0, L, x1);
    s.y += welch_term(x1, x0, x2);
    s.z += welch_term(x2, x1, x3);
    s.w += welch_term(x3, x2, Rn);
  }
  if (g_hi <= F.nblocks) {
    const long long o = (long long)gi * F.K + k0;
    const float inv = (float)F.A;
    *reinterpret_cast<float4*>(F.out + o) = make_float4(s.x / inv, s.y / inv, s.z / inv, s.w / inv);
    if (F.qfmt) {
      quantize_bin(F.qout, F.qfmt, o, s.x, F.A);
      quantize_bin(F.qout, F.qfmt, o + 1, s.y, F.A);
      quantize_bin(F.qout, F.qfmt, o + 2, s.z, F.A);
      quantize_bin(F.qout, F.qfmt, o + 3, s.w, F.A);
    }
  } else {
    *reinterpret_cast<float4*>(F.carry_out + k0) = s;
  }
}

}  // namespace spk
