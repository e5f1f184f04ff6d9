// Hot-path kernels: forward real FFT with fused s16 load and fused Welch,
// the per-size channel kernel (gather x weights -> IFFT -> discard), and the
// Welch record finaliser.
#pragma once
#include "stockham.h"

namespace spk {

constexpr int kLdsBudgetBytes = 150 * 1024;

// ---------------------------------------------------------------------------
// Forward real FFT of N = 2M points per block (spektri/src/dsp/multifft.rs:51-96,
// windowing dsp.rs:70-73, framing main.rs:154-193), Welch accumulation
// (spektri/src/dsp/spectrum.rs:89-116) fused on the spectrum.
//
// Block g covers stream samples g*H - O .. g*H - O + N - 1.  The N real
// samples are packed as z[n] = w (x[2n] + i x[2n+1]) (w = scaling / N), an
// M-point complex FFT Z is taken in three register/LDS passes, and the real
// spectrum is recovered with
//   X[k] = E[k] + W^k O[k],  E = (Z[k] + conj Z[M-k]) / 2,
//   O = -i (Z[k] - conj Z[M-k]) / 2,  W = exp(-2 pi i / N),  k = 0..M.
// The last pass is arranged so that the lanes l and l ^ 32 of a wave own
// mirror bins k and M - k in the same register slot (upper lanes read their
// inputs in reversed order and use conjugate twiddles), and one
// v_permlane32_swap per float gives every lane its mirror values without
// an LDS round trip.
// ---------------------------------------------------------------------------
struct FwdParams {
  const int16_t* src;   // stream samples; sample s >= 0 at src[s]
  const int16_t* head;  // block 0's N samples: O samples of history, then src[0 .. N-O)
  cpx* spec;            // [nblocks][spec_stride]
  long long spec_stride;
  const cpx* tw2;       // pass-2 table, forward sign
  const cpx* tw3;       // pass-3 table, forward sign
  const cpx* twu;       // exp(-2 pi i k / N), k = 0..M/R
  float* welch_acc;     // [ring][welch_stride]
  long long welch_stride;
  int welch_ring;
  int welch_A;
  long long block0;     // stream index of this launch's block 0
  float wscale;
  int nblocks;
  int H, O;
};

template <int M, int R, int R1, int R2>
struct FwdRealCfg {
  static constexpr int T = M / R;
  static constexpr int P3 = R1 * R2;
  static_assert(R1 * R2 * R == M, "radix plan must cover M");
  static_assert(P3 == T, "last pass must have one butterfly per thread");
  static_assert(T >= 64 && T <= 1024, "one wave minimum, 1024 threads maximum");
  static constexpr bool SPLIT = lds_padded_size(M + 1) * 8 > kLdsBudgetBytes;
  static constexpr int LDS_FLOATS = SPLIT ? lds_padded_size(M + 1) : 2 * lds_padded_size(M + 1);
};

__device__ __forceinline__ float swap_hi_lo(float v, float& other) {
  // v_permlane32_swap: returns (A, B) with A = {v.lo | v.lo}, B = {v.hi | v.hi}
  // when both operands are v; A's upper half and B's lower half are the
  // partner lane's value.
  auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v), false, false);
  other = __uint_as_float(r[1]);
  return __uint_as_float(r[0]);
}

template <int M, int R, int R1, int R2>
__global__ __launch_bounds__(M / R) void fwd_real_kernel(FwdParams P) {
  using C = FwdRealCfg<M, R, R1, R2>;
  constexpr int T = C::T;
  constexpr int P3 = C::P3;
  constexpr bool SPLIT = C::SPLIT;
  __shared__ float lds[C::LDS_FLOATS];

  const int t = threadIdx.x;
  const int g = blockIdx.x;
  const float w = P.wscale;
  // Block g starts at stream sample g*H - O; block 0 reads the head copy.
  const int* blk = reinterpret_cast<const int*>(
      g == 0 ? P.head : P.src + ((long long)g * P.H - P.O));

  cpx x[R];
  // Pass 1 loads: point n = b + j M/R1 is the s16 pair at block sample 2n.
  // Rows are addressed as (uniform row pointer) + lane offset so that the
  // lane offsets are shared instead of kept per row.
#pragma unroll
  for (int e = 0; e < R; ++e) {
    const int u = e / R1, j = e % R1;
    const int* row = blk + (T * u + j * (M / R1));
    const int v = row[t];
    x[e].x = (float)(int16_t)v * w;
    x[e].y = (float)(v >> 16) * w;
  }
  st_butterflies<M, R, T, R1, 1, -1>(x, t, nullptr);

  lds_exchange<R, SPLIT>(
      x, lds, [&](int e) { return st_out_addr<M, R, T, R1, 1>(t, e); },
      [&](int e) { return st_in_addr<M, R, T, R2>(t, e); });
  st_butterflies<M, R, T, R2, R1, -1>(x, t, P.tw2);

  // Pass 3 with the mirror-pair lane map.  Lane l < 32 of wave w takes
  // sub-transform k = pi = 32 w + l; lane l + 32 takes its mirror P3 - pi
  // (P3/2 for pi = 0).  An upper lane must end with Z[ksrc + P3 (R-1-q)] in
  // slot q, i.e. the DFT of its twiddled inputs in reversed order.  That is
  // conj(DFT(conj(in_m) w^{m kt})) with kt = P3 - ksrc, so upper lanes read
  // in natural order, conjugate, use the same table rows as lower lanes
  // (column kt), and conjugate the DFT outputs.
  const int wave = t >> 6, lane = t & 63;
  const bool upper = lane >= 32;
  const int pi = 32 * wave + (lane & 31);
  const int ksrc = upper ? (pi == 0 ? P3 / 2 : P3 - pi) : pi;
  const int kt = (upper && pi == 0) ? P3 / 2 : pi;
  const float sg = upper ? -1.f : 1.f;
  lds_exchange<R, SPLIT>(
      x, lds, [&](int e) { return st_out_addr<M, R, T, R2, R1>(t, e); },
      [&](int e) { return lds_pad(ksrc) + pad_off(e * P3); });
#pragma unroll
  for (int e = 0; e < R; ++e) x[e].y *= sg;
  apply_pass_twiddles<R, P3>(x, P.tw3, kt);
  dft<R, -1>(x);
#pragma unroll
  for (int e = 0; e < R; ++e) x[e].y *= sg;
  // Lower lanes now hold Z[pi + P3 q] in slot q, upper lanes
  // Z[ksrc + P3 (R-1-q)], so slot q of lane l and lane l^32 are mirrors
  // (except the pair pi = 0, whose lanes are their own mirrors).
  __syncthreads();  // every lane is done reading pass-3 inputs from LDS

  const cpx wb = P.twu[upper ? ksrc : pi];
  const float h = 0.5f * sg;
  cpx* spec = P.spec + (long long)g * P.spec_stride;
  // Own bin of slot q: lower pi + P3 q, upper ksrc + P3 (R-1-q).
  const int kb = upper ? ksrc + P3 * (R - 1) : pi;
  const int kstep = upper ? -P3 : P3;
  const int ab = lds_pad(upper ? ksrc : pi) + (upper ? pad_off(P3 * (R - 1)) : 0);
  const int astep = upper ? -pad_off(P3) : pad_off(P3);
  static_for<0, R>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    cpx A, B;
    A.x = swap_hi_lo(x[q].x, B.x);
    A.y = swap_hi_lo(x[q].y, B.y);
    if (wave == 0) {
      // lane 0 (k = 0): mirror of Z[P3 q] is its own slot (R - q) mod R;
      // lane 32 (k = P3/2): mirror of its slot q is its own slot R-1-q.
      const cpx m0 = x[(R - q) & (R - 1)];
      const cpx m32 = x[R - 1 - q];
      if (lane == 0) B = m0;
      if (lane == 32) A = m32;
    }
    const cpx E = {0.5f * (A.x + B.x), h * (A.y - B.y)};
    const cpx Od = {0.5f * (A.y + B.y), -h * (A.x - B.x)};
    constexpr float cqx = float(ct_cos2pi(-q, 2 * R));
    constexpr float cqy = float(ct_sin2pi(-q, 2 * R));
    constexpr float crx = float(ct_cos2pi(-(R - 1 - q), 2 * R));
    constexpr float cry = float(ct_sin2pi(-(R - 1 - q), 2 * R));
    const cpx c = upper ? cpx{crx, cry} : cpx{cqx, cqy};
    const cpx wk = cmul(wb, c);
    const cpx X = cadd(E, cmul(wk, Od));
    spec[kb + q * kstep] = X;
    const int a = ab + q * astep;
    if constexpr (SPLIT) {
      lds[a] = X.x;
    } else {
      reinterpret_cast<cpx*>(lds)[a] = X;
    }
    if constexpr (q == 0) {
      if (t == 0) {  // Nyquist bin M = E - O from Z[0]
        const cpx XN = csub(E, Od);
        spec[M] = XN;
        if constexpr (SPLIT) lds[lds_pad(M)] = XN.x;
        else reinterpret_cast<cpx*>(lds)[lds_pad(M)] = XN;
      }
    }
  });

  // Welch: term[k] = |X[k] - (X[k-1] + X[k+1]) / 2|^2 with get_bin
  // neighbours at the edges: X[-1] = conj X[1], X[M+1] = conj X[M-1].
  // Thread t handles bins t + T j (and t = 0 also bin M).
  const long long gs = P.block0 + g;
  const int slot = (int)((gs / P.welch_A) % P.welch_ring);
  float* acc = P.welch_acc + (long long)slot * P.welch_stride;
  const int a0 = lds_pad(t);
  const int am = (t == 0) ? lds_pad(1) : lds_pad(t - 1);
  const int ap = lds_pad(t + 1);
  __syncthreads();
  if constexpr (SPLIT) {
    // Real parts from the LDS plane, imaginary parts straight from this
    // block's spectrum row (written by this workgroup before the barrier,
    // so the reads hit L1/L2).
    const float* spec_im = reinterpret_cast<const float*>(spec) + 1;
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const int o = pad_off(T * j);
      const int kw = t + T * j;
      // bin -1 of row j = 0 is conj X[1]; for j > 0 it is bin T j - 1, which
      // for t = 0 is thread T-1's bin of row j-1.
      const float cm = (t == 0 && j > 0) ? lds[lds_pad(T * j - 1)] : lds[am + o];
      const float cr = lds[a0 + o] - 0.5f * (cm + lds[ap + o]);
      const float* row_im = spec_im + 2 * (T * j);
      const float im_m = (kw == 0) ? -spec_im[2] : row_im[2 * t - 2];
      const float ci = row_im[2 * t] - 0.5f * (im_m + row_im[2 * t + 2]);
      unsafeAtomicAdd(acc + T * j + t, cr * cr + ci * ci);
      // Keep the scheduler from hoisting every row's loads (register spills).
      if ((j & 3) == 3) __builtin_amdgcn_sched_barrier(0);
    }
    if (t == 0) {
      const float cr = lds[lds_pad(M)] - lds[lds_pad(M - 1)];
      const float ci = spec_im[2 * M];
      unsafeAtomicAdd(acc + M, cr * cr + ci * ci);
    }
  } else {
    const cpx* l = reinterpret_cast<const cpx*>(lds);
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const int o = pad_off(T * j);
      cpx cm;
      if (t == 0) {
        if (j > 0) {
          cm = l[lds_pad(T * j - 1)];
        } else {
          cm = l[lds_pad(1)];
          cm.y = -cm.y;
        }
      } else {
        cm = l[am + o];
      }
      const cpx c0 = l[a0 + o];
      const cpx cp = l[ap + o];
      const float cr = c0.x - 0.5f * (cm.x + cp.x);
      const float ci = c0.y - 0.5f * (cm.y + cp.y);
      unsafeAtomicAdd(acc + T * j + t, cr * cr + ci * ci);
    }
    if (t == 0) {
      const cpx c0 = l[lds_pad(M)];
      const cpx cm = l[lds_pad(M - 1)];
      const float cr = c0.x - cm.x;
      const float ci = c0.y;
      unsafeAtomicAdd(acc + M, cr * cr + ci * ci);
    }
  }
}

// ---------------------------------------------------------------------------
// Channel stage, one launch per IFFT size L (spektri/src/dsp/fcfb.rs:154-188):
//   B[i] = W[j] get_bin(first + j), j = (i + L/2) mod L,
//   y = unnormalised inverse FFT_L(B), output y[L/8 .. 7L/8).
// A tile is one (channel, block) pair handled by T = L/R threads; a
// 256-thread workgroup holds 256/T tiles, each with its own LDS region.
// ---------------------------------------------------------------------------
struct ChanDesc {
  long long out_off;  // cpx offset of the channel's output (block 0)
  int f_mod;          // first bin mod N, in [0, N)
  int w_off;          // float offset of the channel's weights
};

struct ChanParams {
  const cpx* spec;
  long long spec_stride;
  int N, K, real_input;
  const float* weights;
  const ChanDesc* chans;
  int nch, nblocks;
  cpx* out;
  long long out_block_stride;  // cpx per block in a channel's output = 3L/4
  const cpx* tw2;
  const cpx* tw3;
};

constexpr int kChanThreads = 256;

template <int L, int R, int R1, int R2, int R3>
__global__ __launch_bounds__(kChanThreads) void chan_kernel(ChanParams P) {
  constexpr int T = L / R;
  static_assert(R1 * R2 * R3 == L, "radix plan must cover L");
  static_assert(T <= kChanThreads && kChanThreads % T == 0, "tile must divide the workgroup");
  constexpr int TPW = kChanThreads / T;
  constexpr int LP = lds_padded_size(L);
  __shared__ float lds_all[TPW * LP * 2];

  const int tl = threadIdx.x / T;
  const int t = threadIdx.x % T;
  float* lds = lds_all + tl * LP * 2;
  const long long ntiles = (long long)P.nch * P.nblocks;
  long long tile = (long long)blockIdx.x * TPW + tl;
  const bool active = tile < ntiles;
  if (!active) tile = 0;
  const int c = (int)(tile % P.nch);
  const int g = (int)(tile / P.nch);
  const ChanDesc d = P.chans[c];
  const cpx* X = P.spec + (long long)g * P.spec_stride;
  const float* W = P.weights + d.w_off;

  cpx x[R];
#pragma unroll
  for (int e = 0; e < R; ++e) {
    const int n = st_in_index<L, R, T, R1>(t, e);
    const int jj = (n + L / 2) & (L - 1);
    int m = d.f_mod + jj;
    if (m >= P.N) m -= P.N;
    const bool mirror = P.real_input && m >= P.K;
    cpx v = X[mirror ? P.N - m : m];
    if (mirror) v.y = -v.y;
    x[e] = cscale(v, W[jj]);
  }
  st_butterflies<L, R, T, R1, 1, +1>(x, t, nullptr);
  lds_exchange<R, false>(
      x, lds, [&](int e) { return st_out_addr<L, R, T, R1, 1>(t, e); },
      [&](int e) { return st_in_addr<L, R, T, R2>(t, e); });
  st_butterflies<L, R, T, R2, R1, +1>(x, t, P.tw2);
  constexpr int RL = (R3 > 1) ? R3 : R2;
  if constexpr (R3 > 1) {
    lds_exchange<R, false>(
        x, lds, [&](int e) { return st_out_addr<L, R, T, R2, R1>(t, e); },
        [&](int e) { return st_in_addr<L, R, T, R3>(t, e); });
    st_butterflies<L, R, T, R3, R1 * R2, +1>(x, t, P.tw3);
  }
  // Last pass radix RL, p = L / RL: slot u*RL + q holds point b + q p.
  constexpr int PL = L / RL;
  cpx* out = P.out + d.out_off + (long long)g * P.out_block_stride;
  if (active) {
#pragma unroll
    for (int e = 0; e < R; ++e) {
      const int u = e / RL, q = e % RL;
      const int idx = t + T * u + q * PL;
      if (idx >= L / 8 && idx < (L / 8) * 7) out[idx - L / 8] = x[e];
    }
  }
}

// ---------------------------------------------------------------------------
// Welch records (spektri/src/dsp/spectrum.rs:113-160): acc / A for every
// group completed in this launch, then the slot is cleared for reuse.
// ---------------------------------------------------------------------------
__global__ void welch_finalize_kernel(float* acc, long long stride, int ring, long long s_first,
                                      int nrec, int K, int A, float* out,
                                      long long out_stride) {
  const long long total = (long long)nrec * K;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    const int r = (int)(i / K), k = (int)(i % K);
    const int slot = (int)((s_first + r) % ring);
    float* a = acc + (long long)slot * stride + k;
    out[(long long)r * out_stride + k] = *a / (float)A;
    *a = 0.f;
  }
}

}  // namespace spk
