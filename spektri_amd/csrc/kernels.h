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
struct ChanDesc {
  long long out_off;  // cpx offset of the channel's output (block 0)
  int f_mod;          // first bin mod N, in [0, N)
  int w_off;          // float offset of the channel's weights
};

// Word 3 of a raw (stride 0) buffer resource on gfx9-family GPUs: 32-bit
// data format, accesses past num_records are dropped (stores) or read 0.
constexpr int kBufferRsrcWord3 = 0x00020000;
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
// Cache-policy bit of a buffer access that makes a load miss the CU's L1
// (sc1, the bit agent-scope atomic loads use).
constexpr int kCpolSc1 = 16;

// Lanes of every wave whose spectrum values are published to the edge
// ring (global scratch, one region per workgroup, L2-resident): the wave
// edges and their inner neighbours, plus wave 0's lanes 33/34 and the
// Nyquist bin.  Index -1 = not published.
constexpr int kEdgeLanes = 10;
__host__ __device__ inline int edge_slot(int lane) {
  switch (lane) {
    case 0: return 0;
    case 1: return 1;
    case 30: return 2;
    case 31: return 3;
    case 32: return 4;
    case 33: return 5;
    case 34: return 6;
    case 62: return 7;
    case 63: return 8;
    default: return -1;
  }
}
constexpr int kEdgeNyq = 9;  // wave 0, slot 0: X[M]

// ---------------------------------------------------------------------------
// Welch records (spektri/src/dsp/spectrum.rs:113-160).  For every group s
// touched by this launch: sum the partial rows of the workgroups that
// processed its blocks (plus the carry of a group begun in an earlier
// launch); a complete group becomes the record sum / A, an incomplete last
// group goes back to the carry row.
// ---------------------------------------------------------------------------
struct FinParams {
  const float* wpart;
  long long wstride;
  int wslots;
  int G;               // workgroups of the forward launch
  int nblocks;
  const long long* pos;  // device stream position of the launch's first block
  long long* pos_next;   // the next call's position slot (pos + nblocks is written there)
  int A, K, M, R;
  // Sums of the Welch group left open at the end of the previous call (read)
  // and of this call (written), in partial-row order, M + 1 floats each.
  // Two buffers alternated per call: the first group reads one while the
  // last group writes the other.
  const float* carry_in;
  float* carry_out;
  float* out;          // records [r][K]
  // input history for the next call: the last hist_words 32-bit words of
  // this call's input
  const int* hist_src;
  int* hist_dst;
  int hist_words;
  int* chan_counters;  // queue counters of chan_batched_kernel, reset here
  int ncounters;
  unsigned char* qout;  // quantised records (qfmt 1: K bytes, 2: K big-endian u16) or null
  int qfmt;
  // Teams per chunk (a power of two; 1 = one team): in the forward kernel's
  // record phase S teams of T / S threads share a chunk of 4 T / S
  // positions, in welch_finalize_kernel S teams of 256 / S threads; team s
  // sums the s-th range of the group's workgroups.
  int split;
  // Chunks per take of the forward kernel's record phase (a power of two):
  // short averages make many small chunks, taken several at a time.
  int take;
};

struct FwdParams {
  const void* src;      // stream samples (s16, or f32 for the f32 loader); sample s >= 0 at src[s]
  const void* hist;     // the O samples before src[0] (block 0 reads them, then src[0 .. N-O))
  cpx* spec;            // per-block spectrum rows (written when keep_spec)
  long long spec_stride;
  int keep_spec;
  const cpx* tw2;       // pass-2 table, forward sign
  const cpx* tw3;       // pass-3 table, forward sign
  const cpx* twu;       // exp(-2 pi i k / N), k = 0..M/R
  float* wpart;         // Welch partial sums [workgroup][slot][wstride], lane-major rows (welch_row_index)
  long long wstride;
  int wslots;
  cpx* edump;           // per workgroup: spectrum values of the edge lanes [workgroup][wave][kEdgeLanes][R]
  const int4* eitems;   // Welch items of the edge bins (edge_items)
  int A;
  const long long* pos;  // device stream position: stream index of this launch's block 0
  float wscale;
  int nblocks;
  int H, O;
  // fused channel stage (FL >= 8): every channel has IFFT size FL, lies in
  // bins [0, K) and no two channels share a bin
  const ChanDesc* fchans;
  int fnch;
  const uint4* fscat;   // per (slot group of 8, lane): LDS cpx index of each slot's bin (u16 x 8)
  int fscat_nyq;        // LDS cpx index of the Nyquist bin
  const float* weights;
  cpx* fout;
  const cpx* ftw2;      // pass-2 table of the FL-point inverse transform
  // For the next call, by workgroup 0: its position slot, its input history
  // (the last O samples of this call) and the batched channel kernel's
  // queue counters (that kernel runs after this one in the same call).
  long long* pos_next;
  const int* hist_src;
  int* hist_dst;
  int hist_words;
  int* chan_counters;
  int ncounters;
  // Welch records of the previous call (when fin.nblocks > 0), finished by
  // the workgroups that are done with their own blocks: chunks of 4 T
  // partial-row positions of one group, taken through fin_ticket (zero
  // between launches; the launch's last taker resets it).
  FinParams fin;
  int* fin_ticket;
  // General fused stage (FL = kFusedGen1 / kFusedGen2, below): the scatter
  // tables of fscat hold fused_targets(FL) tables of M/8 uint4 one after the
  // other, entries LDS cpx index | conj << 15; the Nyquist bin's second
  // target; per (round, thread) slots (see gen_slot) and, beside them, the
  // descriptors of the slots' tiles; pass-2 twiddles of the tiles of 32, 64,
  // 128 and 256 points.
  int fscat_nyq2;
  const struct GenTile* gtiles;
  const int* gslots;
  int grounds;
  const cpx* gtw[4];
};

// A tile of the general fused stage: one channel, its LDS region (cpx
// offset, FL + FL/32 values), weights and output.
struct GenTile {
  long long out_off;
  int lds;
  int w_off;
};
// Thread slot of the general fused stage: bit 31 idle, bits 0-2 size class
// c (FL = 8 << c; every lane of a wave has the same class), bits 3-7 the
// thread's index in the tile, bits 8-30 the tile.
__host__ __device__ constexpr int gen_slot(int tile, int ftt, int cls, bool idle) {
  return (idle ? (int)0x80000000 : 0) | (tile << 8) | (ftt << 3) | cls;
}

// Fused channel tiles: FL points as 16 x (FL/16), 16 points per thread.
template <int FL>
struct FusedCfg {
  static constexpr int RF = 16;
  static constexpr int TF = FL / RF;
  static constexpr int RF1 = 16;
  static constexpr int RF2 = FL / 16;
  static constexpr int STRIDE = FL + (FL >> 5);  // cpx per tile region (>= lds_pad(FL - 1) + 1)
};
template <>
struct FusedCfg<0> {
  static constexpr int RF = 1, TF = 1, RF1 = 1, RF2 = 1, STRIDE = 0;
};
// The general fused stage: channels of any power-of-two IFFT size from 8 to
// 256 (mixed), reaching past DC or Nyquist (get_bin's mirror), overlapping,
// several tiles per thread when there are more than the workgroup holds at
// once; its tiles share the LDS area of the exchange planes.  kFusedGen1
// scatters every bin to at most one tile position, kFusedGen2 to two.
constexpr int kFusedGen1 = 1;
constexpr int kFusedGen2 = 2;
__host__ __device__ constexpr bool fused_general(int FL) { return FL == kFusedGen1 || FL == kFusedGen2; }
__host__ __device__ constexpr int fused_targets(int FL) { return FL == kFusedGen2 ? 2 : 1; }
template <>
struct FusedCfg<kFusedGen1> {
  static constexpr int RF = 1, TF = 1, RF1 = 1, RF2 = 1, STRIDE = 0;
};
template <>
struct FusedCfg<kFusedGen2> {
  static constexpr int RF = 1, TF = 1, RF1 = 1, RF2 = 1, STRIDE = 0;
};
// Tile layout of an FL-point channel in the general fused stage: FL =
// RF1 x RF2 with 16 points per thread (FL / 16 threads), one thread for 8
// and 16 points.
template <int FL>
struct GenTileCfg {
  static constexpr int RF = FL < 16 ? FL : 16;
  static constexpr int TF = FL / RF;
  static constexpr int RF1 = RF;
  static constexpr int RF2 = FL / RF;
  static constexpr int STRIDE = FL + (FL >> 5);
};

template <int M, int R, int R1, int R2, int FL>
struct FwdRealCfg {
  static constexpr int T = M / R;
  static constexpr int W = T / 64;  // waves
  static constexpr int P3 = R1 * R2;
  static_assert(R1 * R2 * R == M, "radix plan must cover M");
  static_assert(P3 == T, "last pass must have one butterfly per thread");
  static_assert(R % 8 == 0, "slots are handled in groups of 8");
  static_assert((M / 4) % T == 0, "block 0's history seam must fall between points");
  static_assert(T >= 64 && T <= 1024, "one wave minimum, 1024 threads maximum");
  static constexpr bool SPLIT = lds_padded_size(M) * 8 > kLdsBudgetBytes;
  static constexpr int PLANE_FLOATS = SPLIT ? lds_padded_size(M) : 2 * lds_padded_size(M);
  static constexpr int TILE_FLOATS = FL > kFusedGen2 ? (T / FusedCfg<FL>::TF) * FusedCfg<FL>::STRIDE * 2 : 0;
  static constexpr int LDS_FLOATS = PLANE_FLOATS > TILE_FLOATS ? PLANE_FLOATS : TILE_FLOATS;
  // After the main area: the pass-3 outputs of lanes 0 and 32 of wave 0
  // (2 R complex), which are their own mirrors.
  static constexpr int STASH_FLOATS = 4 * R;
  // One complex slot that bins outside every fused channel are scattered to.
  static constexpr int DUMMY_FLOATS = 2;
  static constexpr int LDS_DUMMY = (LDS_FLOATS + STASH_FLOATS) / 2;  // cpx index
  static_assert(LDS_DUMMY < 65536, "scatter indices are 16-bit");
  // Welch items whose neighbours are not reachable by DPP (edge_items).
  static constexpr int EDGE_ITEMS = W * 4 * R + R + 1;
  static constexpr int EDGE_PER_THREAD = (EDGE_ITEMS + T - 1) / T;
  static constexpr int EDGE_RING = W * kEdgeLanes * R;  // cpx per workgroup
};

// The bins whose Welch neighbours lie outside their wave: per wave the
// lanes 0, 31, 32, 63 of every slot, wave 0's lane 33, and the Nyquist bin.
// Item i = {edge-ring index of the bin, of its left neighbour, of its right
// neighbour, partial-row index | conj-left << 28 | conj-right << 29}, for
// the edge ring layout [wave][edge_slot(lane)][slot] and the last-pass lane
// map.  Written by the host into a table of W * 4 R + R + 1 entries.
inline void edge_items(int M, int R, int4* out) {
  const int P3 = M / R, W = P3 / 64;
  auto src = [&](int sv, int sln, int sq) { return (sv * kEdgeLanes + edge_slot(sln)) * R + sq; };
  const int n = W * 4 * R + R + 1;
  for (int item = 0; item < n; ++item) {
    int v, ln, q;
    if (item < W * 4 * R) {
      v = item / (4 * R);
      const int which = (item / R) % 4;
      q = item % R;
      ln = which == 0 ? 0 : (which == 1 ? 31 : (which == 2 ? 32 : 63));
    } else if (item < W * 4 * R + R) {
      v = 0;
      ln = 33;
      q = item - W * 4 * R;
    } else {
      v = 0;
      ln = -1;
      q = 0;
    }
    int io, il, ir, ridx, fl = 0;
    if (ln < 0) {  // Nyquist: right neighbour X[M+1] = conj X[M-1]
      io = kEdgeNyq * R;
      il = src(0, 33, 0);
      ir = il;
      fl = 2;
      ridx = M;
    } else {
      io = src(v, ln, q);
      ridx = q * P3 + 64 * v + ln;
      if (ln == 0) {
        ir = src(v, 1, q);
        if (v > 0) il = src(v - 1, 31, q);
        else if (q > 0) il = src(0, 33, R - q);
        else {
          il = src(0, 1, 0);
          fl = 1;  // X[-1] = conj X[1]
        }
      } else if (ln == 31) {
        il = src(v, 30, q);
        ir = (v < W - 1) ? src(v + 1, 0, q) : src(0, 32, R - 1 - q);
      } else if (ln == 32) {
        if (v > 0) {
          il = src(v, 33, q);
          ir = src(v - 1, 63, q);
        } else {
          il = src(W - 1, 31, R - 1 - q);
          ir = src(W - 1, 63, q);
        }
      } else if (ln == 63) {
        il = (v < W - 1) ? src(v + 1, 32, q) : src(0, 32, q);
        ir = src(v, 62, q);
      } else {  // wave 0, lane 33: bin P3 - 1 + P3 (R-1-q)
        il = src(0, 34, q);
        ir = q > 0 ? src(0, 0, R - q) : kEdgeNyq * R;
      }
    }
    out[item] = make_int4(io, il, ir, ridx | (fl << 28));
  }
}

// Lane map of the last forward pass: lane t = 64 v + l of a T = P3 thread
// workgroup holds, in slot q, bin pi + P3 q (l < 32, pi = 32 v + l) or bin
// ksrc + P3 (R-1-q) (l >= 32, ksrc = P3 - pi, P3/2 for pi = 0).  Welch
// partial rows are stored lane-major: bin k at q T + t, the Nyquist bin at M.
__host__ __device__ inline int welch_row_index(int k, int M, int R) {
  const int P3 = M / R;
  if (k >= M) return M;
  const int r = k & (P3 - 1), kq = k / P3;
  int pi, q, hi;
  if (r < P3 / 2) {
    pi = r;
    q = kq;
    hi = 0;
  } else {
    pi = r == P3 / 2 ? 0 : P3 - r;
    q = R - 1 - kq;
    hi = 32;
  }
  return q * P3 + 64 * (pi >> 5) + hi + (pi & 31);
}

// Bin held at partial-row position i (the inverse of welch_row_index).
__host__ __device__ inline int welch_row_bin(int i, int M, int R) {
  const int P3 = M / R;
  if (i >= M) return M;
  const int q = i / P3, t = i % P3;
  const int pi = 32 * (t >> 6) + (t & 31);
  if ((t & 63) < 32) return pi + P3 * q;
  return (pi == 0 ? P3 / 2 : P3 - pi) + P3 * (R - 1 - q);
}

// DPP wave shifts (gfx9 wave_shr:1 / wave_shl:1): the value of lane l-1
// (resp. l+1); lane 0 (resp. 63) keeps its own value.
__device__ __forceinline__ float wave_shr1(float v) {
  return __int_as_float(
      __builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), 0x138, 0xF, 0xF, false));
}
__device__ __forceinline__ float wave_shl1(float v) {
  return __int_as_float(
      __builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), 0x130, 0xF, 0xF, false));
}
// Same shifts where lane 0 (resp. 63) may read anything: one v_mov_dpp.
__device__ __forceinline__ float wave_shr1_any(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x138, 0xF, 0xF, true));
}
__device__ __forceinline__ float wave_shl1_any(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x130, 0xF, 0xF, true));
}

// base[idx] with a 32-bit byte offset, so the access can use a scalar base
// register plus one VGPR offset instead of 64-bit address arithmetic.
template <class T>
__device__ __forceinline__ T& at32(T* base, unsigned idx) {
  return *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + idx * (unsigned)sizeof(T));
}

// 4 |c0 - (cm + cp) / 2|^2, the Welch term at the scale of the forward
// kernel's partial rows.
__device__ __forceinline__ float welch_term4(cpx c0, cpx cm, cpx cp) {
  const float cr = __builtin_fmaf(2.f, c0.x, -(cm.x + cp.x));
  const float ci = __builtin_fmaf(2.f, c0.y, -(cm.y + cp.y));
  return __builtin_fmaf(ci, ci, cr * cr);
}

// Outputs of the channel kernels (not the forward kernel's fused stage) are
// written once and not read again by the path: stored non-temporally, so
// they do not push the spectrum rows, which the channel gathers read many
// times over, out of L2.  (In the fused stage, whose gathers come from LDS,
// non-temporal stores were slower.)
__device__ __forceinline__ void store_out(cpx* p, cpx v) {
  __builtin_nontemporal_store(v.x, &p->x);
  __builtin_nontemporal_store(v.y, &p->y);
}

// Zero up to four device buffers (32-bit words, then the 1-3 tail bytes;
// allocations start 256-byte aligned).  A kernel rather than
// hipMemsetAsync, so that a caller may capture the call into a graph.
struct ZeroParams {
  unsigned char* p[4];
  size_t bytes[4];
};
__global__ void zero_buffers_kernel(ZeroParams P) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  const size_t i0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (int b = 0; b < 4; ++b) {
    unsigned char* p = P.p[b];
    if (!p) continue;
    const size_t nw = P.bytes[b] / 4;
    for (size_t i = i0; i < nw; i += stride) reinterpret_cast<uint32_t*>(p)[i] = 0u;
    for (size_t i = 4 * nw + i0; i < P.bytes[b]; i += stride) p[i] = 0;
  }
}

// Symmetric in the two neighbours.
__device__ __forceinline__ float welch_term(cpx c0, cpx cm, cpx cp) {
  const float cr = c0.x - 0.5f * (cm.x + cp.x);
  const float ci = c0.y - 0.5f * (cm.y + cp.y);
  return cr * cr + ci * ci;
}

// dB quantiser of one accumulated bin, spectrum.rs:128-152 in f32:
// dB = 10 log10(acc) - 10 log10(n); U8 = clamp(2 dB + 250, 0, 255),
// U16 = clamp(20 dB + 4000, 0, 4095) big-endian; NaN and -inf clamp to 0,
// the conversion truncates.
__device__ __forceinline__ void quantize_bin(unsigned char* q, int qfmt, long long idx, float acc, int n) {
  const float db = log10f(acc) * 10.0f + log10f((float)n) * -10.0f;
  if (qfmt == 1) {
    const float v = fminf(fmaxf(db * 2.0f + 250.0f, 0.0f), 255.0f);
    q[idx] = (unsigned char)(unsigned)v;
  } else {
    const float v = fminf(fmaxf(db * 20.0f + 4000.0f, 0.0f), 4095.0f);
    const unsigned o = (unsigned)v;
    q[2 * idx] = (unsigned char)(o >> 8);
    q[2 * idx + 1] = (unsigned char)(o & 0xFF);
  }
}


// Workgroups [w0, w1] of a forward launch over F.nblocks blocks that
// processed blocks of Welch group gi (blk0_mod = stream position of the
// launch's block 0 mod A): the w with floor(w nb / G) <= b < floor((w+1) nb / G).
__device__ __forceinline__ void welch_group_wgs(const FinParams& F, int blk0_mod, int gi, int* w0, int* w1) {
  const int g_lo = gi * F.A - blk0_mod, g_hi = g_lo + F.A;
  const int b_lo = g_lo > 0 ? g_lo : 0;
  const int b_hi = g_hi < F.nblocks ? g_hi : F.nblocks;
  const unsigned nb = (unsigned)F.nblocks, G = (unsigned)F.G;
  *w0 = (int)(((unsigned)(b_lo + 1) * G + nb - 1) / nb) - 1;
  *w1 = (int)(((unsigned)b_hi * G + nb - 1) / nb) - 1;
}

// s + the partial rows of workgroups wa..wb (in order) at positions
// i0 .. i0+3 (or position M alone) of group gi.
__device__ __forceinline__ float4 welch_rows_add(const FinParams& F, int blk0_mod, int gi, int i0, int wa, int wb,
                                                 float4 s) {
  const bool nyq = i0 == F.M;
  const unsigned nb = (unsigned)F.nblocks, G = (unsigned)F.G;
#pragma unroll 4
  for (int wg = wa; wg <= wb; ++wg) {
    const int lo = (int)((unsigned)wg * nb / G);
    const int slot = gi - (blk0_mod + lo) / F.A;
    const float* row = F.wpart + ((long long)wg * F.wslots + slot) * F.wstride;
    if (nyq) {
      s.x += row[i0];
    } else {
      const float4 v = *reinterpret_cast<const float4*>(row + i0);
      s.x += v.x;
      s.y += v.y;
      s.z += v.z;
      s.w += v.w;
    }
  }
  return s;
}

// The carry of a group begun in an earlier call (group 0 only).
__device__ __forceinline__ float4 welch_carry(const FinParams& F, int blk0_mod, int gi, int i0) {
  float4 s = {0.f, 0.f, 0.f, 0.f};
  if (gi == 0 && blk0_mod != 0) {
    if (i0 == F.M) s.x = F.carry_in[i0];
    else s = *reinterpret_cast<const float4*>(F.carry_in + i0);
  }
  return s;
}

// A finished sum: records acc / A (bins by welch_row_bin) of a complete
// group, the next carry of an incomplete one.
__device__ __forceinline__ void welch_emit(const FinParams& F, int blk0_mod, int gi, int i0, float4 s) {
  const bool nyq = i0 == F.M;
  if (gi * F.A - blk0_mod + F.A > F.nblocks) {
    if (nyq) F.carry_out[i0] = s.x;
    else *reinterpret_cast<float4*>(F.carry_out + i0) = s;
    return;
  }
  // (partial rows and carries hold 4x the sums: scaled back exactly)
  const float sv[4] = {0.25f * s.x, 0.25f * s.y, 0.25f * s.z, 0.25f * s.w};
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    if (nyq && j > 0) break;
    const long long o = (long long)gi * F.K + welch_row_bin(i0 + j, F.M, F.R);
    F.out[o] = sv[j] / (float)F.A;
    if (F.qfmt) quantize_bin(F.qout, F.qfmt, o, sv[j], F.A);
  }
}

// One thread's share of a Welch group gi of a launch whose first block is
// at stream position pos: partial-row positions i0 .. i0+3 (or the Nyquist
// position M alone, M % 4 == 0) summed over the workgroups that processed
// the group's blocks, plus the carry of a group begun in an earlier call.  A
// complete group is written as records (bins by welch_row_bin), an
// incomplete one as the next carry.
__device__ __forceinline__ void welch_finalize_quad(const FinParams& F, long long pos, int gi, int i0) {
  const int blk0_mod = (int)(pos % F.A);
  int w0, w1;
  welch_group_wgs(F, blk0_mod, gi, &w0, &w1);
  const float4 s = welch_rows_add(F, blk0_mod, gi, i0, w0, w1, welch_carry(F, blk0_mod, gi, i0));
  welch_emit(F, blk0_mod, gi, i0, s);
}

// One thread's part of a general fused tile (fcfb.rs:154-188): weight the
// gathered bins, inverse FFT of FL points in the tile's LDS region, keep
// [FL/8, 7FL/8).  The tile's threads lie in one wave and only they touch its
// region until the next block's first exchange barrier.
template <int FL>
__device__ __forceinline__ void gen_tile(const FwdParams& P, cpx* ldsc, const GenTile& gd, int ftt, bool act, int g) {
  using TC = GenTileCfg<FL>;
  constexpr int RF = TC::RF, TF = TC::TF;
  const float* wts = P.weights + gd.w_off;
  cpx* tl = ldsc + gd.lds;
  cpx gx[RF];
#pragma unroll
  for (int e = 0; e < RF; ++e) {
    const int n = st_in_index<FL, RF, TF, TC::RF1>(ftt, e);
    gx[e] = cscale(tl[st_in_addr<FL, RF, TF, TC::RF1>(ftt, e)], wts[(n + FL / 2) & (FL - 1)]);
  }
  st_butterflies<FL, RF, TF, TC::RF1, 1, +1>(gx, ftt, nullptr);
  if constexpr (TC::RF2 > 1) {
    tile_exchange<TF>(
        gx, reinterpret_cast<float*>(tl), [&](int e) { return st_out_addr<FL, RF, TF, TC::RF1, 1>(ftt, e); },
        [&](int e) { return st_in_addr<FL, RF, TF, TC::RF2>(ftt, e); });
    constexpr int ci = FL == 32 ? 0 : (FL == 64 ? 1 : (FL == 128 ? 2 : 3));
    st_butterflies<FL, RF, TF, TC::RF2, TC::RF1, +1>(gx, ftt, P.gtw[ci]);
  }
  if (act) {
    cpx* out = P.fout + gd.out_off + (long long)g * (3 * FL / 4);
    constexpr int PL = FL / TC::RF2;
#pragma unroll
    for (int e = 0; e < RF; ++e) {
      const int u = e / TC::RF2, q = e % TC::RF2;
      const int idx = ftt + TF * u + q * PL;
      if (idx >= FL / 8 && idx < (FL / 8) * 7) out[idx - FL / 8] = gx[e];
    }
  }
}

// Write X to one target of the general fused scatter (LDS cpx index in the
// low 15 bits, conjugate when bit 15 is set: get_bin's mirror).
__device__ __forceinline__ void gen_put(cpx* ldsc, unsigned entry, cpx X) {
  cpx v = X;
  v.y = __uint_as_float(__float_as_uint(X.y) ^ ((entry & 0x8000u) << 16));
  ldsc[entry & 0x7FFFu] = v;
}

// Persistent workgroups: workgroup w takes blocks [w nb / G, (w+1) nb / G)
// in order, so its Welch sums stay in a private partial row (slot = Welch
// group - first group of the workgroup) that is written with plain stores.
// The rows of every group are added in the next launch (workgroups that are
// done with their blocks take chunks of the previous call's groups) or by
// welch_finalize_kernel when the records are asked for first.
//
// After the last pass every lane owns R bins of the real spectrum: lower
// lanes pi + P3 q, upper lanes ksrc + P3 (R-1-q) (slot q).  A bin's
// neighbours k +- 1 sit in the adjacent lanes of the same slot, so Welch
// takes them by DPP wave shifts (lower lanes: left = lane-1, right =
// lane+1; upper lanes the other way round).  Lanes where that is wrong (the
// wave edges, and the wrap-around lanes of waves 0 and W-1) are finished
// after a barrier from the edge table.  Channels of the fused stage are
// gathered by scattering each owned bin into an LDS tile buffer through the
// bin -> position map.
// Waves per SIMD the compiler must leave room for (at most 512 / n VGPRs
// per wave): four for the 128- and 256-thread workgroups of M = 2048 and
// 4096 with s16 input, whose fused stages otherwise take a few registers
// more than 128 and so one workgroup per CU less than their LDS admits
// (with the f32 loader the same bound spills); one otherwise.
__host__ __device__ constexpr int fwd_min_waves(int M, bool f32) { return !f32 && M <= 4096 && M >= 2048 ? 4 : 1; }

template <int M, int R, int R1, int R2, int FL, bool F32>
__global__ __launch_bounds__(M / R) __attribute__((amdgpu_waves_per_eu(fwd_min_waves(M, F32))))
void fwd_real_kernel(FwdParams P) {
  using C = FwdRealCfg<M, R, R1, R2, FL>;
  using FC = FusedCfg<FL>;
  constexpr int T = C::T;
  constexpr int P3 = C::P3;
  constexpr bool SPLIT = C::SPLIT;
  __shared__ __align__(16) float lds[C::LDS_FLOATS + C::STASH_FLOATS + C::DUMMY_FLOATS + C::EDGE_PER_THREAD * T + 1];
  cpx* stash = reinterpret_cast<cpx*>(lds + C::LDS_FLOATS);
  float* eacc = lds + C::LDS_FLOATS + C::STASH_FLOATS + C::DUMMY_FLOATS;  // edge item sums
  int* ticket_slot = reinterpret_cast<int*>(eacc + C::EDGE_PER_THREAD * T);

  constexpr int NI = C::EDGE_PER_THREAD;
  const int G = gridDim.x, wg = blockIdx.x;
  const int lo = (int)((long long)wg * P.nblocks / G);
  const int hi = (int)((long long)(wg + 1) * P.nblocks / G);
  const long long block0 = *P.pos;
  // Welch group of each block, carried from block to block in scalars (no
  // 64-bit division per block): ph = stream block index mod A, wslot = its
  // group - the group of this workgroup's first block.
  int ph = (int)((block0 + lo) % P.A), wslot = 0;
  const float w = P.wscale;
  if (wg == 0) {
    if ((int)threadIdx.x < P.ncounters) P.chan_counters[threadIdx.x] = 0;
    if (threadIdx.x == 0 && P.pos_next) *P.pos_next = block0 + P.nblocks;
  }

  // Raw s16 pairs of the block about to be transformed: point n = b + j M/R1
  // of pass 1 is the pair at block sample 2n.  Rows are addressed as (uniform
  // row pointer) + lane offset.  The next block's pairs are loaded while the
  // current block's channel stage runs.
  // A pair is one 32-bit word of two s16 samples, or (F32) one float2.
  using Word = std::conditional_t<F32, float2, int>;
  constexpr int kSampleBytes = F32 ? 4 : 2;
  auto load_block = [&](int g, Word (&vin)[R]) {
    // Block 0 starts with the O = M/2 history samples (M/4 pairs).  Every
    // point's pair range [T u + j M/R1, + T) lies wholly on one side of that
    // seam (M/4 is a multiple of T), so each point reads from one of two
    // uniform row pointers, chosen at compile time.
    // (typed pointers, not integers: the loads keep the global address space)
    const Word* src = static_cast<const Word*>(P.src);
    const Word* blk = reinterpret_cast<const Word*>(static_cast<const char*>(P.src) +
                                                    (g == 0 ? 0 : (long long)g * P.H - P.O) * kSampleBytes);
    const Word* lo_row = g == 0 ? static_cast<const Word*>(P.hist) : blk;
    const Word* hi_row = g == 0 ? src : blk;
    const int hi_off = g == 0 ? -(M / 4) : 0;
#pragma unroll
    for (int e = 0; e < R; ++e) {
      const int u = e / R1, j = e % R1;
      const int base = T * u + j * (M / R1);
      vin[e] = base < M / 4 ? lo_row[base + (int)threadIdx.x] : hi_row[base + hi_off + (int)threadIdx.x];
    }
  };
  // 1024-thread workgroups have no registers to spare for the prefetch
  // (the prefetched words would be spilled, which waits for the loads).
  constexpr bool PREFETCH = T < 1024;
  // The 1024-thread workgroups load the next block's input late in the
  // block, once the spectrum registers are dead: after the edge terms and
  // before the channel tiles (L = 256 stage), after the tiles (general stage
  // with one target per bin; with two, the registers would spill), or while
  // the edge terms wait for their loads (no fused channels).
  constexpr bool PF_LATE = !PREFETCH && FL != kFusedGen2;
  Word vin[R];
  if ((PREFETCH || PF_LATE) && lo < hi) load_block(lo, vin);

  for (int g = lo; g < hi; ++g) {
    if constexpr (!PREFETCH && !PF_LATE) load_block(g, vin);
    // Per-block values that do not depend on g (lane maps, table reads) are
    // recomputed and reloaded every block: kept live across the loop they
    // would cost more registers than the 1024-thread workgroups have.
    // (The opaque value is an integer offset, not the pointers themselves:
    // a pointer passed through asm loses its global address space and its
    // loads become flat loads, which also count against the LDS wait
    // counter.)
    int t = threadIdx.x, zero = 0;
    asm volatile("" : "+v"(t));
    asm volatile("" : "+s"(zero));
    const cpx* tw2 = P.tw2 + zero;
    const cpx* tw3 = P.tw3 + zero;
    const cpx* twu = P.twu + zero;

    // Lane map of the last pass: lane l < 32 of wave v takes sub-transform
    // k = pi = 32 v + l; lane l + 32 takes its mirror P3 - pi (P3/2 for pi = 0).
    const int wave = __builtin_amdgcn_readfirstlane(t >> 6), lane = t & 63;
    const bool upper = lane >= 32;
    const int pi = 32 * wave + (lane & 31);
    const int ksrc = upper ? (pi == 0 ? P3 / 2 : P3 - pi) : pi;
    const int kt = (upper && pi == 0) ? P3 / 2 : pi;
    const float sg = upper ? -1.f : 1.f;

    cpx x[R];
#pragma unroll
    for (int e = 0; e < R; ++e) {
      if constexpr (F32) {
        x[e].x = vin[e].x * w;
        x[e].y = vin[e].y * w;
      } else {
        // the window scale w is applied in the untangle constants
        x[e].x = (float)(int16_t)vin[e];
        x[e].y = (float)(vin[e] >> 16);
      }
    }
    st_butterflies<M, R, T, R1, 1, -1>(x, t, nullptr);

    // The pass-2 and pass-3 twiddles are loaded inside the exchanges before
    // them, while half of the data waits in LDS (for pass 2 when each
    // thread has one butterfly, R2 = R).
    PassTwiddles<R> ptw;
    lds_exchange<R, SPLIT>(
        x, lds, [&](int e) { return st_out_addr<M, R, T, R1, 1>(t, e); },
        [&](int e) { return st_in_addr<M, R, T, R2>(t, e); },
        [&]() {
          if constexpr (R2 == R) ptw = load_pass_twiddles<R, R1>(tw2, t & (R1 - 1));
        });
    // Outputs q >= R2/2 of pass 2 are the pass-3 inputs of the upper lanes
    // (partial-transform index k + R1 q >= P3/2), which take them conjugated.
    if constexpr (R2 == R) {
      apply_loaded_twiddles<R>(x, ptw);
      dft<R, -1, true>(x);
    } else {
      st_butterflies<M, R, T, R2, R1, -1, true>(x, t, tw2);
    }

    // Pass 3: an upper lane must end with Z[ksrc + P3 (R-1-q)] in slot q,
    // i.e. the DFT of its twiddled inputs in reversed order.  That is
    // conj(DFT(conj(in_m) w^{m kt})) with kt = P3 - ksrc, so upper lanes read
    // in natural order, conjugate, use the same table rows as lower lanes
    // (column kt), and conjugate the DFT outputs.
    lds_exchange<R, SPLIT>(
        x, lds, [&](int e) { return st_out_addr<M, R, T, R2, R1>(t, e); },
        [&](int e) { return lds_pad(ksrc) + pad_off(e * P3); },
        [&]() { ptw = load_pass_twiddles<R, P3>(tw3, kt); });
    // Welch partial row of this block's group (lane-major: slot q of lane t
    // at q T + t), and the first slot group's partial sums and scatter
    // indices, loaded ahead of the pass-3 DFT.  The row was written by this
    // workgroup's previous block: it is read at agent scope, which bypasses
    // the CU's L1 and so sees those stores.
    const int slot = wslot;
    const bool fresh = (g == lo) || (ph == 0);
    const __amdgpu_buffer_rsrc_t prow = __builtin_amdgcn_make_buffer_rsrc(
        P.wpart + ((long long)wg * P.wslots + slot) * P.wstride, (short)0, (M + 1) * (int)sizeof(float),
        kBufferRsrcWord3);
    // Lanes whose Welch neighbours are not both reachable by DPP accumulate
    // wrong terms in their partial-row entries here; the edge items
    // overwrite those entries at the end of each Welch group.
    const unsigned tvoff = (unsigned)t * (unsigned)sizeof(float);
    float pv[R];
    uint4 sw[2];
    uint4 swg[fused_targets(FL)][2];
    auto load_group = [&](auto jc) {
      constexpr int j = decltype(jc)::value;
      if (fresh) {
#pragma unroll
        for (int q = 8 * j; q < 8 * j + 8; ++q) pv[q] = 0.f;
      } else {
#pragma unroll
        for (int q = 8 * j; q < 8 * j + 8; ++q)
          pv[q] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(prow, tvoff, q * T * (int)sizeof(float), kCpolSc1));
      }
      if constexpr (FL > kFusedGen2) sw[j & 1] = P.fscat[j * T + t];
      if constexpr (fused_general(FL)) {
#pragma unroll
        for (int d = 0; d < fused_targets(FL); ++d) swg[d][j & 1] = P.fscat[(d * (R / 8) + j) * T + t];
      }
    };
    load_group(std::integral_constant<int, 0>{});

    apply_loaded_twiddles<R>(x, ptw);
    dft<R, -1>(x);
#pragma unroll
    for (int e = 0; e < R; ++e) x[e].y *= sg;
    // Lower lanes now hold Z[pi + P3 q] in slot q, upper lanes
    // Z[ksrc + P3 (R-1-q)], so slot q of lane l and lane l^32 are mirrors
    // (except the pair pi = 0, whose lanes are their own mirrors).  Their
    // slots go to a private LDS stash so every slot register dies once it is
    // untangled.
    if (wave == 0 && (lane & 31) == 0) {
#pragma unroll
      for (int q = 0; q < R; ++q) stash[(upper ? R : 0) + q] = x[q];
    }
    __syncthreads();  // every lane is done reading pass-3 inputs from LDS

    const int es = edge_slot(lane);
    // Edge-lane dump into this workgroup's edge ring through a buffer
    // resource: the other lanes store out of range, which is dropped.
    constexpr int ecount = C::EDGE_RING;
    cpx* ering = P.edump + (long long)wg * ecount;
    const __amdgpu_buffer_rsrc_t edr =
        __builtin_amdgcn_make_buffer_rsrc(ering, (short)0, ecount * (int)sizeof(cpx), kBufferRsrcWord3);
    const unsigned evoff = (unsigned)((wave * kEdgeLanes + (es < 0 ? 0 : es)) * R * (int)sizeof(cpx));
    cpx* spec = P.spec + (long long)g * P.spec_stride;
    const int kb = upper ? ksrc + P3 * (R - 1) : pi;
    const int kstep = upper ? -P3 : P3;
    // Both lanes of a mirror pair form E and O of the lower bin k from
    // A = Z[k] (lower lane's slot) and B = Z[M-k] (upper lane's slot):
    //   X[k] = E + W^k O,  X[M-k] = conj(E - W^k O),
    // so with u = (s/2) W^k (s = +1 lower, -1 upper) every lane computes
    //   X.x = (A.x+B.x)/2 + u.x (A.y+B.y) + u.y (A.x-B.x)
    //   X.y = s ((A.y-B.y)/2 + u.y (A.y+B.y) - u.x (A.x-B.x)).
    // W^k = W^kt (table) * W^(P3 q) (constant per slot).
    const cpx wb = twu[kt];
    // (times the window scale, for the s16 loader, whose samples enter unscaled)
    const float hw = F32 ? 0.5f : 0.5f * w;
    const cpx wbs = {hw * sg * wb.x, hw * sg * wb.y};
    const cpx wh = {hw * wb.x, hw * wb.y};
    const bool lane0 = lane == 0, lane32 = lane == 32;
    // byte address of the mirror lane for ds_bpermute
    const int mirror_addr = (lane ^ 32) << 2;

    // Untangle; Welch terms into the partial row (not for edge lanes); edge
    // dump; channel scatter into the tile buffers.
    static_for<0, R>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      if constexpr (q % 8 == 0 && q + 8 < R) load_group(std::integral_constant<int, q / 8 + 1>{});
      // V: this lane's value, Pm: its mirror's, fetched through the LDS
      // crossbar (ds_bpermute issues no VALU work).  A = Z[k] is V on lower
      // lanes and Pm on upper ones, B = Z[M-k] the other, so with
      // D = V - Pm = s (A - B) the untangle below needs no per-slot sign.
      const cpx V = x[q];
      cpx Pm;
      Pm.x = __int_as_float(__builtin_amdgcn_ds_bpermute(mirror_addr, __float_as_int(V.x)));
      Pm.y = __int_as_float(__builtin_amdgcn_ds_bpermute(mirror_addr, __float_as_int(V.y)));
      if (wave == 0) {
        // lane 0 (k = 0): mirror of Z[P3 q] is its own slot (R - q) mod R;
        // lane 32 (k = P3/2): mirror of its slot q is its own slot R-1-q.
        const cpx s0 = stash[(R - q) & (R - 1)];
        const cpx s32 = stash[R + R - 1 - q];
        Pm = lane0 ? s0 : (lane32 ? s32 : Pm);
      }
      const float S1 = V.x + Pm.x, S2 = V.y + Pm.y, D1 = V.x - Pm.x, D2 = V.y - Pm.y;
      constexpr float cx = float(ct_cos2pi(-q, 2 * R));
      constexpr float cy = float(ct_sin2pi(-q, 2 * R));
      // u = (s/2) W^k (s = +1 lower, -1 upper lanes); ua = Re u, ub = s Im u
      // (slot constants of 0 and 1 are folded: x * 0 is not 0 in IEEE
      // arithmetic, so the compiler keeps such products)
      float ua, ub;
      if constexpr (q == 0) {
        ua = wbs.x;
        ub = wh.y;
      } else if constexpr (2 * q == R) {
        ua = wbs.y;  // cx = 0, cy = -1
        ub = -wh.x;
      } else {
        ua = wbs.x * cx - wbs.y * cy;
        ub = wh.x * cy + wh.y * cx;
      }
      cpx X;
      X.x = hw * S1 + (ua * S2 + ub * D1);
      X.y = hw * D2 + (ub * S2 - ua * D1);
      if (P.keep_spec) at32(spec, (unsigned)(kb + q * kstep)) = X;
      if (es >= 0)
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, X), edr, evoff + q * (int)sizeof(cpx), 0, 0);
      if constexpr (q == 0) {
        if (t == 0) {  // Nyquist bin M = E - O of Z[0]
          const cpx XN = {hw * (S1 - S2), hw * (D2 + D1)};
          if (P.keep_spec) spec[M] = XN;
          ering[kEdgeNyq * R] = XN;
          if constexpr (FL > kFusedGen2) reinterpret_cast<cpx*>(lds)[P.fscat_nyq] = XN;
          if constexpr (fused_general(FL)) {
            gen_put(reinterpret_cast<cpx*>(lds), (unsigned)P.fscat_nyq, XN);
            if constexpr (fused_targets(FL) > 1) gen_put(reinterpret_cast<cpx*>(lds), (unsigned)P.fscat_nyq2, XN);
          }
        }
      }
      // Welch neighbours by DPP: lanes l-1 and l+1 (left/right for lower
      // lanes, right/left for upper lanes; the term is symmetric).
      // The partial rows hold 4 |X - (l + r)/2|^2 = |l + r - 2 X|^2 per bin
      // (a power-of-two scaling, removed exactly when records are formed).
      float dx = wave_shr1_any(X.x) + wave_shl1_any(X.x);
      float dy = wave_shr1_any(X.y) + wave_shl1_any(X.y);
      dx = __builtin_fmaf(-2.f, X.x, dx);
      dy = __builtin_fmaf(-2.f, X.y, dy);
      // (partial rows hold 4x the sums: one product fewer per bin)
      const float acc = __builtin_fmaf(dy, dy, __builtin_fmaf(dx, dx, pv[q]));
      __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(acc), prow, tvoff, q * T * (int)sizeof(float), 0);
      if constexpr (fused_general(FL)) {
        // general scatter: every target of this bin (a dummy slot when none)
#pragma unroll
        for (int d = 0; d < fused_targets(FL); ++d) {
          const uint4 w4 = swg[d][(q / 8) & 1];
          const unsigned wd = (q & 7) < 2 ? w4.x : (q & 7) < 4 ? w4.y : (q & 7) < 6 ? w4.z : w4.w;
          gen_put(reinterpret_cast<cpx*>(lds), (q & 1) ? (wd >> 16) : (wd & 0xFFFFu), X);
        }
      }
      if constexpr (FL > kFusedGen2) {
        // Channel scatter (fcfb.rs:164-172): bin first + j feeds input
        // position (j + FL/2) mod FL of its channel's tile; the index table
        // sends other bins to a dummy slot.
        const uint4 w4 = sw[(q / 8) & 1];
        const unsigned wd = (q & 7) < 2 ? w4.x : (q & 7) < 4 ? w4.y : (q & 7) < 6 ? w4.z : w4.w;
        const unsigned li = (q & 1) ? (wd >> 16) : (wd & 0xFFFFu);
        reinterpret_cast<cpx*>(lds)[li] = X;
      }
      if constexpr ((q & 7) == 7) __builtin_amdgcn_sched_barrier(0);
    });

    // The next block's raw input is loaded while this block's channels run.
    __builtin_amdgcn_sched_barrier(0);
    if (PREFETCH && g + 1 < hi) load_block(g + 1, vin);
    __builtin_amdgcn_sched_barrier(0);

    // Edge items: Welch terms of the bins whose neighbours sit in other
    // waves, from the edge ring after a barrier.  Each item's thread keeps
    // the item's sum in its LDS slot and writes it over the item's
    // partial-row entry when the Welch group ends.  The item table is read
    // before the barrier, the ring after it, and the terms are formed after
    // the fused channel stage so that the loads overlap it.
    const bool slot_end = (g + 1 == hi) || (ph + 1 == P.A);
    int4 eit[NI];
    u32x2 ex[NI], el[NI], er[NI];
    // Threads past the last item repeat the last item's loads and keep
    // their (unused) sums in spare LDS slots; only their row store is skipped.
    auto item_ok = [&](int i) { return i * T + T <= C::EDGE_ITEMS || t + i * T < C::EDGE_ITEMS; };
    auto edge_prep = [&]() {
#pragma unroll
      for (int i = 0; i < NI; ++i) eit[i] = P.eitems[item_ok(i) ? t + i * T : C::EDGE_ITEMS - 1];
    };
    auto edge_issue = [&]() {
#pragma unroll
      for (int i = 0; i < NI; ++i) {
        ex[i] = __builtin_amdgcn_raw_buffer_load_b64(edr, (unsigned)eit[i].x * (unsigned)sizeof(cpx), 0, kCpolSc1);
        el[i] = __builtin_amdgcn_raw_buffer_load_b64(edr, (unsigned)eit[i].y * (unsigned)sizeof(cpx), 0, kCpolSc1);
        er[i] = __builtin_amdgcn_raw_buffer_load_b64(edr, (unsigned)eit[i].z * (unsigned)sizeof(cpx), 0, kCpolSc1);
      }
    };
    auto edge_finish = [&]() {
#pragma unroll
      for (int i = 0; i < NI; ++i) {
        const int fl = eit[i].w >> 28;
        const cpx X = __builtin_bit_cast(cpx, ex[i]);
        cpx L = __builtin_bit_cast(cpx, el[i]), Rn = __builtin_bit_cast(cpx, er[i]);
        if (fl & 1) L.y = -L.y;
        if (fl & 2) Rn.y = -Rn.y;
        const float term = welch_term4(X, L, Rn);
        const float acc = fresh ? term : eacc[t + i * T] + term;
        eacc[t + i * T] = acc;
        if (slot_end && item_ok(i))
          __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(acc), prow,
                                                (unsigned)(eit[i].w & 0x0FFFFFFF) * (unsigned)sizeof(float), 0, 0);
      }
    };
    // Branch-free prefetch of the next block (the last block loads its own
    // input once more): with a skipped load on one path the waits for the
    // loads before it would count on the shorter path.
    auto prefetch_next = [&]() {
      __builtin_amdgcn_sched_barrier(0);
      load_block(g + 1 < hi ? g + 1 : g, vin);
      __builtin_amdgcn_sched_barrier(0);
    };
    if constexpr (FL == 0) {
      edge_prep();
      __syncthreads();  // edge ring complete
      edge_issue();
      if constexpr (PF_LATE) prefetch_next();
      edge_finish();
    }

    // General fused stage: the thread's tile slot of every round (the
    // size class is the same for every lane of a wave).
    if constexpr (fused_general(FL)) {
      cpx* ldsc = reinterpret_cast<cpx*>(lds);
      edge_prep();
      __syncthreads();  // tile buffers and edge ring complete
      edge_issue();
      for (int r = 0; r < P.grounds; ++r) {
        const int sl = P.gslots[r * T + t];
        const GenTile gd = P.gtiles[r * T + t];
        // waves of idle slots only (padding) skip the round: the tile code
        // has no workgroup barrier
        if (__ballot(sl >= 0) == 0) continue;
        const int cls = __builtin_amdgcn_readfirstlane(sl & 7), ftt = (sl >> 3) & 31;
        const bool act = sl >= 0;
        switch (cls) {
          case 0: gen_tile<8>(P, ldsc, gd, ftt, act, g); break;
          case 1: gen_tile<16>(P, ldsc, gd, ftt, act, g); break;
          case 2: gen_tile<32>(P, ldsc, gd, ftt, act, g); break;
          case 3: gen_tile<64>(P, ldsc, gd, ftt, act, g); break;
          case 4: gen_tile<128>(P, ldsc, gd, ftt, act, g); break;
          default: gen_tile<256>(P, ldsc, gd, ftt, act, g); break;
        }
      }
      edge_finish();
      // (after the tiles: their loads, issued behind a prefetch, would be
      // waited for together with it)
      if constexpr (PF_LATE) prefetch_next();
    }

    // Fused channel stage (fcfb.rs:154-188): weight the gathered bins,
    // inverse FFT in the tile's LDS region, keep [FL/8, 7FL/8).
    if constexpr (FL > kFusedGen2) {
      constexpr int TF = FC::TF, RF = FC::RF;
      const int fc = t / TF, ftt = t % TF;
      const bool fact = fc < P.fnch;
      edge_prep();
      const ChanDesc fd = P.fchans[fact ? fc : 0];
      const float* wts = P.weights + fd.w_off;
      float wv[RF];
#pragma unroll
      for (int e = 0; e < RF; ++e) {
        const int n = st_in_index<FL, RF, TF, FC::RF1>(ftt, e);
        wv[e] = wts[(n + FL / 2) & (FL - 1)];
      }
      __syncthreads();  // tile buffers and edge ring complete
      edge_issue();
      int zero2 = 0;
      asm volatile("" : "+s"(zero2));
      const cpx* ftw2 = P.ftw2 + zero2;
      static_assert(RF == FC::RF2, "one second-pass butterfly per thread");
      // (loaded before the next block's input: the loads are waited for in order)
      const PassTwiddles<FC::RF2> ftw = load_pass_twiddles<FC::RF2, FC::RF1>(ftw2, ftt & (FC::RF1 - 1));
      if constexpr (PF_LATE) {
        edge_finish();
        __builtin_amdgcn_sched_barrier(0);
        // Branch-free (the last block loads its own input once more): with
        // a skipped load on one path, the waits for the loads above would
        // count on the shorter path and wait for this whole prefetch, and
        // the registers would stay live across the loop.
        load_block(g + 1 < hi ? g + 1 : g, vin);
        __builtin_amdgcn_sched_barrier(0);
      }
      float* tl = lds + fc * FC::STRIDE * 2;
      cpx gx[RF];
#pragma unroll
      for (int e = 0; e < RF; ++e)
        gx[e] = cscale(reinterpret_cast<const cpx*>(tl)[st_in_addr<FL, RF, TF, FC::RF1>(ftt, e)], wv[e]);
      st_butterflies<FL, RF, TF, FC::RF1, 1, +1>(gx, ftt, nullptr);
      // the tile's TF = 16 threads lie in one wave and only they touch its
      // region until the next block's first exchange barrier
      tile_exchange<TF>(
          gx, tl, [&](int e) { return st_out_addr<FL, RF, TF, FC::RF1, 1>(ftt, e); },
          [&](int e) { return st_in_addr<FL, RF, TF, FC::RF2>(ftt, e); });
      apply_loaded_twiddles<FC::RF2>(gx, ftw);
      dft<FC::RF2, +1>(gx);
      constexpr int PL = FL / FC::RF2;
      if (fact) {
        cpx* out = P.fout + fd.out_off + (long long)g * (3 * FL / 4);
#pragma unroll
        for (int e = 0; e < RF; ++e) {
          const int u = e / FC::RF2, q = e % FC::RF2;
          const int idx = ftt + TF * u + q * PL;
          if (idx >= FL / 8 && idx < (FL / 8) * 7) out[idx - FL / 8] = gx[e];
        }
      }
      if constexpr (!PF_LATE) edge_finish();
    }
    if (++ph == P.A) {
      ph = 0;
      ++wslot;
    }
  }

  // The next call's input history: block 0 (this workgroup's first block)
  // was the only reader of the current one.
  if (wg == 0) {
    for (int i = (int)threadIdx.x; i < P.hist_words; i += T) P.hist_dst[i] = P.hist_src[i];
  }

  // The previous call's Welch records: its partial rows are complete (they
  // were written by the previous launch), so the workgroups that finish
  // their blocks first take them, P.fin.take chunks (of 4 T / S positions of
  // one group) at a time.  Every workgroup ends with exactly one failed take,
  // so the take numbered takes + G - 1 is the launch's last and resets the
  // counter.
  if (P.fin.nblocks > 0) {
    // (its position slot is this launch's pos_next, already rewritten)
    const long long fpos = block0 - P.fin.nblocks;
    const int ngroups = (int)(fpos % P.fin.A + P.fin.nblocks - 1) / P.fin.A + 1;
    // chunks of 4 T / S positions of one group (M = R T positions + the
    // Nyquist position); S teams sum S ranges of the group's workgroups,
    // then team 0 adds the teams' sums in order (S = 1: one team)
    const int S = P.fin.split, TS = T / S;
    const int CPG = (R / 4) * S;
    const int nchunks = ngroups * CPG;
    const int Q = P.fin.take, ntakes = (nchunks + Q - 1) / Q;
    const int blk0_mod = (int)(fpos % P.fin.A);
    const int team = (int)threadIdx.x / TS, u = (int)threadIdx.x % TS;
    float4* stage = reinterpret_cast<float4*>(lds);
    for (;;) {
      __syncthreads();  // the previous ticket (and stage) has been read
      if (threadIdx.x == 0) *ticket_slot = atomicAdd(P.fin_ticket, 1);
      __syncthreads();
      const int ticket = *ticket_slot;
      if (ticket >= ntakes) {
        if (threadIdx.x == 0 && ticket == ntakes + G - 1) *P.fin_ticket = 0;
        break;
      }
      const int cend = min(nchunks, (ticket + 1) * Q);
      for (int ch = ticket * Q; ch < cend; ++ch) {
        if (S > 1 && ch > ticket * Q) __syncthreads();  // the previous chunk's stage has been read
        const int gi = ch / CPG, c = ch % CPG;
        const int i0 = 4 * (c * TS + u);
        if (S == 1) {
          welch_finalize_quad(P.fin, fpos, gi, i0);
        } else {
          int w0, w1;
          welch_group_wgs(P.fin, blk0_mod, gi, &w0, &w1);
          const int nw = w1 - w0 + 1;
          const int wa = w0 + nw * team / S, wb = w0 + nw * (team + 1) / S - 1;
          const float4 z = {0.f, 0.f, 0.f, 0.f};
          stage[team * TS + u] =
              welch_rows_add(P.fin, blk0_mod, gi, i0, wa, wb, team == 0 ? welch_carry(P.fin, blk0_mod, gi, i0) : z);
          // the last chunk of a group also carries the Nyquist position
          float* nstage = reinterpret_cast<float*>(stage + T);
          const bool last = c == CPG - 1;
          if (last && u == 0)
            nstage[team] =
                welch_rows_add(P.fin, blk0_mod, gi, M, wa, wb, team == 0 ? welch_carry(P.fin, blk0_mod, gi, M) : z).x;
          __syncthreads();
          if (team == 0) {
            float4 acc = stage[u];
            for (int k = 1; k < S; ++k) {
              const float4 v = stage[k * TS + u];
              acc.x += v.x;
              acc.y += v.y;
              acc.z += v.z;
              acc.w += v.w;
            }
            welch_emit(P.fin, blk0_mod, gi, i0, acc);
            if (last && u == 0) {
              float n = nstage[0];
              for (int k = 1; k < S; ++k) n += nstage[k];
              welch_emit(P.fin, blk0_mod, gi, M, float4{n, 0.f, 0.f, 0.f});
            }
          }
        }
        if (S == 1 && c == CPG - 1 && threadIdx.x == 0) welch_finalize_quad(P.fin, fpos, gi, M);
      }
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
struct ChanParams {
  const cpx* spec;
  long long spec_stride;
  int N, K, real_input;
  const float* weights;
  const ChanDesc* chans;
  int nch, nblocks;
  cpx* out;  // channel c's block g at out + out_off + g * 3L/4
  const cpx* tw2;
  const cpx* tw3;
};

constexpr int kChanThreads = 256;

// One tile: gather x weights, L-point inverse FFT, discard, for (channel d,
// block g).  T = L/R threads of the workgroup run it on their tile's LDS
// region; inactive tiles compute but do not store (the exchanges hold
// barriers).  tw2/tw3: inverse-sign pass tables of the plan.
template <int L, int R, int R1, int R2, int R3>
__device__ __forceinline__ void chan_tile(const cpx* __restrict__ spec, long long spec_stride, int N, int K,
                                          int real_input, const float* __restrict__ weights, const ChanDesc& d,
                                          int g, bool active, float* lds, int t, const cpx* tw2, const cpx* tw3,
                                          cpx* __restrict__ out_base) {
  constexpr int T = L / R;
  const cpx* X = spec + (long long)g * spec_stride;
  const float* W = weights + d.w_off;
  cpx x[R];
#pragma unroll
  for (int e = 0; e < R; ++e) {
    const int n = st_in_index<L, R, T, R1>(t, e);
    const int jj = (n + L / 2) & (L - 1);
    int m = d.f_mod + jj;
    if (m >= N) m -= N;
    const bool mirror = real_input && m >= K;
    cpx v = X[mirror ? N - m : m];
    if (mirror) v.y = -v.y;
    x[e] = cscale(v, W[jj]);
  }
  st_butterflies<L, R, T, R1, 1, +1>(x, t, nullptr);
  // (a tile of at most 64 threads lies in one wave: wave-local exchanges)
  tile_exchange<T>(
      x, lds, [&](int e) { return st_out_addr<L, R, T, R1, 1>(t, e); },
      [&](int e) { return st_in_addr<L, R, T, R2>(t, e); });
  st_butterflies<L, R, T, R2, R1, +1>(x, t, tw2);
  constexpr int RL = (R3 > 1) ? R3 : R2;
  if constexpr (R3 > 1) {
    tile_exchange<T>(
        x, lds, [&](int e) { return st_out_addr<L, R, T, R2, R1>(t, e); },
        [&](int e) { return st_in_addr<L, R, T, R3>(t, e); });
    st_butterflies<L, R, T, R3, R1 * R2, +1>(x, t, tw3);
  }
  // Last pass radix RL, p = L / RL: slot u*RL + q holds point b + q p.
  constexpr int PL = L / RL;
  cpx* out = out_base + d.out_off + (long long)g * (3 * L / 4);
  if (active) {
#pragma unroll
    for (int e = 0; e < R; ++e) {
      const int u = e / RL, q = e % RL;
      const int idx = t + T * u + q * PL;
      if (idx >= L / 8 && idx < (L / 8) * 7) store_out(out + (idx - L / 8), x[e]);
    }
  }
}

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
  const long long ntiles = (long long)P.nch * P.nblocks;
  long long tile = (long long)blockIdx.x * TPW + tl;
  const bool active = tile < ntiles;
  if (!active) tile = 0;
  const int c = (int)(tile % P.nch);
  const int g = (int)(tile / P.nch);
  const ChanDesc d = P.chans[c];
  chan_tile<L, R, R1, R2, R3>(P.spec, P.spec_stride, P.N, P.K, P.real_input, P.weights, d, g, active,
                              lds_all + tl * LP * 2, t, P.tw2, P.tw3, P.out);
}

// ---------------------------------------------------------------------------
// One tile of the mixed-radix channel transform for IFFT sizes without a
// power-of-two plan (fcfb.rs:154-188 with L any multiple of 4): gather x
// weights into a (natural input order i, get_bin of fftutil.rs:8-22), one
// Stockham stage per prime factor of L ping-ponging between a and b, and the
// outputs y[L/8 .. (L/8)*7) stored when active.  Stage (p, m) computes every
// output o = m (c p + q) + k (k < m, q < p) directly as
//   out[o] = sum_j in[c m + k + j L/p] * w_L^(j (k + q m) L/(m p)),
// merging the inter-stage twiddle and the p-point DFT kernel into one table
// entry per term (tw = w_L^e, e < L, inverse sign).  T threads (t < T) share
// the tile; the stages hold workgroup barriers, so every thread of the
// workgroup makes the same calls.
// ---------------------------------------------------------------------------
constexpr int kGenMaxStages = 16;
constexpr int kGenMaxL = 4096;  // two LDS buffers of L cpx per tile stay within 64 KiB
constexpr int kGenThreads = 256;

__device__ __forceinline__ void gen_chan_tile(const cpx* __restrict__ spec, long long spec_stride, int N, int K,
                                              int real_input, const float* __restrict__ weights, const ChanDesc& d,
                                              int g, bool active, cpx* a, cpx* b, int t, int T, int L, int nstages,
                                              const int* radix, const cpx* __restrict__ tw,
                                              cpx* __restrict__ out_base) {
  const cpx* X = spec + (long long)g * spec_stride;
  const float* W = weights + d.w_off;
  for (int i = t; i < L; i += T) {
    int jj = i + L / 2;
    if (jj >= L) jj -= L;
    int m = d.f_mod + jj;
    if (m >= N) m -= N;
    const bool mirror = real_input && m >= K;
    cpx v = X[mirror ? N - m : m];
    if (mirror) v.y = -v.y;
    a[i] = cscale(v, W[jj]);
  }
  int m = 1;
  for (int s = 0; s < nstages; ++s) {
    __syncthreads();
    const int p = radix[s];
    const int Lp = L / p;
    const int mp = m * p;
    const int scale = L / mp;  // w_mp = w_L^scale
    for (int o = t; o < L; o += T) {
      const int k = o % m;
      const int r = o / m;
      const int q = r % p;
      const int base = (r / p) * m + k;
      const int step = (k + q * m) * scale;  // < L
      cpx acc = a[base];
      int e = step;
      for (int j = 1; j < p; ++j) {
        const cpx v = a[base + j * Lp];
        const cpx w = tw[e];
        acc.x += v.x * w.x - v.y * w.y;
        acc.y += v.x * w.y + v.y * w.x;
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
  if (active) {
    const int lo = L / 8, n = 6 * (L / 8);
    cpx* out = out_base + d.out_off + (long long)g * n;
    for (int i = t; i < n; i += T) store_out(out + i, a[lo + i]);
  }
}

// ---------------------------------------------------------------------------
// Tile shapes of the power-of-two channel transform: shape S of size L is
// the radix plan (R points per thread, passes R1 x R2 x R3, R1 = R), so a
// tile takes L / R threads.  Shape 0 is the default of the batched kernel;
// kPerSizeDefaultShape<L> that of the per-size launch.  A workgroup of the
// batched kernel holds 256 R complex values in LDS: chan_batched_kernel<4096>
// serves the shapes with R <= 16 (four workgroups per CU),
// chan_batched_kernel<8192> those with R <= 32 (two per CU).  The autotuner
// picks one shape per size and form.
// ---------------------------------------------------------------------------
template <int L, int S>
struct TileShape {
  static constexpr bool valid = false;
};
#define SPK_SHAPE(L_, S_, R_, R1_, R2_, R3_)                                              \
  template <>                                                                            \
  struct TileShape<L_, S_> {                                                             \
    static constexpr bool valid = true;                                                  \
    static constexpr int R = R_, R1 = R1_, R2 = R2_, R3 = R3_;                           \
    static_assert(R1 == R && R1 * R2 * R3 == L_ && R % R2 == 0 && R % R3 == 0, "plan"); \
  };
SPK_SHAPE(8, 0, 4, 4, 2, 1)
SPK_SHAPE(16, 0, 4, 4, 4, 1)
SPK_SHAPE(16, 1, 8, 8, 2, 1)
SPK_SHAPE(32, 0, 8, 8, 4, 1)
SPK_SHAPE(32, 1, 4, 4, 4, 2)
SPK_SHAPE(32, 2, 16, 16, 2, 1)
SPK_SHAPE(64, 0, 8, 8, 8, 1)
SPK_SHAPE(64, 1, 16, 16, 4, 1)
SPK_SHAPE(64, 2, 4, 4, 4, 4)
SPK_SHAPE(128, 0, 16, 16, 8, 1)
SPK_SHAPE(128, 1, 8, 8, 8, 2)
SPK_SHAPE(128, 2, 32, 32, 4, 1)
SPK_SHAPE(256, 0, 16, 16, 16, 1)
SPK_SHAPE(256, 1, 8, 8, 8, 4)
SPK_SHAPE(256, 2, 32, 32, 8, 1)
SPK_SHAPE(512, 0, 16, 16, 16, 2)
SPK_SHAPE(512, 1, 32, 32, 16, 1)
SPK_SHAPE(512, 2, 8, 8, 8, 8)
SPK_SHAPE(1024, 0, 16, 16, 16, 4)
SPK_SHAPE(1024, 1, 32, 32, 32, 1)
SPK_SHAPE(2048, 0, 16, 16, 16, 8)
SPK_SHAPE(2048, 1, 32, 32, 32, 2)
SPK_SHAPE(4096, 0, 16, 16, 16, 16)
SPK_SHAPE(4096, 1, 32, 32, 32, 4)
SPK_SHAPE(8192, 0, 32, 32, 16, 16)
#undef SPK_SHAPE
constexpr int kMaxShapes = 3;
template <int L>
constexpr int shape_count() {
  return TileShape<L, 2>::valid ? 3 : (TileShape<L, 1>::valid ? 2 : 1);
}
// The per-size launch's default: the shape with 32 points per thread from
// L = 512 to 2048, shape 0 otherwise.
template <int L>
constexpr int kPerSizeDefaultShape = (L >= 512 && L <= 2048) ? 1 : 0;
constexpr int batch_max_points(int MAXL) { return MAXL > 4096 ? 32 : 16; }
template <int L, int S, int MAXL>
constexpr bool batch_shape_ok() {
  if constexpr (TileShape<L, S>::valid && L <= MAXL) {
    return TileShape<L, S>::R <= batch_max_points(MAXL);
  } else {
    return false;
  }
}

// ---------------------------------------------------------------------------
// Batched persistent channel stage: every channel of every size in one
// launch.  The host cuts the (block, channel) tiles into items: an item
// holds tiles of one IFFT size class from one chunk of consecutive blocks,
// block-major (tile t -> block g0 + t / nch, channel ch0 + t % nch of the
// class-sorted descriptor list).  Chunks are dealt to XCD queues round
// robin; workgroup w serves queue w mod X (the dispatcher places
// workgroup w on XCD w mod X), so one chunk's spectrum rows are read into
// one L2.  Items are taken in order through one atomic counter per queue,
// which welch_finalize_kernel resets for the next launch.
// ---------------------------------------------------------------------------
constexpr int kMaxClasses = 16;
struct BatchItem {
  int cls, ch0, nch, g0, t0, t1;
};
struct BatchParams {
  const cpx* spec;
  long long spec_stride;
  int N, K, real_input;
  const float* weights;
  const ChanDesc* descs;  // class-sorted
  cpx* out;
  const BatchItem* items;
  const int* item_off;  // [X + 1] start of each queue
  int* counters;        // [X]
  int X;
  int cls_L[kMaxClasses];
  // power-of-two classes: tile shape and pass tables; mixed-radix classes
  // (cls_shape < 0): threads per tile, stage radices, and w_L in tw2
  int cls_shape[kMaxClasses];
  const cpx* tw2[kMaxClasses];
  const cpx* tw3[kMaxClasses];
  int gen_T[kMaxClasses];
  int gen_nst[kMaxClasses];
  int gen_radix[kMaxClasses][kGenMaxStages];
};

constexpr int cmax(int a, int b) { return a > b ? a : b; }
// LDS of the batched kernel (floats): 256 threads x R points of the largest
// shape it serves, padded per tile.
template <int L, int S, int MAXL>
constexpr int batch_lds_floats() {
  if constexpr (batch_shape_ok<L, S, MAXL>()) {
    constexpr int T = L / TileShape<L, S>::R;
    return (kChanThreads / T) * lds_padded_size(L) * 2;
  } else {
    return 0;
  }
}
template <int L, int MAXL>
constexpr int batch_lds_floats_all() {
  return cmax(batch_lds_floats<L, 0, MAXL>(), cmax(batch_lds_floats<L, 1, MAXL>(), batch_lds_floats<L, 2, MAXL>()));
}
// LDS floats of chan_batched_kernel<MAXL> (the host sizes mixed-radix tiles by it).
template <int MAXL>
constexpr int batch_kernel_lds_floats() {
  return cmax(cmax(cmax(batch_lds_floats_all<8, MAXL>(), batch_lds_floats_all<16, MAXL>()),
                   cmax(batch_lds_floats_all<32, MAXL>(), batch_lds_floats_all<64, MAXL>())),
              cmax(cmax(cmax(batch_lds_floats_all<128, MAXL>(), batch_lds_floats_all<256, MAXL>()),
                        cmax(batch_lds_floats_all<512, MAXL>(), batch_lds_floats_all<1024, MAXL>())),
                   cmax(cmax(batch_lds_floats_all<2048, MAXL>(), batch_lds_floats_all<4096, MAXL>()),
                        batch_lds_floats_all<8192, MAXL>())));
}

template <int L, int S>
__device__ __forceinline__ void batch_item(const BatchParams& P, const BatchItem& bi, float* lds_all) {
  using BP = TileShape<L, S>;
  constexpr int T = L / BP::R;
  constexpr int TPW = kChanThreads / T;
  constexpr int LP = lds_padded_size(L);
  const int rounds = (bi.t1 - bi.t0 + TPW - 1) / TPW;
  for (int r = 0; r < rounds; ++r) {
    // Lane indices and table pointers are re-derived every round: kept live
    // across rounds (the compiler would hoist the twiddle loads) they cost
    // more registers than four workgroups per CU leave.
    // (an opaque integer offset keeps the tables' global address space)
    int tid = threadIdx.x, zero = 0;
    asm volatile("" : "+v"(tid));
    asm volatile("" : "+s"(zero));
    const cpx* tw2 = P.tw2[bi.cls] + zero;
    const cpx* tw3 = P.tw3[bi.cls] + zero;
    const int tl = tid / T, t = tid % T;
    int tile = bi.t0 + r * TPW + tl;
    const bool active = tile < bi.t1;
    if (!active) tile = bi.t0;
    const ChanDesc d = P.descs[bi.ch0 + tile % bi.nch];
    chan_tile<L, BP::R, BP::R1, BP::R2, BP::R3>(P.spec, P.spec_stride, P.N, P.K, P.real_input, P.weights, d,
                                                bi.g0 + tile / bi.nch, active, lds_all + tl * LP * 2, t, tw2, tw3,
                                                P.out);
  }
}

// A mixed-radix class: gen_T threads per tile, two buffers of L values per
// tile (the host checks they fit the kernel's LDS).
__device__ __forceinline__ void batch_gen_item(const BatchParams& P, const BatchItem& bi, float* lds_all) {
  const int cls = bi.cls, L = P.cls_L[cls], T = P.gen_T[cls];
  const int tpw = kChanThreads / T;
  const int tl = (int)threadIdx.x / T, t = (int)threadIdx.x % T;
  cpx* a = reinterpret_cast<cpx*>(lds_all) + (long long)tl * 2 * L;
  const int rounds = (bi.t1 - bi.t0 + tpw - 1) / tpw;
  for (int r = 0; r < rounds; ++r) {
    int tile = bi.t0 + r * tpw + tl;
    const bool active = tile < bi.t1;
    if (!active) tile = bi.t0;
    const ChanDesc d = P.descs[bi.ch0 + tile % bi.nch];
    __syncthreads();  // the previous round's tiles are done with the buffers
    gen_chan_tile(P.spec, P.spec_stride, P.N, P.K, P.real_input, P.weights, d, bi.g0 + tile / bi.nch, active, a,
                  a + L, t, T, L, P.gen_nst[cls], P.gen_radix[cls], P.tw2[cls], P.out);
  }
}

template <int L, int MAXL>
__device__ __forceinline__ void batch_item_shape(const BatchParams& P, const BatchItem& bi, float* lds_all) {
  switch (P.cls_shape[bi.cls]) {
    case 1:
      if constexpr (batch_shape_ok<L, 1, MAXL>()) batch_item<L, 1>(P, bi, lds_all);
      break;
    case 2:
      if constexpr (batch_shape_ok<L, 2, MAXL>()) batch_item<L, 2>(P, bi, lds_all);
      break;
    default:
      if constexpr (batch_shape_ok<L, 0, MAXL>()) batch_item<L, 0>(P, bi, lds_all);
      break;
  }
}

// MAXL: largest IFFT size the instantiation serves (4096 or 8192).
// (A minimum-occupancy launch bound of four workgroups per CU made this
// kernel 1.7x slower on gfx950; LDS already limits it to four.)
template <int MAXL>
__global__ __launch_bounds__(kChanThreads) void chan_batched_kernel(BatchParams P) {
  __shared__ float lds_all[batch_kernel_lds_floats<MAXL>()];
  __shared__ int s_item;
  const int x = (int)blockIdx.x % P.X;
  const int base = P.item_off[x], n = P.item_off[x + 1] - base;
  for (;;) {
    __syncthreads();  // everyone has read the previous item number
    if (threadIdx.x == 0) s_item = atomicAdd(P.counters + x, 1);
    __syncthreads();
    const int it = s_item;
    if (it >= n) break;
    const BatchItem bi = P.items[base + it];
    if (P.cls_shape[bi.cls] < 0) {
      batch_gen_item(P, bi, lds_all);
      continue;
    }
    switch (P.cls_L[bi.cls]) {
      case 8: batch_item_shape<8, MAXL>(P, bi, lds_all); break;
      case 16: batch_item_shape<16, MAXL>(P, bi, lds_all); break;
      case 32: batch_item_shape<32, MAXL>(P, bi, lds_all); break;
      case 64: batch_item_shape<64, MAXL>(P, bi, lds_all); break;
      case 128: batch_item_shape<128, MAXL>(P, bi, lds_all); break;
      case 256: batch_item_shape<256, MAXL>(P, bi, lds_all); break;
      case 512: batch_item_shape<512, MAXL>(P, bi, lds_all); break;
      case 1024: batch_item_shape<1024, MAXL>(P, bi, lds_all); break;
      case 2048: batch_item_shape<2048, MAXL>(P, bi, lds_all); break;
      case 4096: batch_item_shape<4096, MAXL>(P, bi, lds_all); break;
      default: batch_item_shape<8192, MAXL>(P, bi, lds_all); break;
    }
  }
}

// Grid: x over quads of partial-row positions, y over the Welch groups the
// launch touches (group 0 is the group of the launch's first block).  Row
// y = 0 also does the next call's housekeeping the fields ask for (input
// history, queue counters, position slot).
// Grid: x = chunks of blockDim.x / F.split quads, y = Welch groups.  Each
// workgroup holds F.split teams; team k sums the k-th of F.split contiguous
// ranges of the group's partial rows over the chunk, and team 0 adds the
// teams' sums in order (a walk over every row by one thread is latency-bound
// when a group spans many workgroups).  The last chunk also carries the
// Nyquist position.
__global__ __launch_bounds__(256) void welch_finalize_kernel(FinParams F) {
  __shared__ float4 stage[256];
  __shared__ float nstage[16];
  const int gi = (int)blockIdx.y;
  const int tid = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const long long pos = *F.pos;
  if (gi == 0) {
    for (int i = tid; i < F.hist_words; i += (int)(gridDim.x * blockDim.x)) F.hist_dst[i] = F.hist_src[i];
    if (tid < F.ncounters) F.chan_counters[tid] = 0;
    if (tid == 0 && F.pos_next) *F.pos_next = pos + F.nblocks;
  }
  const int blk0_mod = (int)(pos % F.A);
  // groups beyond the ones this launch touches (the grid is sized for the worst case)
  if (gi > (blk0_mod + F.nblocks - 1) / F.A) return;
  const int S = F.split, TS = (int)blockDim.x / S;
  const int team = (int)threadIdx.x / TS, u = (int)threadIdx.x % TS;
  const int i0 = 4 * ((int)blockIdx.x * TS + u);
  const bool valid = i0 < F.M;
  const bool nyq_chunk = blockIdx.x == gridDim.x - 1;
  int w0, w1;
  welch_group_wgs(F, blk0_mod, gi, &w0, &w1);
  const int nw = w1 - w0 + 1;
  const int wa = w0 + nw * team / S, wb = w0 + nw * (team + 1) / S - 1;
  const float4 z = {0.f, 0.f, 0.f, 0.f};
  const float4 s = valid ? welch_rows_add(F, blk0_mod, gi, i0, wa, wb, team == 0 ? welch_carry(F, blk0_mod, gi, i0) : z)
                         : z;
  if (S == 1) {
    if (valid) welch_emit(F, blk0_mod, gi, i0, s);
    if (nyq_chunk && threadIdx.x == 0) welch_finalize_quad(F, pos, gi, F.M);
    return;
  }
  stage[threadIdx.x] = s;
  if (nyq_chunk && u == 0)
    nstage[team] = welch_rows_add(F, blk0_mod, gi, F.M, wa, wb, team == 0 ? welch_carry(F, blk0_mod, gi, F.M) : z).x;
  __syncthreads();
  if (team == 0) {
    float4 acc = stage[u];
    for (int k = 1; k < S; ++k) {
      const float4 v = stage[k * TS + u];
      acc.x += v.x;
      acc.y += v.y;
      acc.z += v.z;
      acc.w += v.w;
    }
    if (valid) welch_emit(F, blk0_mod, gi, i0, acc);
    if (nyq_chunk && u == 0) {
      float n = nstage[0];
      for (int k = 1; k < S; ++k) n += nstage[k];
      welch_emit(F, blk0_mod, gi, F.M, float4{n, 0.f, 0.f, 0.f});
    }
  }
}

// ---------------------------------------------------------------------------
// Adjoint of the channel stage (no counterpart in the reference; the weights
// become trainable).  For one (channel, block) the stage is
//   B[i] = W[j] G[j],  j = (i + L/2) mod L,  G[j] = get_bin(first + j),
//   y = IFFT_L(B) (unnormalised, + sign),  out[n] = y[n + L/8], n < 3L/4.
// For a real loss with upstream gradient ybar[n] = dL/dRe out[n] + i dL/dIe out[n],
// let Yh = FFT_L(z) (- sign) of z[i] = ybar[i - L/8] on [L/8, 7L/8), 0 else.
// Then dL/dW[j] = sum over blocks of Re(G[j] conj(Yh[i])) and the gradient
// with respect to the gathered bin is Gbar[j] = W[j] Yh[i].
// Grid: x = channel * S + split; every tile of a workgroup walks the blocks
// g = (it S + split) TPW + tile, keeps its W sums in registers, and the
// workgroup adds its tiles in a fixed order into wpart[channel][split][j].
// Packed layouts (channel id order): ybar nblocks * 3L/4 complex per
// channel at nblocks * 3 w_off / 4, xbar nblocks * L complex at
// nblocks * w_off, W sums L floats at w_off.
// ---------------------------------------------------------------------------
struct ChanAdjParams {
  const cpx* spec;
  long long spec_stride;
  int N, K, real_input;
  const float* weights;
  const ChanDesc* chans;
  int nch, nblocks, S;
  const cpx* ybar;      // packed layout (pack_outputs_kernel)
  const long long* pk_off;
  cpx* xbar;     // may be null
  float* wpart;  // [nch][S][L]
  const cpx* tw2;  // forward-sign pass tables
  const cpx* tw3;
};

template <int L, int R, int R1, int R2, int R3>
__global__ __launch_bounds__(kChanThreads) void chan_adjoint_kernel(ChanAdjParams P) {
  constexpr int T = L / R;
  static_assert(R1 * R2 * R3 == L, "radix plan must cover L");
  constexpr int TPW = kChanThreads / T;
  constexpr int LP = lds_padded_size(L);
  static_assert(LP * 2 >= L, "reduction rows fit the tile regions");
  __shared__ float lds_all[TPW * LP * 2];

  const int tl = threadIdx.x / T;
  const int t = threadIdx.x % T;
  float* lds = lds_all + tl * LP * 2;
  const int c = (int)blockIdx.x / P.S, split = (int)blockIdx.x % P.S;
  const ChanDesc d = P.chans[c];
  const float* W = P.weights + d.w_off;
  constexpr int RL = (R3 > 1) ? R3 : R2;
  constexpr int PL = L / RL;

  float wacc[R];
#pragma unroll
  for (int e = 0; e < R; ++e) wacc[e] = 0.f;
  const int iters = (P.nblocks + P.S * TPW - 1) / (P.S * TPW);
  for (int it = 0; it < iters; ++it) {  // uniform trip count: the exchanges hold barriers
    const int g = (it * P.S + split) * TPW + tl;
    const bool act = g < P.nblocks;
    const int gg = act ? g : 0;
    const cpx* yb = P.ybar + (long long)P.nblocks * P.pk_off[c] + (long long)gg * (3 * L / 4);
    cpx x[R];
#pragma unroll
    for (int e = 0; e < R; ++e) {
      const int n = st_in_index<L, R, T, R1>(t, e);
      x[e] = (n >= L / 8 && n < (L / 8) * 7) ? yb[n - L / 8] : cpx{0.f, 0.f};
    }
    st_butterflies<L, R, T, R1, 1, -1>(x, t, nullptr);
    lds_exchange<R, false>(
        x, lds, [&](int e) { return st_out_addr<L, R, T, R1, 1>(t, e); },
        [&](int e) { return st_in_addr<L, R, T, R2>(t, e); });
    st_butterflies<L, R, T, R2, R1, -1>(x, t, P.tw2);
    if constexpr (R3 > 1) {
      lds_exchange<R, false>(
          x, lds, [&](int e) { return st_out_addr<L, R, T, R2, R1>(t, e); },
          [&](int e) { return st_in_addr<L, R, T, R3>(t, e); });
      st_butterflies<L, R, T, R3, R1 * R2, -1>(x, t, P.tw3);
    }
    const cpx* X = P.spec + (long long)gg * P.spec_stride;
    cpx* xb = P.xbar ? P.xbar + (long long)P.nblocks * d.w_off + (long long)gg * L : nullptr;
#pragma unroll
    for (int e = 0; e < R; ++e) {
      const int u = e / RL, q = e % RL;
      const int i = t + T * u + q * PL;
      const int j = (i + L / 2) & (L - 1);
      int m = d.f_mod + j;
      if (m >= P.N) m -= P.N;
      const bool mirror = P.real_input && m >= P.K;
      cpx G = X[mirror ? P.N - m : m];
      if (mirror) G.y = -G.y;
      if (act) {
        wacc[e] += G.x * x[e].x + G.y * x[e].y;
        if (xb) xb[j] = cscale(x[e], W[j]);
      }
    }
  }
  // Add the tiles of this workgroup in tile order.
  __syncthreads();
#pragma unroll
  for (int e = 0; e < R; ++e) {
    const int u = e / RL, q = e % RL;
    lds_all[tl * L + t + T * u + q * PL] = wacc[e];
  }
  __syncthreads();
  for (int i = (int)threadIdx.x; i < L; i += kChanThreads) {
    float sum = 0.f;
    for (int k = 0; k < TPW; ++k) sum += lds_all[k * L + i];
    P.wpart[((long long)c * P.S + split) * L + ((i + L / 2) & (L - 1))] = sum;
  }
}

// wbar[w_off + j] = sum over splits (in order) of wpart[c][s][j].
struct AdjReduceParams {
  const float* wpart;
  const ChanDesc* chans;
  int L, S;
  float* wbar;
};
__global__ void adjoint_reduce_kernel(AdjReduceParams P) {
  const int c = (int)blockIdx.y;
  const int j = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (j >= P.L) return;
  float sum = 0.f;
  for (int s = 0; s < P.S; ++s) sum += P.wpart[((long long)c * P.S + s) * P.L + j];
  P.wbar[P.chans[c].w_off + j] = sum;
}

// Byte o < 24 of a record's metadata (data.rs:110-132: u64 seq, u64
// seconds, u32 nanoseconds, u32 reserved, little-endian).
__device__ __forceinline__ unsigned char metadata_byte(int o, unsigned long long seq, unsigned long long secs,
                                                       unsigned nanos) {
  if (o < 8) return (unsigned char)(seq >> (8 * o));
  if (o < 16) return (unsigned char)(secs >> (8 * (o - 8)));
  if (o < 20) return (unsigned char)(nanos >> (8 * (o - 16)));
  return 0;
}

// Welch records as the reference sends them (spectrum.rs:120-153): per
// record 24 bytes of metadata followed by the quantised body.  Grid: x over
// the bytes of a record, y over records (strided).
struct RecordPackParams {
  const unsigned char* body;  // nrec bodies of body_bytes each
  unsigned char* out;         // nrec records of 24 + body_bytes
  long long body_bytes;
  int nrec;
  unsigned long long first_seq, secs;
  unsigned nanos;
};

__global__ void pack_records_kernel(RecordPackParams P) {
  const long long rec = 24 + P.body_bytes;
  for (int r = (int)blockIdx.y; r < P.nrec; r += (int)gridDim.y) {
    unsigned char* out = P.out + (long long)r * rec;
    const unsigned char* body = P.body + (long long)r * P.body_bytes;
    for (long long o = (long long)blockIdx.x * blockDim.x + threadIdx.x; o < rec;
         o += (long long)gridDim.x * blockDim.x)
      out[o] = o >= 24 ? body[o - 24] : metadata_byte((int)o, P.first_seq + (unsigned long long)r, P.secs, P.nanos);
  }
}

// Channel records as the reference writes them (fcfb.rs:75-89): per live
// channel, in id order, 24 bytes of metadata (the buffer's seq) and the
// call's samples (block-major, 6 (L/8) per block, f32 re, im).  Record r
// starts at byte 24 r + 8 nblocks prefix_r.  Grid: x over samples, y over
// records (strided).
struct ChanRecDesc {
  long long src;     // first sample of the channel's output region
  long long prefix;  // samples per block of the channels before it
  int len;           // samples per block, 6 (L/8)
  int pad;
};
struct ChanRecParams {
  const cpx* src;
  const ChanRecDesc* recs;
  int nrec, nblocks;
  unsigned char* out;
  unsigned long long seq, secs;
  unsigned nanos;
};

__global__ void pack_channel_records_kernel(ChanRecParams P) {
  for (int r = (int)blockIdx.y; r < P.nrec; r += (int)gridDim.y) {
    const ChanRecDesc d = P.recs[r];
    unsigned char* o = P.out + 24LL * r + 8LL * P.nblocks * d.prefix;
    if (blockIdx.x == 0 && threadIdx.x < 24) o[threadIdx.x] = metadata_byte((int)threadIdx.x, P.seq, P.secs, P.nanos);
    cpx* body = reinterpret_cast<cpx*>(o + 24);  // 8-byte aligned: 24 r + 8 k
    const cpx* s = P.src + d.src;
    const long long n = (long long)P.nblocks * d.len;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
      body[i] = s[i];
  }
}

// Copy channel outputs from the capacity layout (out_off) to the packed
// layout: channel c's nblocks * 6 (L/8) samples at nblocks * pk_off[c]
// (pk_off: samples per block of the live channels before it, id order).
// Grid y = channel of the group.
struct PackParams {
  const cpx* src;
  const ChanDesc* chans;
  const long long* pk_off;
  int L, nblocks;
  cpx* dst;
};
__global__ void pack_outputs_kernel(PackParams P) {
  const ChanDesc d = P.chans[blockIdx.y];
  const long long n = (long long)P.nblocks * (6 * (P.L / 8));
  const cpx* s = P.src + d.out_off;
  cpx* o = P.dst + (long long)P.nblocks * P.pk_off[blockIdx.y];
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    o[i] = s[i];
}

}  // namespace spk
