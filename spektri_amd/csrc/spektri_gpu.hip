// C ABI of the FCFB + Welch hot path (declared in include/spektri_gpu.h).
//
// A context owns: the N/4-sample input history, a staging buffer for host
// input, the spectrum rows, the Welch partial rows and carry row, the Welch
// record buffer, per-channel output buffers, and the twiddle tables.
// One process call launches: the persistent forward kernel (fused s16 load,
// FFT, Welch partial sums and, when every channel has L = 256 and they fit,
// the fused channel stage), one channel kernel per distinct IFFT size for
// the channels not fused, and the Welch finaliser.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <exception>
#include <map>
#include <new>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "spektri_gpu.h"
#include "kernels.h"
#include "cplx.h"
#include "convert.h"
#include "chan_generic.h"
#include "gfft.h"

using spk::cpx;

// Scratch of spk_channels_adjoint_noalloc: the per-(channel, split) weight
// sums that spk_channels_adjoint keeps in a context buffer it allocates
// (adj_part).  Every GPU the library runs on holds one copy until the
// process ends.  Host code never uses the name as an address: the device
// address comes from hipGetSymbolAddress, once per device (scratch_address).
constexpr size_t kNoAllocScratchFloats = size_t(2) << 20;  // 8 MiB
static __device__ __align__(256) float scratch[kNoAllocScratchFloats];

#include <mutex>

namespace {

// Device address of `scratch` on the current device, looked up on the
// first context created on that device (nullptr when the lookup fails).
float* scratch_address(int dev) {
  static std::mutex mu;
  static float* addr[256] = {};
  static bool done[256] = {};
  if (dev < 0 || dev >= 256) return nullptr;
  std::lock_guard<std::mutex> lock(mu);
  if (!done[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) == hipSuccess) addr[dev] = static_cast<float*>(p);
    done[dev] = true;
  }
  return addr[dev];
}

const char* kNullCtxMsg = "null context";

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) {
    o.p = nullptr;
    o.bytes = 0;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      if (p) (void)hipFree(p);
      p = o.p;
      bytes = o.bytes;
      o.p = nullptr;
      o.bytes = 0;
    }
    return *this;
  }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  hipError_t alloc(size_t n) {
    if (p) {
      (void)hipFree(p);
      p = nullptr;
      bytes = 0;
    }
    if (n == 0) return hipSuccess;
    hipError_t e = hipMalloc(&p, n);
    if (e == hipSuccess) bytes = n;
    return e;
  }
  template <class T>
  T* as() const {
    return reinterpret_cast<T*>(p);
  }
};

struct Channel {
  int64_t first;
  int L;
  bool removed = false;  // spk_remove_channel: keeps its id, has no output
  std::vector<float> w;
};

struct ChanGroup {  // channels sharing one IFFT size
  int L;
  std::vector<uint32_t> ids;
  DevBuf desc;   // spk::ChanDesc[ids.size()]
  DevBuf tw2, tw3;      // inverse-sign pass tables
  DevBuf atw2, atw3;    // forward-sign pass tables (adjoint)
  // IFFT sizes without a power-of-two plan (spk::chan_generic_kernel):
  bool generic = false;
  std::vector<int> radix;  // stage radices, product L
  DevBuf gtw;              // w_L^e, e = 0..L-1, inverse sign
  DevBuf pk_off;           // long long per channel: packed-layout offset (samples per block before it)
  bool fused = false;      // computed inside the forward kernel
  int ps_shape = 0;        // tile shape of the per-size launch (tw2/tw3 are its tables)
};

// Output samples per block of a channel of IFFT size L: the reference keeps
// buf[L/8 .. L/8*7] with integer division (fcfb.rs:176-177).
inline long long chan_out_len(int L) { return 6LL * (L / 8); }

// Stage radices of the generic channel transform: the prime factors of L
// with pairs of 2 merged into 4.
std::vector<int> generic_radices(int L) {
  std::vector<int> r;
  while (L % 4 == 0) {
    r.push_back(4);
    L /= 4;
  }
  if (L % 2 == 0) {
    r.push_back(2);
    L /= 2;
  }
  for (int p = 3; L > 1; p += 2) {
    while (L % p == 0) {
      r.push_back(p);
      L /= p;
    }
    if ((long long)p * p > L && L > 1) {
      r.push_back(L);
      break;
    }
  }
  return r;
}

struct FwdPlan {
  int M, R, R1, R2;
};
const FwdPlan kFwdPlans[] = {{512, 8, 8, 8},     {1024, 16, 8, 8},     {2048, 16, 8, 16},
                             {4096, 16, 16, 16}, {8192, 32, 16, 16}, {16384, 32, 16, 32},
                             {32768, 32, 32, 32}};

struct ChanPlan {
  int L, R, R1, R2, R3;
};
const ChanPlan kChanPlans[] = {{8, 4, 4, 2, 1},       {16, 4, 4, 4, 1},      {32, 8, 8, 4, 1},
                               {64, 8, 8, 8, 1},      {128, 16, 16, 8, 1},   {256, 16, 16, 16, 1},
                               {512, 32, 32, 16, 1},  {1024, 32, 32, 32, 1}, {2048, 32, 32, 32, 2},
                               {4096, 16, 16, 16, 16}, {8192, 32, 32, 16, 16}};

// Channels that the fused forward kernel can take: IFFT size 256, at most
// one 16-thread tile per channel.
constexpr int kFusedL = 256;

const ChanPlan* find_chan_plan(int L) {
  for (const auto& p : kChanPlans)
    if (p.L == L) return &p;
  return nullptr;
}

// Tile shapes of the power-of-two channel transform (spk::TileShape): the
// radix plan of shape S of IFFT size L.  The batched kernel serves every
// shape: chan_batched_kernel<4096> those of at most 16 points per thread,
// chan_batched_kernel<8192> (twice the LDS) the rest.
struct ShapeInfo {
  int R = 0, R1 = 0, R2 = 0, R3 = 0;
  bool big() const { return R > spk::batch_max_points(4096); }
};
template <int L, int S>
void shape_entry(std::vector<ShapeInfo>& v) {
  if constexpr (spk::TileShape<L, S>::valid) {
    using TS = spk::TileShape<L, S>;
    static_assert(spk::batch_shape_ok<L, S, 8192>(), "every shape is served by chan_batched_kernel<8192>");
    v.push_back(ShapeInfo{TS::R, TS::R1, TS::R2, TS::R3});
  }
}
template <int L>
std::vector<ShapeInfo> shapes_of() {
  std::vector<ShapeInfo> v;
  shape_entry<L, 0>(v);
  shape_entry<L, 1>(v);
  shape_entry<L, 2>(v);
  return v;
}
std::vector<ShapeInfo> tile_shapes(int L) {
  switch (L) {
    case 8: return shapes_of<8>();
    case 16: return shapes_of<16>();
    case 32: return shapes_of<32>();
    case 64: return shapes_of<64>();
    case 128: return shapes_of<128>();
    case 256: return shapes_of<256>();
    case 512: return shapes_of<512>();
    case 1024: return shapes_of<1024>();
    case 2048: return shapes_of<2048>();
    case 4096: return shapes_of<4096>();
    case 8192: return shapes_of<8192>();
    default: return {};
  }
}
int default_shape(int L, bool batched);
int shape_of(const std::map<int, int>& m, int L, bool batched) {
  const auto it = m.find(L);
  return it == m.end() ? default_shape(L, batched) : it->second;
}
int default_shape(int L, bool batched) {
  if (batched) return 0;
  switch (L) {
    case 512: return spk::kPerSizeDefaultShape<512>;
    case 1024: return spk::kPerSizeDefaultShape<1024>;
    case 2048: return spk::kPerSizeDefaultShape<2048>;
    default: return 0;
  }
}
// Pass radices of the mixed-radix forward transform of Nc points: eights,
// then a four or a two, then odd primes up to 31 (the instantiated
// radices); empty when Nc has a larger prime factor.
std::vector<int> gfft_radices(int Nc) {
  std::vector<int> r;
  while (Nc % 8 == 0 && Nc > 1) {
    r.push_back(8);
    Nc /= 8;
  }
  if (Nc % 4 == 0) {
    r.push_back(4);
    Nc /= 4;
  }
  if (Nc % 2 == 0) {
    r.push_back(2);
    Nc /= 2;
  }
  for (int p : {3, 5, 7, 11, 13, 17, 19, 23, 29, 31})
    while (Nc % p == 0) {
      r.push_back(p);
      Nc /= p;
    }
  if (Nc != 1) r.clear();
  return r;
}

const FwdPlan* find_fwd_plan(int M) {
  for (const auto& p : kFwdPlans)
    if (p.M == M) return &p;
  return nullptr;
}

// Pass table for radix r over sub-transforms of size p: entry [(j-1) p + k]
// = exp(sign 2 pi i j k / (p r)), evaluated in double with the exponent
// reduced exactly in integers, rounded to f32.
std::vector<cpx> pass_table(int r, int p, int sign) {
  std::vector<cpx> t;
  if (p <= 1) return t;
  t.resize((size_t)(r - 1) * p);
  const long n = (long)p * r;
  for (int j = 1; j < r; ++j)
    for (int k = 0; k < p; ++k) {
      const long e = ((long)j * k) % n;
      const double a = 2.0 * M_PI * (double)e / (double)n;
      t[(size_t)(j - 1) * p + k] = cpx{(float)std::cos(a), (float)(sign * std::sin(a))};
    }
  return t;
}

// DataFormat codes of spektri/src/dsp/data.rs:81-95.
// Per DataFormat: bytes per raw sample (inputformats.rs:12-27), complex
// flag (:29-44), default scaling (:47-62), and how the transform kernels
// read it: the conversion pass that runs first (0: none, read in place) and
// the bytes per sample of what they read (s16 or f32 real, cf32 complex).
struct FormatInfo {
  uint32_t code;
  bool complex_in;
  int raw_bps;
  int conv;
  int canon_bps;
  float scaling;
};
const FormatInfo* format_info(uint32_t fmt) {
  static const FormatInfo kFormats[] = {
      {0x04, false, 1, spk::kConvS8ToS16, 2, 1.0f / 127.0f},        // S8
      {0x08, false, 2, 0, 2, 1.0f / 32767.0f},                      // S16le
      {0x09, false, 2, spk::kConvS16beToS16, 2, 1.0f / 32767.0f},   // S16be
      {0x1C, false, 4, 0, 4, 1.0f},                                 // F32le
      {0x1D, false, 4, spk::kConvF32beToF32, 4, 1.0f},              // F32be
      {0x24, false, 1, spk::kConvU8ToF32, 4, 1.0f / 127.0f},        // U8
      {0x44, true, 2, spk::kConvS8ToF32, 8, 1.0f / 127.0f},         // Cs8
      {0x48, true, 4, 0, 4, 1.0f / 32767.0f},                      // Cs16le (s16 pairs, read in place)
      {0x49, true, 4, spk::kConvS16beToF32, 8, 1.0f / 32767.0f},    // Cs16be
      {0x5C, true, 8, 0, 8, 1.0f},                                  // Cf32le
      {0x5D, true, 8, spk::kConvF32beToF32, 8, 1.0f},               // Cf32be
      {0x64, true, 2, spk::kConvU8ToF32, 8, 1.0f / 127.0f},         // Cu8
  };
  for (const auto& f : kFormats)
    if (f.code == fmt) return &f;
  return nullptr;
}

}  // namespace

// Batched channel stage plan (chan_batched_kernel): classes, class-sorted
// descriptors, and the per-XCD item queues for one block count.  A captured
// graph holds the plan's device pointers, so plans are kept per block count
// and freed only together with every graph (drop_graphs_and_plans).
struct BatchPlan {
  std::vector<int> cls_L;
  std::vector<int> cls_shape;          // tile shape, or -1 for a mixed-radix class
  DevBuf descs, items, item_off;
  std::vector<DevBuf> tw2, tw3;        // pass tables of the power-of-two classes
  std::vector<const cpx*> twp2, twp3;  // per class: the tables the kernel reads
  std::vector<int> gen_T, gen_nst;     // mixed-radix classes: threads per tile, stages
  std::vector<std::vector<int>> gen_radix;
  int nb = -1;     // block count the items were cut for
  int maxL = 0;
  bool big = false;  // chan_batched_kernel<8192>: a size above 4096 or a shape above 16 points per thread
  size_t nitems = 0;
};
constexpr size_t kMaxCachedPlans = 8;  // graphs and batch plans kept at once
constexpr int kXcdQueues = 8;  // MI355X: 8 XCDs; workgroup w is dispatched to XCD w mod 8

struct spk_ctx {
  spk_config cfg{};
  int N = 0, M = 0, K = 0, H = 0, O = 0, F = 0, A = 0;
  int bps = 0;               // bytes per sample the transform kernels read (s16, f32 or cf32)
  int raw_bps = 0;           // bytes per sample of the caller's input format
  int conv_kind = 0;         // spk::ConvKind run before the transform, 0 = none
  bool f32_in = false;       // real input read as f32 (the forward kernel's f32 loader)
  bool complex_in = false;
  // The two-kernel complex pipeline (row and column transforms of an
  // Nc-point FFT) runs complex input (Nc = N) and real input above the
  // forward kernel's range (Nc = N/2 packed pairs, then an untangle).
  bool two_step = false;
  int Nc = 0;
  bool s16_pairs = false;    // two-step path reading s16 pairs (cs16le, or real s16le)
  float wscale = 0.f;
  int max_blocks = 0;
  const FwdPlan* fplan = nullptr;

  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;

  DevBuf hist, staging, spec, wpart, carry, edump, welch_out, chan_out, weights;
  DevBuf conv;               // converted input (formats with a conversion pass)
  // spk_process_async: two slots of pinned host staging and device staging,
  // a copy stream for the uploads, and per slot the events "upload done"
  // (copy stream) and "kernels reading the slot done" (context stream).
  void* apin[2] = {nullptr, nullptr};
  DevBuf adev[2];
  hipStream_t copy_stream = nullptr;
  hipEvent_t a_up[2] = {nullptr, nullptr}, a_used[2] = {nullptr, nullptr};
  bool a_busy[2] = {false, false};
  int aslot = 0;
  DevBuf fscat;              // fused channel stage: per-lane LDS scatter indices
  int fscat_nyq = 0;
  DevBuf ftw2, ftw3, ftwu;
  DevBuf eitems;             // Welch items of the edge bins (spk::edge_items)
  // complex input: row-transform scratch and tables (N = 64 N1)
  DevBuf cY, ctw2, ctw3, cta, ctb, ctw64;
  DevBuf ctu;                // real two-step input: untangle twiddles W_N^k, k <= N/4
  const ChanPlan* cplan = nullptr;
  // sizes without a two-step plan: mixed-radix passes (gfft.h) over the
  // Nc-point transform, radices in pass order, twiddles W_Nc^m
  bool gfft = false;
  bool real_odd = false;     // odd real N: an N-point complex transform of (x, 0), no untangle
  std::vector<int> gfft_radices;
  DevBuf gtw;
  // four-step form (gfft4_kernel): Nc = N1 N2, stage radices per line length,
  // w_n^e tables of the two line lengths
  bool gfft4 = false;
  int g4N1 = 0, g4N2 = 0;
  std::vector<int> g4r1, g4r2;
  DevBuf g4tw1, g4tw2;
  DevBuf adj_part;           // adjoint per-split weight sums
  float* noalloc_scratch = nullptr;  // this device's `scratch` (spk_channels_adjoint_noalloc)
  int ncu = 0;
  // channel stage launch form and batched tuning (chunk blocks, rounds per item, workgroups per CU)
  int chan_mode = SPK_CHAN_PER_SIZE;
  int tune_chunk = 8, tune_rounds = 2, tune_wgs_per_cu = 0;
  // Tile shape per power-of-two IFFT size (absent: the form's default), for
  // the per-size launch and the batched kernel (spk_set_tile_shape, spk_autotune).
  std::map<int, int> shape_ps, shape_b;
  std::map<int, BatchPlan> batches;  // per block count
  DevBuf chan_counters;      // kXcdQueues ints
  DevBuf pos;                // device stream position (blocks consumed), two long long slots
  int parity = 0;            // slot holding this call's position; the Welch kernel writes the other
  DevBuf qout;               // quantised Welch records (SPK_FLAG_SPECTRUM_U8 / _U16)
  DevBuf qpack;              // the same records with their metadata, packed (spk_spectrum_records_packed)
  DevBuf crecs;              // spk::ChanRecDesc per live channel, id order
  DevBuf cpack;              // channel records packed (spk_channel_records_packed)
  int live_chans = 0;
  long long live_len = 0;    // samples per block over the live channels
  int qfmt = 0;              // 0 none, 1 U8, 2 U16
  bool use_graphs = false;
  // Captured launch sequences per (input pointer, blocks, parity, pending
  // blocks, pending grid).  A channel change marks them stale: the next call
  // with a key captures its sequence again and updates the executable graph
  // in place (hipGraphExecUpdate), instantiating anew only when the launch
  // topology changed.
  struct GraphEntry {
    hipGraphExec_t exec = nullptr;
    bool stale = false;
  };
  std::map<std::tuple<const void*, int, int, int, int>, GraphEntry> graphs;
  uint32_t graphs_instantiated = 0, graphs_updated = 0;
  long long spec_stride = 0, wstride = 0;
  int wslots = 0, welch_out_cap = 0;
  long long wpart_half = 0;  // floats per parity of the partial rows (real input)
  // Real input: the call whose Welch records are not summed yet (its rows
  // are summed by the next forward launch, or by spk::welch_finalize_kernel
  // on flush).  pend_nb == 0: none.
  int pend_nb = 0, pend_G = 0, pend_parity = 0;
  DevBuf fin_ticket;          // chunk counter of the forward kernel's record phase
  int gmax[3] = {0, 0, 0};   // persistent workgroups of the forward kernel: no fused stage / L = 256 / general
  int fmode = 0;             // fused stage of the forward kernel: 0, kFusedL, spk::kFusedGen1 or kFusedGen2
  // general fused stage: tile descriptors, per-(round, thread) slots, pass-2
  // twiddles of the 32..256-point tiles, second Nyquist target
  DevBuf gtiles, gslots, gentw[4];
  size_t weights_live = 0;   // floats of live channels' weights (then 256 zeros)
  int grounds = 0, fscat_nyq2 = 0;
  bool keep_spec = false;
  bool keep_flag = false;  // created with SPK_FLAG_KEEP_SPECTRA (the adjoint's condition)
  bool fused = false;        // channels run inside the forward kernel

  std::vector<Channel> chans;
  std::vector<long long> chan_out_off;  // cpx offsets per channel (capacity layout)
  std::vector<int> chan_w_off;          // float offsets of the live channels' weights
  std::map<int, ChanGroup> groups;
  bool chans_dirty = true;
  // spk_set_weights_device wrote the weights on the device only: the
  // channels' host copies (w) are read back before they are used or changed
  bool w_device_newer = false;
  // weights laid out one copy per channel from now on (spk_set_weights_device
  // addresses them so); w_shared: the current layout shares copies
  bool w_packed = false;
  bool w_shared = false;

  long long blocks_done = 0;  // stream blocks consumed so far
  uint32_t last_blocks = 0;
  uint32_t last_records = 0;
  bool have_output = false;

  // Per-kernel timing: 4 events per call (before the forward kernel, after
  // it, after the channel kernels, after the finaliser), pooled.
  bool timing = false;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;   // events recorded since timing was enabled
  size_t ev_done = 0;   // events already folded into ms_sum
  double ms_sum[3] = {0, 0, 0};
  uint32_t timed_calls = 0;

  std::string err;
};

namespace {

spk_status fail(spk_ctx* c, spk_status s, const std::string& msg) {
  if (c) c->err = msg;
  return s;
}

spk_status hip_fail(spk_ctx* c, hipError_t e, const char* what) {
  return fail(c, SPK_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

// Every extern "C" entry point ends in this handler: a C++ exception (host
// allocation in std containers, std::thread) becomes a status code and
// never crosses the C boundary.
spk_status entry_exception(spk_ctx* c) noexcept {
  try {
    throw;
  } catch (const std::bad_alloc&) {
    if (c) c->err = "host allocation failed";
  } catch (const std::exception& ex) {
    try {
      if (c) c->err = ex.what();
    } catch (...) {
    }
  } catch (...) {
    try {
      if (c) c->err = "unexpected C++ exception";
    } catch (...) {
    }
  }
  return SPK_ERR_DEVICE;
}

#define HIP_TRY(ctx, expr)                              \
  do {                                                  \
    hipError_t _e = (expr);                             \
    if (_e != hipSuccess) return hip_fail(ctx, _e, #expr); \
  } while (0)

template <class T>
hipError_t upload(DevBuf& b, const std::vector<T>& v) {
  hipError_t e = b.alloc(v.size() * sizeof(T));
  if (e != hipSuccess || v.empty()) return e;
  return hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
}

std::vector<float> raised_cosine_weights(int L) {
  // fcfb.rs:197-203 in f32: 0.5 - 0.5 cos(i * (2 pi / L))
  std::vector<float> w(L);
  const float f = (2.0f * (float)M_PI) / (float)L;
  for (int i = 0; i < L; ++i) w[i] = 0.5f - 0.5f * std::cos((float)i * f);
  return w;
}

int fused_lds_dummy(int M);

void drop_graphs(spk_ctx* c) {
  for (auto& kv : c->graphs) (void)hipGraphExecDestroy(kv.second.exec);
  c->graphs.clear();
}
// Graphs hold device pointers of the batch plans: free plans only with them.
void drop_graphs_and_plans(spk_ctx* c) {
  drop_graphs(c);
  c->batches.clear();
}

// LDS dummy slot and tile capacity (cpx) of a forward kernel instantiation
// (s16 loader); -1 when there is none.
int fwd_lds_dummy(int M, int FL);
int fwd_tile_cap(int M, int FL);

// The general fused stage (spk::kFusedGen1 / kFusedGen2): every channel of
// the power-of-two IFFT sizes 8..256 runs inside the forward kernel when
// their tiles fit its LDS, no bin feeds more than two tile positions, and
// they take at most kGenMaxRounds rounds of the workgroup's threads.  Tile
// position i of channel c holds get_bin(first + j), j = (i + L/2) mod L
// (fcfb.rs:164-172): bin m mod N directly when below K, else conj of bin
// N - (m mod N) (fftutil.rs:8-22).  Each bin is scattered to its targets
// through one or two index tables laid out like the scatter table of the
// fixed-size stage.  Leaves the channel set unfused (SPK_OK) otherwise.
constexpr int kGenMaxRounds = 8;
spk_status fuse_general(spk_ctx* c) {
  if (c->f32_in) return SPK_OK;  // the general stage has the s16 loader only
  const int tile_cap = fwd_tile_cap(c->M, spk::kFusedGen1);
  if (tile_cap < 0 || fwd_tile_cap(c->M, spk::kFusedGen2) != tile_cap) return SPK_OK;
  const int M = c->M, R = c->fplan->R, T = M / R, N = c->N, K = c->K;
  std::vector<uint32_t> ids;
  for (const auto& kv : c->groups) {
    const int L = kv.first;
    if (!kv.second.generic && L >= 8 && L <= 256) ids.insert(ids.end(), kv.second.ids.begin(), kv.second.ids.end());
  }
  if (ids.empty()) return SPK_OK;
  // LDS regions, in ascending L (the group map's order)
  std::vector<spk::GenTile> tiles;
  int off = 0;
  for (uint32_t id : ids) {
    const int L = c->chans[id].L;
    tiles.push_back(spk::GenTile{c->chan_out_off[id], off, c->chan_w_off[id]});
    off += L + (L >> 5);
  }
  // Idle lanes (padding a wave to one size class) run the tile code on a
  // spare region at the end of the tile area, with zero weights, and write
  // no output: no lane writes into another tile's region.
  constexpr int kSpare = 256 + (256 >> 5);
  if (off > tile_cap - kSpare) return SPK_OK;
  const int spare = (int)tiles.size();
  tiles.push_back(spk::GenTile{0, tile_cap - kSpare, (int)c->weights_live});
  // targets of every bin 0..M
  const uint16_t dummy = (uint16_t)fwd_lds_dummy(M, spk::kFusedGen1);
  std::vector<uint16_t> tgt((size_t)2 * (M + 1), dummy);
  std::vector<uint8_t> cnt((size_t)M + 1, 0);
  for (size_t ti = 0; ti < ids.size(); ++ti) {
    const Channel& ch = c->chans[ids[ti]];
    const int L = ch.L;
    for (int j = 0; j < L; ++j) {
      long long m = (ch.first + j) % N;
      if (m < 0) m += N;
      const bool conj = m >= K;
      const int src = conj ? (int)(N - m) : (int)m;
      const int li = tiles[ti].lds + spk::lds_pad((j + L / 2) & (L - 1));
      if (cnt[(size_t)src] >= 2) return SPK_OK;  // a bin with three targets: unfused
      tgt[(size_t)cnt[(size_t)src]++ * (M + 1) + src] = (uint16_t)(li | (conj ? 0x8000 : 0));
    }
  }
  int D = 1;
  for (uint8_t k : cnt) D = std::max(D, (int)k);
  // thread slots: per size class, whole waves of one class
  std::vector<int> slots;
  for (int cls = 0; cls < 6; ++cls) {
    const int L = 8 << cls, TF = L < 16 ? 1 : L / 16;
    const size_t start = slots.size();
    for (size_t ti = 0; ti < ids.size(); ++ti)
      if (c->chans[ids[ti]].L == L)
        for (int f = 0; f < TF; ++f) slots.push_back(spk::gen_slot((int)ti, f, cls, false));
    while ((slots.size() - start) % 64) slots.push_back(spk::gen_slot(spare, 0, cls, true));
  }
  const int rounds = (int)((slots.size() + T - 1) / T);
  if (rounds > kGenMaxRounds) return SPK_OK;
  slots.resize((size_t)rounds * T, spk::gen_slot(spare, 0, 0, true));
  // lane-major index tables (the Welch partial-row order of the forward kernel)
  std::vector<uint16_t> tab((size_t)D * M, dummy);
  for (int d = 0; d < D; ++d)
    for (int k = 0; k < M; ++k) {
      const int idx = spk::welch_row_index(k, M, R);
      const int q = idx / T, t = idx % T;
      tab[(size_t)d * M + ((size_t)(q >> 3) * T + t) * 8 + (q & 7)] = tgt[(size_t)d * (M + 1) + k];
    }
  HIP_TRY(c, upload(c->fscat, tab));
  // per slot: its tile's descriptor, loaded beside the slot rather than
  // after it
  std::vector<spk::GenTile> slot_tiles(slots.size());
  for (size_t i = 0; i < slots.size(); ++i) slot_tiles[i] = tiles[(size_t)((slots[i] >> 8) & 0x7FFFFF)];
  HIP_TRY(c, upload(c->gtiles, slot_tiles));
  HIP_TRY(c, upload(c->gslots, slots));
  for (int i = 0; i < 4; ++i)
    if (!c->gentw[i].p) HIP_TRY(c, upload(c->gentw[i], pass_table(2 << i, 16, +1)));
  c->fscat_nyq = tgt[(size_t)M];
  c->fscat_nyq2 = tgt[(size_t)(M + 1) + M];
  c->grounds = rounds;
  c->fmode = D == 1 ? spk::kFusedGen1 : spk::kFusedGen2;
  c->fused = true;
  for (auto& kv : c->groups)
    if (!kv.second.generic && kv.first >= 8 && kv.first <= 256) kv.second.fused = true;
  return SPK_OK;
}

// Lay out channel outputs and descriptors, upload weights and tables.
// Every captured graph and batch plan refers to the old buffers: the plans
// are freed and the graphs marked stale (run_blocks updates a stale graph
// before it is launched again).
spk_status pull_weights(spk_ctx* c);

spk_status rebuild_channels(spk_ctx* c) {
  // Tables are uploaded with null-stream copies, which do not wait for the
  // context stream: no launch of an earlier call may still read them.
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  {
    spk_status st = pull_weights(c);
    if (st != SPK_OK) return st;
  }
  for (auto& kv : c->graphs) kv.second.stale = true;
  c->batches.clear();
  std::vector<float> wall;
  c->chan_out_off.assign(c->chans.size(), -1);
  c->chan_w_off.assign(c->chans.size(), -1);
  long long off = 0;
  std::vector<int>& w_off = c->chan_w_off;
  // Channels of one IFFT size with bit-identical weights share one copy: the
  // channel kernels re-read the weights every block, and a large set (2000
  // wideband channels: 18 MB, four distinct vectors) would not stay in L2.
  // The packed layout (one copy per channel, id order) is kept when the
  // weights are addressed per channel: device uploads and the adjoint
  // (contexts created with SPK_FLAG_KEEP_SPECTRA).
  const bool share = !c->keep_flag && !c->w_packed;
  std::map<std::vector<uint32_t>, int> seen;
  size_t packed = 0;
  for (size_t i = 0; i < c->chans.size(); ++i) {
    const Channel& ch = c->chans[i];
    if (ch.removed) continue;
    packed += (size_t)ch.L;
    c->chan_out_off[i] = off;
    off += (long long)c->max_blocks * chan_out_len(ch.L);
    if (share) {
      std::vector<uint32_t> key(ch.w.size() + 1);
      key[0] = (uint32_t)ch.L;
      std::memcpy(key.data() + 1, ch.w.data(), ch.w.size() * sizeof(float));
      auto it = seen.find(key);
      if (it != seen.end()) {
        w_off[i] = it->second;
        continue;
      }
      seen.emplace(std::move(key), (int)wall.size());
    }
    w_off[i] = (int)wall.size();
    wall.insert(wall.end(), ch.w.begin(), ch.w.end());
  }
  c->w_shared = wall.size() < packed;
  {
    std::vector<spk::ChanRecDesc> recs;
    long long prefix = 0;
    for (size_t i = 0; i < c->chans.size(); ++i) {
      if (c->chans[i].removed) continue;
      const int len = chan_out_len(c->chans[i].L);
      recs.push_back(spk::ChanRecDesc{c->chan_out_off[i], prefix, len, 0});
      prefix += len;
    }
    HIP_TRY(c, upload(c->crecs, recs));
    c->live_chans = (int)recs.size();
    c->live_len = prefix;
  }
  // 256 zeros after the last channel's weights: the idle lanes of the
  // general fused stage read them (fuse_general)
  c->weights_live = wall.size();
  wall.resize(wall.size() + 256, 0.f);
  HIP_TRY(c, upload(c->weights, wall));
  HIP_TRY(c, c->chan_out.alloc((size_t)off * sizeof(cpx)));
  c->groups.clear();
  for (size_t i = 0; i < c->chans.size(); ++i) {
    if (c->chans[i].removed) continue;
    ChanGroup& g = c->groups[c->chans[i].L];
    g.L = c->chans[i].L;
    g.ids.push_back((uint32_t)i);
  }
  // packed layout (spk_channels_output_packed, the adjoint's ybar): the
  // live channels in id order, 6 (L/8) samples per block each
  std::vector<long long> pk_prefix(c->chans.size(), 0);
  {
    long long acc = 0;
    for (size_t i = 0; i < c->chans.size(); ++i) {
      if (c->chans[i].removed) continue;
      pk_prefix[i] = acc;
      acc += chan_out_len(c->chans[i].L);
    }
  }
  for (auto& kv : c->groups) {
    ChanGroup& g = kv.second;
    const ChanPlan* p = find_chan_plan(g.L);
    std::vector<spk::ChanDesc> d;
    std::vector<long long> pk;
    for (uint32_t id : g.ids) {
      const Channel& ch = c->chans[id];
      long long fm = ch.first % c->N;
      if (fm < 0) fm += c->N;
      d.push_back(spk::ChanDesc{c->chan_out_off[id], (int)fm, w_off[id]});
      pk.push_back(pk_prefix[id]);
    }
    HIP_TRY(c, upload(g.desc, d));
    HIP_TRY(c, upload(g.pk_off, pk));
    g.generic = p == nullptr;
    if (g.generic) {
      g.radix = generic_radices(g.L);
      std::vector<cpx> tw((size_t)g.L);
      for (int e = 0; e < g.L; ++e) {
        const double a = 2.0 * M_PI * (double)e / (double)g.L;
        tw[(size_t)e] = cpx{(float)std::cos(a), (float)std::sin(a)};
      }
      HIP_TRY(c, upload(g.gtw, tw));
      continue;
    }
    g.ps_shape = shape_of(c->shape_ps, g.L, false);
    const ShapeInfo sh = tile_shapes(g.L).at((size_t)g.ps_shape);
    HIP_TRY(c, upload(g.tw2, pass_table(sh.R2, sh.R1, +1)));
    HIP_TRY(c, upload(g.tw3, pass_table(sh.R3, sh.R1 * sh.R2, +1)));
    HIP_TRY(c, upload(g.atw2, pass_table(p->R2, p->R1, -1)));
    HIP_TRY(c, upload(g.atw3, pass_table(p->R3, p->R1 * p->R2, -1)));
  }
  // The forward kernel runs the L = 256 channels itself when one 16-thread
  // tile per channel fits in its workgroup (the fixed-size stage), or,
  // otherwise, every channel of a power-of-two IFFT size up to 256 when
  // their tiles fit its LDS (the general stage, fuse_general).
  c->fused = false;
  c->fmode = 0;
  const auto it = c->groups.find(kFusedL);
  c->fused = !c->two_step && c->groups.size() == 1 && it != c->groups.end() &&
             (int)it->second.ids.size() * 16 <= c->M / c->fplan->R;
  if (c->fused) {
    // The fused stage takes channels inside bins [0, K) with no bin shared
    // (get_bin mirrors and overlapping channels go to the channel kernel).
    // Each lane's slot q bin is scattered to LDS index tab[q / 8][lane][q % 8]
    // (its tile position, or the dummy slot).
    const ChanGroup& g = it->second;
    const int M = c->M, R = c->fplan->R, T = M / R;
    const int dummy = fused_lds_dummy(M);
    std::vector<uint16_t> tab((size_t)M, (uint16_t)dummy);
    std::vector<int> owner((size_t)c->K, -1);
    int nyq = dummy;
    for (size_t tile = 0; tile < g.ids.size() && c->fused; ++tile) {
      const Channel& ch = c->chans[g.ids[tile]];
      if (ch.first < 0 || ch.first + kFusedL > c->K) {
        c->fused = false;
        break;
      }
      for (int j = 0; j < kFusedL; ++j) {
        const int k = (int)ch.first + j;
        if (owner[(size_t)k] != -1) {
          c->fused = false;
          break;
        }
        owner[(size_t)k] = (int)tile;
        const int li = (int)tile * spk::FusedCfg<kFusedL>::STRIDE + spk::lds_pad((j + kFusedL / 2) & (kFusedL - 1));
        if (k == M) {
          nyq = li;
        } else {
          const int idx = spk::welch_row_index(k, M, R);
          const int q = idx / T, t = idx % T;
          tab[((size_t)(q >> 3) * T + t) * 8 + (q & 7)] = (uint16_t)li;
        }
      }
    }
    if (c->fused) {
      HIP_TRY(c, upload(c->fscat, tab));
      c->fscat_nyq = nyq;
      c->fmode = kFusedL;
      c->groups.at(kFusedL).fused = true;
    }
  }
  if (!c->fused && !c->two_step) {
    spk_status st = fuse_general(c);
    if (st != SPK_OK) return st;
  }
  c->chans_dirty = false;
  return SPK_OK;
}

template <int M, int R, int R1, int R2, int FL, bool F32>
hipError_t launch_fwd(const spk::FwdParams& P, int G, hipStream_t s) {
  hipLaunchKernelGGL((spk::fwd_real_kernel<M, R, R1, R2, FL, F32>), dim3(G), dim3(M / R), 0, s, P);
  return hipGetLastError();
}

template <int M, int R, int R1, int R2, int FL, bool F32>
int fwd_wgs_per_cu() {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, (const void*)spk::fwd_real_kernel<M, R, R1, R2, FL, F32>,
                                                   M / R, 0) != hipSuccess)
    return 0;
  return n;
}

struct FwdVariant {
  int M, R, FL;
  bool f32;  // f32 loader (real f32 / u8 input) instead of s16 pairs
  hipError_t (*launch)(const spk::FwdParams&, int, hipStream_t);
  int (*occupancy)();
  int lds_dummy;  // LDS cpx index of the fused stage's dummy scatter slot
  int tile_cap;   // cpx of LDS the general fused stage's tiles may use
};
#define SPK_FWD_VARIANT1(M, R, R1, R2, FL, F32) \
  {M, R, FL, F32, &launch_fwd<M, R, R1, R2, FL, F32>, &fwd_wgs_per_cu<M, R, R1, R2, FL, F32>, \
   spk::FwdRealCfg<M, R, R1, R2, FL>::LDS_DUMMY, spk::FwdRealCfg<M, R, R1, R2, FL>::LDS_FLOATS / 2}
#define SPK_FWD_VARIANT(M, R, R1, R2)                                                               \
  SPK_FWD_VARIANT1(M, R, R1, R2, 0, false), SPK_FWD_VARIANT1(M, R, R1, R2, 256, false),                 \
      SPK_FWD_VARIANT1(M, R, R1, R2, 0, true), SPK_FWD_VARIANT1(M, R, R1, R2, 256, true),               \
      SPK_FWD_VARIANT1(M, R, R1, R2, spk::kFusedGen1, false), SPK_FWD_VARIANT1(M, R, R1, R2, spk::kFusedGen2, false)
const FwdVariant kFwdVariants[] = {
    SPK_FWD_VARIANT(512, 8, 8, 8),      SPK_FWD_VARIANT(1024, 16, 8, 8),    SPK_FWD_VARIANT(2048, 16, 8, 16),
    SPK_FWD_VARIANT(4096, 16, 16, 16),  SPK_FWD_VARIANT(8192, 32, 16, 16),  SPK_FWD_VARIANT(16384, 32, 16, 32),
    SPK_FWD_VARIANT(32768, 32, 32, 32),
};
const FwdVariant* find_fwd_variant(int M, int FL, bool f32) {
  for (const auto& v : kFwdVariants)
    if (v.M == M && v.FL == FL && v.f32 == f32) return &v;
  return nullptr;
}
int fwd_lds_dummy(int M, int FL) {
  const FwdVariant* v = find_fwd_variant(M, FL, false);
  return v ? v->lds_dummy : -1;
}
int fwd_tile_cap(int M, int FL) {
  const FwdVariant* v = find_fwd_variant(M, FL, false);
  return v ? v->tile_cap : -1;
}
int fused_lds_dummy(int M) {
  const FwdVariant* v = find_fwd_variant(M, kFusedL, false);
  return v ? v->lds_dummy : 0;
}
template <int L, int R, int R1, int R2, int R3>
void launch_chan(const spk::ChanParams& P, hipStream_t s) {
  constexpr int T = L / R;
  constexpr int TPW = spk::kChanThreads / T;
  const long long tiles = (long long)P.nch * P.nblocks;
  const unsigned grid = (unsigned)((tiles + TPW - 1) / TPW);
  hipLaunchKernelGGL((spk::chan_kernel<L, R, R1, R2, R3>), dim3(grid), dim3(spk::kChanThreads), 0, s, P);
}

template <int L, int S>
bool launch_chan_shape(const spk::ChanParams& P, hipStream_t s) {
  if constexpr (spk::TileShape<L, S>::valid) {
    using TS = spk::TileShape<L, S>;
    launch_chan<L, TS::R, TS::R1, TS::R2, TS::R3>(P, s);
    return true;
  } else {
    return false;
  }
}
template <int L>
bool launch_chan_any(int S, const spk::ChanParams& P, hipStream_t s) {
  switch (S) {
    case 0: return launch_chan_shape<L, 0>(P, s);
    case 1: return launch_chan_shape<L, 1>(P, s);
    case 2: return launch_chan_shape<L, 2>(P, s);
    default: return false;
  }
}
// Per-size launch of IFFT size L in tile shape S (the tables in P must be
// that shape's).
hipError_t dispatch_chan(int L, int S, const spk::ChanParams& P, hipStream_t s) {
  bool ok = false;
  switch (L) {
    case 8: ok = launch_chan_any<8>(S, P, s); break;
    case 16: ok = launch_chan_any<16>(S, P, s); break;
    case 32: ok = launch_chan_any<32>(S, P, s); break;
    case 64: ok = launch_chan_any<64>(S, P, s); break;
    case 128: ok = launch_chan_any<128>(S, P, s); break;
    case 256: ok = launch_chan_any<256>(S, P, s); break;
    case 512: ok = launch_chan_any<512>(S, P, s); break;
    case 1024: ok = launch_chan_any<1024>(S, P, s); break;
    case 2048: ok = launch_chan_any<2048>(S, P, s); break;
    case 4096: ok = launch_chan_any<4096>(S, P, s); break;
    case 8192: ok = launch_chan_any<8192>(S, P, s); break;
    default: break;
  }
  if (!ok) return hipErrorInvalidValue;
  return hipGetLastError();
}

template <int L, int R, int R1, int R2, int R3>
void launch_chan_adjoint(const spk::ChanAdjParams& P, hipStream_t s) {
  hipLaunchKernelGGL((spk::chan_adjoint_kernel<L, R, R1, R2, R3>), dim3(P.nch * P.S), dim3(spk::kChanThreads), 0, s,
                     P);
}

hipError_t dispatch_chan_adjoint(int L, const spk::ChanAdjParams& P, hipStream_t s) {
  switch (L) {
    case 8: launch_chan_adjoint<8, 4, 4, 2, 1>(P, s); break;
    case 16: launch_chan_adjoint<16, 4, 4, 4, 1>(P, s); break;
    case 32: launch_chan_adjoint<32, 8, 8, 4, 1>(P, s); break;
    case 64: launch_chan_adjoint<64, 8, 8, 8, 1>(P, s); break;
    case 128: launch_chan_adjoint<128, 16, 16, 8, 1>(P, s); break;
    case 256: launch_chan_adjoint<256, 16, 16, 16, 1>(P, s); break;
    case 512: launch_chan_adjoint<512, 32, 32, 16, 1>(P, s); break;
    case 1024: launch_chan_adjoint<1024, 32, 32, 32, 1>(P, s); break;
    case 2048: launch_chan_adjoint<2048, 32, 32, 32, 2>(P, s); break;
    case 4096: launch_chan_adjoint<4096, 16, 16, 16, 16>(P, s); break;
    case 8192: launch_chan_adjoint<8192, 32, 32, 16, 16>(P, s); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int MAXL>
int batched_wgs_per_cu() {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, (const void*)spk::chan_batched_kernel<MAXL>,
                                                   spk::kChanThreads, 0) != hipSuccess)
    return 0;
  return n;
}

// Threads per tile of the mixed-radix channel transform: about two outputs
// per thread and stage, 32 to 256.
int gen_tile_threads(int L) {
  int T = 32;
  while (T < 256 && 2 * T < L) T *= 2;
  return T;
}
// A mixed-radix size whose tiles (two buffers of L values each) fit the
// batched kernel's LDS.
bool gen_fits_batch(int L) {
  if (L > spk::kGenMaxL) return false;
  const long long bytes = (long long)(spk::kChanThreads / gen_tile_threads(L)) * 2 * L * (long long)sizeof(cpx);
  return bytes <= (long long)spk::batch_kernel_lds_floats<4096>() * (long long)sizeof(float);
}
// Channels that the batched kernel serves: every power-of-two group except
// the fused one, and the mixed-radix groups whose tiles fit its LDS (up to
// spk::kMaxClasses classes; other groups keep their own launch).
std::vector<int> batched_groups(const spk_ctx* c) {
  std::vector<int> ls, gen;
  for (const auto& kv : c->groups) {
    if (kv.second.fused) continue;
    if (!kv.second.generic) ls.push_back(kv.first);
    else if (gen_fits_batch(kv.first)) gen.push_back(kv.first);
  }
  for (int L : gen)
    if ((int)ls.size() < spk::kMaxClasses) ls.push_back(L);
  return ls;
}
bool in_batch(const spk_ctx* c, int L) {
  const std::vector<int> ls = batched_groups(c);
  return std::find(ls.begin(), ls.end(), L) != ls.end();
}

// One launch of spk::chan_generic_kernel for a group without a
// power-of-two plan.
hipError_t launch_generic(const spk_ctx* c, const ChanGroup& g, int nb, hipStream_t s) {
  if (g.L > spk::kGenMaxL) {
    spk::GenChanParams Q{};
    Q.spec = c->spec.as<cpx>();
    Q.spec_stride = c->spec_stride;
    Q.N = c->N;
    Q.K = c->K;
    Q.real_input = c->complex_in ? 0 : 1;
    Q.weights = c->weights.as<float>();
    Q.chans = g.desc.as<spk::ChanDesc>();
    Q.nch = (int)g.ids.size();
    Q.nblocks = nb;
    Q.out = c->chan_out.as<cpx>();
    Q.tw = g.gtw.as<cpx>();
    Q.L = g.L;
    Q.nstages = (int)g.radix.size();
    for (size_t i = 0; i < g.radix.size(); ++i) Q.radix[i] = g.radix[i];
    Q.tile_threads = spk::kGenBigThreads;
    const unsigned grid = (unsigned)((long long)Q.nch * nb);
    hipLaunchKernelGGL(spk::chan_generic_big_kernel, dim3(grid), dim3(spk::kGenBigThreads), 0, s, Q);
    return hipGetLastError();
  }
  spk::GenChanParams Q{};
  Q.spec = c->spec.as<cpx>();
  Q.spec_stride = c->spec_stride;
  Q.N = c->N;
  Q.K = c->K;
  Q.real_input = c->complex_in ? 0 : 1;
  Q.weights = c->weights.as<float>();
  Q.chans = g.desc.as<spk::ChanDesc>();
  Q.nch = (int)g.ids.size();
  Q.nblocks = nb;
  Q.out = c->chan_out.as<cpx>();
  Q.tw = g.gtw.as<cpx>();
  Q.L = g.L;
  Q.nstages = (int)g.radix.size();
  for (size_t i = 0; i < g.radix.size(); ++i) Q.radix[i] = g.radix[i];
  const int T = gen_tile_threads(g.L);
  Q.tile_threads = T;
  const int tpw = spk::kGenThreads / T;
  const long long tiles = (long long)Q.nch * nb;
  const unsigned grid = (unsigned)((tiles + tpw - 1) / tpw);
  const size_t lds = (size_t)tpw * 2 * g.L * sizeof(cpx);
  hipLaunchKernelGGL(spk::chan_generic_kernel, dim3(grid), dim3(spk::kGenThreads), lds, s, Q);
  return hipGetLastError();
}

// Cut the (block, channel) tiles of nb blocks into per-XCD item queues.
spk_status build_batch(spk_ctx* c, int nb) {
  HIP_TRY(c, hipStreamSynchronize(c->stream));  // as in rebuild_channels
  if (c->batches.size() >= kMaxCachedPlans && !c->batches.count(nb)) drop_graphs_and_plans(c);
  BatchPlan& B = c->batches[nb];
  std::vector<int> ls = batched_groups(c);
  std::sort(ls.begin(), ls.end(), std::greater<int>());  // big tiles first within a chunk
  B.cls_L = ls;
  B.maxL = 0;
  B.big = false;
  std::vector<spk::ChanDesc> descs;
  std::vector<int> ch0(ls.size()), nch(ls.size()), ttile(ls.size());
  const size_t nc = ls.size();
  B.tw2.clear();
  B.tw3.clear();
  B.tw2.resize(nc);
  B.tw3.resize(nc);
  B.cls_shape.assign(nc, -1);
  B.twp2.assign(nc, nullptr);
  B.twp3.assign(nc, nullptr);
  B.gen_T.assign(nc, 0);
  B.gen_nst.assign(nc, 0);
  B.gen_radix.assign(nc, std::vector<int>());
  for (size_t k = 0; k < nc; ++k) {
    const ChanGroup& g = c->groups.at(ls[k]);
    ch0[k] = (int)descs.size();
    nch[k] = (int)g.ids.size();
    for (uint32_t id : g.ids) {
      const Channel& ch = c->chans[id];
      long long fm = ch.first % c->N;
      if (fm < 0) fm += c->N;
      descs.push_back(spk::ChanDesc{c->chan_out_off[id], (int)fm, c->chan_w_off[id]});
    }
    const int L = ls[k];
    if (g.generic) {
      ttile[k] = gen_tile_threads(L);
      B.gen_T[k] = ttile[k];
      B.gen_nst[k] = (int)g.radix.size();
      B.gen_radix[k] = g.radix;
      B.twp2[k] = g.gtw.as<cpx>();
      B.twp3[k] = g.gtw.as<cpx>();
      continue;
    }
    B.maxL = std::max(B.maxL, L);
    const int S = shape_of(c->shape_b, L, true);
    const ShapeInfo sh = tile_shapes(L).at((size_t)S);
    B.big = B.big || L > 4096 || sh.big();
    B.cls_shape[k] = S;
    ttile[k] = L / sh.R;
    HIP_TRY(c, upload(B.tw2[k], pass_table(sh.R2, sh.R1, +1)));
    HIP_TRY(c, upload(B.tw3[k], pass_table(sh.R3, sh.R1 * sh.R2, +1)));
    B.twp2[k] = B.tw2[k].as<cpx>();
    B.twp3[k] = B.tw3[k].as<cpx>();
  }
  std::vector<std::vector<spk::BatchItem>> q(kXcdQueues);
  const int chunk = std::max(1, c->tune_chunk);
  for (int g0 = 0, k = 0; g0 < nb; g0 += chunk, ++k) {
    const int nbk = std::min(chunk, nb - g0);
    auto& qx = q[k % kXcdQueues];
    for (size_t cl = 0; cl < ls.size(); ++cl) {
      const int tpw = spk::kChanThreads / ttile[cl];
      const int per = tpw * std::max(1, c->tune_rounds);
      const int ntiles = nch[cl] * nbk;
      for (int t0 = 0; t0 < ntiles; t0 += per)
        qx.push_back(spk::BatchItem{(int)cl, ch0[cl], nch[cl], g0, t0, std::min(t0 + per, ntiles)});
    }
  }
  std::vector<spk::BatchItem> all;
  std::vector<int> off(kXcdQueues + 1, 0);
  for (int x = 0; x < kXcdQueues; ++x) {
    off[x] = (int)all.size();
    all.insert(all.end(), q[x].begin(), q[x].end());
  }
  off[kXcdQueues] = (int)all.size();
  HIP_TRY(c, upload(B.descs, descs));
  HIP_TRY(c, upload(B.items, all));
  HIP_TRY(c, upload(B.item_off, off));
  B.nitems = all.size();
  B.nb = nb;
  return SPK_OK;
}

spk_status launch_batched(spk_ctx* c, int nb, hipStream_t s) {
  if (!c->batches.count(nb)) {  // (graph capture builds the plan first, in run_blocks)
    spk_status st = build_batch(c, nb);
    if (st != SPK_OK) return st;
  }
  const BatchPlan& B = c->batches.at(nb);
  if (B.cls_L.empty()) return SPK_OK;
  spk::BatchParams P{};
  P.spec = c->spec.as<cpx>();
  P.spec_stride = c->spec_stride;
  P.N = c->N;
  P.K = c->K;
  P.real_input = c->complex_in ? 0 : 1;
  P.weights = c->weights.as<float>();
  P.descs = B.descs.as<spk::ChanDesc>();
  P.out = c->chan_out.as<cpx>();
  P.items = B.items.as<spk::BatchItem>();
  P.item_off = B.item_off.as<int>();
  P.counters = c->chan_counters.as<int>();
  P.X = kXcdQueues;
  for (size_t k = 0; k < B.cls_L.size(); ++k) {
    P.cls_L[k] = B.cls_L[k];
    P.cls_shape[k] = B.cls_shape[k];
    P.tw2[k] = B.twp2[k];
    P.tw3[k] = B.twp3[k];
    P.gen_T[k] = B.gen_T[k];
    P.gen_nst[k] = B.gen_nst[k];
    for (size_t i = 0; i < B.gen_radix[k].size(); ++i) P.gen_radix[k][i] = B.gen_radix[k][i];
  }
  const bool big = B.big;
  int per_cu = big ? batched_wgs_per_cu<8192>() : batched_wgs_per_cu<4096>();
  if (c->tune_wgs_per_cu > 0) per_cu = std::min(per_cu, c->tune_wgs_per_cu);
  int G = std::max(1, per_cu) * c->ncu;
  G = (G + kXcdQueues - 1) / kXcdQueues * kXcdQueues;
  if (big)
    hipLaunchKernelGGL(spk::chan_batched_kernel<8192>, dim3(G), dim3(spk::kChanThreads), 0, s, P);
  else
    hipLaunchKernelGGL(spk::chan_batched_kernel<4096>, dim3(G), dim3(spk::kChanThreads), 0, s, P);
  HIP_TRY(c, hipGetLastError());
  return SPK_OK;
}

// One launch sequence over nb blocks whose samples start at src.
// Enqueue the launch sequence over nb blocks whose samples start at src:
// forward kernel, channel stage, Welch finaliser, stream-position advance.
// Kernels only (no memset or copy nodes), so it can be captured into a
// graph; every per-call quantity the kernels need is in device memory.
template <int N1, int R, int R1, int R2, int R3>
void launch_cplx_rows(const spk::CplxRowParams& P, int nb, bool s16, int ncu, hipStream_t s) {
  // at least four rows per workgroup (rows of 4096 points: 1024 threads),
  // eight for rows of 2048 points (512 threads, 135 KB of LDS: a cache line
  // of s16 pairs then serves four workgroups instead of eight; rx888_256k
  // 0.605 -> 0.569 ms per step)
  constexpr int WGT = N1 == 2048 ? 8 * (N1 / R) : (N1 / R >= spk::kChanThreads ? 4 * (N1 / R) : spk::kChanThreads);
  constexpr int TPW = WGT / (N1 / R);
  if constexpr (N1 >= 2048) {
    if (8LL * P.H >= P.O) {
      // one workgroup per CU: the persistent form, npar workgroups per
      // (XCD, row tile); its prefetched blocks (g >= 8) lie past the history
      constexpr int TILES = 64 / TPW;
      const int grid = 8 * TILES * std::max(1, ncu / (8 * TILES));
      if (s16)
        hipLaunchKernelGGL((spk::cplx_row_pf_kernel<N1, R, R1, R2, R3, true, WGT>), dim3(grid), dim3(WGT), 0, s, P);
      else
        hipLaunchKernelGGL((spk::cplx_row_pf_kernel<N1, R, R1, R2, R3, false, WGT>), dim3(grid), dim3(WGT), 0, s, P);
      return;
    }
  }
  const int grid = 8 * ((nb + 7) / 8) * (64 / TPW);
  if (s16)
    hipLaunchKernelGGL((spk::cplx_row_kernel<N1, R, R1, R2, R3, true, WGT>), dim3(grid), dim3(WGT), 0, s, P);
  else
    hipLaunchKernelGGL((spk::cplx_row_kernel<N1, R, R1, R2, R3, false, WGT>), dim3(grid), dim3(WGT), 0, s, P);
}

hipError_t dispatch_cplx_rows(int N1, const spk::CplxRowParams& P, int nb, bool s16, int ncu, hipStream_t s) {
  switch (N1) {
    case 16: launch_cplx_rows<16, 4, 4, 4, 1>(P, nb, s16, ncu, s); break;
    case 32: launch_cplx_rows<32, 8, 8, 4, 1>(P, nb, s16, ncu, s); break;
    case 64: launch_cplx_rows<64, 8, 8, 8, 1>(P, nb, s16, ncu, s); break;
    case 128: launch_cplx_rows<128, 16, 16, 8, 1>(P, nb, s16, ncu, s); break;
    case 256: launch_cplx_rows<256, 16, 16, 16, 1>(P, nb, s16, ncu, s); break;
    case 512: launch_cplx_rows<512, 32, 32, 16, 1>(P, nb, s16, ncu, s); break;
    case 1024: launch_cplx_rows<1024, 32, 32, 32, 1>(P, nb, s16, ncu, s); break;
    case 2048: launch_cplx_rows<2048, 32, 32, 32, 2>(P, nb, s16, ncu, s); break;
    case 4096: launch_cplx_rows<4096, 16, 16, 16, 16>(P, nb, s16, ncu, s); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

spk_status enqueue_channels(spk_ctx* c, int nb, hipStream_t s);

// Formats without a loader of their own: convert the call's nb * H raw
// samples into c->conv and return it as the transform's source (the input
// itself otherwise).  Launch errors surface at the next hipGetLastError.
const void* enqueue_convert(spk_ctx* c, const void* raw, int nb, hipStream_t s) {
  if (!c->conv_kind) return raw;
  spk::ConvParams P{};
  P.src = static_cast<const uint32_t*>(raw);
  P.dst = c->conv.p;
  const long long bytes = (long long)nb * c->H * c->raw_bps;
  P.words = bytes / 4;
  P.tail = (int)(bytes % 4);
  P.kind = c->conv_kind;
  const long long want = (P.words + 255) / 256 + (P.tail ? 1 : 0);
  const int grid = (int)std::min<long long>(want, (long long)c->ncu * 8);
  hipLaunchKernelGGL(spk::convert_kernel, dim3(grid), dim3(256), 0, s, P);
  return c->conv.p;
}

template <int R>
void launch_gfft_pass(const spk::GfftPassParams& P, int load, dim3 grid, hipStream_t s) {
  switch (load) {
    case 0: hipLaunchKernelGGL((spk::gfft_pass_kernel<R, 0>), grid, dim3(spk::kGfftThreads), 0, s, P); break;
    case 1: hipLaunchKernelGGL((spk::gfft_pass_kernel<R, 1>), grid, dim3(spk::kGfftThreads), 0, s, P); break;
    default: hipLaunchKernelGGL((spk::gfft_pass_kernel<R, 2>), grid, dim3(spk::kGfftThreads), 0, s, P); break;
  }
}

hipError_t dispatch_gfft_pass(int R, const spk::GfftPassParams& P, int load, dim3 grid, hipStream_t s) {
  switch (R) {
    case 2: launch_gfft_pass<2>(P, load, grid, s); break;
    case 3: launch_gfft_pass<3>(P, load, grid, s); break;
    case 4: launch_gfft_pass<4>(P, load, grid, s); break;
    case 5: launch_gfft_pass<5>(P, load, grid, s); break;
    case 7: launch_gfft_pass<7>(P, load, grid, s); break;
    case 8: launch_gfft_pass<8>(P, load, grid, s); break;
    case 11: launch_gfft_pass<11>(P, load, grid, s); break;
    case 13: launch_gfft_pass<13>(P, load, grid, s); break;
    case 17: launch_gfft_pass<17>(P, load, grid, s); break;
    case 19: launch_gfft_pass<19>(P, load, grid, s); break;
    case 23: launch_gfft_pass<23>(P, load, grid, s); break;
    case 29: launch_gfft_pass<29>(P, load, grid, s); break;
    case 31: launch_gfft_pass<31>(P, load, grid, s); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// Four-step split Nc = N1 N2 with both line lengths within one LDS tile
// (N1 the largest such divisor up to sqrt(Nc)); false when there is none.
bool gfft4_split(int Nc, int* N1, int* N2) {
  for (int d = (int)std::sqrt((double)Nc); d >= 2; --d)
    if (Nc % d == 0 && d <= spk::kG4MaxLine && Nc / d <= spk::kG4MaxLine) {
      *N1 = d;
      *N2 = Nc / d;
      return true;
    }
  return false;
}

// Mixed-radix forward transform of nb blocks into the spectrum rows.
// Four-step form: columns (framing and windowing the input as they load
// it) into the scratch rows, then rows into the spectrum rows.  Otherwise
// one pass per radix: the first pass frames and windows the input, passes
// alternate between the scratch rows (cY) and the spectrum rows so that the
// last writes the spectrum rows.
spk_status enqueue_gfft(spk_ctx* c, const void* src, int nb, hipStream_t s) {
  const int pk = c->complex_in ? 1 : 2;
  // real input with an odd hop or overlap (N not a multiple of 8): framed in
  // samples, two per pair (1), or one per point for odd N (2)
  const int samples = c->real_odd ? 2 : (!c->complex_in && (c->H % 2 || c->O % 2) ? 1 : 0);
  const int Hc = samples ? c->H : c->H / pk, Oc = samples ? c->O : c->O / pk;
  if (c->gfft4) {
    spk::Gfft4Params P{};
    P.src = src;
    P.hist = c->hist.p;
    P.Hc = Hc;
    P.Oc = Oc;
    P.samples = samples;
    P.wscale = c->wscale;
    P.tw_full = c->gtw.as<cpx>();
    P.Nc = c->Nc;
    P.N1 = c->g4N1;
    P.N2 = c->g4N2;
    P.nblocks = nb;
    // A: N1-point columns, N2 of them
    P.out = c->cY.as<cpx>();
    P.out_stride = c->Nc;
    P.tw_line = c->g4tw1.as<cpx>();
    P.n = c->g4N1;
    P.nlines = c->g4N2;
    P.nstages = (int)c->g4r1.size();
    for (size_t i = 0; i < c->g4r1.size(); ++i) P.radix[i] = c->g4r1[i];
    // 16 lines while they fit: longer runs of the input pairs
    const int la_lines = c->g4N1 <= spk::kG4MaxLineWide ? spk::kG4LinesWide : spk::kG4Lines;
    const dim3 ga((unsigned)((c->g4N2 + la_lines - 1) / la_lines), (unsigned)nb);
    const size_t la = ((size_t)2 * la_lines * (c->g4N1 + 1) + c->g4N1) * sizeof(cpx);
    if (la_lines == spk::kG4LinesWide) {
      if (c->s16_pairs)
        hipLaunchKernelGGL((spk::gfft4_kernel<1, true, spk::kG4LinesWide>), ga, dim3(spk::kG4Threads), la, s, P);
      else
        hipLaunchKernelGGL((spk::gfft4_kernel<2, true, spk::kG4LinesWide>), ga, dim3(spk::kG4Threads), la, s, P);
    } else {
      if (c->s16_pairs)
        hipLaunchKernelGGL((spk::gfft4_kernel<1, true, spk::kG4Lines>), ga, dim3(spk::kG4Threads), la, s, P);
      else
        hipLaunchKernelGGL((spk::gfft4_kernel<2, true, spk::kG4Lines>), ga, dim3(spk::kG4Threads), la, s, P);
    }
    HIP_TRY(c, hipGetLastError());
    // B: N2-point rows, N1 of them
    P.Y = c->cY.as<cpx>();
    P.out = c->spec.as<cpx>();
    P.out_stride = c->spec_stride;
    P.tw_line = c->g4tw2.as<cpx>();
    P.n = c->g4N2;
    P.nlines = c->g4N1;
    P.nstages = (int)c->g4r2.size();
    for (size_t i = 0; i < c->g4r2.size(); ++i) P.radix[i] = c->g4r2[i];
    const dim3 gb((unsigned)((c->g4N1 + spk::kG4Lines - 1) / spk::kG4Lines), (unsigned)nb);
    const size_t lb = ((size_t)2 * spk::kG4Lines * (c->g4N2 + 1) + c->g4N2) * sizeof(cpx);
    hipLaunchKernelGGL((spk::gfft4_kernel<0, false, spk::kG4Lines>), gb, dim3(spk::kG4Threads), lb, s, P);
    HIP_TRY(c, hipGetLastError());
    return SPK_OK;
  }
  const int np = (int)c->gfft_radices.size();
  spk::GfftPassParams P{};
  P.src = src;
  P.hist = c->hist.p;
  P.Hc = Hc;
  P.Oc = Oc;
  P.samples = samples;
  P.wscale = c->wscale;
  P.tw = c->gtw.as<cpx>();
  P.Nc = c->Nc;
  P.nblocks = nb;
  int Ns = 1;
  for (int i = 0; i < np; ++i) {
    const int R = c->gfft_radices[i];
    const bool to_spec = ((np - 1 - i) % 2) == 0;
    P.out = to_spec ? c->spec.as<cpx>() : c->cY.as<cpx>();
    P.out_stride = to_spec ? c->spec_stride : c->Nc;
    P.Ns = Ns;
    const dim3 grid((unsigned)((c->Nc / R + spk::kGfftThreads - 1) / spk::kGfftThreads), (unsigned)nb);
    HIP_TRY(c, dispatch_gfft_pass(R, P, i == 0 ? (c->s16_pairs ? 1 : 2) : 0, grid, s));
    P.in = P.out;
    P.in_stride = P.out_stride;
    Ns *= R;
  }
  return SPK_OK;
}

// Complex input: row transforms, column transforms (natural-order spectra),
// Welch records, channel stage, stream-position advance.
spk_status enqueue_complex(spk_ctx* c, const void* src, int nb, hipStream_t s, hipEvent_t* ev) {
  const int N1 = c->Nc / 64;
  const int pk = c->complex_in ? 1 : 2;  // real samples per complex sample of the pipeline
  if (ev) HIP_TRY(c, hipEventRecord(ev[0], s));
  src = enqueue_convert(c, src, nb, s);
  if (c->gfft) {
    spk_status st = enqueue_gfft(c, src, nb, s);
    if (st != SPK_OK) return st;
  } else {
  spk::CplxRowParams P{};
  P.src = src;
  P.hist = c->hist.p;
  P.H = c->H / pk;
  P.O = c->O / pk;
  P.wscale = c->wscale;
  P.nblocks = nb;
  P.Y = c->cY.as<cpx>();
  P.tw2 = c->ctw2.as<cpx>();
  P.tw3 = c->ctw3.as<cpx>();
  P.ta = c->cta.as<cpx>();
  P.tb = c->ctb.as<cpx>();
  HIP_TRY(c, dispatch_cplx_rows(N1, P, nb, c->s16_pairs, c->ncu, s));
  spk::CplxColParams Q{};
  Q.Y = c->cY.as<cpx>();
  Q.N1 = N1;
  Q.nblocks = nb;
  Q.spec = c->spec.as<cpx>();
  Q.spec_stride = c->spec_stride;
  Q.tw64 = c->ctw64.as<cpx>();
  if (c->complex_in) {
    // 64 (block, column) pairs per workgroup
    hipLaunchKernelGGL(spk::cplx_col_kernel, dim3((unsigned)(((long long)nb * N1 + 63) / 64)),
                       dim3(spk::kChanThreads), 0, s, Q);
  } else {
    // real input (N1 >= 1024): columns with their mirror columns, untangled in the same launch
    hipLaunchKernelGGL(spk::cplx_col_untangle_kernel, dim3((unsigned)((long long)nb * (N1 / 64))),
                       dim3(spk::kChanThreads), 0, s, Q, c->ctu.as<cpx>());
  }
  HIP_TRY(c, hipGetLastError());
  }
  if (!c->complex_in && !c->real_odd && c->gfft) {
    spk::UntangleParams U{};
    U.spec = c->spec.as<cpx>();
    U.spec_stride = c->spec_stride;
    U.M = c->M;
    U.tw = c->ctu.as<cpx>();
    hipLaunchKernelGGL(spk::untangle_kernel, dim3((unsigned)((c->M / 2 + 1 + 255) / 256), (unsigned)nb), dim3(256), 0,
                       s, U);
    HIP_TRY(c, hipGetLastError());
  }
  if (ev) HIP_TRY(c, hipEventRecord(ev[1], s));
  spk_status st = enqueue_channels(c, nb, s);
  if (st != SPK_OK) return st;
  if (ev) HIP_TRY(c, hipEventRecord(ev[2], s));
  {
    spk::WelchDirectParams W{};
    W.spec = c->spec.as<cpx>();
    W.spec_stride = c->spec_stride;
    W.N = c->N;
    W.K = c->K;
    W.real_input = c->complex_in ? 0 : 1;
    W.nblocks = nb;
    W.A = c->A;
    W.pos = c->pos.as<long long>() + c->parity;
    W.pos_next = c->pos.as<long long>() + (c->parity ^ 1);
    W.carry_in = c->carry.as<float>() + (long long)c->parity * c->spec_stride;
    W.carry_out = c->carry.as<float>() + (long long)(c->parity ^ 1) * c->spec_stride;
    W.out = c->welch_out.as<float>();
    const long long hist_off = ((long long)nb * c->H - c->O) * c->bps;
    W.hist_src = reinterpret_cast<const char*>(src) + hist_off;
    W.hist_dst = c->hist.p;
    W.hist_words = c->O * c->bps / 4;
    // (s16 samples with an odd overlap or start: copied 16 bits at a time)
    W.hist_shorts = (c->O * c->bps) % 4 != 0 || hist_off % 4 != 0 ? c->O * c->bps / 2 : 0;
    W.chan_counters = c->chan_counters.as<int>();
    W.ncounters = kXcdQueues;
    W.qout = c->qout.as<unsigned char>();
    W.qfmt = c->qfmt;
    const int gmax = (c->A - 1 + nb - 1) / c->A + 1;
    if (c->complex_in && !c->gfft)  // four bins per thread; K = N is a multiple of 1024
      hipLaunchKernelGGL(spk::welch_direct_kernel, dim3(c->K / 1024, gmax), dim3(256), 0, s, W);
    else  // any K (complex input: K = N, get_bin wraps without a mirror)
      hipLaunchKernelGGL(spk::welch_direct_real_kernel, dim3((c->K + 1023) / 1024, gmax), dim3(256), 0, s, W);
    HIP_TRY(c, hipGetLastError());
  }
  if (ev) HIP_TRY(c, hipEventRecord(ev[3], s));
  return SPK_OK;
}

// Channel stage for every channel not fused into the forward kernel.
spk_status enqueue_channels(spk_ctx* c, int nb, hipStream_t s) {
  // sizes without a power-of-two plan: their own kernel, unless the batched
  // kernel takes them
  const bool batched = c->chan_mode == SPK_CHAN_BATCHED;
  for (auto& kv : c->groups)
    if (kv.second.generic && !(batched && in_batch(c, kv.first))) HIP_TRY(c, launch_generic(c, kv.second, nb, s));
  if (batched) return launch_batched(c, nb, s);
  for (auto& kv : c->groups) {
    ChanGroup& g = kv.second;
    if (g.fused || g.generic) continue;
    spk::ChanParams Q{};
    Q.spec = c->spec.as<cpx>();
    Q.spec_stride = c->spec_stride;
    Q.N = c->N;
    Q.K = c->K;
    Q.real_input = c->complex_in ? 0 : 1;
    Q.weights = c->weights.as<float>();
    Q.chans = g.desc.as<spk::ChanDesc>();
    Q.nch = (int)g.ids.size();
    Q.nblocks = nb;
    Q.out = c->chan_out.as<cpx>();
    Q.tw2 = g.tw2.as<cpx>();
    Q.tw3 = g.tw3.as<cpx>();
    HIP_TRY(c, dispatch_chan(g.L, g.ps_shape, Q, s));
  }
  return SPK_OK;
}

// Workgroups of the real forward launch over nb blocks.
int fwd_class(int fmode) { return fmode == 0 ? 0 : (fmode == kFusedL ? 1 : 2); }
int fwd_grid(const spk_ctx* c, int nb) { return std::min(nb, c->gmax[fwd_class(c->fmode)]); }

// Welch record parameters of the real-input call over nb blocks launched
// with G workgroups at the given parity (its position slot, partial rows
// and carry buffers).
spk::FinParams fin_params(const spk_ctx* c, int nb, int G, int parity) {
  spk::FinParams F{};
  F.wpart = c->wpart.as<float>() + parity * c->wpart_half;
  F.wstride = c->wstride;
  F.wslots = c->wslots;
  F.G = G;
  F.nblocks = nb;
  F.pos = c->pos.as<long long>() + parity;
  F.A = c->A;
  F.K = c->K;
  F.M = c->M;
  F.R = c->fplan->R;
  F.carry_in = c->carry.as<float>() + (long long)parity * c->spec_stride;
  F.carry_out = c->carry.as<float>() + (long long)(parity ^ 1) * c->spec_stride;
  F.out = c->welch_out.as<float>();
  F.qout = c->qout.as<unsigned char>();
  F.qfmt = c->qfmt;
  // Teams per chunk of the forward kernel's record phase: enough chunks for
  // the G workgroups when a call touches few Welch groups (long averages);
  // from quantities fixed per launch shape, so a captured graph sums in the
  // same order as an eager launch.
  const int T = c->M / c->fplan->R;
  const int ngroups = (c->A - 1 + nb - 1) / c->A + 1;
  const int chunks = ngroups * (c->fplan->R / 4);
  int split = 1;
  while (split < 16 && 2 * split * chunks <= G && T / (2 * split) >= 16) split *= 2;
  F.split = split;
  // Chunks per take: short averages make many chunks of few partial rows
  // each; take several at a time, keeping at least four takes per workgroup.
  const long long nchunks = (long long)chunks * split;
  int take = 1;
  while (take < 16 && nchunks >= 8LL * take * G) take *= 2;
  F.take = take;
  return F;
}

// Sum the pending call's Welch records now (its partial rows are complete
// once the launches before this one on the stream have run).
spk_status flush_records(spk_ctx* c) {
  if (c->two_step || c->pend_nb == 0) return SPK_OK;
  spk::FinParams F = fin_params(c, c->pend_nb, c->pend_G, c->pend_parity);
  // groups touched: at most (A - 1 + nb - 1) / A + 1; rows past the actual count exit
  const int gmax = (c->A - 1 + c->pend_nb - 1) / c->A + 1;
  // teams per workgroup: at least 4 partial rows per team (a group spans at
  // most A G / nb + 2 workgroups of the launch)
  const long long rows = std::min<long long>(c->pend_G, (long long)c->A * c->pend_G / c->pend_nb + 2);
  F.split = 1;
  while (F.split < 16 && 4LL * 2 * F.split <= rows) F.split *= 2;
  const int TS = 256 / F.split;
  hipLaunchKernelGGL(spk::welch_finalize_kernel, dim3((c->M / 4 + TS - 1) / TS, gmax), dim3(256), 0, c->stream, F);
  HIP_TRY(c, hipGetLastError());
  c->pend_nb = 0;
  return SPK_OK;
}

spk_status enqueue_blocks(spk_ctx* c, const void* src, int nb, hipStream_t s, hipEvent_t* ev) {
  if (c->two_step) return enqueue_complex(c, src, nb, s, ev);
  if (ev) HIP_TRY(c, hipEventRecord(ev[0], s));
  src = enqueue_convert(c, src, nb, s);
  spk::FwdParams P{};
  P.src = src;
  P.hist = c->hist.p;
  P.spec = c->spec.as<cpx>();
  P.spec_stride = c->spec_stride;
  // Per-block spectrum rows are needed when kept or read by a channel
  // kernel; otherwise each workgroup reuses one scratch row.
  size_t unfused_groups = 0;
  for (const auto& kv : c->groups) unfused_groups += kv.second.fused ? 0 : 1;
  P.keep_spec = (c->keep_spec || unfused_groups > 0) ? 1 : 0;
  P.tw2 = c->ftw2.as<cpx>();
  P.tw3 = c->ftw3.as<cpx>();
  P.twu = c->ftwu.as<cpx>();
  P.wpart = c->wpart.as<float>() + c->parity * c->wpart_half;
  P.wstride = c->wstride;
  P.wslots = c->wslots;
  P.A = c->A;
  P.pos = c->pos.as<long long>() + c->parity;
  P.wscale = c->wscale;
  P.nblocks = nb;
  P.H = c->H;
  P.O = c->O;
  P.edump = c->edump.as<cpx>();
  P.eitems = c->eitems.as<int4>();
  if (c->fmode == kFusedL) {
    const ChanGroup* fg = &c->groups.at(kFusedL);
    P.fchans = fg->desc.as<spk::ChanDesc>();
    P.fnch = (int)fg->ids.size();
    P.fscat = c->fscat.as<uint4>();
    P.fscat_nyq = c->fscat_nyq;
    P.weights = c->weights.as<float>();
    P.fout = c->chan_out.as<cpx>();
    P.ftw2 = fg->tw2.as<cpx>();
  }
  // Next call's housekeeping (workgroup 0) and the pending call's records.
  P.pos_next = c->pos.as<long long>() + (c->parity ^ 1);
  P.hist_src = reinterpret_cast<const int*>(reinterpret_cast<const char*>(src) + ((long long)nb * c->H - c->O) * c->bps);
  P.hist_dst = c->hist.as<int>();
  P.hist_words = c->O * c->bps / 4;
  P.chan_counters = c->chan_counters.as<int>();
  P.ncounters = kXcdQueues;
  if (c->pend_nb > 0) P.fin = fin_params(c, c->pend_nb, c->pend_G, c->pend_parity);
  P.fin_ticket = c->fin_ticket.as<int>();
  if (spk::fused_general(c->fmode)) {
    P.fscat = c->fscat.as<uint4>();
    P.fscat_nyq = c->fscat_nyq;
    P.fscat_nyq2 = c->fscat_nyq2;
    P.weights = c->weights.as<float>();
    P.fout = c->chan_out.as<cpx>();
    P.gtiles = c->gtiles.as<spk::GenTile>();
    P.gslots = c->gslots.as<int>();
    P.grounds = c->grounds;
    for (int i = 0; i < 4; ++i) P.gtw[i] = c->gentw[i].as<cpx>();
  }
  const FwdVariant* fv = find_fwd_variant(c->M, c->fmode, c->f32_in);
  const int G = fwd_grid(c, nb);
  HIP_TRY(c, fv->launch(P, G, s));
  if (ev) HIP_TRY(c, hipEventRecord(ev[1], s));

  {
    spk_status st = enqueue_channels(c, nb, s);
    if (st != SPK_OK) return st;
  }
  if (ev) HIP_TRY(c, hipEventRecord(ev[2], s));
  // (the records of this call are summed by the next forward launch or on flush)
  if (ev) HIP_TRY(c, hipEventRecord(ev[3], s));
  return SPK_OK;
}

// Whether a call's launch sequence is worth a graph: with channel launches
// after the forward transform.  A real-input call whose channels all run
// inside the forward kernel is two launches (forward, finaliser), and
// replaying it as a graph leaves a longer idle gap between calls than
// launching the two kernels directly.
bool graph_pays(const spk_ctx* c) {
  if (c->two_step) return true;
  for (const auto& kv : c->groups)
    if (!kv.second.fused) return true;
  return false;
}

// One launch sequence over nb blocks whose samples start at src.  With
// graphs enabled (and timing off) the sequence for (src, nb) is captured
// once and replayed when it has channel launches (graph_pays).
spk_status run_blocks(spk_ctx* c, const void* src, int nb) {
  if (c->chans_dirty) {
    spk_status st = rebuild_channels(c);
    if (st != SPK_OK) return st;
  }
  hipStream_t s = c->stream;
  if (c->use_graphs && !c->timing && graph_pays(c)) {
    const auto key = std::make_tuple(src, nb, c->parity, c->pend_nb, c->pend_nb ? c->pend_G : 0);
    auto it = c->graphs.find(key);
    if (it == c->graphs.end() || it->second.stale) {
      if (it == c->graphs.end() && c->graphs.size() >= kMaxCachedPlans) {
        drop_graphs(c);
        it = c->graphs.end();
      }
      if (c->chan_mode == SPK_CHAN_BATCHED && !c->batches.count(nb)) {
        spk_status st = build_batch(c, nb);  // uploads happen outside the capture
        if (st != SPK_OK) return st;
        it = c->graphs.find(key);  // (build_batch may have dropped every graph)
      }
      hipGraph_t graph = nullptr;
      HIP_TRY(c, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      spk_status st = enqueue_blocks(c, src, nb, s, nullptr);
      hipError_t e = hipStreamEndCapture(s, &graph);
      if (st != SPK_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return st;
      }
      HIP_TRY(c, e);
      bool updated = false;
      if (it != c->graphs.end()) {
        hipGraphNode_t err_node = nullptr;
        hipGraphExecUpdateResult res = hipGraphExecUpdateError;
        e = hipGraphExecUpdate(it->second.exec, graph, &err_node, &res);
        if (e == hipSuccess && res == hipGraphExecUpdateSuccess) {
          updated = true;
          it->second.stale = false;
          ++c->graphs_updated;
        } else {
          (void)hipGetLastError();  // a different topology: instantiate anew
          (void)hipGraphExecDestroy(it->second.exec);
          c->graphs.erase(it);
          it = c->graphs.end();
        }
      }
      if (!updated) {
        hipGraphExec_t exec = nullptr;
        e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
          (void)hipGraphDestroy(graph);
          return hip_fail(c, e, "hipGraphInstantiate");
        }
        ++c->graphs_instantiated;
        it = c->graphs.emplace(key, spk_ctx::GraphEntry{exec, false}).first;
      }
      (void)hipGraphDestroy(graph);
    }
    HIP_TRY(c, hipGraphLaunch(it->second.exec, s));
  } else {
    hipEvent_t* ev = nullptr;
    if (c->timing) {
      if (c->ev_used + 4 > c->ev_pool.size()) {
        for (int i = 0; i < 64; ++i) {
          hipEvent_t e;
          HIP_TRY(c, hipEventCreate(&e));
          c->ev_pool.push_back(e);
        }
      }
      ev = &c->ev_pool[c->ev_used];
      c->ev_used += 4;
    }
    spk_status st = enqueue_blocks(c, src, nb, s, ev);
    if (st != SPK_OK) return st;
  }
  if (!c->two_step) {
    c->pend_nb = nb;
    c->pend_G = fwd_grid(c, nb);
    c->pend_parity = c->parity;
  }
  const long long s_first = c->blocks_done / c->A;
  c->last_records = (uint32_t)((c->blocks_done + nb) / c->A - s_first);
  c->blocks_done += nb;
  c->parity ^= 1;
  c->last_blocks = (uint32_t)nb;
  c->have_output = true;
  return SPK_OK;
}

// Rust f64::round: halfway cases away from zero.
double round_half_away(double v) { return std::copysign(std::floor(std::fabs(v) + 0.5), v); }

}  // namespace

extern "C" {

const char* spk_last_error(const spk_ctx* ctx) { return ctx ? ctx->err.c_str() : kNullCtxMsg; }

spk_status spk_create(const spk_config* cfg, spk_ctx** out) {
  try {
  if (!cfg || !out) return SPK_ERR_INVALID;
  *out = nullptr;
  spk_ctx* c = new (std::nothrow) spk_ctx();
  if (!c) return SPK_ERR_DEVICE;
  c->cfg = *cfg;
  const uint32_t N = cfg->fft_size;
  const FormatInfo* fi = format_info(cfg->input_format);
  if (N < 4 || cfg->ffts_per_buf == 0 || cfg->spectrum_averages == 0 || !fi) {
    delete c;
    return SPK_ERR_INVALID;
  }
  const bool cplx = fi->complex_in;
  const bool pow2 = (N & (N - 1)) == 0;
  // real N beyond the forward kernel's plans: the complex pipeline on N/2 pairs
  const bool big_real = !cplx && pow2 && find_fwd_plan((int)N / 2) == nullptr && N > 65536 && N <= 524288;
  // every other size the reference accepts (rustfft plans any N) goes
  // through the mixed-radix passes: at least 16 and at most 2^20, and a
  // transform length without a prime factor above 31 (per-radix passes) or
  // one that splits into two factors of at most 480 (the four-step form,
  // whose lines sum any radix directly).  Even real N runs as N/2 sample
  // pairs (framed in samples when the hop or the overlap is odd), odd real
  // N as an N-point complex transform of the samples.
  const bool planned = pow2 && (cplx ? (N >= 1024 && N <= 262144) : (big_real || find_fwd_plan((int)N / 2) != nullptr));
  const bool real_odd = !cplx && N % 2 != 0;
  if (!planned) {
    const int Nc = cplx || real_odd ? (int)N : (int)N / 2;
    bool ok = N >= 16 && N <= (1u << 20) && !gfft_radices(Nc).empty();
    if (!ok && N >= 16 && N <= (1u << 20)) {
      int a = 0, b = 0;
      ok = gfft4_split(Nc, &a, &b) && (int)generic_radices(a).size() <= spk::kG4MaxStages &&
           (int)generic_radices(b).size() <= spk::kG4MaxStages;
    }
    if (!ok) {
      delete c;
      return SPK_ERR_UNSUPPORTED;
    }
    c->gfft = true;
    c->gfft_radices = gfft_radices(Nc);
  }
  c->N = (int)N;
  c->M = c->N / 2;
  c->complex_in = cplx;
  c->K = cplx ? c->N : c->N / 2 + 1;  // dsp.rs:79
  c->O = c->N / 4;
  c->H = c->N - c->O;
  c->F = (int)cfg->ffts_per_buf;
  c->A = (int)cfg->spectrum_averages;
  c->bps = fi->canon_bps;
  c->raw_bps = fi->raw_bps;
  c->conv_kind = fi->conv;
  c->f32_in = !cplx && fi->canon_bps == 4;
  c->two_step = cplx || big_real || c->gfft;
  c->real_odd = real_odd;
  c->Nc = cplx || real_odd ? c->N : c->N / 2;
  // the row transforms convert s16 pairs as they load them
  if (cfg->input_format == SPK_FMT_CS16LE) c->s16_pairs = true;
  if (big_real || (c->gfft && !cplx)) {
    // the row transforms read pairs of samples as complex samples: s16 pairs
    // for s16le, f32 pairs for every other format
    c->bps = 4;
    switch (cfg->input_format) {
      case SPK_FMT_S8: c->conv_kind = spk::kConvS8ToF32; break;
      case SPK_FMT_S16LE:
        c->conv_kind = 0;
        c->bps = 2;
        c->s16_pairs = true;
        break;
      case SPK_FMT_S16BE: c->conv_kind = spk::kConvS16beToF32; break;
      case SPK_FMT_U8: c->conv_kind = spk::kConvU8ToF32; break;
      case SPK_FMT_F32BE: c->conv_kind = spk::kConvF32beToF32; break;
      default: c->conv_kind = 0; break;  // f32le
    }
  }
  // input_format_scaling (inputformats.rs:47-62)
  const float scaling = cfg->scaling != 0.f ? cfg->scaling : fi->scaling;
  c->wscale = scaling / (float)c->N;  // dsp.rs:70-73
  c->max_blocks = (int)(cfg->max_blocks ? cfg->max_blocks : cfg->ffts_per_buf);
  if (c->max_blocks < c->F) c->max_blocks = c->F;
  c->fplan = c->two_step ? nullptr : find_fwd_plan(c->M);
  c->cplan = c->two_step && !c->gfft ? find_chan_plan(c->Nc / 64) : nullptr;

  auto cleanup = [&](spk_status s) {
    spk_destroy(c);
    return s;
  };
  if (hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking) != hipSuccess) return cleanup(SPK_ERR_DEVICE);
  c->stream = c->own_stream;

  // the two-step pipeline always keeps the spectra (Welch and the channel
  // stage read them)
  c->keep_flag = (cfg->flags & SPK_FLAG_KEEP_SPECTRA) != 0;
  c->keep_spec = c->two_step || c->keep_flag;
  // Persistent forward grid: resident workgroups per CU times CUs.
  int dev = 0, ncu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu <= 0)
    return cleanup(SPK_ERR_DEVICE);
  c->ncu = ncu;
  c->noalloc_scratch = scratch_address(dev);
  // (rows hold the K bins, or the whole transform where it is longer: odd real N)
  c->spec_stride = ((long long)std::max(c->K, c->Nc) + 63) / 64 * 64;
  c->welch_out_cap = c->max_blocks / c->A + 1;
  if (c->chan_counters.alloc(kXcdQueues * sizeof(int)) != hipSuccess ||
      hipMemset(c->chan_counters.p, 0, c->chan_counters.bytes) != hipSuccess ||
      c->pos.alloc(2 * sizeof(long long)) != hipSuccess || hipMemset(c->pos.p, 0, c->pos.bytes) != hipSuccess)
    return cleanup(SPK_ERR_DEVICE);
  c->use_graphs = (cfg->flags & SPK_FLAG_GRAPHS) != 0;
  if ((cfg->flags & SPK_FLAG_SPECTRUM_U8) && (cfg->flags & SPK_FLAG_SPECTRUM_U16)) return cleanup(SPK_ERR_INVALID);
  c->qfmt = (cfg->flags & SPK_FLAG_SPECTRUM_U8) ? 1 : ((cfg->flags & SPK_FLAG_SPECTRUM_U16) ? 2 : 0);
  if (c->qfmt && c->qout.alloc((size_t)c->welch_out_cap * c->K * c->qfmt) != hipSuccess)
    return cleanup(SPK_ERR_DEVICE);
  if (c->two_step) {
    // every spectrum row is kept: the channel stage always reads them
    if ((long long)(c->max_blocks + c->A) * 64 >= (1LL << 31)) return cleanup(SPK_ERR_UNSUPPORTED);
    // the untangle and mixed-radix launches take one grid row per block
    if ((!c->complex_in || c->gfft) && c->max_blocks > 65535) return cleanup(SPK_ERR_UNSUPPORTED);
    if (c->hist.alloc((size_t)c->O * c->bps) != hipSuccess ||
        c->staging.alloc((size_t)c->F * c->H * c->raw_bps) != hipSuccess ||
        (c->conv_kind && c->conv.alloc((size_t)c->max_blocks * c->H * c->bps) != hipSuccess) ||
        c->spec.alloc((size_t)c->max_blocks * c->spec_stride * sizeof(cpx)) != hipSuccess ||
        c->cY.alloc((size_t)c->max_blocks * c->Nc * sizeof(cpx)) != hipSuccess ||
        c->carry.alloc((size_t)2 * c->spec_stride * sizeof(float)) != hipSuccess ||
        c->welch_out.alloc((size_t)c->welch_out_cap * c->K * sizeof(float)) != hipSuccess)
      return cleanup(SPK_ERR_DEVICE);
    if (hipMemset(c->hist.p, 0, c->hist.bytes) != hipSuccess ||
        hipMemset(c->carry.p, 0, c->carry.bytes) != hipSuccess)
      return cleanup(SPK_ERR_DEVICE);
    // W_N^m with m reduced exactly, forward sign
    auto wn = [&](long long m, long long n) {
      const double a = 2.0 * M_PI * (double)(m % n) / (double)n;
      return cpx{(float)std::cos(a), (float)(-std::sin(a))};
    };
    if (!c->complex_in) {
      std::vector<cpx> tu((size_t)c->M / 2 + 1);
      for (size_t k = 0; k < tu.size(); ++k) tu[k] = wn((long long)k, c->N);
      if (upload(c->ctu, tu) != hipSuccess) return cleanup(SPK_ERR_DEVICE);
    }
    if (c->gfft) {
      std::vector<cpx> g((size_t)c->Nc);
      for (int m = 0; m < c->Nc; ++m) g[(size_t)m] = wn(m, c->Nc);
      if (upload(c->gtw, g) != hipSuccess) return cleanup(SPK_ERR_DEVICE);
      c->gfft4 = gfft4_split(c->Nc, &c->g4N1, &c->g4N2);
      if (c->gfft4) {
        c->g4r1 = generic_radices(c->g4N1);
        c->g4r2 = generic_radices(c->g4N2);
        // (lines of at most 480 points have at most 5 stages; more keeps the per-radix passes)
        c->gfft4 = (int)c->g4r1.size() <= spk::kG4MaxStages && (int)c->g4r2.size() <= spk::kG4MaxStages;
      }
      if (c->gfft4) {
        std::vector<cpx> t1((size_t)c->g4N1), t2((size_t)c->g4N2);
        for (int e = 0; e < c->g4N1; ++e) t1[(size_t)e] = wn(e, c->g4N1);
        for (int e = 0; e < c->g4N2; ++e) t2[(size_t)e] = wn(e, c->g4N2);
        if (upload(c->g4tw1, t1) != hipSuccess || upload(c->g4tw2, t2) != hipSuccess) return cleanup(SPK_ERR_DEVICE);
      }
      if (hipDeviceSynchronize() != hipSuccess) return cleanup(SPK_ERR_DEVICE);
      *out = c;
      return SPK_OK;
    }
    const ChanPlan& p = *c->cplan;
    const int N1 = c->Nc / 64;
    std::vector<cpx> ta(64 * 64), tb((size_t)64 * std::max(1, N1 / 64)), t64(64);
    for (int n2 = 0; n2 < 64; ++n2) {
      for (int a = 0; a < 64; ++a) ta[(size_t)n2 * 64 + a] = wn((long long)n2 * a, c->Nc);
      for (int b = 0; b < std::max(1, N1 / 64); ++b) tb[(size_t)n2 * std::max(1, N1 / 64) + b] = wn(64LL * n2 * b, c->Nc);
    }
    for (int m = 0; m < 64; ++m) t64[m] = wn(m, 64);
    if (upload(c->ctw2, pass_table(p.R2, p.R1, -1)) != hipSuccess ||
        upload(c->ctw3, pass_table(p.R3, p.R1 * p.R2, -1)) != hipSuccess || upload(c->cta, ta) != hipSuccess ||
        upload(c->ctb, tb) != hipSuccess || upload(c->ctw64, t64) != hipSuccess)
      return cleanup(SPK_ERR_DEVICE);
    // the null-stream clears above must be done before work on the
    // (non-blocking) context stream
    if (hipDeviceSynchronize() != hipSuccess) return cleanup(SPK_ERR_DEVICE);
    *out = c;
    return SPK_OK;
  }
  for (int f = 0; f < 3; ++f) {
    // class 2: the two general fused instantiations (s16 loader only)
    int per_cu = 1 << 30;
    for (int fl : f == 0 ? std::vector<int>{0} : (f == 1 ? std::vector<int>{kFusedL}
                                                         : std::vector<int>{spk::kFusedGen1, spk::kFusedGen2})) {
      const FwdVariant* v = find_fwd_variant(c->M, fl, c->f32_in);
      if (!v && f == 2) continue;
      per_cu = std::min(per_cu, v ? v->occupancy() : 0);
    }
    if (f == 2 && per_cu == (1 << 30)) per_cu = c->gmax[0] / ncu;  // no general stage for this loader
    if (per_cu <= 0) return cleanup(SPK_ERR_DEVICE);
    c->gmax[f] = per_cu * ncu;
  }
  const int gcap = std::max(c->gmax[0], std::max(c->gmax[1], c->gmax[2]));
  const int gmin = std::min(c->gmax[0], std::min(c->gmax[1], c->gmax[2]));
  // The Welch kernels index blocks times workgroups in 32 bits.
  if ((long long)(c->max_blocks + c->A) * (gcap + 1) >= (1LL << 31)) return cleanup(SPK_ERR_UNSUPPORTED);
  // A workgroup spans at most ceil(ceil(max_blocks / G) / A) + 1 Welch groups.
  const int bpw = (c->max_blocks + gmin - 1) / gmin;
  c->wslots = (bpw + c->A - 1) / c->A + 1;
  c->wstride = c->spec_stride;
  if (c->hist.alloc((size_t)c->O * c->bps) != hipSuccess ||
      c->staging.alloc((size_t)c->F * c->H * c->raw_bps) != hipSuccess ||
      (c->conv_kind && c->conv.alloc((size_t)c->max_blocks * c->H * c->bps) != hipSuccess) ||
      c->spec.alloc((size_t)std::max(c->max_blocks, gcap) * c->spec_stride * sizeof(cpx)) != hipSuccess ||
      c->wpart.alloc((size_t)2 * gcap * c->wslots * c->wstride * sizeof(float)) != hipSuccess ||
      c->edump.alloc((size_t)gcap * (c->M / c->fplan->R / 64) * spk::kEdgeLanes * c->fplan->R * sizeof(cpx)) !=
          hipSuccess ||
      c->carry.alloc((size_t)2 * c->spec_stride * sizeof(float)) != hipSuccess ||
      c->welch_out.alloc((size_t)c->welch_out_cap * c->K * sizeof(float)) != hipSuccess)
    return cleanup(SPK_ERR_DEVICE);
  c->wpart_half = (long long)gcap * c->wslots * c->wstride;
  if (hipMemset(c->hist.p, 0, c->hist.bytes) != hipSuccess || hipMemset(c->carry.p, 0, c->carry.bytes) != hipSuccess ||
      c->fin_ticket.alloc(sizeof(int)) != hipSuccess || hipMemset(c->fin_ticket.p, 0, sizeof(int)) != hipSuccess)
    return cleanup(SPK_ERR_DEVICE);

  const FwdPlan& p = *c->fplan;
  std::vector<cpx> twu((size_t)p.R1 * p.R2 + 1);
  for (size_t k = 0; k < twu.size(); ++k) {
    const double a = 2.0 * M_PI * (double)k / (double)c->N;
    twu[k] = cpx{(float)std::cos(a), (float)(-std::sin(a))};
  }
  std::vector<int4> items((size_t)(c->M / p.R / 64) * 4 * p.R + p.R + 1);
  spk::edge_items(c->M, p.R, items.data());
  if (upload(c->eitems, items) != hipSuccess) return cleanup(SPK_ERR_DEVICE);
  if (upload(c->ftw2, pass_table(p.R2, p.R1, -1)) != hipSuccess ||
      upload(c->ftw3, pass_table(p.R, p.R1 * p.R2, -1)) != hipSuccess || upload(c->ftwu, twu) != hipSuccess)
    return cleanup(SPK_ERR_DEVICE);
  // the null-stream clears above must be done before work on the
  // (non-blocking) context stream
  if (hipDeviceSynchronize() != hipSuccess) return cleanup(SPK_ERR_DEVICE);
  *out = c;
  return SPK_OK;
  } catch (...) {
    return entry_exception(nullptr);
  }
}

spk_status spk_destroy(spk_ctx* c) {
  try {
  if (!c) return SPK_ERR_INVALID;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->copy_stream) {
    (void)hipStreamSynchronize(c->copy_stream);
    (void)hipStreamDestroy(c->copy_stream);
  }
  for (int i = 0; i < 2; ++i) {
    if (c->a_up[i]) (void)hipEventDestroy(c->a_up[i]);
    if (c->a_used[i]) (void)hipEventDestroy(c->a_used[i]);
    if (c->apin[i]) (void)hipHostFree(c->apin[i]);
  }
  drop_graphs(c);
  for (auto e : c->ev_pool) (void)hipEventDestroy(e);
  if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
  delete c;
  return SPK_OK;
  } catch (...) {
    return entry_exception(c);
  }
}

spk_status spk_buffer_size_of(const spk_ctx* c, spk_buffer_size* out) {
  try {
  if (!c || !out) return SPK_ERR_INVALID;
  out->overlap = (uint64_t)c->O;
  out->new_samples = (uint64_t)c->F * c->H;
  out->total = out->overlap + out->new_samples;
  return SPK_OK;
  } catch (...) {
    return entry_exception(const_cast<spk_ctx*>(c));
  }
}

spk_status spk_add_channel(spk_ctx* c, int64_t first_bin, uint32_t bins, const float* weights, uint32_t* id) {
  try {
  if (!c) return SPK_ERR_INVALID;
  if (bins == 0 || bins % 4 != 0) return fail(c, SPK_ERR_INVALID, "IFFT size must be a positive multiple of 4");
  if ((int)bins > c->N || (!find_chan_plan((int)bins) && (int)bins > spk::kGenBigMaxL))
    return fail(c, SPK_ERR_UNSUPPORTED,
                "IFFT size must be <= N and at most 20480 (powers of two: 8192, or 16384)");
  {
    spk_status st = pull_weights(c);
    if (st != SPK_OK) return st;
  }
  Channel ch;
  ch.first = first_bin;
  ch.L = (int)bins;
  ch.w = weights ? std::vector<float>(weights, weights + bins) : raised_cosine_weights((int)bins);
  c->chans.push_back(std::move(ch));
  c->chans_dirty = true;
  if (id) *id = (uint32_t)(c->chans.size() - 1);
  return SPK_OK;
  } catch (...) {
    return entry_exception(c);
  }
}

spk_status spk_remove_channel(spk_ctx* c, uint32_t id) {
  try {
  if (!c) return SPK_ERR_INVALID;
  if (id >= c->chans.size() || c->chans[id].removed) return fail(c, SPK_ERR_INVALID, "no such channel");
  {
    spk_status st = pull_weights(c);
    if (st != SPK_OK) return st;
  }
  c->chans[id].removed = true;
  c->chans[id].w.clear();
  c->chans_dirty = true;
  return SPK_OK;
  } catch (...) {
    return entry_exception(c);
  }
}

spk_status spk_set_channel_weights(spk_ctx* c, uint32_t id, const float* weights) {
  try {
  if (!c || !weights) return SPK_ERR_INVALID;
  if (id >= c->chans.size() || c->chans[id].removed) return fail(c, SPK_ERR_INVALID, "no such channel");
  {
    spk_status st = pull_weights(c);
    if (st != SPK_OK) return st;
  }
  Channel& ch = c->chans[id];
  ch.w.assign(weights, weights + ch.L);
  c->chans_dirty = true;
  return SPK_OK;
  } catch (...) {
    return entry_exception(c);
  }
}

spk_status spk_process(spk_ctx* c, const void* raw_new, size_t nbytes) {
  try {
  if (!c || !raw_new) return SPK_ERR_INVALID;
  const size_t want = (size_t)c->F * c->H * c->raw_bps;
  if (nbytes != want) return fail(c, SPK_ERR_INVALID, "buffer must hold exactly F * 3N/4 samples");
  HIP_TRY(c, hipMemcpyAsync(c->staging.p, raw_new, nbytes, hipMemcpyHostToDevice, c->stream));
  spk_status st = run_blocks(c, c->staging.p, c->F);
  if (st == SPK_OK) st = flush_records(c);
  if (st != SPK_OK) return st;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return SPK_OK;
  } catch (...) {
    return entry_exception(c);
  }
}

}  // extern "C"

namespace {

// Host copy into a pinned slot; large buffers are split over a few threads
// (one thread copies well under the PCIe rate).
void copy_to_pinned(void* dst, const void* src, size_t n) {
  const size_t kPerThread = (size_t)4 << 20;
  unsigned nt = std::min<unsigned>(8u, std::max(1u, std::thread::hardware_concurrency()));
  nt = (unsigned)std::min<size_t>(nt, n / kPerThread);
  if (nt <= 1) {
    std::memcpy(dst, src, n);
    return;
  }
  const size_t chunk = (n / nt + 63) / 64 * 64;
  std::vector<std::thread> th;
  for (unsigned i = 1; i < nt; ++i) {
    const size_t lo = std::min(n, i * chunk), hi = std::min(n, lo + chunk);
    if (lo < hi)
      th.emplace_back([=] { std::memcpy(static_cast<char*>(dst) + lo, static_cast<const char*>(src) + lo, hi - lo); });
  }
  std::memcpy(dst, src, std::min(n, chunk));
  for (auto& t : th) t.join();
}

spk_status async_setup(spk_ctx* c) {
  if (c->copy_stream) return SPK_OK;
  const size_t want = (size_t)c->F * c->H * c->raw_bps;
  HIP_TRY(c, hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking));
  for (int i = 0; i < 2; ++i) {
    HIP_TRY(c, hipHostMalloc(&c->apin[i], want, hipHostMallocDefault));
    HIP_TRY(c, c->adev[i].alloc(want));
    HIP_TRY(c, hipEventCreateWithFlags(&c->a_up[i], hipEventDisableTiming));
    HIP_TRY(c, hipEventCreateWithFlags(&c->a_used[i], hipEventDisableTiming));
  }
  return SPK_OK;
}

}  // namespace

extern "C" {

spk_status spk_input_buffer(spk_ctx* c, void** host_ptr, size_t* nbytes) {
  try {
  if (!c || !host_ptr) return SPK_ERR_INVALID;
  {
    spk_status st = async_setup(c);
    if (st != SPK_OK) return st;
  }
  const int s = c->aslot;
  if (c->a_busy[s]) HIP_TRY(c, hipEventSynchronize(c->a_up[s]));  // its last upload has left it
  *host_ptr = c->apin[s];
  if (nbytes) *nbytes = (size_t)c->F * c->H * c->raw_bps;
  return SPK_OK;
  } catch (...) {
    return entry_exception(c);
  }
}

spk_status spk_process_async(spk_ctx* c, const void* raw_new, size_t nbytes) {
  try {
  if (!c || !raw_new) return SPK_ERR_INVALID;
  const size_t want = (size_t)c->F * c->H * c->raw_bps;
  if (nbytes != want) return fail(c, SPK_ERR_INVALID, "buffer must hold exactly F * 3N/4 samples");
  {
    spk_status st = async_setup(c);
    if (st != SPK_OK) return st;
  }
  const int s = c->aslot;
  // The slot's last upload must have left its pinned memory, and the
  // kernels of the call before last must be done reading its device copy.
  if (c->a_busy[s]) {
    HIP_TRY(c, hipEventSynchronize(c->a_up[s]));
    HIP_TRY(c, hipStreamWaitEvent(c->copy_stream, c->a_used[s], 0));
  }
  // the caller may have filled the slot itself (spk_input_buffer)
  if (raw_new != c->apin[s]) copy_to_pinned(c->apin[s], raw_new, nbytes);
  HIP_TRY(c, hipMemcpyAsync(c->adev[s].p, c->apin[s], nbytes, hipMemcpyHostToDevice, c->copy_stream));
  HIP_TRY(c, hipEventRecord(c->a_up[s], c->copy_stream));
  HIP_TRY(c, hipStreamWaitEvent(c->stream, c->a_up[s], 0));
  c->a_busy[s] = true;
  spk_status st = run_blocks(c, c->adev[s].p, c->F);
  if (st != SPK_OK) return st;
  HIP_TRY(c, hipEventRecord(c->a_used[s], c->stream));
  c->aslot ^= 1;
  return SPK_OK;
  } catch (...) {
    return entry_exception(c);
  }
}

spk_status spk_process_device(spk_ctx* c, const void* d_raw, uint64_t nsamples) {
  try {
  if (!c || !d_raw) return SPK_ERR_INVALID;
  if (reinterpret_cast<uintptr_t>(d_raw) % 4 != 0) return fail(c, SPK_ERR_INVALID, "input must be 4-byte aligned");
  const uint64_t per_buf = (uint64_t)c->F * c->H;
  const uint64_t nb = nsamples / per_buf * c->F;
  if (nb == 0) return fail(c, SPK_ERR_INVALID, "fewer samples than one buffer");
  if (nb > (uint64_t)c->max_blocks) return fail(c, SPK_ERR_CAPACITY, "more blocks than max_blocks");
  return run_blocks(c, d_raw, (int)nb);
  } catch (...) {
    return entry_exception(c);
  }
}

spk_status spk_last_blocks(const spk_ctx* c, uint32_t* nblocks) {
  try {
  if (!c || !nblocks) return SPK_ERR_INVALID;
  *nblocks = c->last_blocks;
  return SPK_OK;
  } catch (...) {
    return entry_exception(const_cast<spk_ctx*>(c));
  }
}

spk_status spk_channel_output_device(spk_ctx* c, uint32_t id, const float** dptr, size_t* n) {
  try {
  if (!c || !dptr || !n) return SPK_ERR_INVALID;
  if (id >= c->chans.size() || c->chans[id].removed) return fail(c, SPK_ERR_INVALID, "no such channel");
  if (!c->have_output || c->chans_dirty) return fail(c, SPK_ERR_STATE, "no output yet");
  *dptr = reinterpret_cast<const float*>(c->chan_out.as<cpx>() + c->chan_out_off[id]);
  *n = (size_t)c->last_blocks * chan_out_len(c->chans[id].L) * 2;
  return SPK_OK;
  } catch (...) {
    return entry_exception(c);
  }
}

spk_status spk_channel_output(spk_ctx* c, uint32_t id, float* out, size_t cap, size_t* n) {
  try {
  if (!c || !out || !n) return SPK_ERR_INVALID;
  const float* d = nullptr;
  spk_status st = spk_channel_output_device(c, id, &d, n);
  if (st != SPK_OK) return st;
  if (*n > cap) return fail(c, SPK_ERR_CAPACITY, "output buffer too small");
  HIP_TRY(c, hipMemcpyAsync(out, d, *n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return SPK_OK;
  } catch (...) {
    return entry_exception(c);
  }
}

spk_status spk_spectrum_records_device(spk_ctx* c, const float** dptr, uint32_t* nrec) {
  try {
  if (!c || !dptr || !nrec) return SPK_ERR_INVALID;
  {
    spk_status st = flush_records(c);
    if (st != SPK_OK) return st;
  }
  *dptr = c->welch_out.as<float>();
  *nrec = c->have_output ? c->last_records : 0;
  return SPK_OK;
  } catch (...) {
    return entry_exception(c);
  }
}

spk_status spk_spectrum_records(spk_ctx* c, float* out, size_t cap, uint32_t* nrec) {
  try {
  if (!c || !nrec) return SPK_ERR_INVALID;
  const float* d = nullptr;
  spk_status st = spk_spectrum_records_device(c, &d, nrec);
  if (st != SPK_OK) return st;
  const size_t n = (size_t)*nrec * c->K;
  if (n == 0) return SPK_OK;
  if (!out || n > cap) return fail(c, SPK_ERR_CAPACITY, "output buffer too small");
  HIP_TRY(c, hipMemcpyAsync(out, d, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return SPK_OK;
  } catch (...) {
    return entry_exception(c);
  }
}

spk_status spk_spectrum_quantized_device(spk_ctx* c, const void** dptr, uint32_t* nrec, uint32_t* bytes_per_bin) {
  try {
  if (!c || !dptr || !nrec || !bytes_per_bin) return SPK_ERR_INVALID;
  if (!c->qfmt) return fail(c, SPK_ERR_STATE, "context created without SPK_FLAG_SPECTRUM_U8 / _U16");
  {
    spk_status st = flush_records(c);
    if (st != SPK_OK) return st;
  }
  *dptr = c->qout.p;
  *nrec = c->have_output ? c->last_records : 0;
  *bytes_per_bin = (uint32_t)c->qfmt;
  return SPK_OK;
  } catch (...) {
    return entry_exception(c);
  }
}

spk_status spk_spectrum_quantized(spk_ctx* c, void* out, size_t cap, uint32_t* nrec) {
  try {
  if (!c || !nrec) return SPK_ERR_INVALID;
  if (!c->qfmt) return fail(c, SPK_ERR_STATE, "context created without SPK_FLAG_SPECTRUM_U8 / _U16");
  {
    spk_status st = flush_records(c);
    if (st != SPK_OK) return st;
  }
  const uint32_t n = c->have_output ? c->last_records : 0;
  const size_t need = (size_t)n * c->K * c->qfmt;
  if (need > cap) return fail(c, SPK_ERR_CAPACITY, "output buffer too small");
  if (need) {
    if (!out) return SPK_ERR_INVALID;
    HIP_TRY(c, hipMemcpyAsync(out, c->qout.p, need, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
  }
  *nrec = n;
  return SPK_OK;
  } catch (...) {
    return entry_exception(c);
  }
}

spk_status spk_spectrum_records_packed(spk_ctx* c, uint64_t first_seq, uint64_t secs, uint32_t nanos, void* out,
                                       size_t cap, uint32_t* nrec, size_t* nbytes) {
  try {
  if (!c || !nrec || !nbytes) return SPK_ERR_INVALID;
  if (!c->qfmt) return fail(c, SPK_ERR_STATE, "context created without SPK_FLAG_SPECTRUM_U8 / _U16");
  {
    spk_status st = flush_records(c);
    if (st != SPK_OK) return st;
  }
  const uint32_t n = c->have_output ? c->last_records : 0;
  const long long body = (long long)c->K * c->qfmt;
  const size_t need = (size_t)n * (size_t)(24 + body);
  *nrec = n;
  *nbytes = need;
  if (need > cap) return fail(c, SPK_ERR_CAPACITY, "output buffer too small");
  if (!need) return SPK_OK;
  if (!out) return SPK_ERR_INVALID;
  if (c->qpack.bytes < need) HIP_TRY(c, c->qpack.alloc((size_t)c->welch_out_cap * (size_t)(24 + body)));
  spk::RecordPackParams P{};
  P.body = c->qout.as<unsigned char>();
  P.out = c->qpack.as<unsigned char>();
  P.body_bytes = body;
  P.nrec = (int)n;
  P.first_seq = first_seq;
  P.secs = secs;
  P.nanos = nanos;
  const unsigned gx = (unsigned)std::min<long long>((24 + body + 255) / 256, 64);
  const unsigned gy = std::min<uint32_t>(n, 65535u);
  hipLaunchKernelGGL(spk::pack_records_kernel, dim3(gx, gy), dim3(256), 0, c->stream, P);
  HIP_TRY(c, hipGetLastError());
  HIP_TRY(c, hipMemcpyAsync(out, c->qpack.p, need, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return SPK_OK;
  } catch (...) {
    return entry_exception(c);
  }
}

spk_status spk_channel_records_packed(spk_ctx* c, uint64_t seq, uint64_t secs, uint32_t nanos, void* out, size_t cap,
                                      uint32_t* nrec, size_t* nbytes) {
  try {
  if (!c || !nrec || !nbytes) return SPK_ERR_INVALID;
  if (!c->have_output || c->chans_dirty) return fail(c, SPK_ERR_STATE, "no output yet");
  const int nb = (int)c->last_blocks;
  const uint32_t n = (uint32_t)c->live_chans;
  const size_t need = (size_t)n * 24 + (size_t)8 * nb * (size_t)c->live_len;
  *nrec = n;
  *nbytes = need;
  if (need > cap) return fail(c, SPK_ERR_CAPACITY, "output buffer too small");
  if (!n) return SPK_OK;
  if (!out) return SPK_ERR_INVALID;
  if (c->cpack.bytes < need)
    HIP_TRY(c, c->cpack.alloc((size_t)n * 24 + (size_t)8 * c->max_blocks * (size_t)c->live_len));
  spk::ChanRecParams P{};
  P.src = c->chan_out.as<cpx>();
  P.recs = c->crecs.as<spk::ChanRecDesc>();
  P.nrec = (int)n;
  P.nblocks = nb;
  P.out = c->cpack.as<unsigned char>();
  P.seq = seq;
  P.secs = secs;
  P.nanos = nanos;
  const long long per_rec = (long long)nb * c->live_len / n + 1;  // mean samples per record
  const unsigned gx = (unsigned)std::max<long long>(1, std::min<long long>((per_rec + 255) / 256, 64));
  const unsigned gy = std::min<uint32_t>(n, 65535u);
  hipLaunchKernelGGL(spk::pack_channel_records_kernel, dim3(gx, gy), dim3(256), 0, c->stream, P);
  HIP_TRY(c, hipGetLastError());
  HIP_TRY(c, hipMemcpyAsync(out, c->cpack.p, need, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return SPK_OK;
  } catch (...) {
    return entry_exception(c);
  }
}

spk_status spk_spectra_device(spk_ctx* c, const float** dptr, uint64_t* stride) {
  try {
  if (!c || !dptr || !stride) return SPK_ERR_INVALID;
  *dptr = c->spec.as<float>();
  *stride = (uint64_t)c->spec_stride;
  return SPK_OK;
  } catch (...) {
    return entry_exception(c);
  }
}

spk_status spk_reset(spk_ctx* c) {
  try {
  if (!c) return SPK_ERR_INVALID;
  // On the context's stream, after every call already enqueued there (no
  // host wait): it does not wait for the null stream.  A kernel, not
  // memsets, so that the call can be captured into a graph.
  spk::ZeroParams Z{};
  const DevBuf* bufs[3] = {&c->hist, &c->carry, &c->pos};
  size_t most = 0;
  for (int i = 0; i < 3; ++i) {
    Z.p[i] = static_cast<unsigned char*>(bufs[i]->p);
    Z.bytes[i] = bufs[i]->bytes;
    most = std::max(most, bufs[i]->bytes);
  }
  const int grid = (int)std::min<size_t>((most / 4 + 255) / 256 + 1, (size_t)c->ncu * 4);
  hipLaunchKernelGGL(spk::zero_buffers_kernel, dim3(grid), dim3(256), 0, c->stream, Z);
  HIP_TRY(c, hipGetLastError());
  c->parity = 0;
  c->pend_nb = 0;  // dropped with the rest of the stream state
  c->blocks_done = 0;
  c->last_blocks = 0;
  c->last_records = 0;
  c->have_output = false;
  return SPK_OK;
  } catch (...) {
    return entry_exception(c);
  }
}

spk_status spk_set_stream(spk_ctx* c, void* stream) {
  try {
  if (!c) return SPK_ERR_INVALID;
  {
    spk_status st = flush_records(c);
    if (st != SPK_OK) return st;
  }
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  c->stream = stream ? reinterpret_cast<hipStream_t>(stream) : c->own_stream;
  drop_graphs(c);
  return SPK_OK;
  } catch (...) {
    return entry_exception(c);
  }
}

spk_status spk_synchronize(spk_ctx* c) {
  try {
  if (!c) return SPK_ERR_INVALID;
  spk_status st = flush_records(c);
  if (st != SPK_OK) return st;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return SPK_OK;
  } catch (...) {
    return entry_exception(c);
  }
}

spk_status spk_flush(spk_ctx* c) {
  try {
  if (!c) return SPK_ERR_INVALID;
  return flush_records(c);
  } catch (...) {
    return entry_exception(c);
  }
}

spk_status spk_enable_timing(spk_ctx* c, int on) {
  try {
  if (!c) return SPK_ERR_INVALID;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  c->timing = on != 0;
  c->ev_used = c->ev_done = 0;
  c->timed_calls = 0;
  for (double& v : c->ms_sum) v = 0;
  return SPK_OK;
  } catch (...) {
    return entry_exception(c);
  }
}

size_t packed_weight_floats(const spk_ctx* c) {
  size_t n = 0;
  for (const auto& ch : c->chans)
    if (!ch.removed) n += (size_t)ch.L;
  return n;
}

namespace {
// The device weights (packed in channel id order, as uploaded by the last
// rebuild) into the channels' host copies, once after spk_set_weights_device.
spk_status pull_weights(spk_ctx* c) {
  if (!c->w_device_newer) return SPK_OK;
  const size_t n = packed_weight_floats(c);
  std::vector<float> hw(n);
  HIP_TRY(c, hipMemcpyAsync(hw.data(), c->weights.p, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  size_t off = 0;
  for (auto& ch : c->chans) {
    if (ch.removed) continue;
    ch.w.assign(hw.begin() + off, hw.begin() + off + ch.L);
    off += (size_t)ch.L;
  }
  c->w_device_newer = false;
  return SPK_OK;
}
}  // namespace

spk_status spk_set_weights_device(spk_ctx* c, const float* w) {
  try {
  if (!c || !w) return SPK_ERR_INVALID;
  if (c->chans.empty()) return fail(c, SPK_ERR_STATE, "no channels");
  if (!c->w_packed) {
    c->w_packed = true;
    if (c->w_shared) c->chans_dirty = true;  // lay the weights out per channel
  }
  if (c->chans_dirty) {
    spk_status st = rebuild_channels(c);
    if (st != SPK_OK) return st;
  }
  const size_t n = packed_weight_floats(c);
  // stream-ordered, no wait: the host copies are read back when needed
  HIP_TRY(c, hipMemcpyAsync(c->weights.p, w, n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
  c->w_device_newer = true;
  return SPK_OK;
  } catch (...) {
    return entry_exception(c);
  }
}

spk_status spk_channels_output_packed(spk_ctx* c, float* dst, size_t cap) {
  try {
  if (!c || !dst) return SPK_ERR_INVALID;
  if (!c->have_output || c->chans_dirty) return fail(c, SPK_ERR_STATE, "no output yet");
  const size_t need = 2 * (size_t)c->last_blocks * (size_t)c->live_len;
  if (cap < need) return fail(c, SPK_ERR_CAPACITY, "destination smaller than the packed outputs");
  for (auto& kv : c->groups) {
    const ChanGroup& g = kv.second;
    spk::PackParams Q{};
    Q.src = c->chan_out.as<cpx>();
    Q.chans = g.desc.as<spk::ChanDesc>();
    Q.pk_off = g.pk_off.as<long long>();
    Q.L = g.L;
    Q.nblocks = (int)c->last_blocks;
    Q.dst = reinterpret_cast<cpx*>(dst);
    const long long n = (long long)c->last_blocks * chan_out_len(g.L);
    const unsigned gx = (unsigned)std::min<long long>((n + 255) / 256, 1024);
    hipLaunchKernelGGL(spk::pack_outputs_kernel, dim3(gx, (unsigned)g.ids.size()), dim3(256), 0, c->stream, Q);
    HIP_TRY(c, hipGetLastError());
  }
  return SPK_OK;
  } catch (...) {
    return entry_exception(c);
  }
}

// Adjoint of one group of IFFT sizes without a power-of-two plan
// (chan_generic_adjoint_kernel, then the split sums in order).
}  // extern "C"

namespace {

// Block splits S of a group's adjoint launch and the floats of per-(channel,
// split) weight sums it writes (nch * S * L).
size_t adjoint_split(const spk_ctx* c, const ChanGroup& g, int nb, int* S_out) {
  const int nch = (int)g.ids.size();
  int S;
  if (g.generic && g.L > spk::kGenMaxL) {
    // wide channels: one (channel, split) tile per 1024-thread workgroup
    S = std::max(1, std::min((2 * c->ncu + nch - 1) / nch, nb));
  } else if (g.generic) {
    int T = 32;
    while (T < 256 && 2 * T < g.L) T *= 2;  // as launch_generic
    const int tpw = spk::kGenThreads / T;
    S = (2 * c->ncu * tpw + nch - 1) / nch;  // enough tiles for every CU
    S = std::max(1, std::min(S, nb));
  } else {
    const int T = g.L / find_chan_plan(g.L)->R;
    const int tpw = spk::kChanThreads / T;
    // splits over blocks: enough workgroups for every CU, at least one tile-row of blocks each
    S = (2 * c->ncu + nch - 1) / nch;
    S = std::max(1, std::min(S, (nb + tpw - 1) / tpw));
  }
  *S_out = S;
  return (size_t)nch * S * g.L;
}

// One group's adjoint: tiles write their weight sums to part[channel][split][j]
// (every entry, so part needs no clearing), then adjoint_reduce_kernel adds
// the splits in order into wbar.  Groups run one after another on the
// context's stream and reuse part.
spk_status group_adjoint(spk_ctx* c, const ChanGroup& g, int nb, const float* ybar, float* wbar, float* xbar,
                         float* part) {
  const int nch = (int)g.ids.size();
  int S = 1;
  (void)adjoint_split(c, g, nb, &S);
  if (g.generic) {
    spk::GenAdjParams Q{};
    Q.spec = c->spec.as<cpx>();
    Q.spec_stride = c->spec_stride;
    Q.N = c->N;
    Q.K = c->K;
    Q.real_input = c->complex_in ? 0 : 1;
    Q.weights = c->weights.as<float>();
    Q.chans = g.desc.as<spk::ChanDesc>();
    Q.pk_off = g.pk_off.as<long long>();
    Q.nch = nch;
    Q.nblocks = nb;
    Q.S = S;
    Q.ybar = reinterpret_cast<const cpx*>(ybar);
    Q.xbar = reinterpret_cast<cpx*>(xbar);
    Q.wpart = part;
    Q.tw = g.gtw.as<cpx>();
    Q.L = g.L;
    Q.nstages = (int)g.radix.size();
    for (size_t i = 0; i < g.radix.size(); ++i) Q.radix[i] = g.radix[i];
    if (g.L > spk::kGenMaxL) {
      Q.tile_threads = spk::kGenBigThreads;
      hipLaunchKernelGGL(spk::chan_generic_big_adjoint_kernel, dim3((unsigned)(nch * S)), dim3(spk::kGenBigThreads),
                         0, c->stream, Q);
    } else {
      int T = 32;
      while (T < 256 && 2 * T < g.L) T *= 2;  // as launch_generic
      const int tpw = spk::kGenThreads / T;
      Q.tile_threads = T;
      const long long tiles = (long long)nch * S;
      const unsigned grid = (unsigned)((tiles + tpw - 1) / tpw);
      const size_t lds = (size_t)tpw * 2 * g.L * sizeof(cpx);
      hipLaunchKernelGGL(spk::chan_generic_adjoint_kernel, dim3(grid), dim3(spk::kGenThreads), lds, c->stream, Q);
    }
    HIP_TRY(c, hipGetLastError());
  } else {
    spk::ChanAdjParams Q{};
    Q.spec = c->spec.as<cpx>();
    Q.spec_stride = c->spec_stride;
    Q.N = c->N;
    Q.K = c->K;
    Q.real_input = c->complex_in ? 0 : 1;
    Q.weights = c->weights.as<float>();
    Q.chans = g.desc.as<spk::ChanDesc>();
    Q.nch = nch;
    Q.nblocks = nb;
    Q.S = S;
    Q.ybar = reinterpret_cast<const cpx*>(ybar);
    Q.pk_off = g.pk_off.as<long long>();
    Q.xbar = reinterpret_cast<cpx*>(xbar);
    Q.wpart = part;
    Q.tw2 = g.atw2.as<cpx>();
    Q.tw3 = g.atw3.as<cpx>();
    HIP_TRY(c, dispatch_chan_adjoint(g.L, Q, c->stream));
  }
  spk::AdjReduceParams Rp{};
  Rp.wpart = part;
  Rp.chans = g.desc.as<spk::ChanDesc>();
  Rp.L = g.L;
  Rp.S = S;
  Rp.wbar = wbar;
  hipLaunchKernelGGL(spk::adjoint_reduce_kernel, dim3((g.L + 255) / 256, nch), dim3(256), 0, c->stream, Rp);
  HIP_TRY(c, hipGetLastError());
  return SPK_OK;
}

// The adjoint over every group, with the weight sums in part (part_floats
// long) or, when part is null, in the context's adj_part (grown as needed).
spk_status channels_adjoint(spk_ctx* c, const float* ybar, float* wbar, float* xbar, float* part,
                            size_t part_floats) {
  if (!c || !ybar || !wbar) return SPK_ERR_INVALID;
  if (!c->have_output || c->chans_dirty) return fail(c, SPK_ERR_STATE, "no output yet");
  if (!c->keep_flag) return fail(c, SPK_ERR_STATE, "the adjoint needs a context created with SPK_FLAG_KEEP_SPECTRA");
  const int nb = (int)c->last_blocks;
  size_t need = 0;
  for (const auto& kv : c->groups) {
    int S = 1;
    need = std::max(need, adjoint_split(c, kv.second, nb, &S));
  }
  if (part) {
    if (need > part_floats)
      return fail(c, SPK_ERR_CAPACITY, "the adjoint's weight sums exceed the fixed scratch size");
  } else {
    if (c->adj_part.bytes < need * sizeof(float)) HIP_TRY(c, c->adj_part.alloc(need * sizeof(float)));
    part = c->adj_part.as<float>();
  }
  for (const auto& kv : c->groups) {
    spk_status st = group_adjoint(c, kv.second, nb, ybar, wbar, xbar, part);
    if (st != SPK_OK) return st;
  }
  return SPK_OK;
}

}  // namespace

extern "C" {

spk_status spk_channels_adjoint(spk_ctx* c, const float* ybar, float* wbar, float* xbar) {
  try {
    return channels_adjoint(c, ybar, wbar, xbar, nullptr, 0);
  } catch (...) {
    return entry_exception(c);
  }
}

spk_status spk_channels_adjoint_noalloc(spk_ctx* c, const float* ybar, float* wbar, float* xbar) {
  try {
    if (c && !c->noalloc_scratch) return fail(c, SPK_ERR_DEVICE, "no device address for the fixed scratch");
    return channels_adjoint(c, ybar, wbar, xbar, c ? c->noalloc_scratch : nullptr, kNoAllocScratchFloats);
  } catch (...) {
    return entry_exception(c);
  }
}

spk_status spk_set_channel_mode(spk_ctx* c, int mode) {
  try {
  if (!c || (mode != SPK_CHAN_PER_SIZE && mode != SPK_CHAN_BATCHED)) return SPK_ERR_INVALID;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  c->chan_mode = mode;
  drop_graphs(c);
  return SPK_OK;
  } catch (...) {
    return entry_exception(c);
  }
}

spk_status spk_autotune(spk_ctx* c, uint32_t nblocks, spk_tune_result* out) {
  try {
  if (!c) return SPK_ERR_INVALID;
  if (c->chans_dirty) {
    spk_status st = rebuild_channels(c);
    if (st != SPK_OK) return st;
  }
  const int nb = nblocks ? (int)nblocks : c->max_blocks;
  if (nb > c->max_blocks || nb > (int)std::max<long long>(c->max_blocks, 1)) return fail(c, SPK_ERR_CAPACITY, "nblocks");
  if (batched_groups(c).empty()) return fail(c, SPK_ERR_STATE, "no channel needs the channel stage");
  hipStream_t s = c->stream;
  HIP_TRY(c, hipStreamSynchronize(s));
  // the plans are cut again for every setting tried
  drop_graphs_and_plans(c);
  hipEvent_t e0, e1;
  HIP_TRY(c, hipEventCreate(&e0));
  HIP_TRY(c, hipEventCreate(&e1));
  const int reps = 3;
  auto time_it = [&](auto&& launch) -> float {
    float best = 1e30f;
    for (int r = 0; r < reps + 1; ++r) {
      if (hipMemsetAsync(c->chan_counters.p, 0, c->chan_counters.bytes, s) != hipSuccess) return -1.f;
      if (hipEventRecord(e0, s) != hipSuccess) return -1.f;
      if (launch() != SPK_OK) return -1.f;
      if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) return -1.f;
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return -1.f;
      if (r > 0) best = std::min(best, ms);  // the first run warms up
    }
    return best;
  };
  auto finish = [&](spk_status st) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return st;
  };
  auto per_size_params = [&](const ChanGroup& g, const DevBuf& t2, const DevBuf& t3) {
    spk::ChanParams Q{};
    Q.spec = c->spec.as<cpx>();
    Q.spec_stride = c->spec_stride;
    Q.N = c->N;
    Q.K = c->K;
    Q.real_input = c->complex_in ? 0 : 1;
    Q.weights = c->weights.as<float>();
    Q.chans = g.desc.as<spk::ChanDesc>();
    Q.nch = (int)g.ids.size();
    Q.nblocks = nb;
    Q.out = c->chan_out.as<cpx>();
    Q.tw2 = t2.as<cpx>();
    Q.tw3 = t3.as<cpx>();
    return Q;
  };
  // 1. Tile shape of every power-of-two size (fcfb.rs:66-100 runs one
  // filter per channel; here one launch per size and shape): each shape's
  // per-size launch over the size's channels is timed; the per-size form
  // takes the fastest, the batched kernel the fastest it serves.
  // (best_small: the fastest shape of at most 16 points per thread, which
  // keeps chan_batched_kernel<4096> and four workgroups per CU)
  std::map<int, int> best_small, best_any;
  for (auto& kv : c->groups) {
    ChanGroup& g = kv.second;
    if (g.fused || g.generic) continue;
    const std::vector<ShapeInfo> shapes = tile_shapes(g.L);
    if (shapes.size() < 2) continue;
    int best_ps = -1, best_b = -1;
    float t_ps = 1e30f, t_b = 1e30f;
    for (size_t S = 0; S < shapes.size(); ++S) {
      DevBuf t2, t3;
      HIP_TRY(c, upload(t2, pass_table(shapes[S].R2, shapes[S].R1, +1)));
      HIP_TRY(c, upload(t3, pass_table(shapes[S].R3, shapes[S].R1 * shapes[S].R2, +1)));
      const spk::ChanParams Q = per_size_params(g, t2, t3);
      const float ms = time_it([&]() -> spk_status {
        return dispatch_chan(g.L, (int)S, Q, s) == hipSuccess ? SPK_OK : SPK_ERR_DEVICE;
      });
      HIP_TRY(c, hipStreamSynchronize(s));  // before the tables are freed
      if (ms < 0.f) return finish(fail(c, SPK_ERR_DEVICE, "autotune launch failed"));
      if (ms < t_ps) {
        t_ps = ms;
        best_ps = (int)S;
      }
      if (!shapes[S].big() && ms < t_b) {
        t_b = ms;
        best_b = (int)S;
      }
    }
    c->shape_ps[g.L] = best_ps;
    best_any[g.L] = best_ps;
    best_small[g.L] = best_b >= 0 ? best_b : best_ps;
    g.ps_shape = best_ps;
    HIP_TRY(c, upload(g.tw2, pass_table(shapes[(size_t)best_ps].R2, shapes[(size_t)best_ps].R1, +1)));
    HIP_TRY(c, upload(g.tw3, pass_table(shapes[(size_t)best_ps].R3,
                                        shapes[(size_t)best_ps].R1 * shapes[(size_t)best_ps].R2, +1)));
  }
  // 2. The per-size form with those shapes: every group the batched kernel
  // would take, in its own launch.
  const float per_size = time_it([&]() -> spk_status {
    for (auto& kv : c->groups) {
      ChanGroup& g = kv.second;
      if (g.fused || !in_batch(c, g.L)) continue;
      if (g.generic) {
        if (launch_generic(c, g, nb, s) != hipSuccess) return SPK_ERR_DEVICE;
        continue;
      }
      if (dispatch_chan(g.L, g.ps_shape, per_size_params(g, g.tw2, g.tw3), s) != hipSuccess) return SPK_ERR_DEVICE;
    }
    return SPK_OK;
  });
  // 3. The batched form: the shapes of either LDS tier (every size's
  // fastest shape, or its fastest of at most 16 points per thread), then
  // the work split (chunk of blocks per item, tile rounds per item,
  // workgroups per CU).  One-block chunks keep the fewest spectrum rows per
  // XCD live, which suits channel sets that read each row many times over.
  const int chunks[] = {1, 2, 4, 8, 16};
  const int rounds[] = {1, 2, 4};
  const int wpcu[] = {0, 2};
  float best = 1e30f;
  int bc = c->tune_chunk, br = c->tune_rounds, bw = c->tune_wgs_per_cu;
  std::map<int, int> bshape = c->shape_b;
  for (const std::map<int, int>* tier : {&best_small, &best_any}) {
    if (tier == &best_any && best_any == best_small) break;
    for (const auto& kv : *tier) c->shape_b[kv.first] = kv.second;
    for (int ch : chunks)
      for (int rd : rounds)
        for (int w : wpcu) {
          c->tune_chunk = ch;
          c->tune_rounds = rd;
          c->tune_wgs_per_cu = w;
          c->batches.clear();
          const float ms = time_it([&]() { return launch_batched(c, nb, s); });
          if (ms < 0.f) return finish(fail(c, SPK_ERR_DEVICE, "autotune launch failed"));
          if (ms < best) {
            best = ms;
            bc = ch;
            br = rd;
            bw = w;
            bshape = c->shape_b;
          }
        }
  }
  c->shape_b = bshape;
  finish(SPK_OK);
  HIP_TRY(c, hipMemsetAsync(c->chan_counters.p, 0, c->chan_counters.bytes, s));
  HIP_TRY(c, hipStreamSynchronize(s));
  c->tune_chunk = bc;
  c->tune_rounds = br;
  c->tune_wgs_per_cu = bw;
  // The batched form must win by 2 %: measured in place, right after the
  // forward kernel, its launch has cost up to ~3 % more than alone (the
  // per-size launches cost the same in place), and the timings here vary by
  // about 1.5 % from run to run.
  c->chan_mode = best < 0.98f * per_size ? SPK_CHAN_BATCHED : SPK_CHAN_PER_SIZE;
  drop_graphs_and_plans(c);
  if (out) {
    out->per_size_ms = per_size;
    out->batched_ms = best;
    out->chunk_blocks = (uint32_t)bc;
    out->rounds_per_item = (uint32_t)br;
    out->wgs_per_cu = (uint32_t)bw;
    out->mode = c->chan_mode;
  }
  return SPK_OK;
  } catch (...) {
    return entry_exception(c);
  }
}

spk_status spk_set_tile_shape(spk_ctx* c, uint32_t bins, int form, uint32_t shape) {
  try {
  if (!c || (form != SPK_CHAN_PER_SIZE && form != SPK_CHAN_BATCHED)) return SPK_ERR_INVALID;
  const std::vector<ShapeInfo> shapes = tile_shapes((int)bins);
  if (shape >= shapes.size()) return fail(c, SPK_ERR_INVALID, "no such tile shape for this IFFT size");
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  (form == SPK_CHAN_BATCHED ? c->shape_b : c->shape_ps)[(int)bins] = (int)shape;
  drop_graphs_and_plans(c);
  c->chans_dirty = true;  // the per-size tables are uploaded again
  return SPK_OK;
  } catch (...) {
    return entry_exception(c);
  }
}

spk_status spk_tile_shape(spk_ctx* c, uint32_t bins, int form, uint32_t* shape, uint32_t* threads_per_tile,
                          uint32_t* points_per_thread) {
  try {
  if (!c || (form != SPK_CHAN_PER_SIZE && form != SPK_CHAN_BATCHED)) return SPK_ERR_INVALID;
  const std::vector<ShapeInfo> shapes = tile_shapes((int)bins);
  if (shapes.empty()) return fail(c, SPK_ERR_INVALID, "no tile shapes for this IFFT size");
  const int S = shape_of(form == SPK_CHAN_BATCHED ? c->shape_b : c->shape_ps, (int)bins, form == SPK_CHAN_BATCHED);
  if (shape) *shape = (uint32_t)S;
  if (threads_per_tile) *threads_per_tile = (uint32_t)(bins / (uint32_t)shapes[(size_t)S].R);
  if (points_per_thread) *points_per_thread = (uint32_t)shapes[(size_t)S].R;
  return SPK_OK;
  } catch (...) {
    return entry_exception(c);
  }
}

spk_status spk_graph_info(spk_ctx* c, uint32_t* live, uint32_t* instantiated, uint32_t* updated) {
  try {
    if (!c) return SPK_ERR_INVALID;
    if (live) *live = (uint32_t)c->graphs.size();
    if (instantiated) *instantiated = c->graphs_instantiated;
    if (updated) *updated = c->graphs_updated;
    return SPK_OK;
  } catch (...) {
    return entry_exception(c);
  }
}

spk_status spk_fused_channels(spk_ctx* c, uint32_t* n) {
  try {
  if (!c || !n) return SPK_ERR_INVALID;
  if (c->chans_dirty) {
    spk_status st = rebuild_channels(c);
    if (st != SPK_OK) return st;
  }
  uint32_t nf = 0;
  for (const auto& kv : c->groups)
    if (kv.second.fused) nf += (uint32_t)kv.second.ids.size();
  *n = nf;
  return SPK_OK;
  } catch (...) {
    return entry_exception(c);
  }
}

spk_status spk_last_timing(spk_ctx* c, float ms[3]) {
  try {
  if (!c || !ms) return SPK_ERR_INVALID;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  for (; c->ev_done + 4 <= c->ev_used; c->ev_done += 4) {
    hipEvent_t* e = &c->ev_pool[c->ev_done];
    for (int i = 0; i < 3; ++i) {
      float v = 0;
      HIP_TRY(c, hipEventElapsedTime(&v, e[i], e[i + 1]));
      c->ms_sum[i] += v;
    }
    ++c->timed_calls;
  }
  if (c->ev_done == c->ev_used) c->ev_used = c->ev_done = 0;  // recycle the pool
  for (int i = 0; i < 3; ++i) ms[i] = c->timed_calls ? (float)(c->ms_sum[i] / c->timed_calls) : 0.f;
  return SPK_OK;
  } catch (...) {
    return entry_exception(c);
  }
}

spk_status spk_freq_to_bins(double fs_in, double fc_in, uint32_t fft_size, double fs_out, double fc_out,
                            uint32_t* bins, int64_t* first) {
  try {
  if (!bins || !first || fft_size == 0) return SPK_ERR_INVALID;
  const double spacing = fs_in / (double)fft_size;
  const int64_t b = (int64_t)(round_half_away(fs_out / spacing / 4.0) * 4.0);
  const int64_t center = (int64_t)(round_half_away((fc_out - fc_in) / spacing / 4.0) * 4.0);
  if (b <= 0) return SPK_ERR_INVALID;
  *bins = (uint32_t)b;
  *first = center - b / 2;
  return SPK_OK;
  } catch (...) {
    return entry_exception(nullptr);
  }
}

spk_status spk_bins_to_freq(double fs_in, double fc_in, uint32_t fft_size, uint32_t bins, int64_t first,
                            double* fs_out, double* fc_out) {
  try {
  if (!fs_out || !fc_out || fft_size == 0) return SPK_ERR_INVALID;
  const double spacing = fs_in / (double)fft_size;
  *fs_out = spacing * (double)bins;
  *fc_out = fc_in + spacing * (double)(first + (int64_t)(bins / 2));
  return SPK_OK;
  } catch (...) {
    return entry_exception(nullptr);
  }
}

spk_status spk_freq_to_bins_exact(double fs_in, double fc_in, uint32_t fft_size, double fs_out, double fc_out,
                                  uint32_t* bins, int64_t* first) {
  try {
  uint32_t b = 0;
  int64_t f = 0;
  spk_status st = spk_freq_to_bins(fs_in, fc_in, fft_size, fs_out, fc_out, &b, &f);
  if (st != SPK_OK) return st;
  double fso = 0, fco = 0;
  spk_bins_to_freq(fs_in, fc_in, fft_size, b, f, &fso, &fco);
  if (fso != fs_out || fco != fc_out) return SPK_ERR_INVALID;
  *bins = b;
  *first = f;
  return SPK_OK;
  } catch (...) {
    return entry_exception(nullptr);
  }
}

}  // extern "C"
